"""The Fortran drop-in (vpm_lib over the C ABI) running the reference application's loop
(ring_driver.f90: particles.bin -> vpm(0) -> remesh -> [vpm(2), Euler, vpm(0), remesh]),
checked against the oracle on the same input: UP/GP of the first call within relative L2
1e-12, remeshed particle counts and node positions bit-exact."""
import os
import subprocess

import numpy as np
import pytest

from vpm_free_amd import cases, io

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vpm_free_amd", "lib", "ring_driver")


def rel_l2(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.mark.parametrize("h,sigma,opts", [
    (0.125, 0.2, {}),
    # the shipped pm.inp's solver options: 6 x 1 x 1 yaps blocks, tree boundary sums, 3 levels
    (1 / 16, 0.1, dict(blocks=(6, 1, 1), iyntree=1, ilevmax=3)),
])
def test_fortran_driver_matches_oracle(oracle, tmp_path, h, sigma, opts):
    assert os.path.exists(DRIVER), "run python -m vpm_free_amd.build first"
    dt, nsteps = 0.01, 2
    xp, qp = cases.single_ring(h=h, sigma=sigma)
    vref = h ** 3
    io.write_pm_inp(tmp_path / "pm.inp", h, **opts)
    io.write_particles_bin(tmp_path / "particles.bin", xp, -qp[:, :3] / vref, vref)
    print("case: fortran ring_driver", nsteps, "steps", opts, flush=True)
    r = subprocess.run([DRIVER, str(nsteps), str(dt)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    # oracle: same sequence
    xr, w, vr = io.read_particles_bin(tmp_path / "particles.bin")
    q = io.particles_to_qp(xr, w, vr)
    p = oracle.Params.make(h, **opts)
    g = oracle.define_sizes(p, xr, ntime=0)
    x, q = oracle.remesh(p, g, xr, q)
    first = None
    for i in range(1, nsteps + 1):
        up, gp = oracle.vpm2(p, g, x, q)
        if first is None:
            first = (up, gp)
        x = x + up * dt
        q = q.copy()
        q[:, :3] -= gp * dt
        oracle.define_sizes(p, x, ntime=i, geom=g)
        x, q = oracle.remesh(p, g, x, q)
    raw = np.fromfile(tmp_path / "first_call.bin", dtype=np.uint8)
    n1 = raw[:4].view(np.int32)[0]
    arr = raw[4:].view(np.float64)
    up_f, gp_f = arr[:3 * n1].reshape(n1, 3), arr[3 * n1:].reshape(n1, 3)
    assert rel_l2(up_f, first[0]) <= 1e-12 and rel_l2(gp_f, first[1]) <= 1e-12
    raw = np.fromfile(tmp_path / "final.bin", dtype=np.uint8)
    n2 = raw[:4].view(np.int32)[0]
    arr = raw[4:].view(np.float64)
    xf, qf = arr[:3 * n2].reshape(n2, 3), arr[3 * n2:].reshape(n2, 4)
    assert n2 == x.shape[0]
    assert np.array_equal(xf, x)
    assert rel_l2(qf, q) <= 1e-12
    # writeline ran on every vpm(2) call: 9 history files with nsteps records
    hist = io.read_hist(tmp_path / "01hist.bin")
    assert [t for t, _ in hist["lines"]] == list(range(1, nsteps + 1))
