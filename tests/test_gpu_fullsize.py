"""GPU parity on the headline workload: two coaxial leapfrogging rings (R=1 and 1.2,
sigma=0.08, separation 0.5) in [-2,2]^3 at h=1/64 -- N=273 grid nodes per axis, a 289^3
yaps block, a 43^3 coarse grid, about 1.4M particles after the initial remesh -- one
vpm(WhatToDo=2) (vpm.f90:230-313) through the C ABI against the FP64 oracle on the same
particles.  The plain FP64 form must match the RHS, stream function and velocity grids and UP
to 1e-12 relative L2, the stretching grid and GP -- second differences of the stream
function -- to 3e-12: at this size the oracle's own FP64 rounding (it restates the MKL
solver: sine transforms along x and y, tridiagonal elimination along z, whose low modes have
condition numbers near (m/pi)^2 = 7.4e3) puts its second differences 1.3e-12 from an exact
sine-transform solve of the same system (and its stream function 4.1e-13), so 1e-12 there
would measure the oracle's rounding, not the library's.  The mixed form must match every
field to 1e-5 (stream function, velocity and stretching grids, UP and GP).

The oracle runs its direct boundary sums over every boundary node (about 2.5e11
source-target pairs), so this module takes a minute or two of host time."""
import time

import numpy as np
import pytest

from vpm_free_amd import cases

pytestmark = pytest.mark.gpu

TOL = {"rhs": 1e-12, "psi": 1e-12, "vel": 1e-12, "up": 1e-12, "sol": 3e-12, "gp": 3e-12}
MIXED_TOL = 1e-5
H = 1.0 / 64


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def leapfrog(vpmlib, oracle):
    """Particles after the initial remesh (as the bench starts its loop) and the oracle's
    fields of one vpm(2) on them."""
    V = vpmlib
    t0 = time.perf_counter()
    xp, qp = cases.leapfrog_rings(h=H)
    say(f"case: leapfrog rings h=1/64, {xp.shape[0]} particles before remesh")
    with V.VPM(V.Params.make(H)) as c:
        c.upload(xp, qp)
        c.vpm_resident(0, 0)
        c.remesh_resident(1)
        xr, qr, _, _ = c.download()
        geom = bytes(c.geometry)
    po = oracle.Params.make(H)
    go = oracle.define_sizes(po, xp)
    # the remeshed particles keep the box (idefine = 1); the oracle's geometry is the library's
    assert bytes(go)[: 8 * 9 + 4 * 9] == geom[: 8 * 9 + 4 * 9]
    assert list(go.nn) == [273, 273, 273] and list(go.nnbl)[:3] == [289, 289, 289]
    say(f"  {xr.shape[0]} particles after remesh; oracle vpm(2) on {oracle.num_threads()} threads ...")
    up_r, gp_r, f = oracle.vpm2(po, go, xr, qr, fields=True)
    say(f"  oracle done ({time.perf_counter() - t0:.1f}s)")
    return xr, qr, up_r, gp_r, f


def _gpu_fields(V, xp, qp, precision):
    with V.VPM(V.Params.make(H)) as c:
        c.set_precision(precision)
        up, gp = np.zeros_like(xp), np.zeros_like(xp)
        # the box of the remeshed particles' first call is the one of the original cloud
        # only through idefine = 1; mode 0 with NTIME = 0 on the original cloud fixes it
        xo, qo = cases.leapfrog_rings(h=H)
        c.vpm(xo, qo, None, None, xo.shape[0], 3, 0, 0)
        c.vpm(xp, qp, up, gp, xp.shape[0], 3, 2, 1)
        return {
            "rhs": c.get_grid(0),
            "psi": c.get_grid(1),
            "vel": c.get_grid(3),
            "sol": c.get_grid(2),
            "up": up,
            "gp": gp,
        }


def test_leapfrog_fullsize_fp64(vpmlib, leapfrog):
    xr, qr, up_r, gp_r, f = leapfrog
    say("case: leapfrog 256^3 class, plain FP64 form vs oracle")
    g = _gpu_fields(vpmlib, xr, qr, 0)
    ref = {"rhs": f["rhs"], "psi": f["psi"], "vel": f["vel"], "sol": f["sol"], "up": up_r, "gp": gp_r}
    errs = {k: rel_l2(g[k], ref[k]) for k in ref}
    say("  rel L2:", errs)
    for k, v in errs.items():
        assert v <= TOL[k], (k, v)


def test_leapfrog_fullsize_mixed(vpmlib, leapfrog):
    xr, qr, up_r, gp_r, f = leapfrog
    say("case: leapfrog 256^3 class, mixed form vs FP64 oracle")
    g = _gpu_fields(vpmlib, xr, qr, 1)
    ref = {"psi": f["psi"], "vel": f["vel"], "sol": f["sol"], "up": up_r, "gp": gp_r}
    errs = {k: rel_l2(g[k], ref[k]) for k in ref}
    say("  rel L2 vs FP64:", errs)
    for k, v in errs.items():
        assert v <= MIXED_TOL, (k, v)
    assert errs["up"] > 0.0  # the FP32 passes ran
