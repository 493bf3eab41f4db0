"""The allocation-free diagnostics entry point (vpmx_diagnostics_noalloc) through the C ABI
against vpmx_diagnostics on the same particles: two calls in a row on one context, the
second with fewer particles, each bit-identical to the allocating entry point (same kernels,
same summation order) and within 1e-12 of the oracle's FP64 sums."""
import numpy as np
import pytest

from vpm_free_amd import cases

pytestmark = pytest.mark.gpu


def test_diagnostics_noalloc_matches(vpmlib, oracle):
    import hip_mem as M
    V = vpmlib
    xa, qa = cases.single_ring(h=1 / 16, sigma=0.1)
    rng = np.random.default_rng(7)
    xb = rng.uniform(-1, 1, (4000, 3))
    qb = np.concatenate([rng.standard_normal((4000, 3)), np.full((4000, 1), 0.1)], axis=1)
    with V.VPM(V.Params.make(1 / 16)) as c:
        out_d = M.malloc(6 * 8)
        try:
            for x, q in ((xa, qa), (xb, qb)):  # the second call has fewer particles
                n = x.shape[0]
                print("case: diagnostics_noalloc n=%d" % n, flush=True)
                c.upload(x, q)
                circ, imp = c.diagnostics()
                xd = M.to_device(np.ascontiguousarray(x))
                qd = M.to_device(np.ascontiguousarray(q))
                try:
                    c.diagnostics_noalloc(xd, qd, n, q.shape[1], out_d)
                    M.synchronize()
                    got = np.empty(6)
                    M.to_host(got, out_d)
                finally:
                    M.free(xd)
                    M.free(qd)
                assert np.array_equal(got[:3], circ) and np.array_equal(got[3:], imp), (got, circ, imp)
                c_r, i_r = oracle.diagnostics(x, q)
                # (a closed ring's circulation is zero up to rounding: scale by the sum of |QP|)
                scale = np.abs(q[:, :3]).sum() * (1.0 + np.abs(x).max())
                assert np.abs(got[:3] - c_r).max() <= 1e-12 * scale
                assert np.abs(got[3:] - i_r).max() <= 1e-12 * scale
            # sizes outside 0..2^31-1 are rejected before anything is queued
            with pytest.raises(V.VpmError):
                c.diagnostics_noalloc(out_d, out_d, -1, 4, out_d)
            with pytest.raises(V.VpmError):
                c.diagnostics_noalloc(out_d, out_d, 2 ** 31, 4, out_d)
        finally:
            M.free(out_d)
