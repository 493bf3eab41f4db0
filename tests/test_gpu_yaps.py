"""The yaps composition checks of test_oracle_yaps.py on the HIP library: with vorticity
blobs inside different blocks, the several-block solution (plain FP64 form; even blocks, and
blocks of uneven extents) converges to the one-block solution as h shrinks, and equals the
oracle's several-block solution to 1e-12."""
import numpy as np
import pytest

import test_oracle_yaps as Y

pytestmark = pytest.mark.gpu


def gpu_solve(V, h, layout, blocks, weights=None):
    xp, qp = Y.blob_particles(h, layout)
    with V.VPM(V.Params.make(h, blocks=blocks, ncoarse=Y.NCOARSE)) as c:
        if weights is not None:
            c.set_block_weights(*weights)
        up, gp = np.zeros((xp.shape[0], 3)), np.zeros((xp.shape[0], 3))
        c.vpm(xp, qp, up, gp, xp.shape[0], 3, 0, 0)
        c.vpm(xp, qp, up, gp, xp.shape[0], 3, 2, 1)
        g = c.geometry
        return (list(g.xbound)[:3], list(g.nn), c.get_grid(1), np.moveaxis(c.get_grid(3), 0, -1))


@pytest.mark.parametrize("layout", list(Y.BLOBS))
def test_gpu_blocks_converge_to_one_block(vpmlib, oracle, layout):
    d = []
    for h in Y.HS:
        print("case: yaps blocks", layout, "h=1/%d" % round(1 / h), flush=True)
        multi = gpu_solve(vpmlib, h, layout, layout)
        single = gpu_solve(vpmlib, h, layout, (1, 1, 1))
        d.append(Y.difference(multi, single, h))
        ref = Y.solve(h, layout, layout)
        for k, a, b in (("psi", multi[2], ref[2]), ("vel", multi[3], ref[3])):
            e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            assert e <= 1e-12, (layout, h, k, e)
    print(layout, d, flush=True)
    for k in ("psi", "vel"):
        assert d[1][k] <= 1e-2, (layout, k, d)
        assert d[0][k] / d[1][k] >= 1.7, (layout, k, d)


@pytest.mark.parametrize("layout", list(Y.UNEVEN))
def test_gpu_uneven_blocks(vpmlib, oracle, layout):
    w = Y.UNEVEN[layout]
    d = []
    for h in Y.HS:
        print("case: uneven yaps blocks", layout, w, "h=1/%d" % round(1 / h), flush=True)
        multi = gpu_solve(vpmlib, h, layout, layout, w)
        single = gpu_solve(vpmlib, h, layout, (1, 1, 1))
        d.append(Y.difference(multi, single, h))
        ref = Y.solve(h, layout, layout, 0, w)
        assert multi[1] == ref[1] and multi[0] == ref[0]
        for k, a, b in (("psi", multi[2], ref[2]), ("vel", multi[3], ref[3])):
            e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            assert e <= 1e-12, (layout, h, k, e)
    print(layout, d, flush=True)
    for k in ("psi", "vel"):
        assert d[1][k] <= 2e-2, (layout, k, d)
        assert d[0][k] / d[1][k] >= 1.7, (layout, k, d)
