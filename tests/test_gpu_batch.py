"""Batched clouds (vpmx_batch_*): independent rings stepped together -- queued on the clouds'
streams, or captured once as a hipGraph and replayed -- give every cloud exactly the particles
(positions, strengths, counts: bit for bit, FP64 form) of the same steps run alone on one
context, through remeshing steps that change the particle counts."""
import numpy as np
import pytest

from vpm_free_amd import cases

pytestmark = pytest.mark.gpu
DT = 0.01


def solo(V, h, xp, qp, steps):
    with V.VPM(V.Params.make(h)) as c:
        c.upload(xp, qp)
        c.vpm_resident(0, 0)
        for s in range(steps):
            c.step_rk2(DT, ntime=s + 1)
        x, q, _, _ = c.download()
    return x, q


def batched(V, clouds, modes, precision=0):
    with V.Batch(len(clouds)) as b:
        for i, (h, xp, qp) in enumerate(clouds):
            c = b.cloud(i, V.Params.make(h))
            c.set_precision(precision)
            c.upload(xp, qp)
            c.vpm_resident(0, 0)
        for g in modes:
            b.step_rk2(DT, graph=g)
        return [b.cloud(i).download()[:2] for i in range(len(clouds))]


@pytest.mark.parametrize("modes", [(False, False), (True, True), (False, True, True)])
def test_batch_matches_solo_runs(vpmlib, modes):
    V = vpmlib
    clouds = cases.batch_rings(seed=11, count=4, grids=(32, 48))
    print("case: batch of", len(clouds), "rings, steps (graph?)", modes, flush=True)
    got = batched(V, clouds, modes)
    for i, (h, xp, qp) in enumerate(clouds):
        x_r, q_r = solo(V, h, xp, qp, len(modes))
        x, q = got[i]
        assert x.shape == x_r.shape, (i, x.shape, x_r.shape)
        assert np.array_equal(x, x_r) and np.array_equal(q, q_r), i


def test_batch_full_size_graph_equals_queued(vpmlib):
    """The 64-ring batch (grids 32^3..128^3) in the mixed form: a captured-and-replayed step
    gives the same particles as the queued one, bit for bit, and each ring's total
    circulation (zero for a closed ring) stays small after two steps: the remesh threshold
    and the stretching term leave up to 1.5e-4 of the sum of |QP| (measured; bar 1e-2)."""
    V = vpmlib
    clouds = cases.batch_rings(seed=3, count=64)
    cells = sum(int(round(2 * (np.abs(xp).max()) / h)) ** 3 for h, xp, _ in clouds)
    print("case: 64-ring batch, about %.1fM cells" % (cells / 1e6), flush=True)
    a = batched(V, clouds, (False, False), precision=1)
    b = batched(V, clouds, (True, True), precision=1)
    for i in range(len(clouds)):
        assert np.array_equal(a[i][0], b[i][0]) and np.array_equal(a[i][1], b[i][1]), i
        circ = -a[i][1][:, :3].sum(axis=0)
        scale = np.abs(a[i][1][:, :3]).sum()
        assert np.abs(circ).max() <= 1e-2 * scale, i
