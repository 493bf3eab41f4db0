"""Pins of the oracle's tree boundary sums (build_level_nbound_3d + tree_calc_3d,
pinfdomain.f90:875-1530, pmbound.f90:1139-1250).  Runs without a GPU.

* levmax 0: the tree is the face cells themselves, so the sums are the direct sums.
* Refining every cell (factor 1e30 instead of 10) reaches every level-0 cell the links
  lead to.  Where each level's finer rows hold at most one cell beyond the whole groups,
  the links lead to every level-0 cell except, on every face whose two extents are both
  not multiples of 2^lev, the finer corner cell that the closing row of dummies does not
  link (its corner test can never hold).  The sums must then equal the direct sums
  without exactly those cells — derived here from the geometry, not from the links.
* With the reference's refinement rule, the sums on a grid whose faces divide into whole
  groups equal an independent walk over the geometric quadtree of each face (groups
  formed from index ranges, no link tables).
"""
import math

import numpy as np
import pytest


def _face_dims(nn):
    nnx, nny, nnz = nn[0] - 1, nn[1] - 1, nn[2] - 1
    return [(nny, nnz)] * 2 + [(nnx, nnz)] * 2 + [(nnx, nny)] * 2


def _unreached(nn, levmax):
    out = []
    for a, b in _face_dims(nn):
        m = np.zeros((b, a), bool)
        for lev in range(1, levmax + 1):
            s = 1 << lev
            if a % s and b % s:
                m[(b // s) * s:, (a // s) * s:] = True
        out.append(m.ravel())
    return np.concatenate(out)


def _one_extra_cell(nn, levmax):
    for a, b in _face_dims(nn):
        for lev in range(1, levmax + 1):
            s = 1 << lev
            if a % s > s // 2 or b % s > s // 2 or a <= s or b <= s:
                return False
    return True


def _sources(oracle, nn, seed):
    rng = np.random.default_rng(seed)
    xb = np.array([-1.0, -1.1, -0.9, 0, 0, 0])
    dpm = np.array([0.07, 0.06, 0.065])
    f = rng.standard_normal((nn[2], nn[1], nn[0], 3))
    src, cxyz, ds = oracle.normalderiv(xb, dpm, nn, f)
    return xb, dpm, src, cxyz, ds


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nn", [(33, 29, 31), (43, 43, 43), (25, 37, 21)])
def test_tree_level0_is_direct(oracle, nn):
    nn = np.array(nn, np.int32)
    xb, dpm, src, cxyz, ds = _sources(oracle, nn, 1)
    zero = np.zeros((nn[2], nn[1], nn[0], 3))
    direct = oracle.boundsum(xb, dpm, nn, src, cxyz, ds, zero)
    tree = oracle.boundsum_tree(xb, dpm, nn, src, cxyz, ds, zero, 0)
    assert _rel(tree, direct) < 1e-13


@pytest.mark.parametrize("nn,levmax,unreached", [
    ((33, 29, 31), 2, 0),     # every face extent a multiple of 4
    ((43, 43, 43), 2, 24),    # 42 = 10*4 + 2: one dummy row and column, corner unreached
    ((43, 43, 43), 3, 24),
    ((25, 37, 21), 3, 32),
    ((19, 23, 27), 2, 24),
    ((45, 41, 39), 2, 0),
])
def test_tree_refine_all_reaches_linked_cells(oracle, nn, levmax, unreached):
    nn = np.array(nn, np.int32)
    assert _one_extra_cell(nn, levmax)
    xb, dpm, src, cxyz, ds = _sources(oracle, nn, 2)
    zero = np.zeros((nn[2], nn[1], nn[0], 3))
    miss = _unreached(nn, levmax)
    assert int(miss.sum()) == unreached
    src_kept = src.copy()
    src_kept[miss] = 0.0
    want = oracle.boundsum(xb, dpm, nn, src_kept, cxyz, ds, zero)
    try:
        oracle.set_tree_factor(1e30)
        got = oracle.boundsum_tree(xb, dpm, nn, src, cxyz, ds, zero, levmax)
    finally:
        oracle.set_tree_factor(10.0)
    assert _rel(got, want) < 1e-13
    if unreached:  # the unreached cells do change the sums
        full = oracle.boundsum(xb, dpm, nn, src, cxyz, ds, zero)
        assert _rel(got, full) > 1e-6


def _geometric_tree_sums(xb, dpm, nn, src, cxyz, ds, levmax, factor):
    """Walk of each face's quadtree: a level-lev group is a 2^lev x 2^lev block of face cells
    (faces divisible into whole groups), centroid = mean of the cell centroids, area = sum,
    source = area-weighted mean; refine while r < factor * sqrt(area) above level 0."""
    faces = []
    off = 0
    for a, b in _face_dims(nn):
        faces.append((off, a, b))
        off += a * b
    lg = math.log((math.sqrt(2.0) - 1) / (math.sqrt(2.0) + 1))

    def group(o, a, lev, ga, gb):
        s = 1 << lev
        xc = yc = zc = sc = 0.0
        q = [0.0, 0.0, 0.0]
        for mm in range(s):
            for n in range(s):
                l = o + (gb * s + mm) * a + ga * s + n
                xc = cxyz[l, 0] + xc
                yc = cxyz[l, 1] + yc
                zc = cxyz[l, 2] + zc
                sc = ds[l] + sc
                for c in range(3):
                    q[c] = src[l, c] * ds[l] + q[c]
        fl = float(s * s)
        return xc / fl, yc / fl, zc / fl, sc, [v / sc for v in q]

    cache = {}

    def g(fi, lev, ga, gb):
        key = (fi, lev, ga, gb)
        if key not in cache:
            o, a, _ = faces[fi]
            cache[key] = group(o, a, lev, ga, gb)
        return cache[key]

    def visit(fi, lev, ga, gb, X, Y, Z, S):
        x, y, z, dsv, q = g(fi, lev, ga, gb)
        XR, YR, ZR = x - X, y - Y, z - Z
        r = math.sqrt(XR * XR + YR * YR + ZR * ZR)
        if r < factor * math.sqrt(dsv) and lev > 0:
            for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):  # link order: nj, nj+1, nj+nn+1, nj+nn
                visit(fi, lev - 1, 2 * ga + da, 2 * gb + db, X, Y, Z, S)
        elif r > 1e-5:
            for c in range(3):
                S[c] = S[c] - q[c] * dsv / r
        else:
            for c in range(3):
                S[c] = S[c] + q[c] * 2 * math.sqrt(dsv) * lg

    NX, NY, NZ = (int(v) for v in nn)
    out = np.zeros((NZ, NY, NX, 3))
    targets = [(i, j, k) for i in (1, NX) for k in range(1, NZ + 1) for j in range(1, NY + 1)]
    targets += [(i, j, k) for j in (1, NY) for k in range(1, NZ + 1) for i in range(2, NX)]
    targets += [(i, j, k) for k in (1, NZ) for j in range(2, NY) for i in range(2, NX)]
    s = 1 << levmax
    for (i, j, k) in targets:
        X = xb[0] + (i - 1) * dpm[0]
        Y = xb[1] + (j - 1) * dpm[1]
        Z = xb[2] + (k - 1) * dpm[2]
        S = [0.0, 0.0, 0.0]
        for fi, (o, a, b) in enumerate(faces):
            for gb in range(b // s):
                for ga in range(a // s):
                    visit(fi, levmax, ga, gb, X, Y, Z, S)
        out[k - 1, j - 1, i - 1] = [v / (4.0 * math.pi) for v in S]
    return out


@pytest.mark.parametrize("factor", [1.0, 10.0])
def test_tree_matches_geometric_quadtree(oracle, factor):
    nn = np.array([33, 17, 25], np.int32)  # 32 x 16 x 24 cells: whole 8 x 8 groups on every face
    levmax = 3
    xb, dpm, src, cxyz, ds = _sources(oracle, nn, 3)
    zero = np.zeros((nn[2], nn[1], nn[0], 3))
    want = _geometric_tree_sums(xb, dpm, nn, src, cxyz, ds, levmax, factor)
    try:
        oracle.set_tree_factor(factor)
        got = oracle.boundsum_tree(xb, dpm, nn, src, cxyz, ds, zero, levmax)
    finally:
        oracle.set_tree_factor(10.0)
    assert _rel(got, want) < 1e-13
    # groups were accepted (the sums differ from the direct ones), fewer with factor 10
    direct = oracle.boundsum(xb, dpm, nn, src, cxyz, ds, zero)
    err = _rel(got, direct)
    assert err > 1e-8
    assert err < (0.5 if factor == 1.0 else 0.05)
