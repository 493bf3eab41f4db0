"""Pins of the oracle's panel-source boundary sums (pm.inp ibctyp = 3): FSOUR_A4
(pmbound.f90:1541-1632), calc_boundinf_3d_s (:520-648) and tree_calc_3d_s (:1346-1481).
Runs without a GPU.

* FSOUR_A4 near field against the closed form of the potential of a uniform rectangle,
  sum over corners of p ln(q + r) + q ln(p + r) - z atan(pq / (z r)) — a different formula
  from the reference's edge sums.  In the panel's plane the reference offsets the field
  point by 1e-14 (logs) and 1e-10 (angles), so there the bar is 5e-11 absolute.
* FSOUR_A4 far field (RO > 10 DIAG): -AREA / (4 pi RO).
* calc_boundinf_3d_s against a numpy sum of the closed form in the frame of the target
  plane (the reference integrates every source panel as if it lay in the target's plane,
  with the panel extents of pmbound.f90:537-620).
* tree_calc_3d_s: with one level it takes every cell as a top-level cell, whose term is
  FSOUR_A4 times the source of the first top-level cell (the reference passes nv = 1); with
  every group refined it reaches the level-0 cells except those below dummies of kind -2,
  which it takes as cells of zero area.
"""
import math

import numpy as np
import pytest

# (E1 axis, E2 axis, normal axis, spacing index of the E1 extent, of the E2 extent) per |iplane|
FRAME = {1: (0, 2, 1, 0, 2), 2: (2, 1, 0, 1, 2), 3: (0, 1, 2, 0, 1)}


def _rect_integral(X, Y, Z, a, b):
    """Integral of 1/|x - y| over the rectangle |u| <= a/2, |v| <= b/2 in the plane z = 0,
    field point (X, Y, Z); vectorised."""
    X, Y, Z = (np.asarray(v, dtype=np.float64) for v in (X, Y, Z))

    def log_sum(u, v, r):
        # ln(u + r) with r = |(u, v, Z)|; for u < 0 as ln(v^2 + Z^2) - ln(r - u) (no cancellation)
        return np.where(u >= 0, np.log(np.where(u >= 0, u + r, 1.0)),
                        np.log(v * v + Z * Z) - np.log(np.where(u >= 0, 1.0, r - u)))

    def F(p, q):
        r = np.sqrt(p * p + q * q + Z * Z)
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = np.where(p == 0, 0.0, p * log_sum(q, p, r))
            t2 = np.where(q == 0, 0.0, q * log_sum(p, q, r))
            t3 = np.where((Z == 0) | (r == 0), 0.0, Z * np.arctan(p * q / np.where(Z == 0, 1.0, Z * r)))
        return t1 + t2 - t3

    p1, p2 = -a / 2 - X, a / 2 - X
    q1, q2 = -b / 2 - Y, b / 2 - Y
    return F(p2, q2) - F(p1, q2) - F(p2, q1) + F(p1, q1)


def test_rect_integral_against_quadrature():
    """The closed form itself, against Gauss-Legendre quadrature off the plane."""
    g, w = np.polynomial.legendre.leggauss(80)
    a, b = 0.07, 0.05
    for X, Y, Z in [(0.01, 0.02, 0.04), (0.1, -0.03, 0.02), (-0.02, 0.0, 0.08), (0.3, 0.2, -0.05)]:
        u = g * a / 2
        v = g * b / 2
        U, V = np.meshgrid(u, v, indexing="ij")
        W = np.outer(w, w) * (a / 2) * (b / 2)
        q = np.sum(W / np.sqrt((X - U) ** 2 + (Y - V) ** 2 + Z * Z))
        assert abs(_rect_integral(X, Y, Z, a, b) - q) < 1e-10 * abs(q)


def _fis_expected(xo, rg, iplane, dpm):
    e1, e2, en, ia, ib = FRAME[abs(iplane)]
    d = np.asarray(xo, float) - np.asarray(rg, float)
    X, Y, Z = d[e1], d[e2], d[en]
    a, b = dpm[ia], dpm[ib]
    return -_rect_integral(X, Y, Z, a, b) / (4 * math.pi)


@pytest.mark.parametrize("iplane", [1, -1, 2, -2, 3, -3])
def test_fsour_near_field_closed_form(oracle, iplane):
    dpm = np.array([0.07, 0.06, 0.05])
    rg = np.array([0.3, -0.2, 0.1])
    rng = np.random.default_rng(abs(iplane) * 7 + (iplane > 0))
    e1, e2, en, ia, ib = FRAME[abs(iplane)]
    for _ in range(40):
        xo = rg + rng.uniform(-0.25, 0.25, 3)
        want = _fis_expected(xo, rg, iplane, dpm)
        got = oracle.fsour_a4(xo, rg, iplane, dpm)
        assert abs(got - want) < 1e-11 * abs(want), (xo, got, want)
    # in the panel's plane: lattice offsets (edges, corners, their extensions)
    for kx in (-3, -1, 0, 1, 2):
        for ky in (-2, -1, 0, 1, 3):
            xo = rg.copy()
            xo[e1] += (kx + 0.5) * dpm[ia]
            xo[e2] += (ky + 0.5) * dpm[ib]
            want = _fis_expected(xo, rg, iplane, dpm)
            got = oracle.fsour_a4(xo, rg, iplane, dpm)
            # the 1e-10 offset adds at most 1e-10 * 2 pi / (4 pi) to the angle term
            assert abs(got - want) < 5e-11 + 1e-12 * abs(want), (kx, ky, got, want)


def test_fsour_far_field(oracle):
    dpm = np.array([0.07, 0.06, 0.05])
    rg = np.zeros(3)
    diag = math.hypot(dpm[1], dpm[2])
    for ro in (10.001 * diag, 25 * diag):
        xo = np.array([ro, 0.0, 0.0])
        got = oracle.fsour_a4(xo, rg, 2, dpm)
        assert got == -(dpm[1] * dpm[2] / (4.0 * (4.0 * math.atan(1.0)))) / ro
        assert abs(got - _fis_expected(xo, rg, 2, dpm)) < 1e-3 * abs(got)  # monopole limit


def _sources(oracle, nn, seed, dpm, xb):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((nn[2], nn[1], nn[0], 3))
    return oracle.normalderiv(xb, dpm, nn, f)


def _targets(nn):
    NX, NY, NZ = (int(v) for v in nn)
    t = [(i, j, k, 2 if i == 1 else -2) for i in (1, NX) for k in range(1, NZ + 1) for j in range(1, NY + 1)]
    t += [(i, j, k, 1 if j == 1 else -1) for j in (1, NY) for k in range(1, NZ + 1) for i in range(2, NX)]
    t += [(i, j, k, 3 if k == 1 else -3) for k in (1, NZ) for j in range(2, NY) for i in range(2, NX)]
    return t


def test_panel_sums_closed_form(oracle):
    nn = np.array([11, 9, 10], np.int32)
    dpm = np.array([0.07, 0.06, 0.065])
    xb = np.array([0.4, -0.9, 0.3, 0, 0, 0])
    src, cxyz, ds = _sources(oracle, nn, 5, dpm, xb)
    zero = np.zeros((nn[2], nn[1], nn[0], 3))
    got = oracle.boundsum_s(xb, dpm, nn, src, cxyz, zero)
    want = np.zeros_like(zero)
    for (i, j, k, pl) in _targets(nn):
        xo = np.array([xb[0] + (i - 1) * dpm[0], xb[1] + (j - 1) * dpm[1], xb[2] + (k - 1) * dpm[2]])
        e1, e2, en, ia, ib = FRAME[abs(pl)]
        d = xo[None, :] - cxyz
        ro = np.sqrt(d[:, e1] ** 2 + d[:, e2] ** 2 + d[:, en] ** 2)
        diag = math.hypot(dpm[ia], dpm[ib])
        area = dpm[ia] * dpm[ib]
        fis = np.where(ro / diag > 10, -area / (4 * math.pi) / ro,
                       -_rect_integral(d[:, e1], d[:, e2], d[:, en], dpm[ia], dpm[ib]) / (4 * math.pi))
        want[k - 1, j - 1, i - 1] = fis @ src
    assert np.linalg.norm(got - want) < 1e-9 * np.linalg.norm(want)
    point = oracle.boundsum(xb, dpm, nn, src, cxyz, ds, zero)
    assert np.linalg.norm(got - point) > 1e-4 * np.linalg.norm(point)  # panels are not points


def test_panel_tree_one_level(oracle):
    nn = np.array([21, 17, 19], np.int32)
    dpm = np.array([0.05, 0.05, 0.05])  # equal spacings: cell areas equal the target-plane area
    xb = np.array([0.4, -0.9, 0.3, 0, 0, 0])
    src, cxyz, ds = _sources(oracle, nn, 6, dpm, xb)
    zero = np.zeros((nn[2], nn[1], nn[0], 3))
    got = oracle.boundsum_tree_s(xb, dpm, nn, src, cxyz, ds, zero, 0)
    first = np.broadcast_to((src[0] * ds[0]) / ds[0], src.shape)
    want = oracle.boundsum_s(xb, dpm, nn, first, cxyz, zero)
    assert np.linalg.norm(got - want) < 1e-13 * np.linalg.norm(want)


def _unreached_s(nn, levmax):
    """Level-0 cells below dummies of kind -2 (the closing row of groups of a face)."""
    nnx, nny, nnz = nn[0] - 1, nn[1] - 1, nn[2] - 1
    out = []
    for a, b in [(nny, nnz)] * 2 + [(nnx, nnz)] * 2 + [(nnx, nny)] * 2:
        m = np.zeros((b, a), bool)
        for lev in range(1, levmax + 1):
            s = 1 << lev
            if b % s:
                m[(b // s) * s:, :] = True
        out.append(m.ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("nn,levmax,unreached", [((33, 17, 25), 3, 0), ((43, 43, 43), 2, 504),
                                                 ((25, 37, 21), 3, 672)])
def test_panel_tree_refine_all(oracle, nn, levmax, unreached):
    nn = np.array(nn, np.int32)
    dpm = np.array([0.07, 0.06, 0.065])
    xb = np.array([3.0, 2.5, 4.0, 0, 0, 0])  # top-level dummies sit at the origin: keep it far away
    src, cxyz, ds = _sources(oracle, nn, 7, dpm, xb)
    zero = np.zeros((nn[2], nn[1], nn[0], 3))
    miss = _unreached_s(nn, levmax)
    assert int(miss.sum()) == unreached
    kept = src.copy()
    kept[miss] = 0.0
    want = oracle.boundsum(xb, dpm, nn, kept, cxyz, ds, zero)
    try:
        oracle.set_tree_factor(1e30)
        got = oracle.boundsum_tree_s(xb, dpm, nn, src, cxyz, ds, zero, levmax)
    finally:
        oracle.set_tree_factor(10.0)
    assert np.linalg.norm(got - want) < 1e-13 * np.linalg.norm(want)
