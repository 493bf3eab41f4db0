"""The identities behind the fine solves' spectral shortcuts, checked in numpy on random sine
series (no GPU, no oracle): a zero-BC solution is a sine series psi(p, q) = sum S(p, kx)
S(q, ky) W(kx, ky) (S(p, k) = sin(pi p k / (m + 1)), p the interior index), so

* the one-sided normal derivative (1/h)(25/12 f0 - 4 f1 + 3 f2 - 4/3 f3 + 1/4 f4) on the low
  x face (f0 = 0 there) is sum_kx w(kx) (S W)(kx, .) with w(k) = (1/h) sum_s a_s S(s, k), and
  on the high face the same with (-1)^(k+1) w(k) (calc_normalderiv_3d, pinfdomain.f90:294-547);
* the value on an inner plane p0 is sum_kx S(p0, kx) W(kx, .), and on the mirrored plane
  m + 1 - p0 the same with (-1)^(k+1) S(p0, k) (the inner box's faces of yaps3d.f90:340-395);
* the extended DST-I matrix (rows 0 .. m+1) is zero on its first and last rows, so products with
  it leave the face edges zero.
A dropped stencil term, a wrong parity or an off-by-one plane index breaks these at O(1)."""
import numpy as np


def S_ext(m):
    r = np.arange(m + 2)[:, None]
    k = np.arange(1, m + 1)[None, :]
    return np.sin(np.pi * r * k / (m + 1))


A = np.array([25 / 12, -4.0, 3.0, -4 / 3, 0.25])


def field(rng, mx, my):
    W = rng.standard_normal((mx, my)) / (1 + np.arange(mx))[:, None] / (1 + np.arange(my))[None, :]
    return W, S_ext(mx) @ W @ S_ext(my).T  # nodes 0..mx+1 x 0..my+1, zero on the edges


def test_normal_derivative_from_spectrum():
    rng = np.random.default_rng(4)
    for mx, my, h in ((13, 11, 0.1), (30, 17, 1 / 32), (41, 41, 1 / 8)):
        W, psi = field(rng, mx, my)
        assert np.all(psi[0] == 0) and np.abs(psi[-1]).max() <= 1e-13 * np.abs(psi).max()
        lo = (A @ psi[0:5]) / h
        hi = (A @ psi[::-1][0:5]) / h
        k = np.arange(1, mx + 1)
        w = sum(A[s] * np.sin(np.pi * k * s / (mx + 1)) for s in range(1, 5)) / h
        sgn = (-1.0) ** (k + 1)
        scale = np.abs(lo).max()
        assert np.abs(S_ext(my) @ (w @ W) - lo).max() <= 1e-12 * scale
        assert np.abs(S_ext(my) @ ((sgn * w) @ W) - hi).max() <= 1e-12 * scale
        # the parity is what carries the high face: without it the sums differ at O(1)
        assert np.abs(S_ext(my) @ (w @ W) - hi).max() > 1e-3 * scale


def test_inner_plane_values_from_spectrum():
    rng = np.random.default_rng(5)
    for mx, my, p0 in ((287, 9, 8), (31, 7, 3), (20, 5, 1)):
        W, psi = field(rng, mx, my)
        k = np.arange(1, mx + 1)
        row = np.sin(np.pi * p0 * k / (mx + 1))
        sgn = (-1.0) ** (k + 1)
        scale = np.abs(psi).max()
        assert np.abs(S_ext(my) @ (row @ W) - psi[p0]).max() <= 1e-12 * scale
        assert np.abs(S_ext(my) @ ((sgn * row) @ W) - psi[mx + 1 - p0]).max() <= 1e-12 * scale
        assert np.abs(S_ext(my) @ (row @ W) - psi[p0 + 1]).max() > 1e-6 * scale


def test_extended_sine_matrix_edges():
    for m in (4, 41, 287):
        Se = S_ext(m)
        assert Se.shape == (m + 2, m)
        assert np.abs(Se[0]).max() == 0.0
        assert np.abs(Se[-1]).max() < 1e-12
        # its interior rows are the DST-I matrix, which is its own inverse up to 2/(m+1)
        S = Se[1:-1]
        assert np.abs(S @ S * (2.0 / (m + 1)) - np.eye(m)).max() < 1e-11
