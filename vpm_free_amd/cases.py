"""Synthetic particle clouds of the benchmark workloads (no datasets are read).

Conventions of the reference's particles.bin (test.f90:24-36): vorticity w is the input,
the stored strengths are QP(1:3) = -w * Vref and QP(4) = Vref (the particle volume).
"""
import numpy as np


def ring_vorticity(pts, R=1.0, sigma=0.1, gamma=1.0, center=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0)):
    """Gaussian-core vortex ring w = G/(pi s^2) exp(-rho^2/s^2) e_theta; rho = distance to the
    centre line.  Returns (N, 3) vorticity at pts."""
    pts = np.asarray(pts, dtype=np.float64)
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    d = pts - np.asarray(center, dtype=np.float64)
    z = d @ a
    radial = d - np.outer(z, a)
    r = np.linalg.norm(radial, axis=1)
    rho2 = (r - R) ** 2 + z ** 2
    mag = gamma / (np.pi * sigma ** 2) * np.exp(-rho2 / sigma ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        er = np.where(r[:, None] > 0, radial / np.maximum(r, 1e-300)[:, None], 0.0)
    et = np.cross(a[None, :], er)
    return mag[:, None] * et


def cell_centres(lo, hi, h):
    n = int(round((hi - lo) / h))
    c = lo + (np.arange(n) + 0.5) * h
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def particles_from_vorticity(pts, w, vol, threshold=1e-6, box_corners=None):
    """Keep points with |w| > threshold; QP(1:3) = -w*vol, QP(4) = vol.  With box_corners
    (lo, hi), 8 zero-strength particles at the box corners are added so that the box of
    define_sizes is [lo, hi] +/- interf_iproj*h (they carry no vorticity; remesh drops them)."""
    keep = np.linalg.norm(w, axis=1) > threshold
    xp = pts[keep]
    qp = np.zeros((xp.shape[0], 4))
    qp[:, :3] = -w[keep] * vol
    qp[:, 3] = vol
    if box_corners is not None:
        lo, hi = box_corners
        cs = np.array([[x, y, z] for z in (lo, hi) for y in (lo, hi) for x in (lo, hi)], dtype=np.float64)
        cq = np.zeros((8, 4))
        cq[:, 3] = vol
        xp = np.concatenate([cs, xp])
        qp = np.concatenate([cq, qp])
    return np.ascontiguousarray(xp), np.ascontiguousarray(qp)


def single_ring(h=1.0 / 16, R=1.0, sigma=0.1, gamma=1.0, box=2.0, threshold=1e-6):
    """One thin-core ring in [-box, box]^3, particles at cell centres where |w| > threshold."""
    pts = cell_centres(-box, box, h)
    w = ring_vorticity(pts, R=R, sigma=sigma, gamma=gamma)
    return particles_from_vorticity(pts, w, h ** 3, threshold, box_corners=(-box, box))


def leapfrog_rings(h=1.0 / 64, R1=1.0, R2=1.2, sigma=0.08, sep=0.5, gamma=1.0, box=2.0, threshold=1e-6):
    """Two coaxial rings (radii R1, R2, axial separation sep) in [-box, box]^3."""
    pts = cell_centres(-box, box, h)
    w = ring_vorticity(pts, R=R1, sigma=sigma, gamma=gamma, center=(0, 0, -0.5 * sep))
    w += ring_vorticity(pts, R=R2, sigma=sigma, gamma=gamma, center=(0, 0, 0.5 * sep))
    return particles_from_vorticity(pts, w, h ** 3, threshold, box_corners=(-box, box))


def random_blobs(n, h, seed, lo=-1.0, hi=1.0):
    """Uniform positions in [lo, hi]^3, strengths N(0,1) per component times h^3."""
    rng = np.random.default_rng(seed)
    xp = rng.uniform(lo, hi, size=(n, 3))
    qp = np.zeros((n, 4))
    qp[:, :3] = rng.standard_normal((n, 3)) * h ** 3
    qp[:, 3] = h ** 3
    return xp, qp


def batch_rings(seed=3, count=64, grids=(32, 48, 64, 96, 128), threshold=1e-6):
    """Independent single-ring clouds: radius U[0.5, 1.5], core sigma U[0.05, 0.15], axis
    uniform on the sphere, each in a box [-L, L]^3 with L = R + 4 sigma, cut into N cells per
    axis with N drawn from `grids` with probability proportional to 1/N^3 (every size class
    holds a similar share of the cells: about 6-7M cells for 64 clouds), h = 2L/N.  Returns a
    list of (h, xp, qp)."""
    rng = np.random.default_rng(seed)
    g = np.asarray(grids, dtype=np.float64)
    prob = (1.0 / g ** 3) / np.sum(1.0 / g ** 3)
    out = []
    for _ in range(count):
        R = rng.uniform(0.5, 1.5)
        sigma = rng.uniform(0.05, 0.15)
        a = rng.standard_normal(3)
        a /= np.linalg.norm(a)
        N = int(rng.choice(grids, p=prob))
        L = R + 4.0 * sigma
        h = 2.0 * L / N
        pts = cell_centres(-L, L, h)
        w = ring_vorticity(pts, R=R, sigma=sigma, axis=a)
        xp, qp = particles_from_vorticity(pts, w, h ** 3, threshold, box_corners=(-L, L))
        out.append((h, xp, qp))
    return out


def blob_cloud(n=10_000_000, h=1.0 / 64, half=4.0, seed=5):
    """Random vortex blobs: positions uniform in the box [-half, half]^3 (its inner part: half
    minus the projection margin of 2h), strengths N(0, 1) per component, QP(1:3) = -w h^3,
    QP(4) = h^3; 8 zero-strength corner particles fix the box at [-half, half]^3."""
    rng = np.random.default_rng(seed)
    lo, hi = -half + 2 * h, half - 2 * h
    xp = rng.uniform(lo, hi, size=(n, 3))
    w = rng.standard_normal((n, 3))
    vol = h ** 3
    qp = np.empty((n, 4))
    qp[:, :3] = -w * vol
    qp[:, 3] = vol
    cs = np.array([[x, y, z] for z in (-half, half) for y in (-half, half) for x in (-half, half)], dtype=np.float64)
    cq = np.zeros((8, 4))
    cq[:, 3] = vol
    return np.concatenate([cs, xp]), np.concatenate([cq, qp])
