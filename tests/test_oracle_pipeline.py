"""Pins of the oracle's composed routines and of the library's host logic (no GPU).

  geometry     hand-derived sizes (the shipped pm.inp, the 64^3 and 256^3 ring boxes),
               divisibility and containment invariants; the product's host define_sizes
               must be bit-identical
  remesh       brute-force threshold scan, k->j->i order, exact conservation of total
               strength and first moments (M4' reproduces linear functions)
  yaps3d       free-space potential of a Gaussian blob on one block (discretisation level)
  vpm mode 2   Hill's spherical vortex closed form (test_applic.f90:2-77 restated from the
               mathematics): velocity and stretching
  C ABI        libvpmx.so loads on a host without a GPU and exports every declared symbol
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from vpm_free_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# ------------------------------------------------------------------ pm.inp and geometry
def test_read_shipped_pm_inp(oracle, vpmlib):
    path = os.path.join(GOLDEN, "pm.inp")
    for mod in (oracle, vpmlib):
        p = mod.Params()
        rd = oracle.lib().orc_read_pm_inp if mod is oracle else vpmlib.load_library().vpmx_read_pm_inp
        assert rd(path.encode(), C.byref(p)) == 0
        assert list(p.dpm) == [4.0, 4.0, 4.0]
        assert (p.interf_iproj, p.ibctyp, p.idvpm, p.epsvol, p.ncoarse) == (4, 2, 1, 0.0, 8)
        assert (p.nbi, p.nbj, p.nbk, p.nremesh, p.ncell_rem) == (6, 1, 1, 1, 1)
        assert (p.iyntree, p.ilevmax, p.ompthreads, p.idefine, p.iynslice, p.ipmwrite) == (1, 3, 1, 1, 0, 7)
        assert list(p.ipmwstart[:7]) == [100, 180, 360, 760, 1120, 3600, 7200]
        assert list(p.ipmwsteps[:7]) == [1] * 7


def test_pm_inp_ipmwrite_limit(oracle, vpmlib, tmp_path):
    src = open(os.path.join(GOLDEN, "pm.inp")).read().splitlines()
    src[12] = "11   ! IPMWRITE"
    f = tmp_path / "pm.inp"
    f.write_text("\n".join(src) + "\n")
    assert oracle.lib().orc_read_pm_inp(str(f).encode(), C.byref(oracle.Params())) != 0
    with pytest.raises(vpmlib.VpmError):
        vpmlib.read_pm_inp(str(f))


@pytest.mark.parametrize("h,box,nn,nnbl,nc", [(1 / 16, 2.0, 81, 97, 19), (1 / 64, 2.0, 273, 289, 43),
                                              (1 / 32, 2.0, 145, 161, 27), (0.125, 2.0, 49, 65, 15)])
def test_define_sizes_ring_boxes(oracle, vpmlib, h, box, nn, nnbl, nc):
    """Box [-box,box]^3 of particles +/- 4h, padded so (NN-1) is a multiple of 8
    (definepm itype 3 adds a full 8 when it already is), 8 dummy cells per block side,
    coarse spacing 8h with 4 dummy coarse cells."""
    pts = np.array([[-box] * 3, [box] * 3])
    g = oracle.define_sizes(oracle.Params.make(h), pts)
    assert list(g.nn) == [nn] * 3 and list(g.nn_bl) == [1, 1, 1, nn, nn, nn]
    assert list(g.nnbl[:3]) == [nnbl] * 3 and list(g.nnbl_bl[:6]) == [9, 9, 9, nnbl - 8, nnbl - 8, nnbl - 8]
    assert list(g.nn_coarse) == [nc] * 3
    assert (g.nn[0] - 1) % 8 == 0
    assert g.xbound[0] <= -box - 4 * h and g.xbound[3] >= box + 4 * h
    assert g.xbound_bl[0] == g.xbound[0] - 8 * h
    gv = vpmlib.define_sizes(vpmlib.Params.make(h), pts)
    assert bytes(gv) == bytes(g)


def test_define_sizes_random_clouds_bit_identical(oracle, vpmlib):
    rng = np.random.default_rng(0)
    for trial in range(40):
        h = rng.choice([0.1, 1 / 16, 0.03, 0.2])
        n = 50
        xp = rng.normal(rng.uniform(-3, 3, 3), rng.uniform(0.1, 2, 3), (n, 3))
        p_o = oracle.Params.make(h, ncoarse=int(rng.choice([4, 8])), interf_iproj=int(rng.choice([2, 3, 4])))
        p_v = vpmlib.Params.make(h, ncoarse=p_o.ncoarse, interf_iproj=p_o.interf_iproj)
        g_o = oracle.define_sizes(p_o, xp)
        g_v = vpmlib.define_sizes(p_v, xp)
        assert bytes(g_o) == bytes(g_v), trial
        for d in range(3):
            assert (g_o.nn[d] - 1) % p_o.ncoarse == 0
            assert g_o.xbound[d] <= xp[:, d].min() - p_o.interf_iproj * h * (1 - 1e-12)
            assert g_o.xbound[d + 3] >= xp[:, d].max() + p_o.interf_iproj * h * (1 - 1e-12)


def test_define_sizes_uneven_blocks(oracle, vpmlib):
    """Blocks of uneven extents: along each axis the inner boxes tile the global grid, each
    extent is a whole number of coarse cells, cut i is ncoarse * floor(units * W_<i / W), and
    equal weights give the even split bit for bit; the library's host geometry is the
    oracle's."""
    rng = np.random.default_rng(5)
    for trial in range(30):
        h = float(rng.choice([0.1, 1 / 16, 0.05]))
        blocks = tuple(int(v) for v in rng.integers(1, 4, 3))
        nc = int(rng.choice([4, 8]))
        xp = rng.normal(0.0, rng.uniform(0.5, 2, 3), (40, 3))
        w = [rng.integers(1, 6, blocks[d]).astype(np.int32) for d in range(3)]
        p_o = oracle.Params.make(h, blocks=blocks, ncoarse=nc)
        p_v = vpmlib.Params.make(h, blocks=blocks, ncoarse=nc)
        oracle.set_block_weights(*w)
        try:
            g_o = oracle.define_sizes(p_o, xp)
        finally:
            oracle.set_block_weights(None)
        g_v = vpmlib.define_sizes(p_v, xp, weights=w)
        assert bytes(g_o) == bytes(g_v), trial
        nb = blocks[0] * blocks[1] * blocks[2]
        for d in range(3):
            units = (g_o.nn[d] - 1) // nc
            W = int(w[d].sum())
            stride = [1, blocks[0], blocks[0] * blocks[1]][d]
            pos = g_o.xbound[d]
            for i in range(blocks[d]):
                b = i * stride
                lo, hi = g_o.nnbl_bl[6 * b + d], g_o.nnbl_bl[6 * b + d + 3]
                ext = hi - lo
                cut0 = nc * ((units * int(w[d][:i].sum())) // W)
                cut1 = nc * ((units * int(w[d][:i + 1].sum())) // W)
                assert ext == cut1 - cut0 and ext % nc == 0 and ext >= nc, (trial, d, i)
                # the inner box starts where the previous one ended
                x_in = g_o.xbound_bl[6 * b + d] + (lo - 1) * h
                assert abs(x_in - pos) < 1e-9 * max(1.0, abs(pos)), (trial, d, i)
                pos = g_o.xbound_bl[6 * b + d] + (hi - 1) * h
            last = g_o.xbound[d] + (g_o.nn[d] - 1) * h  # the global grid's last node
            assert abs(pos - last) < 1e-9 * max(1.0, abs(pos))
        assert g_o.nblocks == nb
        # equal weights: the reference's even split
        eq = [np.full(blocks[d], 3, np.int32) for d in range(3)]
        assert bytes(vpmlib.define_sizes(p_v, xp, weights=eq)) == bytes(vpmlib.define_sizes(p_v, xp))


def test_define_sizes_box_fixed_after_first_call(oracle, vpmlib):
    """idefine=1: the box is computed at NTIME=0 only (vpm.f90:500)."""
    xp = np.array([[-1.0, -1, -1], [1, 1, 1]])
    for mod in (oracle, vpmlib):
        p = mod.Params.make(0.1)
        g = mod.define_sizes(p, xp, ntime=0)
        before = bytes(g)
        mod.define_sizes(p, xp * 3, ntime=5, geom=g)
        assert bytes(g) == before
        p0 = mod.Params.make(0.1, idefine=0)
        g2 = mod.define_sizes(p0, xp, ntime=0)
        mod.define_sizes(p0, xp * 3, ntime=5, geom=g2)
        assert g2.nn[0] > g.nn[0]


# ------------------------------------------------------------------ remesh
def test_remesh_bruteforce_and_conservation(oracle):
    h = 0.1
    xp, qp = cases.single_ring(h=h, R=0.8, sigma=0.2, box=1.5)
    rng = np.random.default_rng(2)
    xp = xp + rng.uniform(-0.4, 0.4, xp.shape) * h
    p = oracle.Params.make(h)
    g = oracle.define_sizes(p, xp)
    xr, qr = oracle.remesh(p, g, xp, qp)
    # brute force: project, scan the window, keep |w| >= real(1e-9,4)
    NN = np.array(g.nn)
    D = np.array([(g.xbound[d + 3] - g.xbound[d]) / (NN[d] - 1) for d in range(3)])
    rhs = oracle.project(g.xbound, D, NN, g.nn_bl, xp, qp)
    w = np.sqrt((rhs[..., :3] ** 2).sum(-1))
    lo, hi = np.array(g.nn_bl[:3]) + 2, np.array(g.nn_bl[3:]) - 2
    win = w[lo[2] - 1:hi[2], lo[1] - 1:hi[1], lo[0] - 1:hi[0]]
    keep = win >= np.float64(np.float32(1e-9))
    assert xr.shape[0] == keep.sum()
    kk, jj, ii = np.nonzero(keep)  # C order == k -> j -> i loop order
    want = np.stack([g.xbound[0] + D[0] * (ii + lo[0] - 1), g.xbound[1] + D[1] * (jj + lo[1] - 1),
                     g.xbound[2] + D[2] * (kk + lo[2] - 1)], axis=1)
    assert np.array_equal(xr, want)
    dv = D[0] * D[1] * D[2]
    assert np.all(qr[:, 3] == dv)
    # conservation: strength and first moments up to what the dropped nodes carried
    c0, i0 = oracle.diagnostics(xp, qp)
    c1, i1 = oracle.diagnostics(xr, qr)
    full = np.zeros(w.shape, bool)
    full[lo[2] - 1:hi[2], lo[1] - 1:hi[1], lo[0] - 1:hi[0]] = keep
    dropped = -(rhs[..., :3][~full]).sum(0) * dv
    assert np.allclose(c0 - c1, dropped, rtol=0, atol=1e-14)
    assert np.abs(dropped).max() < 1e-10
    gx = [g.xbound[d] + D[d] * np.arange(NN[d]) for d in range(3)]
    Z, Y, X = np.meshgrid(gx[2], gx[1], gx[0], indexing="ij")
    pos = np.stack([X, Y, Z], -1)[~full]
    wd = -rhs[..., :3][~full] * dv
    imp_dropped = 0.5 * np.cross(pos, wd).sum(0)
    assert np.allclose(i0 - i1, imp_dropped, rtol=0, atol=1e-13)


# ------------------------------------------------------------------ yaps3d
def test_yaps_single_block_free_space(oracle):
    """One block, RHS = -Gaussian: the composed fine/coarse/inner solve reproduces the
    free-space potential erf(r/s)/(4 pi r) in the inner box."""
    from scipy.special import erf
    h = 1 / 16
    p = oracle.Params.make(h)
    g = oracle.define_sizes(p, np.array([[-1.0] * 3, [1.0] * 3]))
    nb = list(g.nnbl[:3])
    x = [g.xbound_bl[d] + np.arange(nb[d]) * h for d in range(3)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    s = 0.2
    r = np.sqrt(X ** 2 + (Y - 0.03) ** 2 + Z ** 2)
    w = np.exp(-r ** 2 / s ** 2) / (np.pi ** 1.5 * s ** 3)
    psi = erf(r / s) / (4 * np.pi * r)
    bl = g.nnbl_bl
    inner = (slice(bl[2] - 1, bl[5]), slice(bl[1] - 1, bl[4]), slice(bl[0] - 1, bl[3]))
    rhs = np.zeros(w.shape + (3,))
    rhs[inner + (0,)] = -w[inner]
    rhs[inner + (2,)] = 0.5 * w[inner]
    sol = np.zeros_like(rhs)
    r_ = oracle.lib().orc_yaps3d(C.byref(p), C.byref(g), 3, oracle._d(rhs), oracle._d(sol))
    assert r_ == 0
    e0 = np.abs(sol[inner + (0,)] - psi[inner]).max() / psi[inner].max()
    e2 = np.abs(sol[inner + (2,)] + 0.5 * psi[inner]).max() / (0.5 * psi[inner].max())
    assert e0 < 2e-2 and e2 < 2e-2, (e0, e2)
    assert np.abs(sol[inner + (1,)]).max() < 1e-14


# ------------------------------------------------------------------ Hill's vortex
def hill(pts, a=1.0, U=1.0):
    """Hill's spherical vortex translating with +U z in fluid at rest at infinity:
    vorticity (15U/(2a^2)) (-y, x, 0) inside r<a; velocity and (w.grad)u closed forms."""
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    rho2 = x * x + y * y
    r2 = rho2 + z * z
    inside = r2 < a * a
    k = 15 * U / (2 * a * a)
    w = np.where(inside[:, None], k * np.stack([-y, x, 0 * x], 1), 0.0)
    uz_in = 0.5 * U * (5 - 6 * rho2 / a ** 2 - 3 * z * z / a ** 2)
    fr_in = 1.5 * U * z / a ** 2  # u_rho / rho
    r5 = np.maximum(r2, 1e-300) ** 2.5
    uz_out = 0.5 * U * a ** 3 * (2 * r2 - 3 * rho2) / r5
    fr_out = 1.5 * U * a ** 3 * z / r5
    uz = np.where(inside, uz_in, uz_out)
    fr = np.where(inside, fr_in, fr_out)
    u = np.stack([fr * x, fr * y, uz], 1)
    # (w . grad) u = (w_phi u_rho / rho) e_phi inside
    st = np.where(inside[:, None], (k * fr_in)[:, None] * np.stack([-y, x, 0 * x], 1), 0.0)
    return w, u, st


def _yaps_blocks(oracle, p, g, rhs):
    """rhsscat_3d into every block, orc_yaps3d over all blocks; the blocks' solutions."""
    nb = g.nblocks
    dims = [tuple(g.nnbl[3 * b:3 * b + 3]) for b in range(nb)]
    bls = [tuple(g.nnbl_bl[6 * b:6 * b + 6]) for b in range(nb)]
    rb = []
    for b in range(nb):
        nx, ny, nz = dims[b]
        bl = bls[b]
        bi, bj, bk = b % g.nbi + 1, (b // g.nbi) % g.nbj + 1, b // (g.nbi * g.nbj) + 1
        x0 = (bi - 1) * (bl[3] - bl[0]) + g.nn_bl[0] - 1
        y0 = (bj - 1) * (bl[4] - bl[1]) + g.nn_bl[1] - 1
        z0 = (bk - 1) * (bl[5] - bl[2]) + g.nn_bl[2] - 1
        a = np.zeros((nz, ny, nx, 3))
        a[bl[2] - 1:bl[5], bl[1] - 1:bl[4], bl[0] - 1:bl[3]] = \
            rhs[z0:z0 + bl[5] - bl[2] + 1, y0:y0 + bl[4] - bl[1] + 1, x0:x0 + bl[3] - bl[0] + 1, :3]
        if bi > 1:
            a[:, :, bl[0] - 1] = 0
        if bj > 1:
            a[:, bl[1] - 1, :] = 0
        if bk > 1:
            a[bl[2] - 1] = 0
        rb.append(a)
    rflat = np.ascontiguousarray(np.concatenate([a.ravel() for a in rb]))
    sflat = np.zeros_like(rflat)
    assert oracle.lib().orc_yaps3d(C.byref(p), C.byref(g), 3, oracle._d(rflat), oracle._d(sflat)) == 0
    out, off = [], 0
    for b in range(nb):
        nx, ny, nz = dims[b]
        out.append(sflat[off:off + nx * ny * nz * 3].reshape(nz, ny, nx, 3))
        off += nx * ny * nz * 3
    return out, bls


@pytest.mark.parametrize("blocks", [(2, 1, 1), (2, 2, 1), (3, 1, 2)])
def test_yaps_blocks_agree_on_shared_faces(oracle, blocks):
    """Several yaps blocks: two blocks that share an inner-box face reach its nodes by
    different routes (own fine solution + the neighbour's through BBound, the neighbour's
    coarse samples removed) and must give the same values; a wrong neighbour offset or a
    one-sided term breaks the symmetry.  Against one block, the several-block answer differs
    only by the coarse correction's accuracy."""
    h = 0.125
    xp, qp = cases.single_ring(h=h, sigma=0.2)
    p = oracle.Params.make(h, blocks=blocks)
    g = oracle.define_sizes(p, xp)
    rhs = oracle.project(g.xbound, g.dpm, g.nn, g.nn_bl, xp, qp)
    sols, bls = _yaps_blocks(oracle, p, g, rhs)
    pairs = 0
    for b in range(g.nblocks):
        bi, bj, bk = b % g.nbi, (b // g.nbi) % g.nbj, b // (g.nbi * g.nbj)
        A, bl = sols[b], bls[b]
        if bi + 1 < g.nbi:  # x neighbour
            Bv, b2 = sols[b + 1], bls[b + 1]
            fa, fb = A[bl[2] - 1:bl[5], bl[1] - 1:bl[4], bl[3] - 1], Bv[b2[2] - 1:b2[5], b2[1] - 1:b2[4], b2[0] - 1]
            assert np.linalg.norm(fa - fb) <= 1e-14 * np.linalg.norm(fa)
            pairs += 1
        if bj + 1 < g.nbj:  # y neighbour
            Bv, b2 = sols[b + g.nbi], bls[b + g.nbi]
            fa, fb = A[bl[2] - 1:bl[5], bl[4] - 1, bl[0] - 1:bl[3]], Bv[b2[2] - 1:b2[5], b2[1] - 1, b2[0] - 1:b2[3]]
            assert np.linalg.norm(fa - fb) <= 1e-14 * np.linalg.norm(fa)
            pairs += 1
        if bk + 1 < g.nbk:  # z neighbour
            Bv, b2 = sols[b + g.nbi * g.nbj], bls[b + g.nbi * g.nbj]
            fa, fb = A[bl[5] - 1, bl[1] - 1:bl[4], bl[0] - 1:bl[3]], Bv[b2[2] - 1, b2[1] - 1:b2[4], b2[0] - 1:b2[3]]
            assert np.linalg.norm(fa - fb) <= 1e-14 * np.linalg.norm(fa)
            pairs += 1
    assert pairs >= 1
    up, gp = oracle.vpm2(p, g, xp, qp)
    p1 = oracle.Params.make(h)
    up1, gp1 = oracle.vpm2(p1, oracle.define_sizes(p1, xp), xp, qp)
    err = np.linalg.norm(up - up1) / np.linalg.norm(up1)
    assert 1e-6 < err < 5e-2


def test_vpm2_hills_vortex(oracle):
    h = 1 / 16
    pts = cases.cell_centres(-1.0, 1.0, h)
    w, _, _ = hill(pts)
    keep = np.linalg.norm(w, axis=1) > 0
    xp = np.ascontiguousarray(pts[keep])
    qp = np.zeros((xp.shape[0], 4))
    qp[:, :3] = -w[keep] * h ** 3
    qp[:, 3] = h ** 3
    p = oracle.Params.make(h)
    g = oracle.define_sizes(p, xp)
    up, gp = oracle.vpm2(p, g, xp, qp)
    _, u, st = hill(xp)
    eu = np.linalg.norm(up - u) / np.linalg.norm(u)
    assert eu < 2e-2, eu
    # stretching: interior of the vortex (the vorticity jumps at r=a)
    core = np.linalg.norm(xp, axis=1) < 0.8
    es = np.linalg.norm(gp[core] / h ** 3 - st[core]) / np.linalg.norm(st[core])
    assert es < 5e-2, es


# ------------------------------------------------------------------ C ABI
def test_library_exports_every_declared_symbol(vpmlib):
    hdr = open(os.path.join(ROOT, "include", "vpmx.h")).read()
    names = set(re.findall(r"\b(vpmx_[a-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 20
    lib = vpmlib.load_library()
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_create_without_gpu_fails_loudly(vpmlib):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(vpmlib.VpmError):
        vpmlib.VPM(vpmlib.Params.make(0.1))


def test_io_formats_roundtrip(oracle, vpmlib, tmp_path):
    from vpm_free_amd import io
    rng = np.random.default_rng(1)
    xp, w = rng.standard_normal((17, 3)), rng.standard_normal((17, 3))
    io.write_particles_bin(tmp_path / "particles.bin", xp, w, 0.5 ** 3)
    x2, w2, v = io.read_particles_bin(tmp_path / "particles.bin")
    assert np.array_equal(x2, xp) and np.array_equal(w2, w) and v == 0.125
    io.write_pm_inp(tmp_path / "pm.inp", 1 / 16, ncoarse=4, blocks=(1, 1, 1), ipmwrite=2, windows=[(5, 1), (9, 3)])
    p = vpmlib.read_pm_inp(str(tmp_path / "pm.inp"))
    assert list(p.dpm) == [1 / 16] * 3 and p.ncoarse == 4 and p.ipmwrite == 2
    assert list(p.ipmwstart[:2]) == [5, 9] and list(p.ipmwsteps[:2]) == [1, 3]
