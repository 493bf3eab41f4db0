"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product never imports this module.

Grids are numpy arrays of shape (nz, ny, nx, nc) -- the memory layout of the
reference's RHS_pm(nc, nx, ny, nz).  Particles: XP (NVR, 3), QP (NVR, neq+1).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_PROJLIB = os.path.join(HERE, "_ref", "libref_projlib.so")
MAXB = 64


class Params(C.Structure):
    _fields_ = [
        ("dpm", C.c_double * 3), ("interf_iproj", C.c_int32), ("ibctyp", C.c_int32), ("idvpm", C.c_int32),
        ("epsvol", C.c_double), ("ncoarse", C.c_int32), ("nbi", C.c_int32), ("nbj", C.c_int32),
        ("nbk", C.c_int32), ("nremesh", C.c_int32), ("ncell_rem", C.c_int32), ("iyntree", C.c_int32),
        ("ilevmax", C.c_int32), ("ompthreads", C.c_int32), ("idefine", C.c_int32), ("iynslice", C.c_int32),
        ("ipmwrite", C.c_int32), ("ipmwstart", C.c_int32 * 10), ("ipmwsteps", C.c_int32 * 10),
    ]

    @classmethod
    def make(cls, dpm, interf_iproj=4, ibctyp=2, idvpm=1, epsvol=0.0, ncoarse=8, blocks=(1, 1, 1),
             nremesh=1, ncell_rem=1, iyntree=0, ilevmax=1, ompthreads=1, idefine=1, iynslice=0, ipmwrite=0):
        p = cls()
        d = np.broadcast_to(np.asarray(dpm, dtype=np.float64), (3,))
        p.dpm[:] = [float(v) for v in d]
        p.interf_iproj, p.ibctyp, p.idvpm, p.epsvol = interf_iproj, ibctyp, idvpm, epsvol
        p.ncoarse = ncoarse
        p.nbi, p.nbj, p.nbk = blocks
        p.nremesh, p.ncell_rem, p.iyntree, p.ilevmax = nremesh, ncell_rem, iyntree, ilevmax
        p.ompthreads, p.idefine, p.iynslice, p.ipmwrite = ompthreads, idefine, iynslice, ipmwrite
        return p


class Geom(C.Structure):
    _fields_ = [
        ("xbound", C.c_double * 6), ("dpm", C.c_double * 3), ("nn", C.c_int32 * 3), ("nn_bl", C.c_int32 * 6),
        ("nblocks", C.c_int32), ("nbi", C.c_int32), ("nbj", C.c_int32), ("nbk", C.c_int32),
        ("xbound_bl", C.c_double * (6 * MAXB)), ("nnbl", C.c_int32 * (3 * MAXB)),
        ("nnbl_bl", C.c_int32 * (6 * MAXB)), ("xbound_coarse", C.c_double * 6), ("dpm_coarse", C.c_double * 3),
        ("nn_coarse", C.c_int32 * 3), ("nn_bl_coarse", C.c_int32 * 6), ("ilevmax", C.c_int32),
        ("defined", C.c_int32),
    ]


_lib = None
D = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int32)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/source"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        i32, i64 = C.c_int32, C.c_int64
        sig = {
            "orc_read_pm_inp": (i32, [C.c_char_p, C.POINTER(Params)]),
            "orc_projection_fun": (C.c_double, [C.c_int, C.c_double]),
            "orc_definepm": (None, [C.c_int, D, D, C.c_int, I32P, I32P, I32P]),
            "orc_define_sizes": (i32, [C.POINTER(Params), C.POINTER(Geom), D, i64, C.c_int]),
            "orc_project": (i32, [D, D, I32P, I32P, D, D, i64, C.c_int, C.c_int, C.c_int, D]),
            "orc_solve_window": (None, [D, D, I32P, C.c_int, C.c_int, I32P, D, D, C.c_int, D]),
            "orc_pmesh_inf": (None, [D, D, I32P, C.c_int, D, D]),
            "orc_normalderiv": (i64, [D, D, I32P, C.c_int, D, D, D, D]),
            "orc_boundsum": (None, [D, D, I32P, C.c_int, i64, D, D, D, D]),
            "orc_yaps3d": (i32, [C.POINTER(Params), C.POINTER(Geom), C.c_int, D, D]),
            "orc_velocity": (None, [C.POINTER(Geom), C.c_int, D, D, D, D, D, C.c_int]),
            "orc_m2p": (None, [C.POINTER(Geom), C.c_int, C.c_int, D, D, D, D, D, D, D, D, i64]),
            "orc_vpm2": (i32, [C.POINTER(Params), C.POINTER(Geom), C.c_int, D, D, D, D, i64, D, D, D, D]),
            "orc_remesh": (i64, [C.POINTER(Params), C.POINTER(Geom), C.c_int, C.c_int, D, D, i64, D, D]),
            "orc_diagnostics": (None, [C.c_int, D, D, i64, D, D]),
            "orc_dst1": (None, [C.c_int, D, D]),
            "orc_num_threads": (C.c_int, []),
            "orc_set_threads": (None, [C.c_int]),
            "orc_timers": (None, [D, C.c_int]),
            "orc_set_boundsum_stride": (None, [C.c_int]),
            "orc_diffuse": (None, [C.POINTER(Geom), C.c_int, C.c_double, D, D]),
            "orc_set_tree": (None, [C.c_int]),
            "orc_set_yaps_plant": (None, [C.c_int]),
            "orc_set_block_weights": (None, [I32P, I32P, I32P, C.c_int, C.c_int, C.c_int]),
            "orc_set_tree_factor": (None, [C.c_double]),
            "orc_boundsum_tree": (None, [D, D, I32P, C.c_int, i64, D, D, D, D, C.c_int]),
            "orc_vpm45": (i32, [C.POINTER(Params), C.POINTER(Geom), C.c_int, C.c_int, C.c_double, D, D, D, D, i64,
                                D, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(D) if a is not None else None


def _i(a):
    return a.ctypes.data_as(I32P)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def timers(reset=False):
    """Seconds spent in [projection, Poisson solves, boundary sums, whole vpm2/remesh]."""
    t = np.zeros(4)
    lib().orc_timers(_d(t), 1 if reset else 0)
    return t


def set_boundsum_stride(s):
    """Evaluate only every s-th boundary target (bounded timing samples of big cases)."""
    lib().orc_set_boundsum_stride(int(s))


def projection_fun(itype, x):
    return lib().orc_projection_fun(int(itype), float(x))


def definepm(itype, xbound, dpm, ndum, nsize):
    xb = _f(xbound).copy()
    dp = _f(dpm).copy()
    ns = np.asarray(nsize, dtype=np.int32)
    nn = np.zeros(3, np.int32)
    nnb = np.zeros(6, np.int32)
    lib().orc_definepm(itype, _d(xb), _d(dp), ndum, _i(ns), _i(nn), _i(nnb))
    return xb, dp, nn, nnb


def define_sizes(params, xp, ntime=0, geom=None):
    g = geom if geom is not None else Geom()
    xp = _f(xp)
    r = lib().orc_define_sizes(C.byref(params), C.byref(g), _d(xp), xp.shape[0], ntime)
    if r != 0:
        raise RuntimeError(f"orc_define_sizes: {r}")
    return g


def project(xbound, dpm, nn, nn_bl, xp, qp, itype=4, idvpm=1):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    nnb = np.asarray(nn_bl, np.int32)
    xp, qp = _f(xp), _f(qp)
    nc = qp.shape[1]
    out = np.zeros((nn[2], nn[1], nn[0], nc))
    lib().orc_project(_d(xb), _d(dp), _i(nn), _i(nnb), _d(xp), _d(qp), xp.shape[0], nc, itype, idvpm, _d(out))
    return out


def solve_window(xbound, dpm, nn, win, rhs, sol=None, zero_bc=False):
    """All components of rhs (nz,ny,nx,nc); returns the new solution array."""
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    w = np.asarray(win, np.int32)
    rhs = _f(rhs)
    nc = rhs.shape[-1]
    out = np.array(sol if sol is not None else np.zeros_like(rhs), dtype=np.float64, order="C")
    for c in range(nc):
        if zero_bc:
            lib().orc_solve_window(_d(xb), _d(dp), _i(nn), nc, c, _i(w), _d(rhs), None, 1, _d(out))
        else:
            lib().orc_solve_window(_d(xb), _d(dp), _i(nn), nc, c, _i(w), _d(rhs), _d(out), 0, None)
    return out


def pmesh_inf(xbound, dpm, nn, rhs, sol):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    rhs = _f(rhs)
    out = np.array(sol, dtype=np.float64, order="C")
    lib().orc_pmesh_inf(_d(xb), _d(dp), _i(nn), rhs.shape[-1], _d(rhs), _d(out))
    return out


def normalderiv(xbound, dpm, nn, sol0):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    s0 = _f(sol0)
    nc = s0.shape[-1]
    mx = 2 * (nn[0] * nn[1] + nn[0] * nn[2] + nn[1] * nn[2])
    src = np.zeros((mx, nc))
    cxyz = np.zeros((mx, 3))
    ds = np.zeros(mx)
    nb = lib().orc_normalderiv(_d(xb), _d(dp), _i(nn), nc, _d(s0), _d(src), _d(cxyz), _d(ds))
    return src[:nb].copy(), cxyz[:nb].copy(), ds[:nb].copy()


def boundsum(xbound, dpm, nn, src, cxyz, ds, sol):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    src, cxyz, ds = _f(src), _f(cxyz), _f(ds)
    out = np.array(sol, dtype=np.float64, order="C")
    lib().orc_boundsum(_d(xb), _d(dp), _i(nn), src.shape[1], src.shape[0], _d(src), _d(cxyz), _d(ds), _d(out))
    return out


def vpm2(params, geom, xp, qp, neqpm=3, fields=False):
    """vpm(WhatToDo=2): returns UP, GP (and RHS, PSI, SOL, vel grids when fields=True)."""
    xp, qp = _f(xp), _f(qp)
    n = xp.shape[0]
    up, gp = np.zeros((n, 3)), np.zeros((n, 3))
    nx, ny, nz = geom.nn
    rhs = psi = sol = vel = None
    if fields:
        rhs = np.zeros((nz, ny, nx, neqpm + 1))
        psi = np.zeros((nz, ny, nx, neqpm))
        sol = np.zeros((nz, ny, nx, neqpm))
        vel = np.zeros((3, nz, ny, nx))
    r = lib().orc_vpm2(C.byref(params), C.byref(geom), neqpm, _d(xp), _d(qp), _d(up), _d(gp), n, _d(rhs), _d(psi),
                       _d(sol), _d(vel))
    if r != 0:
        raise RuntimeError(f"orc_vpm2: {r}")
    if fields:
        return up, gp, {"rhs": rhs, "psi": psi, "sol": sol, "vel": vel}
    return up, gp


def set_yaps_plant(mode):
    """Plant a mistake in the yaps composition (tests of the tests): 0 none, 1 no TSC
    correction, 2 its sign flipped, 3 no neighbour fine values, 4 nnb = 0 everywhere, 5 the
    coarse Laplacian over the footprint boundary too."""
    lib().orc_set_yaps_plant(int(mode))


def set_block_weights(wx=None, wy=None, wz=None):
    """Uneven yaps blocks for later define_sizes calls: extents in proportion to the weights
    (None: the reference's even split)."""
    if wx is None:
        lib().orc_set_block_weights(None, None, None, 0, 0, 0)
        return
    a = [np.ascontiguousarray(w, dtype=np.int32) for w in (wx, wy, wz)]
    lib().orc_set_block_weights(_i(a[0]), _i(a[1]), _i(a[2]), a[0].size, a[1].size, a[2].size)


def set_tree(levmax):
    """Boundary sums of pmesh_inf by the tree with this levmax (>= 0), or direct (-1)."""
    lib().orc_set_tree(int(levmax))


def set_tree_factor(f):
    lib().orc_set_tree_factor(float(f))


def boundsum_tree(xbound, dpm, nn, src, cxyz, ds, sol, levmax):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    src, cxyz, ds = _f(src), _f(cxyz), _f(ds)
    out = np.array(sol, dtype=np.float64, order="C")
    lib().orc_boundsum_tree(_d(xb), _d(dp), _i(nn), src.shape[1], src.shape[0], _d(src), _d(cxyz), _d(ds), _d(out),
                            int(levmax))
    return out


def set_ibctyp(ibctyp):
    """Boundary sums of pmesh_inf with point sources (2) or panels (3, FSOUR_A4)."""
    lib().orc_set_ibctyp(int(ibctyp))


def fsour_a4(xo, rg, iplane, dpm, area=0.0):
    lib().orc_fsour_a4.restype = C.c_double
    return lib().orc_fsour_a4(_d(_f(xo)), _d(_f(rg)), int(iplane), _d(_f(dpm)), C.c_double(area))


def boundsum_s(xbound, dpm, nn, src, cxyz, sol):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    src, cxyz = _f(src), _f(cxyz)
    out = np.array(sol, dtype=np.float64, order="C")
    lib().orc_boundsum_s(_d(xb), _d(dp), _i(nn), src.shape[1], src.shape[0], _d(src), _d(cxyz), _d(out))
    return out


def boundsum_tree_s(xbound, dpm, nn, src, cxyz, ds, sol, levmax):
    xb, dp = _f(xbound), _f(dpm)
    nn = np.asarray(nn, np.int32)
    src, cxyz, ds = _f(src), _f(cxyz), _f(ds)
    out = np.array(sol, dtype=np.float64, order="C")
    lib().orc_boundsum_tree_s(_d(xb), _d(dp), _i(nn), src.shape[1], src.shape[0], _d(src), _d(cxyz), _d(ds),
                              _d(out), int(levmax))
    return out


def diffuse(geom, ni, rhs):
    """diffuse_vort_3d on rhs (G x 4, node-major, x fastest); returns SOL (G x 3)."""
    rhs = _f(rhs)
    G = rhs.shape[0] if rhs.ndim == 2 else int(np.prod(rhs.shape[:-1]))
    out = np.zeros((G, 3))
    lib().orc_diffuse(C.byref(geom), 3, ni, _d(rhs), _d(out))
    return out


def vpm45(params, geom, xp, qp, mode, ni=0.0, vel=None, neqpm=3):
    """vpm WhatToDo 4 (stretching from the velocity grids `vel`, 3 x G) or 5 (diffusion).
    Returns UP, GP and the SOL grid."""
    xp, qp = _f(xp), _f(qp)
    n = xp.shape[0]
    up, gp = np.zeros((n, 3)), np.zeros((n, 3))
    G = int(geom.nn[0]) * int(geom.nn[1]) * int(geom.nn[2])
    sol = np.zeros((G, neqpm))
    v = _f(vel) if vel is not None else None
    r = lib().orc_vpm45(C.byref(params), C.byref(geom), neqpm, mode, ni, _d(xp), _d(qp), _d(up), _d(gp),
                        n, _d(v) if v is not None else None, _d(sol))
    if r != 0:
        raise RuntimeError("orc_vpm45 failed")
    return up, gp, sol


def remesh(params, geom, xp, qp, neqpm=3, mrem=1):
    xp, qp = _f(xp), _f(qp)
    n = lib().orc_remesh(C.byref(params), C.byref(geom), neqpm, mrem, _d(xp), _d(qp), xp.shape[0], None, None)
    if n < 0:
        raise RuntimeError("orc_remesh")
    xo, qo = np.zeros((n, 3)), np.zeros((n, neqpm + 1))
    lib().orc_remesh(C.byref(params), C.byref(geom), neqpm, mrem, _d(xp), _d(qp), xp.shape[0], _d(xo), _d(qo))
    return xo, qo


def diagnostics(xp, qp, neqpm=3):
    xp, qp = _f(xp), _f(qp)
    circ, imp = np.zeros(3), np.zeros(3)
    lib().orc_diagnostics(neqpm, _d(xp), _d(qp), xp.shape[0], _d(circ), _d(imp))
    return circ, imp


def dst1(x):
    x = _f(x)
    out = np.zeros_like(x)
    lib().orc_dst1(x.shape[0], _d(x), _d(out))
    return out


def step_rk2(params, geom, xp, qp, dt, uinf=(0.0, 0.0, 0.0), ntime=1, mrem=1, remesh_step=True):
    """One RK2 (midpoint) step of the application: 2 x vpm(2), Euler-form updates
    (test.f90:104-117), vpm(0) and remesh_particles_3d(1)."""
    uinf = np.asarray(uinf, dtype=np.float64)
    x0, q0 = _f(xp), _f(qp)
    up, gp = vpm2(params, geom, x0, q0)
    xm = x0 + (up + uinf) * (0.5 * dt)
    qm = q0.copy()
    qm[:, :3] = q0[:, :3] - gp * (0.5 * dt)
    up, gp = vpm2(params, geom, xm, qm)
    x1 = x0 + (up + uinf) * dt
    q1 = q0.copy()
    q1[:, :3] = q0[:, :3] - gp * dt
    if not remesh_step:
        return x1, q1
    define_sizes(params, x1, ntime=ntime, geom=geom)
    return remesh(params, geom, x1, q1, mrem=mrem)
