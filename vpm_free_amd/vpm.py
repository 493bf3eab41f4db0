"""Python host mirror of the reference's vpm() interface over the libvpmx C ABI.

The reference entry points (vpm.f90:54-475 `vpm`, vpm_remesh.f90:5-225
`remesh_particles_3d`, the pm.inp parse at vpm.f90:132-152) are exposed with the same
names, argument meaning and call modes.  Particle arrays are numpy float64 arrays laid
out like the Fortran arrays: XP(3,NVR) is a C-contiguous array of shape (NVR, 3),
QP(neqpm+1,NVR) has shape (NVR, neqpm+1).

Everything runs in libvpmx.so on the GPU; there is no CPU fallback.  If the library is
not built, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libvpmx.so")

MAXB = 64
OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_UNSUP, ERR_IO, ERR_NODEV = range(7)
STAGES = ["Proj", "Poisson fine", "Bounds", "pmesh_coarse", "pm_bc2", "pmesh_final", "fd", "back", "remeshing"]
KERNELS = ["boundsum_fine", "p2m", "m2p", "dst_y_fine", "dst_z_fine"]


class VpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vpmx error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """pm.inp contents (vpm.f90:132-152)."""
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
    """Geometry of define_sizes (vpm.f90:477-660)."""
    _fields_ = [
        ("xbound", C.c_double * 6), ("dpm", C.c_double * 3), ("nn", C.c_int32 * 3), ("nn_bl", C.c_int32 * 6),
        ("nblocks", C.c_int32), ("nbi", C.c_int32), ("nbj", C.c_int32), ("nbk", C.c_int32),
        ("xbound_bl", C.c_double * (6 * MAXB)), ("nnbl", C.c_int32 * (3 * MAXB)),
        ("nnbl_bl", C.c_int32 * (6 * MAXB)), ("xbound_coarse", C.c_double * 6), ("dpm_coarse", C.c_double * 3),
        ("nn_coarse", C.c_int32 * 3), ("nn_bl_coarse", C.c_int32 * 6), ("ilevmax", C.c_int32),
        ("defined", C.c_int32),
    ]


_lib = None


def load_library():
    """Load libvpmx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m vpm_free_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(LIB_PATH)
    d, i32, i64 = C.POINTER(C.c_double), C.c_int32, C.c_int64
    vp = C.c_void_p
    sig = {
        "vpmx_read_pm_inp": [C.c_char_p, C.POINTER(Params)],
        "vpmx_define_sizes": [C.POINTER(Params), C.POINTER(Geom), d, i64, i32],
        "vpmx_define_sizes_weighted": [C.POINTER(Params), C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                       C.POINTER(Geom), d, i64, i32],
        "vpmx_set_block_weights": [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)],
        "vpmx_create": [C.POINTER(vp), i32],
        "vpmx_set_params": [vp, C.POINTER(Params)],
        "vpmx_get_geometry": [vp, C.POINTER(Geom)],
        "vpmx_set_precision": [vp, i32],
        "vpmx_set_option": [vp, C.c_char_p, i32],
        "vpmx_tuning_report": [vp, C.c_char_p, i64],
        "vpmx_vpm": [vp, d, d, d, d, i64, i32, i32, i32, C.c_double],
        "vpmx_remesh_count": [vp, d, d, i64, i32, i32, C.POINTER(i64)],
        "vpmx_remesh_fetch": [vp, d, d],
        "vpmx_get_grid": [vp, i32, d, i64],
        "vpmx_particles_upload": [vp, d, d, i64, i32],
        "vpmx_particles_count": [vp, C.POINTER(i64)],
        "vpmx_particles_download": [vp, d, d, d, d],
        "vpmx_vpm_resident": [vp, i32, i32],
        "vpmx_remesh_resident": [vp, i32],
        "vpmx_step_rk2": [vp, C.c_double, d, i32, i32, i32],
        "vpmx_diagnostics": [vp, d, d],
        "vpmx_batch_create": [C.POINTER(vp), i32, i32],
        "vpmx_batch_destroy": [vp],
        "vpmx_batch_cloud": [vp, i32, C.POINTER(vp)],
        "vpmx_batch_step_rk2": [vp, C.c_double, d, i32, i32, i32],
        "vpmx_batch_times": [vp, d, d],
        "vpmx_diagnostics_noalloc": [vp, vp, vp, i64, i32, vp],
        "vpmx_adjoint_probe_energy": [vp, d, d, i64, i32, d, i64, d, d, d],
        "vpmx_adjoint_rollout": [vp, d, d, i64, i32, i32, C.c_double, d, i32, d, i64, d, d, d],
        "vpmx_stage_solve": [vp, d, d, C.POINTER(i32), i32, C.POINTER(i32), d, d, i32, d],
        "vpmx_stage_pmesh_inf": [vp, d, d, C.POINTER(i32), i32, d, d],
        "vpmx_stage_pmesh_inf_lev": [vp, d, d, C.POINTER(i32), i32, d, d, i32, i32, i32],
        "vpmx_stage_project": [vp, d, d, C.POINTER(i32), C.POINTER(i32), d, d, i64, i32, i32, i32, d],
        "vpmx_stage_times": [vp, d, i32],
        "vpmx_kernel_times": [vp, d, C.POINTER(i64), i32, i32],
        "vpmx_kernel_work": [vp, d, i32],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = i32
    lib.vpmx_destroy.argtypes = [vp]
    lib.vpmx_destroy.restype = None
    lib.vpmx_batch_destroy.restype = None
    lib.vpmx_last_error.argtypes = [vp]
    lib.vpmx_last_error.restype = C.c_char_p
    lib.vpmx_batch_last_error.restype = C.c_char_p
    lib.vpmx_batch_last_error.argtypes = [vp]
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def read_pm_inp(path):
    """Parse pm.inp exactly as vpm.f90:132-152 reads it."""
    lib = load_library()
    p = Params()
    r = lib.vpmx_read_pm_inp(str(path).encode(), C.byref(p))
    if r != OK:
        raise VpmError(r, f"cannot read {path}")
    return p


def _weights(w):
    return None if w is None else np.ascontiguousarray(w, dtype=np.int32)


def _wp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def define_sizes(params, xp, ntime=0, geom=None, weights=None):
    """define_sizes (vpm.f90:477-660) on the host; returns the (updated) Geom.  weights
    (wx, wy, wz): yaps blocks of uneven extents in proportion to them (None: even split)."""
    lib = load_library()
    g = geom if geom is not None else Geom()
    xp = _f64(xp)
    w = [None, None, None] if weights is None else [_weights(v) for v in weights]
    r = lib.vpmx_define_sizes_weighted(C.byref(params), _wp(w[0]), _wp(w[1]), _wp(w[2]), C.byref(g), _dp(xp),
                                       xp.shape[0] if xp.ndim == 2 else 0, ntime)
    if r != OK:
        raise VpmError(r, "define_sizes")
    return g


def geom_dict(g):
    """Plain-Python view of a Geom (first block only when BLOCKS=1)."""
    nb = g.nblocks
    return {
        "xbound": list(g.xbound), "dpm": list(g.dpm), "nn": list(g.nn), "nn_bl": list(g.nn_bl),
        "nblocks": nb, "xbound_bl": [list(g.xbound_bl[6 * b:6 * b + 6]) for b in range(nb)],
        "nnbl": [list(g.nnbl[3 * b:3 * b + 3]) for b in range(nb)],
        "nnbl_bl": [list(g.nnbl_bl[6 * b:6 * b + 6]) for b in range(nb)],
        "xbound_coarse": list(g.xbound_coarse), "dpm_coarse": list(g.dpm_coarse),
        "nn_coarse": list(g.nn_coarse), "nn_bl_coarse": list(g.nn_bl_coarse), "ilevmax": g.ilevmax,
    }


class VPM:
    """One device context: vpm() modes, remesh_particles_3d(), device-resident stepping."""

    @classmethod
    def _borrow(cls, lib, handle, params):
        """A view of a context owned elsewhere (a batch's cloud); closing it does nothing."""
        c = cls.__new__(cls)
        c._lib, c._h, c._owned, c.neqpm = lib, handle, False, 3
        c.params = params
        c._check(c._lib.vpmx_set_params(c._h, C.byref(params)))
        return c

    def __init__(self, params, device=0):
        self._lib = load_library()
        h = C.c_void_p()
        r = self._lib.vpmx_create(C.byref(h), device)
        if r != OK:
            raise VpmError(r, "vpmx_create failed (no GPU?)")
        self._h = h
        self.params = params
        self._check(self._lib.vpmx_set_params(self._h, C.byref(params)))
        self.neqpm = 3

    def set_block_weights(self, wx=None, wy=None, wz=None):
        """Uneven yaps blocks for the next vpm mode 0 (None: the even split)."""
        self._w = [_weights(wx), _weights(wy), _weights(wz)]  # kept alive for the call
        self._check(self._lib.vpmx_set_block_weights(self._h, _wp(self._w[0]), _wp(self._w[1]), _wp(self._w[2])))

    def set_option(self, name, value):
        """boundsum (0 lattice, 1 direct, 2 spectral), dst_reuse (0/1), face_derivs (-1/0/1), fine_partial (-1/0/1), tuning (0/1),
        tune_variant (-1 or a variant)."""
        self._check(self._lib.vpmx_set_option(self._h, name.encode(), int(value)))

    def tuning_report(self):
        buf = C.create_string_buffer(1 << 16)
        self._check(self._lib.vpmx_tuning_report(self._h, buf, len(buf)))
        return buf.value.decode()

    def set_precision(self, mode):
        """0: plain FP64 form; 1: mixed form (FP32 Poisson passes, FP64 boundary sums)."""
        self._check(self._lib.vpmx_set_precision(self._h, int(mode)))

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                self._lib.vpmx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, r):
        if r != OK:
            raise VpmError(r, self._lib.vpmx_last_error(self._h).decode(errors="replace"))

    # -- reference interface -------------------------------------------------------
    def vpm(self, XP, QP, UP, GP, NVR, neqpm, WhatToDo, NTIME=0, NI=0.0):
        """vpm(XP,QP,UP,GP,NVR,neqpm,WhatToDo,...): UP/GP are overwritten in mode 2."""
        XP = _f64(XP)
        QP = _f64(QP)
        for a in (UP, GP):
            if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous):
                raise ValueError("UP/GP must be C-contiguous float64 arrays")
        self._check(self._lib.vpmx_vpm(self._h, _dp(XP), _dp(QP), _dp(UP), _dp(GP), int(NVR), int(neqpm),
                                       int(WhatToDo), int(NTIME), float(NI)))

    def remesh_particles_3d(self, XP, QP, mrem=1):
        """remesh_particles_3d(1): returns the new (XP, QP)."""
        XP = _f64(XP)
        QP = _f64(QP)
        n = C.c_int64()
        self._check(self._lib.vpmx_remesh_count(self._h, _dp(XP), _dp(QP), XP.shape[0], QP.shape[1] - 1, mrem,
                                                C.byref(n)))
        xo = np.empty((n.value, 3))
        qo = np.empty((n.value, QP.shape[1]))
        self._check(self._lib.vpmx_remesh_fetch(self._h, _dp(xo), _dp(qo)))
        return xo, qo

    @property
    def geometry(self):
        g = Geom()
        self._check(self._lib.vpmx_get_geometry(self._h, C.byref(g)))
        return g

    def get_grid(self, which):
        """0: RHS_pm(4,N) ; 1: SOL_pm after the solve ; 2: SOL_pm stretching ; 3: vel (3,N)."""
        g = self.geometry
        nx, ny, nz = g.nn
        if which == 0:
            out = np.empty((nz, ny, nx, 4))
        elif which in (1, 2):
            out = np.empty((nz, ny, nx, 3))
        elif which == 3:
            out = np.empty((3, nz, ny, nx))
        else:
            raise ValueError(which)
        self._check(self._lib.vpmx_get_grid(self._h, which, _dp(out), out.size))
        return out

    def stage_times(self):
        ms = np.zeros(len(STAGES))
        self._check(self._lib.vpmx_stage_times(self._h, _dp(ms), len(STAGES)))
        return dict(zip(STAGES, ms.tolist()))

    def kernel_times(self, reset=False):
        """{class: (total device ms, launches)} from the dispatches' own events."""
        ms = np.zeros(len(KERNELS))
        n = (C.c_int64 * len(KERNELS))()
        self._check(self._lib.vpmx_kernel_times(self._h, _dp(ms), n, len(KERNELS), 1 if reset else 0))
        return {k: (ms[i], n[i]) for i, k in enumerate(KERNELS)}

    def kernel_work(self):
        """{class: algorithmic work of its last launch} (FP64 flops for boundsum_fine)."""
        w = np.zeros(len(KERNELS))
        self._check(self._lib.vpmx_kernel_work(self._h, _dp(w), len(KERNELS)))
        return dict(zip(KERNELS, w.tolist()))

    # -- device-resident particles -------------------------------------------------
    def upload(self, XP, QP):
        XP, QP = _f64(XP), _f64(QP)
        self._check(self._lib.vpmx_particles_upload(self._h, _dp(XP), _dp(QP), XP.shape[0], QP.shape[1] - 1))

    def count(self):
        n = C.c_int64()
        self._check(self._lib.vpmx_particles_count(self._h, C.byref(n)))
        return n.value

    def download(self):
        n = self.count()
        xp, qp, up, gp = np.empty((n, 3)), np.empty((n, self.neqpm + 1)), np.empty((n, 3)), np.empty((n, 3))
        self._check(self._lib.vpmx_particles_download(self._h, _dp(xp), _dp(qp), _dp(up), _dp(gp)))
        return xp, qp, up, gp

    def vpm_resident(self, WhatToDo, NTIME=0):
        self._check(self._lib.vpmx_vpm_resident(self._h, WhatToDo, NTIME))

    def remesh_resident(self, mrem=1):
        self._check(self._lib.vpmx_remesh_resident(self._h, mrem))

    def step_rk2(self, dt, uinf=(0.0, 0.0, 0.0), ntime=1, mrem=1, remesh=True):
        u = np.asarray(uinf, dtype=np.float64)
        self._check(self._lib.vpmx_step_rk2(self._h, float(dt), _dp(u), ntime, mrem, 1 if remesh else 0))

    def diagnostics(self):
        circ, imp = np.zeros(3), np.zeros(3)
        self._check(self._lib.vpmx_diagnostics(self._h, _dp(circ), _dp(imp)))
        return circ, imp

    def adjoint_probe_energy(self, XP, QP, probes):
        """J = 1/2 sum |u(probe)|^2 of the vpm(2) velocity grid (plain FP64 form) and its gradient
        with respect to QP(:, 0:3) and XP: returns (J, dJ/dQP, dJ/dXP)."""
        XP, QP, pr = _f64(XP), _f64(QP), _f64(np.asarray(probes).reshape(-1, 3))
        n = XP.shape[0]
        dq, dx = np.zeros((n, 3)), np.zeros((n, 3))
        J = np.zeros(1)
        self._check(self._lib.vpmx_adjoint_probe_energy(self._h, _dp(XP), _dp(QP), n, QP.shape[1] - 1, _dp(pr),
                                                        pr.shape[0], _dp(J), _dp(dq), _dp(dx)))
        return float(J[0]), dq, dx

    def adjoint_rollout(self, XP, QP, probes, nsteps, dt, uinf=(0.0, 0.0, 0.0), ckpt=2):
        """J = 1/2 sum |u(probe)|^2 after nsteps RK2 steps with remesh (plain FP64 form) and its
        gradient with respect to the starting QP(:, 0:3) and XP: returns (J, dJ/dQP, dJ/dXP)."""
        XP, QP, pr = _f64(XP), _f64(QP), _f64(np.asarray(probes).reshape(-1, 3))
        n = XP.shape[0]
        dq, dx = np.zeros((n, 3)), np.zeros((n, 3))
        J = np.zeros(1)
        u = _f64(np.asarray(uinf, dtype=np.float64).reshape(3))
        self._check(self._lib.vpmx_adjoint_rollout(self._h, _dp(XP), _dp(QP), n, QP.shape[1] - 1, int(nsteps),
                                                   float(dt), _dp(u), int(ckpt), _dp(pr), pr.shape[0], _dp(J),
                                                   _dp(dq), _dp(dx)))
        return float(J[0]), dq, dx

    def diagnostics_noalloc(self, xp_dev, qp_dev, n, qstride, out_dev):
        """vpmx_diagnostics_noalloc on device addresses (integers); out_dev receives 6 doubles
        once the context's stream has run the call."""
        self._check(self._lib.vpmx_diagnostics_noalloc(self._h, C.c_void_p(xp_dev), C.c_void_p(qp_dev), int(n),
                                                        int(qstride), C.c_void_p(out_dev)))

    # -- single routines of the path -------------------------------------------------
    def stage_solve(self, xbound, dpm, nn, nc, win, rhs, sol=None, zero_bc=False):
        xb, dp = _f64(xbound), _f64(dpm)
        nn_a, nn_p = _ip(nn)
        w_a, w_p = _ip(win)
        rhs = _f64(rhs)
        out = np.array(sol if sol is not None else np.zeros_like(rhs), dtype=np.float64, order="C")
        if zero_bc:
            self._check(self._lib.vpmx_stage_solve(self._h, _dp(xb), _dp(dp), nn_p, nc, w_p, _dp(rhs), None, 1,
                                                   _dp(out)))
        else:
            self._check(self._lib.vpmx_stage_solve(self._h, _dp(xb), _dp(dp), nn_p, nc, w_p, _dp(rhs), _dp(out), 0,
                                                   None))
        return out

    def stage_pmesh_inf(self, xbound, dpm, nn, rhs, sol, itree=0, levmax=1, ibctyp=2):
        """pmesh(iynbc=1) (pmlib.f90:69) with its ibctyp / itree / levmax arguments."""
        xb, dp = _f64(xbound), _f64(dpm)
        nn_a, nn_p = _ip(nn)
        rhs = _f64(rhs)
        out = np.array(sol, dtype=np.float64, order="C")
        self._check(self._lib.vpmx_stage_pmesh_inf_lev(self._h, _dp(xb), _dp(dp), nn_p, 3, _dp(rhs), _dp(out),
                                                       int(ibctyp), int(itree), int(levmax)))
        return out

    def stage_project(self, xbound, dpm, nn, nn_bl, XP, QP, itype=4, idvpm=1):
        xb, dp = _f64(xbound), _f64(dpm)
        nn_a, nn_p = _ip(nn)
        nb_a, nb_p = _ip(nn_bl)
        XP, QP = _f64(XP), _f64(QP)
        nx, ny, nz = [int(v) for v in nn]
        out = np.empty((nz, ny, nx, QP.shape[1]))
        self._check(self._lib.vpmx_stage_project(self._h, _dp(xb), _dp(dp), nn_p, nb_p, _dp(XP), _dp(QP),
                                                 XP.shape[0], QP.shape[1], itype, idvpm, _dp(out)))
        return out


class Batch:
    """Many independent clouds stepped together (vpmx_batch_*); cloud(i) is a VPM view of
    cloud i's context for its set-up (parameters, particles, geometry, precision)."""

    def __init__(self, nclouds, device=0):
        self._lib = load_library()
        h = C.c_void_p()
        r = self._lib.vpmx_batch_create(C.byref(h), device, int(nclouds))
        if r != OK:
            raise VpmError(r, "vpmx_batch_create failed")
        self._h = h
        self.n = int(nclouds)
        self._views = {}

    def _check(self, r):
        if r != OK:
            raise VpmError(r, self._lib.vpmx_batch_last_error(self._h).decode(errors="replace"))

    def cloud(self, i, params=None):
        if i not in self._views:
            h = C.c_void_p()
            self._check(self._lib.vpmx_batch_cloud(self._h, int(i), C.byref(h)))
            self._views[i] = VPM._borrow(self._lib, h, params)
        elif params is not None:
            self._views[i].params = params
            self._views[i]._check(self._lib.vpmx_set_params(self._views[i]._h, C.byref(params)))
        return self._views[i]

    def step_rk2(self, dt, uinf=(0.0, 0.0, 0.0), mrem=1, remesh=True, graph=False):
        u = np.asarray(uinf, dtype=np.float64)
        self._check(self._lib.vpmx_batch_step_rk2(self._h, float(dt), _dp(u), int(mrem), 1 if remesh else 0,
                                                  1 if graph else 0))

    def times(self):
        """(device ms of the last step, host ms spent queueing it)."""
        a, b = np.zeros(1), np.zeros(1)
        self._check(self._lib.vpmx_batch_times(self._h, _dp(a), _dp(b)))
        return float(a[0]), float(b[0])

    def close(self):
        if getattr(self, "_h", None):
            for v in self._views.values():
                v._h = None
            self._lib.vpmx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
