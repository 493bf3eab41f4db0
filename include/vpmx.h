/*
 * vpmx.h -- C ABI of the MI355X (gfx950) VPM core.
 *
 * Drop-in boundary for the reference's vpm() path (papis84/vpm_free).  The Fortran
 * module vpm_free_amd/fortran/vpm_lib.f90 keeps the reference's own entry points
 * (vpm(), remesh_particles_3d(), the pm.inp parse and the public modules) and calls
 * these functions through ISO_C_BINDING; Python reaches them through ctypes.
 *
 * Conventions
 *   - FP64 at the ABI.  Particle arrays are the reference's Fortran arrays as plain
 *     pointers: XP(3,NVR), QP(neqpm+1,NVR), UP(3,NVR), GP(3,NVR), column-major.
 *   - Grids returned by vpmx_get_grid use the reference layout: RHS_pm(neqpm+1,NX,NY,NZ),
 *     SOL_pm(neqpm,NX,NY,NZ) component-fastest, velvr*_pm(NX,NY,NZ).
 *   - Every call returns int32 status, 0 = ok; vpmx_last_error() gives the message.
 *     Failures the reference turns into STOP ('WRONG IDVPM', 'No such projection
 *     function', IPMWRITE>10, ...) come back as status codes.
 *   - One host thread per context, one HIP stream per context; calls are synchronous
 *     at return.  One GPU, one process: the reference's MPI ranks collapse into one
 *     device (BLOCKS = NBI*NBJ*NBK blocks on that device).
 */
#ifndef VPMX_H
#define VPMX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPMX_MAXB 64

enum {
    VPMX_OK = 0,
    VPMX_ERR_ARG = 1,      /* bad argument / parameter outside what the reference accepts   */
    VPMX_ERR_STATE = 2,    /* call order (e.g. mode 2 before any mode 0 geometry)          */
    VPMX_ERR_HIP = 3,      /* HIP runtime error                                              */
    VPMX_ERR_UNSUP = 4,    /* configuration the reference accepts but this build does not    */
    VPMX_ERR_IO = 5,       /* pm.inp missing or malformed                                     */
    VPMX_ERR_NODEV = 6     /* no GPU / library loaded without a device                        */
};

/* pm.inp contents, parsed as vpm.f90:132-152 reads them (one record per line). */
typedef struct {
    double dpm[3];                 /* DXpm DYpm DZpm                  */
    int32_t interf_iproj;          /* projection function 2..6 (4=M4')*/
    int32_t ibctyp;                /* 2 point-source, 3 panel Green sums */
    int32_t idvpm;                 /* 1 constant volume, 0 variable     */
    double epsvol;
    int32_t ncoarse;
    int32_t nbi, nbj, nbk;         /* yaps blocks                       */
    int32_t nremesh, ncell_rem;
    int32_t iyntree, ilevmax;
    int32_t ompthreads, idefine, iynslice, ipmwrite;
    int32_t ipmwstart[10], ipmwsteps[10];
} vpmx_params;

/* Geometry of define_sizes (vpm.f90:477-660), kept by the context between calls. */
typedef struct {
    double xbound[6], dpm[3];      /* Xbound(6), Dpm(3)                         */
    int32_t nn[3], nn_bl[6];       /* NN(3), NN_bl(6) (1-based node indices)    */
    int32_t nblocks, nbi, nbj, nbk;
    double xbound_bl[6 * VPMX_MAXB];
    int32_t nnbl[3 * VPMX_MAXB], nnbl_bl[6 * VPMX_MAXB];
    double xbound_coarse[6], dpm_coarse[3];
    int32_t nn_coarse[3], nn_bl_coarse[6];
    int32_t ilevmax;
    int32_t defined;               /* 1 once the box is fixed (idefine=1)       */
} vpmx_geom;

typedef struct vpmx_ctx vpmx_ctx;

/* ---------------------------------------------------------------- host-only calls */

/* Parse pm.inp (vpm.f90:132-152).  No device needed. */
int32_t vpmx_read_pm_inp(const char *path, vpmx_params *out);

/* define_sizes (vpm.f90:477-660) + definepm (pmlib.f90:216-432) on the host, bit-exact
 * FP64 sequence.  `g` carries the fixed box between calls (idefine=1 semantics:
 * recomputed only when ntime==0, or every call when idefine==0).  No device needed. */
int32_t vpmx_define_sizes(const vpmx_params *p, vpmx_geom *g, const double *xp, int64_t nvr, int32_t ntime);

/* define_sizes with yaps blocks of uneven extents (no counterpart in the reference, whose
 * split is even, NXbl = NXB/NBI, vpm.f90:556-596): along each axis the block extents are in
 * proportion to the weights wx[NBI], wy[NBJ], wz[NBK] (positive), in whole coarse cells
 * (ncoarse fine cells each, at least one per block); equal weights give the reference's even
 * split bit for bit, and null weight arrays mean equal weights.  The box, the coarse grid and
 * every other rule are those of vpmx_define_sizes. */
int32_t vpmx_define_sizes_weighted(const vpmx_params *p, const int32_t *wx, const int32_t *wy, const int32_t *wz,
                                   vpmx_geom *g, const double *xp, int64_t nvr, int32_t ntime);

/* ------------------------------------------------------------------ context */
int32_t vpmx_create(vpmx_ctx **out, int32_t device);
void vpmx_destroy(vpmx_ctx *ctx);
const char *vpmx_last_error(const vpmx_ctx *ctx);
int32_t vpmx_set_params(vpmx_ctx *ctx, const vpmx_params *p);
/* Uneven yaps blocks for the context's later define_sizes (vpm mode 0): weights as in
 * vpmx_define_sizes_weighted, NBI/NBJ/NBK entries of the current parameters; null arrays
 * restore the even split.  No counterpart in the reference (its split is even). */
int32_t vpmx_set_block_weights(vpmx_ctx *ctx, const int32_t *wx, const int32_t *wy, const int32_t *wz);
int32_t vpmx_get_geometry(const vpmx_ctx *ctx, vpmx_geom *out);
/* Arithmetic form (no reference counterpart: the reference is FP64 throughout).
 * 0 = plain FP64 form (default).  1 = mixed form: in FP32 the projection's weights and sums,
 * the sine-transform passes of the fine-block and coarse infinite-domain solves, the forward
 * passes and diagonal of the inner-box solve, and the interpolation's node values and sums; in
 * FP64 the inner-box solve's inverse passes, the Green's-function boundary sums, the normal
 * derivatives, curl and stretching, remesh and diagnostics, and every particle array and grid
 * in HBM.  VPMX_PRECISION=mixed in the environment selects 1 at vpmx_create. */
int32_t vpmx_set_precision(vpmx_ctx *ctx, int32_t mode);
/* Evaluation choices that leave the results unchanged beyond rounding (no counterpart in the
 * reference): "boundsum" -- the Green's-function boundary sums by 2 spectral convolution
 * (default), 0 direct sums over lattice tables, 1 direct sums over LDS source tiles;
 * "dst_reuse" -- 1 (default): the second fine solve of pmesh starts from the first one's
 * forward transforms, 0: both solves in full; "face_derivs" -- the normal derivatives of the
 * zero-BC solve of pmesh (calc_normalderiv_3d): -1 (default) by grid size, 1 from its sine
 * spectrum after the z pass (no inverse y and x passes), 0 from the solved field;
 * "fine_partial" -- the second fine solve of a single yaps block: -1 (default) by grid size, 1 only
 * the nodes that yaps reads (the inner box's faces and the coarse samples, from the spectrum after
 * the z pass; the block's other nodes are left unset), 0 every node; "tuning" -- 1 (default): the autotuner picks
 * kernel variants and persistent grid sizes by timing them once per launch key, 0: fixed
 * defaults (clears the choices made so far); "adjoint_boundsum" -- the adjoint's transposed
 * boundary sums by 2 spectral correlation on the forward's plan (default), 1 direct sums;
 * "adjoint_m2p_scatter" -- 1: the adjoint's transposed interpolation by FP64 atomic scatters,
 * 0 (default): by the sorted gather projection; "tune_variant" -- -1 (default): the tuner (or
 * the defaults) decide, v >= 0: every launch takes its variant min(v, last) untimed (a check
 * that the variants meant to give the same bits do; clears the choices made so far). */
int32_t vpmx_set_option(vpmx_ctx *ctx, const char *name, int32_t value);
/* The autotuner's choices so far, one line per launch key ("key -> choice (time)"), into buf
 * (NUL-terminated, cut to capacity). */
int32_t vpmx_tuning_report(const vpmx_ctx *ctx, char *buf, int64_t capacity);

/* vpm(XP,QP,UP,GP,NVR,neqpm,WhatToDo,...) (vpm.f90:54-475) on host arrays.
 * WhatToDo 0: geometry (define_sizes), velocity grids zeroed; 2: project, yaps solve,
 * velocity + stretching, interpolate (UP, GP overwritten); 1: project + solve + velocity
 * (grids kept); 3: project + solve; 4: project, stretching from the velocity grids of the
 * last 0/1/5 call, interpolate (vpm.f90:316-343); 5: project, diffusion -NI lap(w)
 * (vpm_gcalc.f90:138-210), UP = 0 and GP = interpolated diffusion * QP(4)
 * (vpm_time.f90:109-151), velocity grids zeroed (vpm.f90:345-373). */
int32_t vpmx_vpm(vpmx_ctx *ctx, double *xp, double *qp, double *up, double *gp, int64_t nvr, int32_t neqpm,
                 int32_t whattodo, int32_t ntime, double ni);

/* remesh_particles_3d(iflag=1) (vpm_remesh.f90:5-225), ncell_rem=1, two-phase so that
 * the caller owns the reallocation of XPR/QPR: _count projects and compacts on the
 * device and returns the new count, _fetch copies the particles out (k->j->i node order). */
int32_t vpmx_remesh_count(vpmx_ctx *ctx, const double *xp, const double *qp, int64_t nvr, int32_t neqpm,
                          int32_t mrem, int64_t *nvr_out);
int32_t vpmx_remesh_fetch(vpmx_ctx *ctx, double *xp_out, double *qp_out);

/* Grids of the last mode-1/2/3 call, in the reference layout.
 * which: 0 RHS_pm(neqpm+1,N) after project_vol3d, 1 SOL_pm(neqpm,N) after the yaps solve
 * (stream function), 2 SOL_pm after calc_velocity_serial_3d (stretching),
 * 3 velvrx_pm,velvry_pm,velvrz_pm (3 consecutive N grids). */
int32_t vpmx_get_grid(vpmx_ctx *ctx, int32_t which, double *dst, int64_t capacity);

/* ------------------------------------------------ device-resident particle state */
/* Particles stay in HBM between calls; the step API runs the application's step
 * (test.f90:83-150 shape) without host round trips. */
int32_t vpmx_particles_upload(vpmx_ctx *ctx, const double *xp, const double *qp, int64_t nvr, int32_t neqpm);
int32_t vpmx_particles_count(const vpmx_ctx *ctx, int64_t *nvr);
int32_t vpmx_particles_download(vpmx_ctx *ctx, double *xp, double *qp, double *up, double *gp);
/* vpm(WhatToDo) on the resident particles (modes 0..3 as vpmx_vpm). */
int32_t vpmx_vpm_resident(vpmx_ctx *ctx, int32_t whattodo, int32_t ntime);
/* remesh_particles_3d(1) on the resident particles (replaces them). */
int32_t vpmx_remesh_resident(vpmx_ctx *ctx, int32_t mrem);
/* One RK2 (midpoint) step: 2 x vpm(2) + updates x += (u+uinf) dt, QP(1:3) -= GP dt,
 * then vpm(0) and remesh_particles_3d(1) when do_remesh != 0. */
int32_t vpmx_step_rk2(vpmx_ctx *ctx, double dt, const double uinf[3], int32_t ntime, int32_t mrem,
                      int32_t do_remesh);
/* FP64 diagnostics of the resident particles: circ = -sum QP(1:3),
 * imp = 1/2 sum x x (-QP(1:3)). */
int32_t vpmx_diagnostics(vpmx_ctx *ctx, double circ[3], double imp[3]);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the diagnostics of
 * vpmx_diagnostics on caller-provided DEVICE arrays xp (3 x n), qp (qstride x n, qstride >= 3)
 * into the device array out[6] = (circ[0..2], imp[0..2]), with the partial sums in a fixed
 * device array of the library (48 KiB per GPU, held until the process ends) instead of a
 * buffer it allocates.  n must be 0..2^31-1.  The call allocates nothing and does not wait:
 * its two kernels are queued on the context's stream, and out is ready after the next
 * synchronizing call on the context or a device synchronization.  Same summation order and
 * bits as vpmx_diagnostics.  Every call on one GPU (any context) uses that GPU's one array:
 * calls that overlap on one GPU race. */
int32_t vpmx_diagnostics_noalloc(vpmx_ctx *ctx, const double *xp, const double *qp, int64_t n, int32_t qstride,
                                 double *out);

/* ------------------------------------------------------------- adjoint
 * No counterpart in the reference (it has no derivative of vpm()).  For the particles (xp, qp as
 * vpmx_vpm, host arrays) and nprobe probe points (3 doubles each), evaluates
 *   J = 1/2 sum_k |u(p_k)|^2,  u = the velocity grid of vpm(WhatToDo=2) (vpm.f90:230-313)
 * interpolated at the probes with the M4' weights of the projection, in the plain FP64 form, and
 * the gradient of J with respect to QP(1:3) (dq, nvr x 3) and XP (dx, nvr x 3) by the transposed
 * stages of the solve.  Geometry from an earlier vpm(WhatToDo=0).  One yaps block (NBI = NBJ =
 * NBK = 1), ibctyp 2 without the tree, interf_iproj 4: other set-ups return VPMX_ERR_UNSUP.
 * Probes closer than two cells to the edge of the grid see the truncated stencil. */
int32_t vpmx_adjoint_probe_energy(vpmx_ctx *ctx, const double *xp, const double *qp, int64_t nvr, int32_t neqpm,
                                  const double *probes, int64_t nprobe, double *J, double *dq, double *dx);

/* The same J after nsteps RK2 steps of the particles (the step of vpmx_step_rk2 with mrem = 1:
 * two vpm(2) evaluations, the Euler-form updates of test.f90:104-117 with velocity dt*(UP+uinf)
 * and strength -dt*GP, remesh_particles_3d(1) after each step), in the plain FP64 form, and its
 * gradient with respect to the starting QP(1:3) and XP (dq, dx: nvr x 3).  The remesh's
 * threshold mask is held fixed (it is piecewise constant in the particles), so the remeshed
 * particles' positions (grid nodes) carry no gradient.  The box must stay fixed (idefine = 1,
 * geometry from vpm(WhatToDo=0) at NTIME = 0); steps run with NTIME = 1..nsteps.  The particles
 * are kept every ckpt steps; the backward pass recomputes ckpt steps at a time and keeps their
 * grids (about 6 grids and 7 particle arrays per step).  ckpt = 0 keeps every step's grids from
 * the forward pass instead (no recomputation; HBM for all nsteps records).  The context's resident particles are
 * left in an unspecified state.  Same set-up limits as vpmx_adjoint_probe_energy. */
int32_t vpmx_adjoint_rollout(vpmx_ctx *ctx, const double *xp, const double *qp, int64_t nvr, int32_t neqpm,
                             int32_t nsteps, double dt, const double *uinf, int32_t ckpt, const double *probes,
                             int64_t nprobe, double *J, double *dq, double *dx);

/* ------------------------------------------------------------- batched clouds
 * No counterpart in the reference, which runs independent clouds as one vpm() call after
 * another (vpm.f90:54-475; per cloud the application's RK2 step with remesh_particles_3d,
 * test.f90:83-150).  A batch owns one context per cloud; set each cloud up with the calls above
 * (vpmx_set_params, vpmx_particles_upload, vpmx_vpm_resident(ctx, 0, 0) for its geometry with
 * idefine = 1, vpmx_set_precision), then step all clouds together.  The first step prepares the
 * batch: every cloud's particle arrays get the capacity of its remesh window and the count
 * moves to the device, one step of every cloud runs and is undone (all buffers and plans reach
 * their sizes); after that a cloud's particles change only through the batch (its upload and
 * remesh calls return VPMX_ERR_STATE) and can be read with vpmx_particles_download. */
typedef struct vpmx_batch vpmx_batch;
int32_t vpmx_batch_create(vpmx_batch **out, int32_t device, int32_t nclouds);
void vpmx_batch_destroy(vpmx_batch *b);
const char *vpmx_batch_last_error(const vpmx_batch *b);
int32_t vpmx_batch_cloud(vpmx_batch *b, int32_t i, vpmx_ctx **ctx);
/* One RK2 step of every cloud (vpmx_step_rk2's step, remesh_particles_3d(mrem) when
 * do_remesh).  use_graph = 0: each cloud's launches queued on its own stream, forked from and
 * joined to the batch's stream; use_graph = 1: that whole step captured once as a hipGraph
 * (again when dt, uinf, mrem or do_remesh change) and replayed.  Synchronous at return. */
int32_t vpmx_batch_step_rk2(vpmx_batch *b, double dt, const double uinf[3], int32_t mrem, int32_t do_remesh,
                            int32_t use_graph);
/* Device time of the last batch step (events on the batch's stream around it) and the host
 * time spent queueing it. */
int32_t vpmx_batch_times(const vpmx_batch *b, double *step_ms, double *enqueue_ms);

/* ----------------------------------------------------- stage entry points
 * Single routines of the path on host arrays in the reference layout, for parity
 * testing of each kernel in isolation. */
/* solve_eq_3d / solve_eq_0_3d (pmsolve.f90:284-432) on window win (1-based, NN_bl
 * order) of a grid (nn, xbound, dpm) with nc components: zero_bc=1 writes the zero-BC
 * solution into out, else solves in place in sol with its window faces as BCs. */
int32_t vpmx_stage_solve(vpmx_ctx *ctx, const double xbound[6], const double dpm[3], const int32_t nn[3],
                         int32_t nc, const int32_t win[6], const double *rhs, double *sol, int32_t zero_bc,
                         double *out);
/* pmesh(iynbc=1, ibctyp=2, direct sums) (pmlib.f90:69-192): sol += boundary sums, then
 * the Dirichlet solve over the whole grid. */
int32_t vpmx_stage_pmesh_inf(vpmx_ctx *ctx, const double xbound[6], const double dpm[3], const int32_t nn[3],
                             int32_t nc, const double *rhs, double *sol);
/* The same with pmesh's ibctyp / itree / levmax arguments (pmlib.f90:69): ibctyp 2 point
 * sources (calc_boundinf_3d), 3 constant-source panels (calc_boundinf_3d_s, FSOUR_A4,
 * pmbound.f90:520-648, 1541-1632); itree = 1 takes the sums from the face-cell tree of
 * build_level_nbound_3d + tree_calc_3d / tree_calc_3d_s (pinfdomain.f90:875-1530,
 * pmbound.f90:980-1481) with levmax levels. */
int32_t vpmx_stage_pmesh_inf_lev(vpmx_ctx *ctx, const double xbound[6], const double dpm[3], const int32_t nn[3],
                                 int32_t nc, const double *rhs, double *sol, int32_t ibctyp, int32_t itree,
                                 int32_t levmax);
/* project_particles_3D + project_vol3d (pmproject.f90:43-153) on grid (nn,xbound,dpm). */
int32_t vpmx_stage_project(vpmx_ctx *ctx, const double xbound[6], const double dpm[3], const int32_t nn[3],
                           const int32_t nn_bl[6], const double *xp, const double *qp, int64_t nvr,
                           int32_t ncomp, int32_t itype, int32_t idvpm, double *rhs);

/* Timing of the last vpm/step call per stage (ms of device time between events recorded on
 * the library's stream at the stage's start and end, summed over the stage's runs; read back
 * when the call ends, so timing adds no stream drain), in the order of the reference's timer
 * labels: Proj, Poisson fine, Bounds, pmesh_coarse, pm_bc2, pmesh_final, fd, back, remeshing. */
int32_t vpmx_stage_times(const vpmx_ctx *ctx, double *ms, int32_t n);
/* Accumulated device time (ms, from each dispatch's start/stop events) and launch counts
 * of the timed kernel classes: 0 fine-block boundary sum (Bounds3d, pmbound.f90:80-157),
 * 1 projection (pmproject.f90:43-153), 2 interpolation (vpm_time.f90:30-107), 3 the y passes
 * and 4 the z passes (forward, diagonal, inverse) of the fine block's Poisson solves
 * (pmsolve.f90:284-432).  reset != 0 zeroes them. */
int32_t vpmx_kernel_times(vpmx_ctx *ctx, double *ms_total, int64_t *launches, int32_t n, int32_t reset);
/* Algorithmic work of the last launch of each timed class: FP64 flops for the boundary sum
 * (spectral form: complex products + transforms; lattice form: 6 per pair + 20 per
 * distinct 1/r), bytes read + written for the Poisson passes. */
int32_t vpmx_kernel_work(const vpmx_ctx *ctx, double *work, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
