// Context, pipeline orchestration and the extern "C" entry points of libvpmx.
//
// vpm(WhatToDo=2) (vpm.f90:171-313) on one device, one yaps block:
//   project_particles (P2M into the block RHS; rhsscat_3d's window and zero halo)
//   yaps3d (yaps3d.f90:1-453): fine infinite-domain pmesh on the block, coarse sampling,
//     coarse Laplacian, coarse infinite-domain pmesh, TSC face correction, inner solve
//   solget_3d is a view: the global grid is the inner box of the block arrays
//   calc_velocity_serial_3d(1), back_to_particles_3D(itype=1)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

int fail(Ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

void *ensure(Ctx *c, DBuf &b, size_t bytes, int *st) {
    if (bytes == 0) bytes = 8;
    if (b.bytes >= bytes) return b.p;
    // growing: an eighth more than asked, so slowly growing particle counts do not drain the
    // stream and reallocate on every call
    if (b.p) bytes += bytes / 8;
    if (b.p) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    c->alloc_gen++;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        *st = fail(c, VPMX_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
        b.p = nullptr;
        return nullptr;
    }
    b.bytes = bytes;
    return b.p;
}

static void release(DBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

static GView view(double *p, const int n[3], int ox = 0, int oy = 0, int oz = 0, const int *alloc = nullptr) {
    const int *a = alloc ? alloc : n;
    GView v;
    v.p = p;
    v.py = a[0];
    v.pz = (int64_t)a[0] * a[1];
    v.cs = (int64_t)a[0] * a[1] * a[2];
    v.ox = ox;
    v.oy = oy;
    v.oz = oz;
    return v;
}

static Grid make_grid(const int32_t n[3], const double x0[3], const double h[3]) {
    Grid g;
    for (int d = 0; d < 3; d++) {
        g.n[d] = n[d];
        g.x0[d] = x0[d];
        g.h[d] = h[d];
    }
    return g;
}

// MKL-style spacing of a solve window: h = (xmax_calc - xmin_calc)/(n-1) (pmsolve.f90:297-305).
static void window_h(const double x0[3], const double dpm[3], const int32_t win[6], double h[3]) {
    for (int d = 0; d < 3; d++) {
        double mn = x0[d] + (double)(win[d] - 1) * dpm[d];
        double mx = x0[d] + (double)(win[d + 3] - 1) * dpm[d];
        h[d] = (mx - mn) / (double)(win[d + 3] - win[d]);
    }
}

// ---------------------------------------------------------------- timing
double wall_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}
// (stage times are read back when the outermost call's TimingScope ends)
static void collect_times(Ctx *, const bool *) {}

// ---------------------------------------------------------------- pmesh (iynbc=1)
// pmesh(iynbc=1) on grid g: SOL_0 = zero-BC solve, sources from its normal derivatives,
// SOL(faces) += boundary sums (ibctyp 2 point sources, 3 panels; tree form with levmax
// levels when levmax >= 0, pmlib.f90:69 ibctyp/itree/levmax), then the Dirichlet solve of
// the whole grid.
constexpr int64_t FACE_SPECTRAL_MIN = 96 * 96 * 96;  // grid nodes from which the spectral derivatives pay
static int pmesh_inf(Ctx *c, const Grid &g, GView rhs, GView sol, DBuf &s0buf, bool fine, int levmax, int ibctyp,
                     const SolvePartial *part = nullptr) {
    int st = VPMX_OK;
    const int n[3] = {g.n[0], g.n[1], g.n[2]};
    int64_t G = (int64_t)n[0] * n[1] * n[2];
    double *s0 = (double *)ensure(c, s0buf, sizeof(double) * 3 * (size_t)G, &st);
    if (st) return st;
    GView s0v = view(s0, n);
    int32_t win[6] = {1, 1, 1, n[0], n[1], n[2]};
    double h[3];
    window_h(g.x0, g.h, win, h);
    const int ru = c->dst_reuse ? 1 : 0;
    int64_t nbmax = 2 * ((int64_t)n[0] * n[1] + (int64_t)n[0] * n[2] + (int64_t)n[1] * n[2]);
    double *sig = (double *)ensure(c, c->src_sig, sizeof(double) * 3 * (size_t)nbmax, &st); if (st) return st;
    double *cxyz = (double *)ensure(c, c->src_cxyz, sizeof(double) * 3 * (size_t)nbmax, &st); if (st) return st;
    double *wgt = (double *)ensure(c, c->src_w, sizeof(double) * 4 * (size_t)nbmax, &st); if (st) return st;
    int64_t nb = 0;
    // SOL_0 is read only through its normal derivatives at the faces: on large grids they come
    // straight from its spectrum (no inverse y and x passes), else from the solved field
    const bool spectral = c->face_derivs == 1 || (c->face_derivs < 0 && G >= FACE_SPECTRAL_MIN && n[0] <= 322);
    if (spectral) {
        double *fd = normal_derivs_table(c, 3, n, &st);
        if (st) return st;
        VPMX_TRY(poisson_dirichlet(c, 3, n, h, rhs, nullptr, s0v, ru, SOLVE_CTX, fd));  // keeps its forward transforms
        VPMX_TRY(normal_derivs_cells(c, 3, g, fd, sig, cxyz, wgt, &nb));
    } else {
        VPMX_TRY(poisson_dirichlet(c, 3, n, h, rhs, nullptr, s0v, ru));  // keeps its forward transforms
        VPMX_TRY(normal_derivs(c, 3, g, s0v, sig, cxyz, wgt, &nb));
    }
    // SOL_0 is dead once the sources exist: scratch of the panel sums
    VPMX_TRY(boundary_sums(c, 3, g, nb, sig, cxyz, wgt, sol, fine, levmax, ibctyp, &s0v));
    // same RHS: reuses them (part: only the nodes the caller reads)
    VPMX_TRY(poisson_dirichlet(c, 3, n, h, rhs, &sol, sol, 2 * ru, SOLVE_CTX, nullptr, part));
    return VPMX_OK;
}

// Coarse-map table on the device: uploaded only when its contents or buffer change (the
// geometry is usually the same from call to call), from a copy the context keeps, so no
// call has to wait for the upload before its host vector goes away.
static int upload_cmap(Ctx *c, const std::vector<int> &cmap, int **out) {
    int st = VPMX_OK;
    int *d = (int *)ensure(c, c->cmap, sizeof(int) * cmap.size(), &st);
    if (st) return st;
    if (d != c->cmap_dev || cmap != c->cmap_host) {
        VPMX_HIP(c, hipStreamSynchronize(c->stream));  // an earlier upload may still read cmap_host
        c->cmap_host = cmap;
        c->cmap_dev = d;
        VPMX_HIP(c, hipMemcpyAsync(d, c->cmap_host.data(), sizeof(int) * cmap.size(), hipMemcpyHostToDevice,
                                   c->stream));
    }
    *out = d;
    return VPMX_OK;
}

// The inner-box solve gives the stream function whose second differences are the stretching
// term: in the mixed form its passes after the diagonal run in FP64 (an FP32 inverse transform
// leaves white noise that the stencils amplify by 1/h^2); its forward passes, and the fine and
// coarse infinite-domain solves, run in FP32.
static int final_form(const Ctx *c) { return c->precision == 1 ? SOLVE_INV64 : SOLVE_FP64; }

// ---------------------------------------------------------------- yaps (1 block)
static int yaps_single(Ctx *c, bool *used) {
    int st = VPMX_OK;
    const vpmx_geom &G = c->geom;
    const vpmx_params &P = c->prm;
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    int64_t B = (int64_t)nb[0] * nb[1] * nb[2];
    double *rhs_bl = (double *)c->rhs_bl.p;
    double *sol_bl = (double *)ensure(c, c->sol_bl, sizeof(double) * 3 * (size_t)B, &st); if (st) return st;
    GView rv = view(rhs_bl, nb), sv = view(sol_bl, nb);
    Grid gb = make_grid(G.nnbl, &G.xbound_bl[0], G.dpm);
    // 1. fine infinite-domain solve over the whole block
    mark(c, ST_FINE, 0);
    // the boundary sums accumulate on the faces; the second solve writes every interior node
    VPMX_TRY(zero_faces(c, nb, sv));
    c->dst_timed_m = gb.n[1] - 2;  // the block's y lines are the timed sine-transform passes
    // tree levels: ilevmax on the block, ilevmax-1 on the coarse grid (yaps3d.f90:46-48, 286-292)
    const bool tree = P.iyntree == 1;
    if (tree && G.ilevmax < 1)
        return fail(c, VPMX_ERR_UNSUP, "tree = 1 needs at least one level on the coarse grid (ilevmax >= 1)");
    // the block solution is read only on the inner box's faces and at the coarse samples (the
    // final solve overwrites the rest of the inner box, and the margin is never read): on large
    // blocks with equal x and y margins only those nodes are evaluated
    const int32_t *wi = G.nnbl_bl;
    SolvePartial part;
    for (int q = 0; q < 2; q++) {
        part.xs[q] = wi[3 * q];
        part.ys[q] = wi[3 * q + 1];
        part.zs[q] = wi[3 * q + 2];
    }
    part.step = P.ncoarse;
    const bool part_ok = part.xs[0] >= 2 && part.ys[0] >= 2 && part.zs[0] >= 2 && part.xs[1] <= nb[0] - 1 &&
                         part.ys[1] <= nb[1] - 1 && part.zs[1] <= nb[2] - 1 && part.xs[0] < part.xs[1] &&
                         part.ys[0] < part.ys[1] && part.zs[0] < part.zs[1] && part.xs[1] == nb[0] + 1 - part.xs[0] &&
                         part.ys[1] == nb[1] + 1 - part.ys[0] && nb[0] <= 322 && P.ncoarse >= 1;
    const bool use_part = part_ok && (c->fine_partial == 1 || (c->fine_partial < 0 && B >= FACE_SPECTRAL_MIN));
    VPMX_TRY(pmesh_inf(c, gb, rv, sv, c->sol0, true, tree ? G.ilevmax : -1, P.ibctyp, use_part ? &part : nullptr));
    mark(c, ST_FINE, 1);
    used[ST_FINE] = used[ST_BOUNDS] = true;
    // 2. coarse correction
    mark(c, ST_COARSE, 0);
    const int cn[3] = {G.nn_coarse[0], G.nn_coarse[1], G.nn_coarse[2]};
    int64_t GC = (int64_t)cn[0] * cn[1] * cn[2];
    double *samp = (double *)ensure(c, c->samp, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    double *rhsc = (double *)ensure(c, c->rhs_c, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    double *solc = (double *)ensure(c, c->sol_c, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    const int ncs = P.ncoarse;
    int ns[3];
    std::vector<int> cmap;
    int fp_lo[3], fp_hi[3];
    for (int d = 0; d < 3; d++) {
        ns[d] = (nb[d] - 1) / ncs + 1;
        for (int a = 0; a < ns[d]; a++) {
            int i = 1 + a * ncs;
            double X = G.xbound_bl[d] + (double)(i - 1) * G.dpm[d];
            int node = (int)std::lround((X - G.xbound_coarse[d]) / G.dpm_coarse[d]) + 1;
            if (node < 1 || node > cn[d]) return fail(c, VPMX_ERR_STATE, "coarse sample outside the coarse grid");
            cmap.push_back(node);
        }
        fp_lo[d] = cmap[cmap.size() - ns[d]];
        fp_hi[d] = cmap.back();
        // the reference records NN_coarse_map at i = 1 and i = NNbl; with (NNbl-1) a multiple
        // of ncoarse the last sample is i = NNbl
    }
    int *cmap_d = nullptr;
    VPMX_TRY(upload_cmap(c, cmap, &cmap_d));
    GView smv = view(samp, cn), rcv = view(rhsc, cn), scv = view(solc, cn);
    VPMX_TRY(fill(c, samp, 3 * GC, 0.0));
    VPMX_TRY(coarse_sample(c, 3, ns, ncs, cmap_d, sv, smv, cn));
    VPMX_TRY(coarse_laplacian(c, 3, cn, fp_lo, fp_hi, G.dpm_coarse, smv, rcv));
    VPMX_TRY(fill(c, solc, 3 * GC, 0.0));
    Grid gc = make_grid(G.nn_coarse, G.xbound_coarse, G.dpm_coarse);
    VPMX_TRY(pmesh_inf(c, gc, rcv, scv, c->sol0_c, false, tree ? G.ilevmax - 1 : -1, P.ibctyp));
    mark(c, ST_COARSE, 1);
    used[ST_COARSE] = true;
    // 3. TSC correction of the inner-box faces
    mark(c, ST_BC, 0);
    InterpArgs ia;
    for (int d = 0; d < 6; d++) ia.nbl[d] = G.nnbl_bl[d];
    for (int d = 0; d < 3; d++) {
        ia.xbl[d] = G.xbound_bl[d];
        ia.h[d] = G.dpm[d];
        ia.xc[d] = G.xbound_coarse[d];
        ia.hc[d] = G.dpm_coarse[d];
        ia.cn[d] = cn[d];
        ia.fp_lo[d] = fp_lo[d];
        ia.fp_hi[d] = fp_hi[d];
    }
    VPMX_TRY(face_interp(c, 3, ia, scv, smv, sv));
    mark(c, ST_BC, 1);
    used[ST_BC] = true;
    // 4. final solve on the inner box
    mark(c, ST_FINAL, 0);
    const int32_t *w = G.nnbl_bl;
    int wn[3] = {w[3] - w[0] + 1, w[4] - w[1] + 1, w[5] - w[2] + 1};
    double h[3];
    window_h(&G.xbound_bl[0], G.dpm, w, h);
    GView rin = view(rhs_bl, wn, w[0] - 1, w[1] - 1, w[2] - 1, nb);
    GView sin = view(sol_bl, wn, w[0] - 1, w[1] - 1, w[2] - 1, nb);
    VPMX_TRY(poisson_dirichlet(c, 3, wn, h, rin, &sin, sin, 0, final_form(c)));
    mark(c, ST_FINAL, 1);
    used[ST_FINAL] = true;
    return VPMX_OK;
}

// ---------------------------------------------------------------- yaps (several blocks)
// Offsets of the blocks' arrays (3 components of each block's NNbl nodes, back to back) in
// the blocks' RHS / SOL buffers; returns the total.  Blocks may differ in size (uneven
// splits); the reference's even blocks are all the first one's size.
static int64_t block_offsets(const vpmx_geom &G, std::vector<int64_t> &off) {
    off.resize((size_t)G.nblocks + 1);
    off[0] = 0;
    for (int b = 0; b < G.nblocks; b++)
        off[b + 1] = off[b] + 3 * (int64_t)G.nnbl[3 * b] * G.nnbl[3 * b + 1] * G.nnbl[3 * b + 2];
    return off[G.nblocks];
}
static const int32_t *block_n(const vpmx_geom &G, int b) { return &G.nnbl[3 * b]; }

// yaps3d.f90:1-453 with BLOCKS = NBI*NBJ*NBK blocks on this device, the work of every rank
// in turn: all fine infinite-domain solves; the coarse problem from every block's samples
// (Laplacians summed in block order); the coarse solve; interp_stencil on every block's
// inner-box faces in one launch, reading every block's fine solution before any face is
// written (a block's faces are its neighbours' BBound nodes); the inner-box solves.
// c->rhs_bl holds the blocks' RHS back to back (rhsscat_3d done by the caller).
static int yaps_multi(Ctx *c, bool *used) {
    int st = VPMX_OK;
    const vpmx_geom &G = c->geom;
    const vpmx_params &P = c->prm;
    const int NB = G.nblocks;
    std::vector<int64_t> boff;
    const int64_t tot = block_offsets(G, boff);
    double *rhs_bl = (double *)c->rhs_bl.p;
    double *sol_bl = (double *)ensure(c, c->sol_bl, sizeof(double) * (size_t)tot, &st); if (st) return st;
    auto bview = [&](double *base, int b) { return view(base + boff[b], block_n(G, b)); };
    const bool tree = P.iyntree == 1;
    if (tree && G.ilevmax < 1)
        return fail(c, VPMX_ERR_UNSUP, "tree = 1 needs at least one level on the coarse grid (ilevmax >= 1)");
    // 1. fine infinite-domain solves
    mark(c, ST_FINE, 0);
    VPMX_TRY(fill(c, sol_bl, tot, 0.0));
    c->dst_timed_m = G.nnbl[1] - 2;
    for (int b = 0; b < NB; b++) {
        Grid gb = make_grid(block_n(G, b), &G.xbound_bl[6 * b], G.dpm);
        VPMX_TRY(pmesh_inf(c, gb, bview(rhs_bl, b), bview(sol_bl, b), c->sol0, true, tree ? G.ilevmax : -1, P.ibctyp));
    }
    mark(c, ST_FINE, 1);
    used[ST_FINE] = used[ST_BOUNDS] = true;
    // 2. samples of every block, coarse RHS, coarse solve
    mark(c, ST_COARSE, 0);
    const int cn[3] = {G.nn_coarse[0], G.nn_coarse[1], G.nn_coarse[2]};
    const int64_t GC = (int64_t)cn[0] * cn[1] * cn[2];
    double *samp = (double *)ensure(c, c->samp, sizeof(double) * 3 * (size_t)GC * NB, &st); if (st) return st;
    double *rhsc = (double *)ensure(c, c->rhs_c, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    double *solc = (double *)ensure(c, c->sol_c, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    const int ncs = P.ncoarse;
    // per block: samples every ncoarse nodes along each axis (their coarse nodes in cmap)
    std::vector<int> ns(3 * (size_t)NB), cmoff((size_t)NB + 1, 0);
    for (int b = 0; b < NB; b++) {
        for (int d = 0; d < 3; d++) ns[3 * b + d] = (block_n(G, b)[d] - 1) / ncs + 1;
        cmoff[b + 1] = cmoff[b] + ns[3 * b] + ns[3 * b + 1] + ns[3 * b + 2];
    }
    std::vector<int> cmap((size_t)cmoff[NB]);
    std::vector<MbBlock> blk(NB);
    int64_t nface_tot = 0;
    for (int b = 0; b < NB; b++) {
        MbBlock &M = blk[b];
        M.off = boff[b];
        for (int d = 0; d < 6; d++) M.nbl[d] = G.nnbl_bl[6 * b + d];
        int off = cmoff[b];
        for (int d = 0; d < 3; d++) {
            M.n[d] = block_n(G, b)[d];
            M.xbl[d] = G.xbound_bl[6 * b + d];
            for (int a = 0; a < ns[3 * b + d]; a++) {
                const int i = 1 + a * ncs;
                const double X = G.xbound_bl[6 * b + d] + (double)(i - 1) * G.dpm[d];
                const int node = (int)std::lround((X - G.xbound_coarse[d]) / G.dpm_coarse[d]) + 1;
                if (node < 1 || node > cn[d]) return fail(c, VPMX_ERR_STATE, "coarse sample outside the coarse grid");
                if (a > 0 && node != cmap[(size_t)off + a - 1] + 1)
                    return fail(c, VPMX_ERR_STATE, "block samples are not consecutive coarse nodes");
                cmap[(size_t)off + a] = node;
            }
            M.fp_lo[d] = cmap[(size_t)off];
            M.fp_hi[d] = cmap[(size_t)off + ns[3 * b + d] - 1];
            off += ns[3 * b + d];
        }
        M.bi = b % G.nbi + 1;
        M.bj = (b / G.nbi) % G.nbj + 1;
        M.bk = b / (G.nbi * G.nbj) + 1;
        const int64_t fx = M.nbl[3] - M.nbl[0] + 1, fy = M.nbl[4] - M.nbl[1] + 1, fz = M.nbl[5] - M.nbl[2] + 1;
        M.f0 = nface_tot;
        nface_tot += 2 * (fy * fz + (fx - 2) * fz + (fx - 2) * (fy - 2));
    }
    int *cmap_d = nullptr;
    VPMX_TRY(upload_cmap(c, cmap, &cmap_d));
    MbBlock *blk_d = (MbBlock *)ensure(c, c->mb_blocks, sizeof(MbBlock) * (size_t)NB, &st); if (st) return st;
    // the block table, uploaded when it or its buffer changes (from a copy the context keeps,
    // so no call waits for the upload)
    const size_t tb = sizeof(MbBlock) * (size_t)NB;
    if (blk_d != c->mb_dev || c->mb_host.size() != tb || std::memcmp(c->mb_host.data(), blk.data(), tb) != 0) {
        VPMX_HIP(c, hipStreamSynchronize(c->stream));  // an earlier upload may still read mb_host
        c->mb_host.assign((const char *)blk.data(), (const char *)blk.data() + tb);
        c->mb_dev = blk_d;
        VPMX_HIP(c, hipMemcpyAsync(blk_d, c->mb_host.data(), tb, hipMemcpyHostToDevice, c->stream));
    }
    VPMX_TRY(fill(c, samp, 3 * GC * NB, 0.0));
    VPMX_TRY(fill(c, rhsc, 3 * GC, 0.0));
    auto cview = [&](double *base, int b) { return view(base + (int64_t)b * 3 * GC, cn); };
    for (int b = 0; b < NB; b++) {
        const int *cm = cmap_d + cmoff[b];
        VPMX_TRY(coarse_sample(c, 3, &ns[3 * b], ncs, cm, bview(sol_bl, b), cview(samp, b), cn));
        VPMX_TRY(coarse_laplacian(c, 3, cn, blk[b].fp_lo, blk[b].fp_hi, G.dpm_coarse, cview(samp, b), view(rhsc, cn),
                                  true));
    }
    VPMX_TRY(fill(c, solc, 3 * GC, 0.0));
    Grid gc = make_grid(G.nn_coarse, G.xbound_coarse, G.dpm_coarse);
    VPMX_TRY(pmesh_inf(c, gc, view(rhsc, cn), view(solc, cn), c->sol0_c, false, tree ? G.ilevmax - 1 : -1, P.ibctyp));
    mark(c, ST_COARSE, 1);
    used[ST_COARSE] = true;
    // 3. interp_stencil on every block's inner-box faces
    mark(c, ST_BC, 0);
    MbArgs a;
    for (int d = 0; d < 3; d++) {
        a.h[d] = G.dpm[d];
        a.xc[d] = G.xbound_coarse[d];
        a.hc[d] = G.dpm_coarse[d];
        a.cn[d] = cn[d];
    }
    a.nbi = G.nbi;
    a.nbj = G.nbj;
    a.nbk = G.nbk;
    a.nblocks = NB;
    a.sstride = 3 * GC;
    a.ntot = nface_tot;
    double *tmp = (double *)ensure(c, c->mb_tmp, sizeof(double) * 3 * (size_t)nface_tot, &st); if (st) return st;
    VPMX_TRY(face_interp_blocks(c, a, blk_d, view(solc, cn), samp, sol_bl, tmp));
    mark(c, ST_BC, 1);
    used[ST_BC] = true;
    // 4. inner-box solves
    mark(c, ST_FINAL, 0);
    for (int b = 0; b < NB; b++) {
        const int32_t *w = &G.nnbl_bl[6 * b];
        int wn[3] = {w[3] - w[0] + 1, w[4] - w[1] + 1, w[5] - w[2] + 1};
        double h[3];
        window_h(&G.xbound_bl[6 * b], G.dpm, w, h);
        GView rin = view(rhs_bl + boff[b], wn, w[0] - 1, w[1] - 1, w[2] - 1, block_n(G, b));
        GView sin = view(sol_bl + boff[b], wn, w[0] - 1, w[1] - 1, w[2] - 1, block_n(G, b));
        VPMX_TRY(poisson_dirichlet(c, 3, wn, h, rin, &sin, sin, 0, final_form(c)));
    }
    mark(c, ST_FINAL, 1);
    used[ST_FINAL] = true;
    return VPMX_OK;
}

// Global origin (1-based global node) of block b's inner box (rhsscat_3d / solget_3d): its
// first inner node's position on the global lattice (the reference's
// (nb_i - 1)(NXf - NXs) + NN_bl(1) for its even blocks, vpm_mpi.f90:199-238).
static void block_origin(const vpmx_geom &G, int b, int g0[3]) {
    const int32_t *w = &G.nnbl_bl[6 * b];
    for (int d = 0; d < 3; d++) {
        const double x = G.xbound_bl[6 * b + d] + (double)(w[d] - 1) * G.dpm[d];
        g0[d] = (int)std::lround((x - G.xbound[d]) / G.dpm[d]) + 1;
    }
}

// The global RHS / SOL grids: the block arrays' inner box for one block, own grids otherwise.
static GView global_rhs(Ctx *c) {
    const vpmx_geom &G = c->geom;
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    if (G.nblocks > 1) return view((double *)c->rhs_g.p, gn);
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    const int32_t *w = G.nnbl_bl;
    return view((double *)c->rhs_bl.p, gn, w[0] - 1, w[1] - 1, w[2] - 1, nb);
}
static GView global_psi(Ctx *c) {
    const vpmx_geom &G = c->geom;
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    if (G.nblocks > 1) return view((double *)c->sol_g.p, gn);
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    const int32_t *w = G.nnbl_bl;
    return view((double *)c->sol_bl.p, gn, w[0] - 1, w[1] - 1, w[2] - 1, nb);
}

static int check_ready(Ctx *c, int neqpm) {
    if (!c->have_prm) return fail(c, VPMX_ERR_STATE, "no parameters: call vpmx_set_params or vpm mode 0 first");
    if (!c->geom.defined) return fail(c, VPMX_ERR_STATE, "no geometry: vpm(WhatToDo=0) must run first");
    if (c->geom.nblocks < 1 || c->geom.nblocks > VPMX_MAXB ||
        c->geom.nbi * c->geom.nbj * c->geom.nbk != c->geom.nblocks)
        return fail(c, VPMX_ERR_ARG, "NBI*NBJ*NBK must be 1..64 yaps blocks");
    if (neqpm != 3) return fail(c, VPMX_ERR_UNSUP, "neqpm must be 3");
    const vpmx_params &P = c->prm;
    if (P.interf_iproj < 2 || P.interf_iproj > 4)
        return fail(c, VPMX_ERR_ARG, "No such projection function (interf_iproj must be 2, 3 or 4)");
    if (P.idvpm != 1) {
        if (P.idvpm != 0) return fail(c, VPMX_ERR_ARG, "WRONG IDVPM");
        return fail(c, VPMX_ERR_UNSUP, "IDVPM=0 (variable volume) is not supported by this build");
    }
    if (P.ibctyp != 2 && P.ibctyp != 3)
        return fail(c, VPMX_ERR_UNSUP, "ibctyp must be 2 (point sources) or 3 (constant-source panels)");
    return VPMX_OK;
}

// Projection + yaps solve (+ velocity/stretching + interpolation for mode 2) on the
// particles (xp, qp) already on the device.
static int run_vpm(Ctx *c, int whattodo, const double *xp, const double *qp, double *up, double *gp, int64_t n,
                   double ni = 0.0) {
    int st = VPMX_OK;
    VPMX_TRY(check_ready(c, c->neqpm));
    if (n == 0) return VPMX_OK;  // vpm.f90:172
    const vpmx_geom &G = c->geom;
    const vpmx_params &P = c->prm;
    TimingScope timing(c);
    bool used[ST_COUNT] = {};
    c->last_pairs = 0.0;
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    const bool multi = G.nblocks > 1;
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    const int64_t GG = (int64_t)gn[0] * gn[1] * gn[2];
    std::vector<int64_t> boff;
    const int64_t btot = block_offsets(G, boff);
    double *rhs_bl = (double *)ensure(c, c->rhs_bl, sizeof(double) * (size_t)btot, &st); if (st) return st;
    mark(c, ST_PROJ, 0);
    Grid gg = make_grid(G.nn, G.xbound, G.dpm);
    VPMX_TRY(sort_particles(c, gg, xp, n, c->sort));
    const double DV = G.dpm[0] * G.dpm[1] * G.dpm[2];
    if (!multi) {
        // project_particles (vpm.f90:408-444) straight into the block layout of rhsscat_3d
        int wlo[3] = {G.nnbl_bl[0] - G.nn_bl[0], G.nnbl_bl[1] - G.nn_bl[1], G.nnbl_bl[2] - G.nn_bl[2]};
        VPMX_TRY(p2m_gather(c, gg, c->sort, xp, qp, c->neqpm + 1, 3, P.interf_iproj, DV, view(rhs_bl, nb), nb, wlo,
                            K_P2M, c->precision == 1, true, true));
    } else {
        // the global grid, then rhsscat_3d into every block (vpm_mpi.f90:199-238)
        double *rhs_g = (double *)ensure(c, c->rhs_g, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
        const int wlo[3] = {0, 0, 0};
        VPMX_TRY(p2m_gather(c, gg, c->sort, xp, qp, c->neqpm + 1, 3, P.interf_iproj, DV, view(rhs_g, gn), gn, wlo,
                            K_P2M, c->precision == 1));
        c->zt_valid = false;  // the blocks' RHS come from the global grid
        for (int b = 0; b < G.nblocks; b++) {
            int g0[3];
            block_origin(G, b, g0);
            const int zmin[3] = {b % G.nbi > 0, (b / G.nbi) % G.nbj > 0, b / (G.nbi * G.nbj) > 0};
            VPMX_TRY(rhsscat_block(c, block_n(G, b), &G.nnbl_bl[6 * b], g0, zmin, view(rhs_g, gn),
                                   view(rhs_bl + boff[b], block_n(G, b))));
        }
    }
    mark(c, ST_PROJ, 1);
    used[ST_PROJ] = true;
    if (whattodo == 4 || whattodo == 5) {
        // no Poisson solve (vpm.f90:316-373): mode 4 stretching from the velocity grids of the
        // last mode 0/1/5 call, mode 5 diffusion of the projected vorticity
        const bool fresh = c->vel.bytes < sizeof(double) * 3 * (size_t)GG;
        double *vel = (double *)ensure(c, c->vel, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
        double *str = (double *)ensure(c, c->str, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
        if (fresh || whattodo == 5) VPMX_TRY(fill(c, vel, 3 * GG, 0.0));  // velvr*_pm = 0
        GView rg = global_rhs(c);
        mark(c, ST_FD, 0);
        if (whattodo == 4) VPMX_TRY(stretch_kernel(c, gn, G.dpm, rg, view(vel, gn), view(str, gn)));
        else VPMX_TRY(diffuse_kernel(c, gn, G.dpm, ni, rg, view(str, gn)));
        mark(c, ST_FD, 1);
        used[ST_FD] = true;
        mark(c, ST_BACK, 0);
        VPMX_TRY(m2p_gather(c, gg, c->sort, xp, qp, c->neqpm + 1, P.interf_iproj, view(vel, gn), view(str, gn), up, gp,
                            n, whattodo == 5 ? 2 : 1));
        mark(c, ST_BACK, 1);
        used[ST_BACK] = true;
        c->have_grids = true;
        collect_times(c, used);
        return VPMX_OK;
    }
    if (!multi) {
        VPMX_TRY(yaps_single(c, used));
    } else {
        VPMX_TRY(yaps_multi(c, used));
        // solget_3d: inner boxes into the global grid in block order
        double *sol_g = (double *)ensure(c, c->sol_g, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
        for (int b = 0; b < G.nblocks; b++) {
            int g0[3];
            block_origin(G, b, g0);
            VPMX_TRY(solget_block(c, &G.nnbl_bl[6 * b], g0, view((double *)c->sol_bl.p + boff[b], block_n(G, b)),
                                  view(sol_g, gn)));
        }
    }
    c->have_grids = true;
    if (whattodo == 3) { collect_times(c, used); return VPMX_OK; }
    // calc_velocity_serial_3d on the global grid
    double *vel = (double *)ensure(c, c->vel, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    double *str = (double *)ensure(c, c->str, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    GView psi = global_psi(c);
    GView rg = global_rhs(c);
    mark(c, ST_FD, 0);
    if (whattodo == 2) VPMX_TRY(curl_stretch_kernel(c, gn, G.dpm, psi, rg, view(vel, gn), view(str, gn)));
    else VPMX_TRY(curl_kernel(c, gn, G.dpm, psi, view(vel, gn)));
    mark(c, ST_FD, 1);
    used[ST_FD] = true;
    if (whattodo == 2) {
        mark(c, ST_BACK, 0);
        VPMX_TRY(m2p_gather(c, gg, c->sort, xp, qp, c->neqpm + 1, P.interf_iproj, view(vel, gn), view(str, gn), up, gp,
                            n));
        mark(c, ST_BACK, 1);
        used[ST_BACK] = true;
    }
    collect_times(c, used);
    return VPMX_OK;
}

// remesh_particles_3d(1) on device particles; result in c->rm_xp/rm_qp/rm_count.
static int run_remesh(Ctx *c, const double *xp, const double *qp, int64_t n, int mrem) {
    int st = VPMX_OK;
    VPMX_TRY(check_ready(c, c->neqpm));
    const vpmx_geom &G = c->geom;
    const vpmx_params &P = c->prm;
    if (P.ncell_rem != 1) return fail(c, VPMX_ERR_UNSUP, "ncell_rem must be 1");
    if (mrem < 1) return fail(c, VPMX_ERR_ARG, "mrem must be >= 1");
    TimingScope timing(c);
    bool used[ST_COUNT] = {};
    mark(c, ST_REMESH, 0);
    int32_t NN[3], NNb[6];
    double D[3];
    for (int d = 0; d < 3; d++) NN[d] = G.nn[d] * mrem;
    for (int d = 0; d < 6; d++) NNb[d] = G.nn_bl[d] * mrem;
    for (int d = 0; d < 3; d++) D[d] = (G.xbound[d + 3] - G.xbound[d]) / (double)(NN[d] - 1);
    double DV = D[0] * D[1] * D[2];
    int64_t GG = (int64_t)NN[0] * NN[1] * NN[2];
    double *grid = (double *)ensure(c, c->rm_grid, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    Grid gr = make_grid(NN, G.xbound, D);
    VPMX_TRY(sort_particles(c, gr, xp, n, c->sort));
    int wlo[3] = {0, 0, 0};
    // FP64 in both forms (the flags and positions are exact); nodes of tiles no particle
    // reaches stay unwritten, and the flag pass reads the tile map instead
    VPMX_TRY(p2m_gather(c, gr, c->sort, xp, qp, c->neqpm + 1, 3, P.interf_iproj, DV, view(grid, NN), NN, wlo, -1,
                        false, false));
    int half = P.interf_iproj / 2;
    int lo[3] = {NNb[0] + half, NNb[1] + half, NNb[2] + half}, hi[3] = {NNb[3] - half, NNb[4] - half, NNb[5] - half};
    const double thr = (double)1e-09f;  // default-REAL literal (vpm_remesh.f90:184)
    int64_t cnt = 0;
    double *xo = nullptr, *qo = nullptr;
    VPMX_TRY(remesh_compact(c, view(grid, NN), NN, lo, hi, thr, G.xbound, D, DV, c->neqpm, &xo, &qo, c->rm_x, c->rm_q,
                            &cnt));
    mark(c, ST_REMESH, 1);
    used[ST_REMESH] = true;
    collect_times(c, used);
    c->rm_xp = xo;
    c->rm_qp = qo;
    c->rm_count = cnt;
    return VPMX_OK;
}

static int upload_particles(Ctx *c, const double *xp, const double *qp, int64_t n, int neqpm) {
    int st = VPMX_OK;
    if (c->dn) return fail(c, VPMX_ERR_STATE, "a prepared batch's cloud: its particles change only through the batch");
    if (n < 0 || n > 0x7fffffffLL) return fail(c, VPMX_ERR_ARG, "particle count out of range");
    int nr = neqpm + 1;
    double *dx = (double *)ensure(c, c->xp, sizeof(double) * 3 * (size_t)(n ? n : 1), &st); if (st) return st;
    double *dq = (double *)ensure(c, c->qp, sizeof(double) * nr * (size_t)(n ? n : 1), &st); if (st) return st;
    ensure(c, c->up, sizeof(double) * 3 * (size_t)(n ? n : 1), &st); if (st) return st;
    ensure(c, c->gp, sizeof(double) * 3 * (size_t)(n ? n : 1), &st); if (st) return st;
    if (n) {
        VPMX_HIP(c, hipMemcpyAsync(dx, xp, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
        VPMX_HIP(c, hipMemcpyAsync(dq, qp, sizeof(double) * nr * n, hipMemcpyHostToDevice, c->stream));
    }
    c->np = n;
    c->neqpm = neqpm;
    return VPMX_OK;
}

// define_sizes with the context's block weights (even split when none are set).
static int define_sizes_ctx(Ctx *c, const double *xp, int64_t n, int ntime) {
    const int32_t *w[3];
    for (int d = 0; d < 3; d++) w[d] = c->block_w[d].empty() ? nullptr : c->block_w[d].data();
    const int r = vpmx_define_sizes_weighted(&c->prm, w[0], w[1], w[2], &c->geom, xp, n, ntime);
    return r == VPMX_OK ? r : fail(c, r, "define_sizes failed (box, blocks or block weights)");
}

static int define_from_device(Ctx *c, int ntime) {
    // define_sizes needs the particle bounding box only when it redefines the box
    const vpmx_params &P = c->prm;
    bool redefine = (ntime == 0 && P.idefine == 1) || P.idefine == 0 || !c->geom.defined;
    if (!redefine) return define_sizes_ctx(c, nullptr, 0, ntime);
    if (c->np <= 0) return fail(c, VPMX_ERR_ARG, "define_sizes: no particles");
    double mn[3], mx[3];
    VPMX_TRY(particle_bbox(c, (const double *)c->xp.p, c->np, mn, mx));
    // the bbox is all define_sizes reads from the particles: two corner points suffice
    double pts[6] = {mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]};
    return define_sizes_ctx(c, pts, 2, ntime);
}

static int replace_with_remesh(Ctx *c) {
    // swap the compacted particles in as the resident set
    std::swap(c->xp, c->rm_x);
    std::swap(c->qp, c->rm_q);
    c->np = c->rm_count;
    int st = VPMX_OK;
    ensure(c, c->up, sizeof(double) * 3 * (size_t)(c->np ? c->np : 1), &st); if (st) return st;
    ensure(c, c->gp, sizeof(double) * 3 * (size_t)(c->np ? c->np : 1), &st); if (st) return st;
    c->rm_count = -1;
    c->rm_xp = c->rm_qp = nullptr;
    return VPMX_OK;
}

// ---------------------------------------------------------------- adjoint
// pmesh(iynbc=1)^T on grid g: the adjoint pb of the infinite-domain solution (on every node)
// to the adjoint rb of its right-hand side (interior nodes; faces 0):
//   rb = D0(pb_in) + D0(A^T BS^T (pb_faces - F^T D0(pb_in)))
// (D0: zero-BC solve, symmetric; F: fold of face values; BS: boundary sums; A: normal
// derivatives times -area/4pi).  t1, t2: scratch grids of g's size.
static int pmesh_inf_adj(Ctx *c, const Grid &g, GView pb, GView rb, DBuf &t1, DBuf &t2, int slot) {
    int st = VPMX_OK;
    const int n[3] = {g.n[0], g.n[1], g.n[2]};
    const int64_t G = (int64_t)n[0] * n[1] * n[2];
    double *y1 = (double *)ensure(c, t1, sizeof(double) * 3 * (size_t)G, &st); if (st) return st;
    double *w2 = (double *)ensure(c, t2, sizeof(double) * 3 * (size_t)G, &st); if (st) return st;
    int32_t win[6] = {1, 1, 1, n[0], n[1], n[2]};
    double h[3];
    window_h(g.x0, g.h, win, h);
    VPMX_TRY(poisson_dirichlet(c, 3, n, h, pb, nullptr, view(y1, n), 0, SOLVE_FP64));
    VPMX_TRY(fold_adj(c, n, h, pb, view(y1, n), view(w2, n)));  // gbar on the faces of w2
    const int64_t nbmax = 2 * ((int64_t)n[0] * n[1] + (int64_t)n[0] * n[2] + (int64_t)n[1] * n[2]);
    double *sig = (double *)ensure(c, c->src_sig, sizeof(double) * 3 * (size_t)nbmax, &st); if (st) return st;
    double *cxyz = (double *)ensure(c, c->src_cxyz, sizeof(double) * 3 * (size_t)nbmax, &st); if (st) return st;
    double *wgt = (double *)ensure(c, c->src_w, sizeof(double) * 4 * (size_t)nbmax, &st); if (st) return st;
    double *abar = (double *)ensure(c, c->bs_part, sizeof(double) * 3 * (size_t)nbmax, &st); if (st) return st;
    int64_t nb = 0;
    // (the face-cell geometry: centroids and areas; the sources it also forms are not used)
    VPMX_TRY(normal_derivs(c, 3, g, view(y1, n), sig, cxyz, wgt, &nb));
    // BS^T: the spectral form on the forward's plan, else the direct sums
    const int sp = c->adj_bs == 2 ? boundary_sums_spectral_adj(c, g, nb, view(w2, n), abar, slot) : VPMX_ERR_UNSUP;
    if (sp == VPMX_ERR_UNSUP) VPMX_TRY(bsum_adj(c, g, nb, cxyz, view(w2, n), abar));
    else VPMX_TRY(sp);
    VPMX_TRY(nderiv_adj(c, g, nb, abar, wgt + 3 * nb, view(w2, n)));  // writes every node of w2
    VPMX_TRY(poisson_dirichlet(c, 3, n, h, view(w2, n), nullptr, rb, 0, SOLVE_FP64));
    VPMX_TRY(axpy_view(c, n, view(y1, n), rb));
    return VPMX_OK;
}

// The transposed stages of one yaps block after the curl: psibar, the adjoint of the stream
// function on the global grid (the block's inner window), to the adjoint of the block's RHS on
// every block node (returned in c->adj_b).  Transposes, last stage first, of: the inner-box
// solve with the corrected faces, the TSC face correction, the coarse infinite-domain solve,
// the coarse Laplacian and sampling, and the block's infinite-domain solve (yaps3d.f90:1-453).
static int yaps_adj(Ctx *c, GView pbg, double **wb_out) {
    int st = VPMX_OK;
    const vpmx_geom &G = c->geom;
    const vpmx_params &P = c->prm;
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    const int64_t B = (int64_t)nb[0] * nb[1] * nb[2];
    const int32_t *w = G.nnbl_bl;
    const int wn[3] = {w[3] - w[0] + 1, w[4] - w[1] + 1, w[5] - w[2] + 1};
    if (wn[0] != gn[0] || wn[1] != gn[1] || wn[2] != gn[2])
        return fail(c, VPMX_ERR_STATE, "adjoint: the block's inner window is not the global grid");
    double *PB = (double *)ensure(c, c->adj_a, sizeof(double) * 3 * (size_t)B, &st); if (st) return st;
    double *WB = (double *)ensure(c, c->adj_b, sizeof(double) * 3 * (size_t)B, &st); if (st) return st;
    double *YB = (double *)ensure(c, c->adj_c, sizeof(double) * 3 * (size_t)B, &st); if (st) return st;
    double *FB = (double *)ensure(c, c->adj_d, sizeof(double) * 3 * (size_t)B, &st); if (st) return st;
    auto inner = [&](double *base) { return view(base, wn, w[0] - 1, w[1] - 1, w[2] - 1, nb); };
    VPMX_TRY(fill(c, PB, 3 * B, 0.0));
    VPMX_TRY(fill(c, WB, 3 * B, 0.0));
    VPMX_TRY(fill(c, YB, 3 * B, 0.0));
    VPMX_TRY(fill(c, FB, 3 * B, 0.0));
    VPMX_TRY(axpy_view(c, wn, pbg, inner(PB)));
    // final inner-box solve^T: wbar += D0(psibar_in); fbar = psibar_faces - F^T D0(psibar_in)
    double hI[3];
    window_h(&G.xbound_bl[0], G.dpm, w, hI);
    VPMX_TRY(poisson_dirichlet(c, 3, wn, hI, inner(PB), nullptr, inner(YB), 0, SOLVE_FP64));
    VPMX_TRY(axpy_view(c, wn, inner(YB), inner(WB)));
    VPMX_TRY(fold_adj(c, wn, hI, inner(PB), inner(YB), inner(FB)));
    // face correction^T: phibar (PB reused) = fbar on the inner faces; psibar_c, smpbar by TSC^T
    VPMX_TRY(fill(c, PB, 3 * B, 0.0));
    const int cn[3] = {G.nn_coarse[0], G.nn_coarse[1], G.nn_coarse[2]};
    const int64_t GC = (int64_t)cn[0] * cn[1] * cn[2];
    double *cb = (double *)ensure(c, c->adj_e, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    double *sb = (double *)ensure(c, c->adj_f, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    double *rc = (double *)ensure(c, c->adj_g, sizeof(double) * 3 * (size_t)GC, &st); if (st) return st;
    VPMX_TRY(fill(c, cb, 3 * GC, 0.0));
    VPMX_TRY(fill(c, sb, 3 * GC, 0.0));
    // the forward's coarse map and footprint (yaps_single)
    const int ncs = P.ncoarse;
    int ns[3], fp_lo[3], fp_hi[3];
    std::vector<int> cmap;
    for (int d = 0; d < 3; d++) {
        ns[d] = (nb[d] - 1) / ncs + 1;
        for (int a = 0; a < ns[d]; a++) {
            const double X = G.xbound_bl[d] + (double)(a * ncs) * G.dpm[d];
            cmap.push_back((int)std::lround((X - G.xbound_coarse[d]) / G.dpm_coarse[d]) + 1);
        }
        fp_lo[d] = cmap[cmap.size() - ns[d]];
        fp_hi[d] = cmap.back();
    }
    int *cmap_d = nullptr;
    VPMX_TRY(upload_cmap(c, cmap, &cmap_d));
    InterpArgs ia;
    for (int d = 0; d < 6; d++) ia.nbl[d] = w[d];
    for (int d = 0; d < 3; d++) {
        ia.xbl[d] = G.xbound_bl[d];
        ia.h[d] = G.dpm[d];
        ia.xc[d] = G.xbound_coarse[d];
        ia.hc[d] = G.dpm_coarse[d];
        ia.cn[d] = cn[d];
        ia.fp_lo[d] = fp_lo[d];
        ia.fp_hi[d] = fp_hi[d];
    }
    VPMX_TRY(face_interp_adj(c, ia, view(FB, nb), view(PB, nb), view(cb, cn), view(sb, cn)));
    // coarse solve^T, coarse Laplacian^T, sampling^T
    const Grid gc = make_grid(G.nn_coarse, G.xbound_coarse, G.dpm_coarse);
    VPMX_TRY(pmesh_inf_adj(c, gc, view(cb, cn), view(rc, cn), c->adj_h, c->adj_i, 1));
    VPMX_TRY(coarse_laplacian_adj(c, cn, fp_lo, fp_hi, G.dpm_coarse, view(rc, cn), view(sb, cn)));
    VPMX_TRY(coarse_sample_adj(c, ns, ncs, cmap_d, view(sb, cn), view(PB, nb)));
    // fine block solve^T
    const Grid gb = make_grid(G.nnbl, &G.xbound_bl[0], G.dpm);
    VPMX_TRY(pmesh_inf_adj(c, gb, view(PB, nb), view(YB, nb), c->adj_h, c->adj_i, 0));
    VPMX_TRY(axpy_view(c, nb, view(YB, nb), view(WB, nb)));
    *wb_out = WB;
    return VPMX_OK;
}

static int adjoint_supported(Ctx *c) {
    const vpmx_geom &G = c->geom;
    const vpmx_params &P = c->prm;
    if (G.nblocks != 1) return fail(c, VPMX_ERR_UNSUP, "adjoint: one yaps block only");
    if (P.iyntree != 0 || P.ibctyp != 2 || P.interf_iproj != 4)
        return fail(c, VPMX_ERR_UNSUP, "adjoint: direct point-source boundary sums and the M4' projection only");
    return VPMX_OK;
}

// J = 1/2 sum_k |u(p_k)|^2 for the velocity grid of vpm(2) on the particles (xp, qp: device),
// and its gradient with respect to QP(1:3) (dq) and XP (dx), device arrays of 3 per particle.
// Runs the forward (projection, yaps solve, curl) itself, in the plain FP64 form.
static int adjoint_probe_energy(Ctx *c, const double *xp, const double *qp, int64_t n, const double *pts, int64_t np,
                                double *J, double *dq, double *dx) {
    int st = VPMX_OK;
    VPMX_TRY(check_ready(c, c->neqpm));
    VPMX_TRY(adjoint_supported(c));
    const vpmx_geom &G = c->geom;
    const int saved_prec = c->precision;
    c->precision = 0;
    int r = run_vpm(c, 1, xp, qp, nullptr, nullptr, n);  // projection, yaps solve, velocity grid
    c->precision = saved_prec;
    VPMX_TRY(r);
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    const int64_t GG = (int64_t)gn[0] * gn[1] * gn[2];
    const Grid gg = make_grid(G.nn, G.xbound, G.dpm);
    // velocity at the probes and J
    double *u = (double *)ensure(c, c->stage_c, sizeof(double) * 3 * (size_t)(np > 0 ? np : 1), &st); if (st) return st;
    VPMX_TRY(probe_interp(c, gg, pts, np, view((double *)c->vel.p, gn), u));
    std::vector<double> uh((size_t)(3 * np));
    if (np) VPMX_HIP(c, hipMemcpyAsync(uh.data(), u, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    double j = 0.0;
    for (double v : uh) j += 0.5 * v * v;
    *J = j;
    // ubar = M^T u on the global grid, psibar = curl^T ubar
    double *ub = (double *)ensure(c, c->stage_a, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    double *pbg = (double *)ensure(c, c->stage_b, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    VPMX_TRY(fill(c, ub, 3 * GG, 0.0));
    VPMX_TRY(probe_scatter(c, gg, pts, np, u, view(ub, gn)));
    VPMX_TRY(curl_adj(c, gn, G.dpm, view(ub, gn), view(pbg, gn)));
    double *WB = nullptr;
    VPMX_TRY(yaps_adj(c, view(pbg, gn), &WB));
    // the projection^T over the whole block, as the projection writes it
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    const Grid gb = make_grid(G.nnbl, &G.xbound_bl[0], G.dpm);
    const double DV = G.dpm[0] * G.dpm[1] * G.dpm[2];
    VPMX_TRY(grad_particles(c, gb, xp, qp, c->neqpm + 1, n, view(WB, nb), 1.0 / DV, dq, dx));
    return VPMX_OK;
}

// One vpm(WhatToDo=2) evaluation as the adjoint needs it: the particles and the grids the
// forward formed (the block RHS, the velocity and stretching grids on the global grid).
struct StageRef {
    const double *x, *q;
    int64_t n;
    const double *R, *u, *s;
};

// vpm(2)^T at one evaluation: upstream adjoints ubar_p = au*us_p of UP and gbar_p = ag*gs_p of
// GP (3 per particle) to the particles' adjoints, added to xacc and qacc (3 per particle):
// M2P^T, stretching^T, curl^T, yaps^T, projection^T.
static int vpm2_adj(Ctx *c, const StageRef &S, const double *us, double au, const double *gs, double ag, double *xacc,
                    double *qacc) {
    int st = VPMX_OK;
    const vpmx_geom &G = c->geom;
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    const int64_t GG = (int64_t)gn[0] * gn[1] * gn[2];
    const int nb[3] = {G.nnbl[0], G.nnbl[1], G.nnbl[2]};
    const int32_t *w = G.nnbl_bl;
    const Grid gg = make_grid(G.nn, G.xbound, G.dpm);
    const int qs = c->neqpm + 1;
    double *ubg = (double *)ensure(c, c->adj_u, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    double *sbg = (double *)ensure(c, c->adj_s, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    double *rs = (double *)ensure(c, c->adj_r, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    double *pbg = (double *)ensure(c, c->stage_b, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    VPMX_TRY(m2p_adj(c, gg, S.x, S.q, qs, S.n, us, au, gs, ag, view((double *)S.u, gn), view((double *)S.s, gn),
                     view(ubg, gn), view(sbg, gn), xacc));
    const GView Rg = view((double *)S.R, gn, w[0] - 1, w[1] - 1, w[2] - 1, nb);
    VPMX_TRY(stretch_adj(c, gn, G.dpm, Rg, view((double *)S.u, gn), view(sbg, gn), view(rs, gn), view(ubg, gn)));
    VPMX_TRY(curl_adj(c, gn, G.dpm, view(ubg, gn), view(pbg, gn)));
    double *WB = nullptr;
    VPMX_TRY(yaps_adj(c, view(pbg, gn), &WB));
    VPMX_TRY(axpy_view(c, gn, view(rs, gn), view(WB, gn, w[0] - 1, w[1] - 1, w[2] - 1, nb)));
    const Grid gb = make_grid(G.nnbl, &G.xbound_bl[0], G.dpm);
    const double DV = G.dpm[0] * G.dpm[1] * G.dpm[2];
    VPMX_TRY(grad_particles(c, gb, S.x, S.q, qs, S.n, view(WB, nb), 1.0 / DV, qacc, xacc, true));
    return VPMX_OK;
}

static int save_buf(Ctx *c, DBuf &dst, const void *src, size_t bytes) {
    int st = VPMX_OK;
    void *d = ensure(c, dst, bytes, &st);
    if (st) return st;
    if (bytes) VPMX_HIP(c, hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return VPMX_OK;
}

// One RK2 step of the resident particles in the plain FP64 form (the sequence of
// step_rk2_enqueue with mrem = 1), keeping what the step's adjoint needs when rec is given:
// b[0..5] particles at the start, the midpoint and the end (before the remesh), b[6..8] and
// b[9..11] the grids of the two evaluations, b[12] the remeshed positions.
static int rk2_step_rec(Ctx *c, double dt, const double uinf[3], int ntime, Ctx::AdjStep *rec) {
    int st = VPMX_OK;
    const int64_t n = c->np;
    const int nr = c->neqpm + 1;
    const vpmx_geom &G = c->geom;
    const size_t xb = sizeof(double) * 3 * (size_t)n, qb = sizeof(double) * nr * (size_t)n;
    const size_t gb = sizeof(double) * 3 * (size_t)G.nn[0] * G.nn[1] * G.nn[2];
    const size_t bb = sizeof(double) * 3 * (size_t)G.nnbl[0] * G.nnbl[1] * G.nnbl[2];
    double *x0 = (double *)ensure(c, c->x0, sizeof(double) * 3 * (size_t)(n ? n : 1), &st); if (st) return st;
    double *q0 = (double *)ensure(c, c->q0, sizeof(double) * nr * (size_t)(n ? n : 1), &st); if (st) return st;
    double *xp = (double *)c->xp.p, *qp = (double *)c->qp.p, *up = (double *)c->up.p, *gp = (double *)c->gp.p;
    if (n > 0) {
        VPMX_HIP(c, hipMemcpyAsync(x0, xp, xb, hipMemcpyDeviceToDevice, c->stream));
        VPMX_HIP(c, hipMemcpyAsync(q0, qp, qb, hipMemcpyDeviceToDevice, c->stream));
    }
    if (rec) {
        rec->n0 = n;
        VPMX_TRY(save_buf(c, rec->b[0], xp, xb));
        VPMX_TRY(save_buf(c, rec->b[1], qp, qb));
    }
    VPMX_TRY(run_vpm(c, 2, xp, qp, up, gp, n));
    if (rec) {
        VPMX_TRY(save_buf(c, rec->b[6], c->rhs_bl.p, bb));
        VPMX_TRY(save_buf(c, rec->b[7], c->vel.p, gb));
        VPMX_TRY(save_buf(c, rec->b[8], c->str.p, gb));
    }
    VPMX_TRY(euler_update(c, xp, qp, nr, up, gp, n, 0.5 * dt, uinf, x0, q0));
    if (rec) {
        VPMX_TRY(save_buf(c, rec->b[2], xp, xb));
        VPMX_TRY(save_buf(c, rec->b[3], qp, qb));
    }
    VPMX_TRY(run_vpm(c, 2, xp, qp, up, gp, n));
    if (rec) {
        VPMX_TRY(save_buf(c, rec->b[9], c->rhs_bl.p, bb));
        VPMX_TRY(save_buf(c, rec->b[10], c->vel.p, gb));
        VPMX_TRY(save_buf(c, rec->b[11], c->str.p, gb));
    }
    VPMX_TRY(euler_update(c, xp, qp, nr, up, gp, n, dt, uinf, x0, q0));
    if (rec) {
        VPMX_TRY(save_buf(c, rec->b[4], xp, xb));
        VPMX_TRY(save_buf(c, rec->b[5], qp, qb));
    }
    const vpmx_geom g0 = c->geom;
    VPMX_TRY(define_from_device(c, ntime));
    const vpmx_geom &g1 = c->geom;
    auto same = [](const void *a, const void *b, size_t n) { return std::memcmp(a, b, n) == 0; };
    if (!same(g0.xbound, g1.xbound, sizeof g0.xbound) || !same(g0.dpm, g1.dpm, sizeof g0.dpm) ||
        !same(g0.nn, g1.nn, sizeof g0.nn) || !same(g0.nnbl, g1.nnbl, sizeof g0.nnbl) ||
        !same(g0.xbound_bl, g1.xbound_bl, sizeof g0.xbound_bl) || !same(g0.nnbl_bl, g1.nnbl_bl, sizeof g0.nnbl_bl) ||
        !same(g0.xbound_coarse, g1.xbound_coarse, sizeof g0.xbound_coarse) ||
        !same(g0.nn_coarse, g1.nn_coarse, sizeof g0.nn_coarse))
        return fail(c, VPMX_ERR_UNSUP, "adjoint roll-out: the box changed during the roll-out (use idefine = 1)");
    VPMX_TRY(run_remesh(c, (const double *)c->xp.p, (const double *)c->qp.p, n, 1));
    VPMX_TRY(replace_with_remesh(c));
    if (rec) {
        rec->n1 = c->np;
        VPMX_TRY(save_buf(c, rec->b[12], c->xp.p, sizeof(double) * 3 * (size_t)c->np));
    }
    return VPMX_OK;
}

// The adjoint of one recorded step: (xbar, qbar) of the remeshed particles (rec.n1, 3 per
// particle) in adj_xb / adj_qb become those of the step's starting particles (rec.n0).
// remesh^T with its threshold mask held fixed (the remeshed particles sit on the kept nodes;
// their positions carry no adjoint), then the two stages' vpm(2)^T.
static int rk2_step_adj(Ctx *c, const Ctx::AdjStep &rec, double dt) {
    int st = VPMX_OK;
    const vpmx_geom &G = c->geom;
    const int64_t n = rec.n0;
    const int qs = c->neqpm + 1;
    const size_t pb = sizeof(double) * 3 * (size_t)(n > 0 ? n : 1);
    // remesh^T: q' = sum_p W(x_p - X) q_p on the remesh grid (vpm_remesh.f90:137-195, ncell_rem 1)
    int32_t NN[3];
    double D[3];
    for (int d = 0; d < 3; d++) {
        NN[d] = G.nn[d];
        D[d] = (G.xbound[d + 3] - G.xbound[d]) / (double)(NN[d] - 1);
    }
    const Grid gr = make_grid(NN, G.xbound, D);
    const int64_t GG = (int64_t)NN[0] * NN[1] * NN[2];
    double *grid = (double *)ensure(c, c->adj_r, sizeof(double) * 3 * (size_t)GG, &st); if (st) return st;
    VPMX_TRY(fill(c, grid, 3 * GG, 0.0));
    VPMX_TRY(scatter_nodes(c, gr, (const double *)rec.b[12].p, (const double *)c->adj_qb.p, rec.n1, view(grid, NN)));
    double *XB = (double *)ensure(c, c->adj_xb2, pb, &st); if (st) return st;
    double *QB = (double *)ensure(c, c->adj_qb2, pb, &st); if (st) return st;
    VPMX_TRY(grad_particles(c, gr, (const double *)rec.b[4].p, (const double *)rec.b[5].p, qs, n, view(grid, NN), 1.0, QB,
                            XB, false));
    std::swap(c->adj_xb, c->adj_xb2);
    std::swap(c->adj_qb, c->adj_qb2);
    // x1 = x0 + dt (u_m + uinf), q1 = q0 - dt g_m at the midpoint (x_m, q_m) = (x0, q0) + dt/2 (u_0 + uinf, -g_0)
    double *XM = (double *)ensure(c, c->adj_xm, pb, &st); if (st) return st;
    double *QM = (double *)ensure(c, c->adj_qm, pb, &st); if (st) return st;
    XB = (double *)c->adj_xb.p;
    QB = (double *)c->adj_qb.p;
    VPMX_TRY(fill(c, XM, 3 * n, 0.0));
    VPMX_TRY(fill(c, QM, 3 * n, 0.0));
    const StageRef Sm{(const double *)rec.b[2].p, (const double *)rec.b[3].p, n, (const double *)rec.b[9].p,
                      (const double *)rec.b[10].p, (const double *)rec.b[11].p};
    VPMX_TRY(vpm2_adj(c, Sm, XB, dt, QB, -dt, XM, QM));
    VPMX_TRY(axpy(c, 3 * n, 1.0, XM, XB));
    VPMX_TRY(axpy(c, 3 * n, 1.0, QM, QB));
    const StageRef S0{(const double *)rec.b[0].p, (const double *)rec.b[1].p, n, (const double *)rec.b[6].p,
                      (const double *)rec.b[7].p, (const double *)rec.b[8].p};
    VPMX_TRY(vpm2_adj(c, S0, XM, 0.5 * dt, QM, -0.5 * dt, XB, QB));
    return VPMX_OK;
}

// J of the probes after nsteps RK2 steps (each with its remesh) of the resident particles, and
// its gradient with respect to the starting QP(1:3) and XP (in adj_qb / adj_xb, 3 per particle).
// Checkpoints every ckpt steps; the backward pass recomputes one segment of steps at a time,
// keeping its records.  ckpt = 0: the forward pass keeps every step's record (no recomputation;
// about 6 grids and 7 particle arrays of HBM per step).
static int adjoint_rollout(Ctx *c, int nsteps, double dt, const double uinf[3], int ckpt, const double *pts,
                           int64_t np, double *J) {
    int st = VPMX_OK;
    VPMX_TRY(check_ready(c, c->neqpm));
    VPMX_TRY(adjoint_supported(c));
    if (c->prm.ncell_rem != 1) return fail(c, VPMX_ERR_UNSUP, "ncell_rem must be 1");
    const int nr = c->neqpm + 1;
    const bool keep_all = ckpt <= 0;  // every step's record from the forward pass, no recomputation
    if (keep_all) ckpt = nsteps > 0 ? nsteps : 1;
    const int nseg = keep_all ? 0 : (nsteps + ckpt - 1) / ckpt;
    if ((int)c->adj_ck.size() < nseg) c->adj_ck.resize(nseg);
    if ((int)c->adj_steps.size() < ckpt) c->adj_steps.resize(ckpt);
    std::vector<int64_t> ckn(nseg), fwd_n(nsteps);
    for (int k = 0; k < nsteps && keep_all; k++) VPMX_TRY(rk2_step_rec(c, dt, uinf, k + 1, &c->adj_steps[k]));
    // the last segment keeps its records from this pass (the backward pass starts there)
    const int klast = keep_all ? 0 : (nseg - 1) * ckpt;
    for (int k = 0; k < nsteps && !keep_all; k++) {
        if (k % ckpt == 0) {
            ckn[k / ckpt] = c->np;
            VPMX_TRY(save_buf(c, c->adj_ck[k / ckpt][0], c->xp.p, sizeof(double) * 3 * (size_t)c->np));
            VPMX_TRY(save_buf(c, c->adj_ck[k / ckpt][1], c->qp.p, sizeof(double) * nr * (size_t)c->np));
        }
        VPMX_TRY(rk2_step_rec(c, dt, uinf, k + 1, k >= klast ? &c->adj_steps[k - klast] : nullptr));
        fwd_n[k] = c->np;
    }
    const int64_t nK = c->np;
    const size_t pbK = sizeof(double) * 3 * (size_t)(nK > 0 ? nK : 1);
    double *XB = (double *)ensure(c, c->adj_xb, pbK, &st); if (st) return st;
    double *QB = (double *)ensure(c, c->adj_qb, pbK, &st); if (st) return st;
    VPMX_TRY(adjoint_probe_energy(c, (const double *)c->xp.p, (const double *)c->qp.p, nK, pts, np, J, QB, XB));
    for (int k = nsteps - 1; k >= 0 && keep_all; k--) VPMX_TRY(rk2_step_adj(c, c->adj_steps[k], dt));
    for (int s = nseg - 1; s >= 0; s--) {
        const int k0 = s * ckpt, k1 = std::min(nsteps, k0 + ckpt);
        if (s == nseg - 1) {  // recorded by the forward pass
            for (int k = k1 - 1; k >= k0; k--) VPMX_TRY(rk2_step_adj(c, c->adj_steps[k - k0], dt));
            continue;
        }
        const int64_t n0 = ckn[s];
        double *dx = (double *)ensure(c, c->xp, sizeof(double) * 3 * (size_t)(n0 ? n0 : 1), &st); if (st) return st;
        double *dq = (double *)ensure(c, c->qp, sizeof(double) * nr * (size_t)(n0 ? n0 : 1), &st); if (st) return st;
        if (n0) {
            VPMX_HIP(c, hipMemcpyAsync(dx, c->adj_ck[s][0].p, sizeof(double) * 3 * n0, hipMemcpyDeviceToDevice, c->stream));
            VPMX_HIP(c, hipMemcpyAsync(dq, c->adj_ck[s][1].p, sizeof(double) * nr * n0, hipMemcpyDeviceToDevice,
                                       c->stream));
        }
        c->np = n0;
        ensure(c, c->up, sizeof(double) * 3 * (size_t)(n0 ? n0 : 1), &st); if (st) return st;
        ensure(c, c->gp, sizeof(double) * 3 * (size_t)(n0 ? n0 : 1), &st); if (st) return st;
        for (int k = k0; k < k1; k++) {
            VPMX_TRY(rk2_step_rec(c, dt, uinf, k + 1, &c->adj_steps[k - k0]));
            if (c->np != fwd_n[k]) return fail(c, VPMX_ERR_STATE, "adjoint roll-out: the recomputed steps differ");
        }
        for (int k = k1 - 1; k >= k0; k--) VPMX_TRY(rk2_step_adj(c, c->adj_steps[k - k0], dt));
    }
    return VPMX_OK;
}

// One RK2 (midpoint) step of the resident particles, queued on the context's stream: two
// vpm(2) evaluations with the Euler updates of test.f90:104-117, then vpm(0) and
// remesh_particles_3d(1) when do_remesh.  In the device-count mode the launches are sized for
// the capacity, the count stays on the device and the remesh writes the resident arrays in
// place; the ordinary mode swaps the remeshed particles in (their count is read back).
int step_rk2_enqueue(Ctx *c, double dt, const double uinf[3], int ntime, int mrem, int do_remesh) {
    int st = VPMX_OK;
    const int64_t n = c->dn ? c->cap : c->np;
    int nr = c->neqpm + 1;
    double *x0 = (double *)ensure(c, c->x0, sizeof(double) * 3 * (size_t)(n ? n : 1), &st); if (st) return st;
    double *q0 = (double *)ensure(c, c->q0, sizeof(double) * nr * (size_t)(n ? n : 1), &st); if (st) return st;
    double *xp = (double *)c->xp.p, *qp = (double *)c->qp.p, *up = (double *)c->up.p, *gp = (double *)c->gp.p;
    TimingScope timing(c);  // the stage times of the whole step
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    VPMX_HIP(c, hipStreamIsCapturing(c->stream, &cap));
    const bool swap = !c->dn && cap == hipStreamCaptureStatusNone;
    if (swap) {
        // the step's start particles become (x0, q0) by exchanging the buffers (both hold n
        // particles), and the first update writes the resident arrays from them
        std::swap(c->x0, c->xp);
        std::swap(c->q0, c->qp);
        x0 = (double *)c->x0.p;
        q0 = (double *)c->q0.p;
        xp = (double *)c->xp.p;
        qp = (double *)c->qp.p;
    } else if (n > 0) {
        // (a batch's cloud, or under capture: the buffers' roles stay fixed for replays)
        VPMX_HIP(c, hipMemcpyAsync(x0, xp, sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, c->stream));
        VPMX_HIP(c, hipMemcpyAsync(q0, qp, sizeof(double) * nr * n, hipMemcpyDeviceToDevice, c->stream));
    }
    // stage 1 at (x0, q0); if it fails, the resident arrays get their particles back
    const int r1 = run_vpm(c, 2, x0, q0, up, gp, n);
    if (r1 != VPMX_OK) {
        if (swap) {
            std::swap(c->x0, c->xp);
            std::swap(c->q0, c->qp);
        }
        return r1;
    }
    VPMX_TRY(euler_update(c, xp, qp, nr, up, gp, n, 0.5 * dt, uinf, x0, q0));
    // stage 2 at the midpoint
    VPMX_TRY(run_vpm(c, 2, xp, qp, up, gp, n));
    VPMX_TRY(euler_update(c, xp, qp, nr, up, gp, n, dt, uinf, x0, q0));
    if (do_remesh) {
        VPMX_TRY(define_from_device(c, ntime));
        VPMX_TRY(run_remesh(c, xp, qp, n, mrem));
        if (!c->dn) VPMX_TRY(replace_with_remesh(c));
    }
    return VPMX_OK;
}

// The particles of the remesh window of the context's geometry (mrem-fold grid): the most a
// remesh can keep.
int64_t remesh_window_nodes(const Ctx *c, int mrem) {
    const vpmx_geom &G = c->geom;
    const int half = c->prm.interf_iproj / 2;
    int64_t n = 1;
    for (int d = 0; d < 3; d++) {
        const int lo = G.nn_bl[d] * mrem + half, hi = G.nn_bl[d + 3] * mrem - half;
        n *= hi >= lo ? (int64_t)(hi - lo + 1) : 0;
    }
    return n;
}

}  // namespace vpmx

using namespace vpmx;

static Ctx *C(vpmx_ctx *p) { return static_cast<Ctx *>(p); }
static const Ctx *C(const vpmx_ctx *p) { return static_cast<const Ctx *>(p); }

extern "C" {

int32_t vpmx_create(vpmx_ctx **out, int32_t device) {
    if (!out) return VPMX_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return VPMX_ERR_NODEV;
    if (device < 0 || device >= ndev) return VPMX_ERR_ARG;
    Ctx *c = new Ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return VPMX_ERR_HIP;
    }
    c->kev_ok = true;
    if (scratch_init(c) != VPMX_OK) c->scratch_err = c->err;  // only the allocation-free calls need it
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        c->num_cu = ncu;
    int mlds = 0;
    if (hipDeviceGetAttribute(&mlds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && mlds > 0)
        c->max_lds = mlds;
    const char *pm = getenv("VPMX_PRECISION");
    if (pm && std::string(pm) == "mixed") c->precision = 1;

    *out = c;
    return VPMX_OK;
}

void vpmx_destroy(vpmx_ctx *p) {
    Ctx *c = C(p);
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DBuf *bufs[] = {&c->xp, &c->qp, &c->up, &c->gp, &c->x0, &c->q0, &c->rhs_bl, &c->sol_bl, &c->sol0, &c->vel,
                    &c->str, &c->src_sig, &c->src_cxyz, &c->src_w, &c->bs_part, &c->lat_pairs, &c->lat_pairs_c, &c->panel_a, &c->rhs_g, &c->sol_g, &c->mb_blocks, &c->mb_tmp, &c->tile_flag, &c->tile_pos, &c->tile_list, &c->samp, &c->rhs_c, &c->sol_c,
                    &c->sol0_c, &c->cmap, &c->dst_work, &c->lam[0], &c->lam[1], &c->lam[2], &c->scan_tmp,
                    &c->red_tmp, &c->rm_flag, &c->rm_pos, &c->rm_x, &c->rm_q, &c->rm_grid, &c->stage_a, &c->stage_b,
                    &c->stage_c, &c->sort.keys, &c->sort.perm, &c->sort.count, &c->sort.start, &c->sort.fill,
                    &c->sort.xs, &c->sort.qs, &c->sort.rowcnt, &c->dst_save, &c->fold_tab, &c->zt_prev, &c->dst_work64,
                    &c->tune_scratch, &c->adj_a, &c->adj_b, &c->adj_c, &c->adj_d, &c->adj_e, &c->adj_f,
                    &c->adj_g, &c->adj_h, &c->adj_i, &c->adj_p, &c->adj_q, &c->adj_u, &c->adj_s, &c->adj_r,
                    &c->adj_xb, &c->adj_qb, &c->adj_xb2, &c->adj_qb2, &c->adj_xm, &c->adj_qm, &c->adj_pu,
                    &c->adj_ps, &c->nd_face, &c->fd_spec, &c->fp_spec};
    for (DBuf *b : bufs) release(*b);
    for (auto &r : c->adj_steps)
        for (DBuf &b : r.b) release(b);
    for (auto &ck : c->adj_ck)
        for (DBuf &b : ck) release(b);
    free_bspec(c);
    free_tree(c);
    free_fft_plans(c);
    for (int k = 0; k < K_COUNT; k++) {
        for (hipEvent_t e : c->kev0[k]) (void)hipEventDestroy(e);
        for (hipEvent_t e : c->kev1[k]) (void)hipEventDestroy(e);
    }
    for (int k = 0; k < ST_COUNT; k++) {
        for (hipEvent_t e : c->sev0[k]) (void)hipEventDestroy(e);
        for (hipEvent_t e : c->sev1[k]) (void)hipEventDestroy(e);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *vpmx_last_error(const vpmx_ctx *p) { return p ? C(p)->err.c_str() : "null context"; }

int32_t vpmx_set_params(vpmx_ctx *p, const vpmx_params *prm) {
    Ctx *c = C(p);
    if (!c || !prm) return VPMX_ERR_ARG;
    if (prm->ipmwrite > 10) return fail(c, VPMX_ERR_ARG, "IPMWRITE > 10");
    if (c->have_prm && (prm->nbi != c->prm.nbi || prm->nbj != c->prm.nbj || prm->nbk != c->prm.nbk))
        for (int d = 0; d < 3; d++) c->block_w[d].clear();  // the weights were for other block counts
    c->prm = *prm;
    c->have_prm = true;
    return VPMX_OK;
}

int32_t vpmx_set_block_weights(vpmx_ctx *p, const int32_t *wx, const int32_t *wy, const int32_t *wz) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    if (!c->have_prm) return fail(c, VPMX_ERR_STATE, "block weights: set the parameters first");
    const int32_t *w[3] = {wx, wy, wz};
    const int n[3] = {c->prm.nbi, c->prm.nbj, c->prm.nbk};
    for (int d = 0; d < 3; d++) {
        c->block_w[d].clear();
        if (!w[d]) continue;
        for (int i = 0; i < n[d]; i++)
            if (w[d][i] < 1) return fail(c, VPMX_ERR_ARG, "block weights must be positive");
        c->block_w[d].assign(w[d], w[d] + n[d]);
    }
    return VPMX_OK;
}

int32_t vpmx_get_geometry(const vpmx_ctx *p, vpmx_geom *out) {
    if (!p || !out) return VPMX_ERR_ARG;
    *out = C(p)->geom;
    return VPMX_OK;
}

int32_t vpmx_set_option(vpmx_ctx *p, const char *name, int32_t value) {
    Ctx *c = C(p);
    if (!c || !name) return VPMX_ERR_ARG;
    const std::string k(name);
    if (k == "boundsum") {
        if (value < 0 || value > 2) return fail(c, VPMX_ERR_ARG, "boundsum: 0 lattice, 1 direct, 2 spectral");
        c->bs_mode = value;
    } else if (k == "adjoint_boundsum") {
        if (value != 1 && value != 2) return fail(c, VPMX_ERR_ARG, "adjoint_boundsum: 1 direct, 2 spectral");
        c->adj_bs = value;
    } else if (k == "adjoint_m2p_scatter") {
        c->adj_m2p_scatter = value != 0;
    } else if (k == "dst_reuse") {
        c->dst_reuse = value != 0;
    } else if (k == "fine_partial") {
        if (value < -1 || value > 1) return fail(c, VPMX_ERR_ARG, "fine_partial: -1 by grid size, 0 every node, 1 faces and samples");
        c->fine_partial = value;
    } else if (k == "face_derivs") {
        if (value < -1 || value > 1) return fail(c, VPMX_ERR_ARG, "face_derivs: -1 by grid size, 0 solved field, 1 spectral");
        c->face_derivs = value;
    } else if (k == "tuning") {
        c->tuning = value != 0;
        c->tuned.clear();
    } else if (k == "tune_variant") {
        if (value < -1) return fail(c, VPMX_ERR_ARG, "tune_variant: -1 (tuner) or a variant index");
        c->tune_variant = value;
        c->tuned.clear();
    } else {
        return fail(c, VPMX_ERR_ARG, "unknown option " + k);
    }
    return VPMX_OK;
}

int32_t vpmx_tuning_report(const vpmx_ctx *p, char *buf, int64_t capacity) {
    if (!p || !buf || capacity < 1) return VPMX_ERR_ARG;
    const Ctx *c = C(p);
    std::string out;
    for (const auto &kv : c->tuned) out += kv.first + " -> " + kv.second.what + "\n";
    const size_t n = std::min((size_t)(capacity - 1), out.size());
    std::memcpy(buf, out.data(), n);
    buf[n] = 0;
    return VPMX_OK;
}

int32_t vpmx_set_precision(vpmx_ctx *p, int32_t mode) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    if (mode != 0 && mode != 1) return fail(c, VPMX_ERR_ARG, "precision: 0 plain FP64 form, 1 mixed form");
    c->precision = mode;
    return VPMX_OK;
}

int32_t vpmx_vpm(vpmx_ctx *p, double *xp, double *qp, double *up, double *gp, int64_t nvr, int32_t neqpm,
                 int32_t whattodo, int32_t ntime, double ni) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    VPMX_HIP(c, hipSetDevice(c->device));
    if (!c->have_prm) return fail(c, VPMX_ERR_STATE, "no parameters set");
    c->ni = ni;
    if (whattodo == 0) {
        VPMX_TRY(define_sizes_ctx(c, xp, nvr, ntime));
        c->have_grids = false;
        // velvr*_pm = 0 after WhatToDo = 0 (vpm.f90:154-166)
        if (c->vel.p) VPMX_TRY(fill(c, (double *)c->vel.p, (int64_t)(c->vel.bytes / sizeof(double)), 0.0));
        VPMX_HIP(c, hipStreamSynchronize(c->stream));
        return VPMX_OK;
    }
    if (whattodo < 1 || whattodo > 5) return fail(c, VPMX_ERR_ARG, "WhatToDo must be 0..5");
    if (nvr == 0) return VPMX_OK;
    VPMX_TRY(upload_particles(c, xp, qp, nvr, neqpm));
    VPMX_TRY(run_vpm(c, whattodo, (const double *)c->xp.p, (const double *)c->qp.p, (double *)c->up.p,
                     (double *)c->gp.p, nvr, ni));
    if (whattodo == 2 || whattodo == 4 || whattodo == 5) {
        if (up) VPMX_HIP(c, hipMemcpyAsync(up, c->up.p, sizeof(double) * 3 * nvr, hipMemcpyDeviceToHost, c->stream));
        if (gp) VPMX_HIP(c, hipMemcpyAsync(gp, c->gp.p, sizeof(double) * 3 * nvr, hipMemcpyDeviceToHost, c->stream));
    }
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_remesh_count(vpmx_ctx *p, const double *xp, const double *qp, int64_t nvr, int32_t neqpm, int32_t mrem,
                          int64_t *nvr_out) {
    Ctx *c = C(p);
    if (!c || !nvr_out) return VPMX_ERR_ARG;
    VPMX_HIP(c, hipSetDevice(c->device));
    VPMX_TRY(upload_particles(c, xp, qp, nvr, neqpm));
    VPMX_TRY(run_remesh(c, (const double *)c->xp.p, (const double *)c->qp.p, nvr, mrem));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    *nvr_out = c->rm_count;
    return VPMX_OK;
}

int32_t vpmx_remesh_fetch(vpmx_ctx *p, double *xp_out, double *qp_out) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    if (c->rm_count < 0) return fail(c, VPMX_ERR_STATE, "no remesh result to fetch");
    int nr = c->neqpm + 1;
    if (c->rm_count > 0) {
        if (xp_out)
            VPMX_HIP(c, hipMemcpyAsync(xp_out, c->rm_xp, sizeof(double) * 3 * c->rm_count, hipMemcpyDeviceToHost,
                                       c->stream));
        if (qp_out)
            VPMX_HIP(c, hipMemcpyAsync(qp_out, c->rm_qp, sizeof(double) * nr * c->rm_count, hipMemcpyDeviceToHost,
                                       c->stream));
    }
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_get_grid(vpmx_ctx *p, int32_t which, double *dst, int64_t capacity) {
    Ctx *c = C(p);
    if (!c || !dst) return VPMX_ERR_ARG;
    if (!c->have_grids) return fail(c, VPMX_ERR_STATE, "no grids: run vpm mode 1, 2 or 3 first");
    const vpmx_geom &G = c->geom;
    const int gn[3] = {G.nn[0], G.nn[1], G.nn[2]};
    int64_t GG = (int64_t)gn[0] * gn[1] * gn[2];
    int st = VPMX_OK;
    int ncd = which == 0 ? c->neqpm + 1 : 3;
    if (capacity < ncd * GG) return fail(c, VPMX_ERR_ARG, "get_grid: destination too small");
    double *stage = (double *)ensure(c, c->stage_a, sizeof(double) * ncd * (size_t)GG, &st);
    if (st) return st;
    const double DV = G.dpm[0] * G.dpm[1] * G.dpm[2];
    if (which == 0) {
        VPMX_TRY(soa_to_aos(c, 3, 4, gn, global_rhs(c), DV, stage));
    } else if (which == 1) {
        VPMX_TRY(soa_to_aos(c, 3, 3, gn, global_psi(c), 0.0, stage));
    } else if (which == 2) {
        if (!c->str.p) return fail(c, VPMX_ERR_STATE, "no stretching grid");
        VPMX_TRY(soa_to_aos(c, 3, 3, gn, view((double *)c->str.p, gn), 0.0, stage));
    } else if (which == 3) {
        if (!c->vel.p) return fail(c, VPMX_ERR_STATE, "no velocity grid");
        VPMX_HIP(c, hipMemcpyAsync(stage, c->vel.p, sizeof(double) * 3 * GG, hipMemcpyDeviceToDevice, c->stream));
    } else {
        return fail(c, VPMX_ERR_ARG, "get_grid: which must be 0..3");
    }
    VPMX_HIP(c, hipMemcpyAsync(dst, stage, sizeof(double) * ncd * GG, hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_particles_upload(vpmx_ctx *p, const double *xp, const double *qp, int64_t nvr, int32_t neqpm) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    if (neqpm != 3) return fail(c, VPMX_ERR_UNSUP, "neqpm must be 3");
    VPMX_HIP(c, hipSetDevice(c->device));
    VPMX_TRY(upload_particles(c, xp, qp, nvr, neqpm));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_particles_count(const vpmx_ctx *p, int64_t *nvr) {
    if (!p || !nvr) return VPMX_ERR_ARG;
    *nvr = C(p)->np;
    return VPMX_OK;
}

int32_t vpmx_particles_download(vpmx_ctx *p, double *xp, double *qp, double *up, double *gp) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    int64_t n = c->np;
    int nr = c->neqpm + 1;
    if (n > 0) {
        if (xp) VPMX_HIP(c, hipMemcpyAsync(xp, c->xp.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        if (qp) VPMX_HIP(c, hipMemcpyAsync(qp, c->qp.p, sizeof(double) * nr * n, hipMemcpyDeviceToHost, c->stream));
        if (up) VPMX_HIP(c, hipMemcpyAsync(up, c->up.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
        if (gp) VPMX_HIP(c, hipMemcpyAsync(gp, c->gp.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    }
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_vpm_resident(vpmx_ctx *p, int32_t whattodo, int32_t ntime) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    VPMX_HIP(c, hipSetDevice(c->device));
    if (!c->have_prm) return fail(c, VPMX_ERR_STATE, "no parameters set");
    if (whattodo == 0) {
        VPMX_TRY(define_from_device(c, ntime));
        c->have_grids = false;
        if (c->vel.p) VPMX_TRY(fill(c, (double *)c->vel.p, (int64_t)(c->vel.bytes / sizeof(double)), 0.0));
        VPMX_HIP(c, hipStreamSynchronize(c->stream));
        return VPMX_OK;
    }
    if (whattodo < 1 || whattodo > 5) return fail(c, VPMX_ERR_ARG, "WhatToDo must be 0..5");
    VPMX_TRY(run_vpm(c, whattodo, (const double *)c->xp.p, (const double *)c->qp.p, (double *)c->up.p,
                     (double *)c->gp.p, c->np, c->ni));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}


int32_t vpmx_remesh_resident(vpmx_ctx *p, int32_t mrem) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    if (c->dn) return fail(c, VPMX_ERR_STATE, "a prepared batch's cloud: remesh through vpmx_batch_step_rk2");
    VPMX_HIP(c, hipSetDevice(c->device));
    VPMX_TRY(run_remesh(c, (const double *)c->xp.p, (const double *)c->qp.p, c->np, mrem));
    VPMX_TRY(replace_with_remesh(c));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_step_rk2(vpmx_ctx *p, double dt, const double uinf[3], int32_t ntime, int32_t mrem, int32_t do_remesh) {
    Ctx *c = C(p);
    if (!c || !uinf) return VPMX_ERR_ARG;
    VPMX_HIP(c, hipSetDevice(c->device));
    if (c->dn) return fail(c, VPMX_ERR_STATE, "this context belongs to a batch: step it with vpmx_batch_step_rk2");
    VPMX_TRY(step_rk2_enqueue(c, dt, uinf, ntime, mrem, do_remesh));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_diagnostics(vpmx_ctx *p, double circ[3], double imp[3]) {
    Ctx *c = C(p);
    if (!c || !circ || !imp) return VPMX_ERR_ARG;
    VPMX_HIP(c, hipSetDevice(c->device));
    if (c->np <= 0) {
        for (int d = 0; d < 3; d++) circ[d] = imp[d] = 0.0;
        return VPMX_OK;
    }
    return diagnostics(c, (const double *)c->xp.p, (const double *)c->qp.p, c->neqpm + 1, c->np, circ, imp);
}

int32_t vpmx_diagnostics_noalloc(vpmx_ctx *p, const double *xp, const double *qp, int64_t nvr, int32_t qstride,
                                 double *out) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    if (!c->scratch) return fail(c, VPMX_ERR_STATE, "allocation-free diagnostics unavailable: " + c->scratch_err);
    VPMX_HIP(c, hipSetDevice(c->device));
    return diagnostics_noalloc(c, xp, qp, qstride, nvr, out);
}

int32_t vpmx_adjoint_probe_energy(vpmx_ctx *p, const double *xp, const double *qp, int64_t nvr, int32_t neqpm,
                                  const double *probes, int64_t nprobe, double *J, double *dq, double *dx) {
    Ctx *c = C(p);
    if (!c || !J || !dq || !dx || (nprobe > 0 && !probes)) return VPMX_ERR_ARG;
    if (nprobe < 0 || nprobe > 0x7fffffffLL) return fail(c, VPMX_ERR_ARG, "probe count out of range");
    VPMX_HIP(c, hipSetDevice(c->device));
    if (!c->have_prm) return fail(c, VPMX_ERR_STATE, "no parameters set");
    VPMX_TRY(upload_particles(c, xp, qp, nvr, neqpm));
    int st = VPMX_OK;
    double *pd = (double *)ensure(c, c->adj_p, sizeof(double) * 3 * (size_t)(nprobe > 0 ? nprobe : 1), &st);
    if (st) return st;
    double *g = (double *)ensure(c, c->adj_q, sizeof(double) * 6 * (size_t)(nvr > 0 ? nvr : 1), &st);
    if (st) return st;
    if (nprobe) VPMX_HIP(c, hipMemcpyAsync(pd, probes, sizeof(double) * 3 * nprobe, hipMemcpyHostToDevice, c->stream));
    VPMX_TRY(adjoint_probe_energy(c, (const double *)c->xp.p, (const double *)c->qp.p, nvr, pd, nprobe, J, g,
                                  g + 3 * nvr));
    if (nvr) {
        VPMX_HIP(c, hipMemcpyAsync(dq, g, sizeof(double) * 3 * nvr, hipMemcpyDeviceToHost, c->stream));
        VPMX_HIP(c, hipMemcpyAsync(dx, g + 3 * nvr, sizeof(double) * 3 * nvr, hipMemcpyDeviceToHost, c->stream));
    }
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_adjoint_rollout(vpmx_ctx *p, const double *xp, const double *qp, int64_t nvr, int32_t neqpm,
                             int32_t nsteps, double dt, const double *uinf, int32_t ckpt, const double *probes,
                             int64_t nprobe, double *J, double *dq, double *dx) {
    Ctx *c = C(p);
    if (!c || !J || !dq || !dx || !uinf || (nprobe > 0 && !probes)) return VPMX_ERR_ARG;
    if (nprobe < 0 || nprobe > 0x7fffffffLL) return fail(c, VPMX_ERR_ARG, "probe count out of range");
    if (nsteps < 0 || ckpt < 0) return fail(c, VPMX_ERR_ARG, "nsteps >= 0 and ckpt >= 0");
    VPMX_HIP(c, hipSetDevice(c->device));
    if (!c->have_prm) return fail(c, VPMX_ERR_STATE, "no parameters set");
    VPMX_TRY(upload_particles(c, xp, qp, nvr, neqpm));
    int st = VPMX_OK;
    double *pd = (double *)ensure(c, c->adj_p, sizeof(double) * 3 * (size_t)(nprobe > 0 ? nprobe : 1), &st);
    if (st) return st;
    if (nprobe) VPMX_HIP(c, hipMemcpyAsync(pd, probes, sizeof(double) * 3 * nprobe, hipMemcpyHostToDevice, c->stream));
    const int saved_prec = c->precision;
    c->precision = 0;
    const int r = adjoint_rollout(c, nsteps, dt, uinf, ckpt, pd, nprobe, J);
    c->precision = saved_prec;
    VPMX_TRY(r);
    if (nvr) {
        VPMX_HIP(c, hipMemcpyAsync(dq, c->adj_qb.p, sizeof(double) * 3 * nvr, hipMemcpyDeviceToHost, c->stream));
        VPMX_HIP(c, hipMemcpyAsync(dx, c->adj_xb.p, sizeof(double) * 3 * nvr, hipMemcpyDeviceToHost, c->stream));
    }
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_kernel_times(vpmx_ctx *p, double *ms_total, int64_t *launches, int32_t n, int32_t reset) {
    Ctx *c = C(p);
    if (!c) return VPMX_ERR_ARG;
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    resolve_kernel_times(c);
    for (int k = 0; k < n && k < K_COUNT; k++) {
        if (ms_total) ms_total[k] = c->k_ms[k];
        if (launches) launches[k] = c->k_launches[k];
    }
    if (reset)
        for (int k = 0; k < K_COUNT; k++) {
            c->k_ms[k] = 0.0;
            c->k_launches[k] = 0;
        }
    return VPMX_OK;
}

int32_t vpmx_kernel_work(const vpmx_ctx *p, double *work, int32_t n) {
    if (!p || !work) return VPMX_ERR_ARG;
    const Ctx *c = C(p);
    for (int k = 0; k < n && k < K_COUNT; k++) work[k] = c->k_work[k];
    return VPMX_OK;
}

int32_t vpmx_stage_times(const vpmx_ctx *p, double *ms, int32_t n) {
    if (!p || !ms) return VPMX_ERR_ARG;
    const Ctx *c = C(p);
    for (int i = 0; i < n && i < ST_COUNT; i++) ms[i] = c->stage_ms[i];
    return VPMX_OK;
}

int32_t vpmx_stage_solve(vpmx_ctx *p, const double xbound[6], const double dpm[3], const int32_t nn[3], int32_t nc,
                         const int32_t win[6], const double *rhs, double *sol, int32_t zero_bc, double *out) {
    Ctx *c = C(p);
    if (!c || !xbound || !dpm || !nn || !win || !rhs) return VPMX_ERR_ARG;
    if (nc < 1 || nc > 4) return fail(c, VPMX_ERR_ARG, "nc must be 1..4");
    if (zero_bc ? !out : !sol) return fail(c, VPMX_ERR_ARG, "missing output array");
    for (int d = 0; d < 3; d++)
        if (win[d] < 1 || win[d + 3] > nn[d] || win[d + 3] - win[d] < 2) return fail(c, VPMX_ERR_ARG, "bad window");
    VPMX_HIP(c, hipSetDevice(c->device));
    int st = VPMX_OK;
    const int n[3] = {nn[0], nn[1], nn[2]};
    int64_t G = (int64_t)n[0] * n[1] * n[2];
    double *ra = (double *)ensure(c, c->stage_a, sizeof(double) * nc * (size_t)G, &st); if (st) return st;
    double *rs = (double *)ensure(c, c->stage_b, sizeof(double) * nc * (size_t)G, &st); if (st) return st;
    double *ss = (double *)ensure(c, c->stage_c, sizeof(double) * nc * (size_t)G, &st); if (st) return st;
    VPMX_HIP(c, hipMemcpyAsync(ra, rhs, sizeof(double) * nc * G, hipMemcpyHostToDevice, c->stream));
    VPMX_TRY(aos_to_soa(c, nc, n, ra, view(rs, n)));
    const double *src = zero_bc ? out : sol;
    VPMX_HIP(c, hipMemcpyAsync(ra, src, sizeof(double) * nc * G, hipMemcpyHostToDevice, c->stream));
    VPMX_TRY(aos_to_soa(c, nc, n, ra, view(ss, n)));
    int wn[3] = {win[3] - win[0] + 1, win[4] - win[1] + 1, win[5] - win[2] + 1};
    double h[3];
    window_h(xbound, dpm, win, h);
    GView rv = view(rs, wn, win[0] - 1, win[1] - 1, win[2] - 1, n);
    GView sv = view(ss, wn, win[0] - 1, win[1] - 1, win[2] - 1, n);
    VPMX_TRY(poisson_dirichlet(c, nc, wn, h, rv, zero_bc ? nullptr : &sv, sv));
    VPMX_TRY(soa_to_aos(c, nc, nc, n, view(ss, n), 0.0, ra));
    VPMX_HIP(c, hipMemcpyAsync(zero_bc ? out : sol, ra, sizeof(double) * nc * G, hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_stage_pmesh_inf(vpmx_ctx *p, const double xbound[6], const double dpm[3], const int32_t nn[3],
                             int32_t nc, const double *rhs, double *sol) {
    return vpmx_stage_pmesh_inf_lev(p, xbound, dpm, nn, nc, rhs, sol, 2, 0, 1);
}

int32_t vpmx_stage_pmesh_inf_lev(vpmx_ctx *p, const double xbound[6], const double dpm[3], const int32_t nn[3],
                                 int32_t nc, const double *rhs, double *sol, int32_t ibctyp, int32_t itree,
                                 int32_t levmax) {
    Ctx *c = C(p);
    if (!c || !xbound || !dpm || !nn || !rhs || !sol) return VPMX_ERR_ARG;
    if (nc != 3) return fail(c, VPMX_ERR_ARG, "nc must be 3");
    VPMX_HIP(c, hipSetDevice(c->device));
    int st = VPMX_OK;
    const int n[3] = {nn[0], nn[1], nn[2]};
    int64_t G = (int64_t)n[0] * n[1] * n[2];
    double *ra = (double *)ensure(c, c->stage_a, sizeof(double) * nc * (size_t)G, &st); if (st) return st;
    double *rs = (double *)ensure(c, c->stage_b, sizeof(double) * nc * (size_t)G, &st); if (st) return st;
    double *ss = (double *)ensure(c, c->stage_c, sizeof(double) * nc * (size_t)G, &st); if (st) return st;
    VPMX_HIP(c, hipMemcpyAsync(ra, rhs, sizeof(double) * nc * G, hipMemcpyHostToDevice, c->stream));
    VPMX_TRY(aos_to_soa(c, nc, n, ra, view(rs, n)));
    VPMX_HIP(c, hipMemcpyAsync(ra, sol, sizeof(double) * nc * G, hipMemcpyHostToDevice, c->stream));
    VPMX_TRY(aos_to_soa(c, nc, n, ra, view(ss, n)));
    Grid g = make_grid(nn, xbound, dpm);
    if (ibctyp != 2 && ibctyp != 3) return fail(c, VPMX_ERR_ARG, "ibctyp must be 2 or 3");
    VPMX_TRY(pmesh_inf(c, g, view(rs, n), view(ss, n), c->sol0, true, itree == 1 ? levmax : -1, ibctyp));
    VPMX_TRY(soa_to_aos(c, nc, nc, n, view(ss, n), 0.0, ra));
    VPMX_HIP(c, hipMemcpyAsync(sol, ra, sizeof(double) * nc * G, hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

int32_t vpmx_stage_project(vpmx_ctx *p, const double xbound[6], const double dpm[3], const int32_t nn[3],
                           const int32_t nn_bl[6], const double *xp, const double *qp, int64_t nvr, int32_t ncomp,
                           int32_t itype, int32_t idvpm, double *rhs) {
    Ctx *c = C(p);
    if (!c || !xbound || !dpm || !nn || !xp || !qp || !rhs) return VPMX_ERR_ARG;
    (void)nn_bl;
    if (ncomp != 4) return fail(c, VPMX_ERR_ARG, "ncomp must be 4 (3 strengths + volume)");
    if (itype < 2 || itype > 6) return fail(c, VPMX_ERR_ARG, "No such projection function");
    if (idvpm != 1) return fail(c, VPMX_ERR_UNSUP, "IDVPM=0 is not supported by this build");
    VPMX_HIP(c, hipSetDevice(c->device));
    int st = VPMX_OK;
    VPMX_TRY(upload_particles(c, xp, qp, nvr, ncomp - 1));
    const int n[3] = {nn[0], nn[1], nn[2]};
    int64_t G = (int64_t)n[0] * n[1] * n[2];
    double *gs = (double *)ensure(c, c->stage_b, sizeof(double) * 3 * (size_t)G, &st); if (st) return st;
    double *ga = (double *)ensure(c, c->stage_a, sizeof(double) * 4 * (size_t)G, &st); if (st) return st;
    Grid g = make_grid(nn, xbound, dpm);
    VPMX_TRY(sort_particles(c, g, (const double *)c->xp.p, nvr, c->sort));
    const double DV = dpm[0] * dpm[1] * dpm[2];
    int wlo[3] = {0, 0, 0};
    VPMX_TRY(p2m_gather(c, g, c->sort, (const double *)c->xp.p, (const double *)c->qp.p, ncomp, 3, itype, DV,
                        view(gs, n), n, wlo));
    VPMX_TRY(soa_to_aos(c, 3, 4, n, view(gs, n), DV, ga));
    VPMX_HIP(c, hipMemcpyAsync(rhs, ga, sizeof(double) * 4 * G, hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}

}  // extern "C"
