// Internal declarations shared by the HIP translation units of libvpmx.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/vpmx.h"

namespace vpmx {

// Node grid view of a component-planar (SoA) FP64 array:
// value(c, i, j, k) = p[c * cs + (k + oz) * pz + (j + oy) * py + (i + ox)], 0-based node indices.
struct GView {
    double *p;
    int64_t cs, py, pz;
    int ox, oy, oz;
    __host__ __device__ inline int64_t at(int i, int j, int k) const {
        return (int64_t)(k + oz) * pz + (int64_t)(j + oy) * py + (i + ox);
    }
};

// Work-group order for XCD locality: the hardware hands work-group b (linear id, x fastest) to
// XCD b % 8, so within each round of R work-groups (R a multiple of 8) the XCD of b takes the
// contiguous slice [(b % 8) R / 8, (b % 8 + 1) R / 8) of the round instead of every 8th index:
// neighbouring tiles then run at the same time on one XCD and share its L2 (halo values, the
// partial cache lines of their rows).  A bijection of [0, grid); the last partial round keeps
// launch order.
__device__ __forceinline__ int64_t xcd_slice(int64_t b, int64_t grid, int64_t R) {
    const int64_t base = b / R * R;
    if ((R & 7) || base + R > grid) return b;
    const int64_t o = b - base;
    return base + (o & 7) * (R >> 3) + (o >> 3);
}

// Geometry of a uniform node grid.
struct Grid {
    int n[3];
    double x0[3];
    double h[3];
};

// Device scratch that only grows.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Ctx;

// ---------------------------------------------------------------- error helpers
int fail(Ctx *c, int code, const std::string &msg);
#define VPMX_HIP(c, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::vpmx::fail((c), VPMX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define VPMX_TRY(expr)           \
    do {                         \
        int s_ = (expr);         \
        if (s_ != VPMX_OK) return s_; \
    } while (0)

void *ensure(Ctx *c, DBuf &b, size_t bytes, int *st);

// ---------------------------------------------------------------- launchers
// Particles / projection (k_particles.hip)
struct SortState {
    DBuf keys, perm, count, start, fill;
    DBuf xs, qs;  // particle positions / strengths in cell order (projection)
    DBuf rowcnt;  // particles per cell row (projection)
    int64_t ncells = 0;
    int cn[3] = {0, 0, 0};  // padded cell grid dims
    bool count_zero = false;  // every word of count is zero (a completed sort leaves it so)
};
// Sort particles (xp AoS 3xn) into cells of grid g; cells are node indices inode in [-1, n+1].
int sort_particles(Ctx *c, const Grid &g, const double *xp, int64_t n, SortState &s);
// Gather-form M-type projection: out(view over nodes of `g`, covering node window
// [win0, win1) in each dim) = sum_p W * qp(c, p) / div for c < ncomp; nodes of the view
// outside the window are written 0.
// f32: weights and sums in FP32 (mixed form), the grid stays FP64.
// zero_empty false: the node-tile path leaves the nodes of tiles that no particle reaches
// unwritten and keeps the tile map (Ctx::tile_map) for a caller that reads only the rest.
// keep_zeros: nothing but this call writes `out` between such calls (the block RHS), so the
// tiles left empty by the last one are still zero and only newly emptied tiles are cleared.
int p2m_gather(Ctx *c, const Grid &g, const SortState &s, const double *xp, const double *qp, int qstride,
               int ncomp, int itype, double div, GView out, const int vn[3], const int wlo[3], int kclass = -1,
               bool f32 = false, bool zero_empty = true, bool keep_zeros = false);
// itypeb 1: UP = sum W vel, GP = (sum W sol) QP(4); itypeb 2 (diffusion): UP = 0, GP as above
int m2p_gather(Ctx *c, const Grid &g, const SortState &s, const double *xp, const double *qp, int qstride,
               int itype, GView vel, GView sol, double *up, double *gp, int64_t n, int itypeb = 1);
int euler_update(Ctx *c, double *xp, double *qp, int qstride, const double *up, const double *gp, int64_t n,
                 double dt, const double uinf[3], const double *x0, const double *q0);
int particle_bbox(Ctx *c, const double *xp, int64_t n, double mn[3], double mx[3]);
// One RK2 step of the context's resident particles queued on its stream (no drain at the end).
int step_rk2_enqueue(Ctx *c, double dt, const double uinf[3], int ntime, int mrem, int do_remesh);
int64_t remesh_window_nodes(const Ctx *c, int mrem);

// Address of the fixed device array of the allocation-free entry points on the context's
// device (once per device), and diagnostics with its partials there (device output out[6]).
int scratch_init(Ctx *c);
int diagnostics_noalloc(Ctx *c, const double *xp, const double *qp, int qs, int64_t n, double *out);
int diagnostics(Ctx *c, const double *xp, const double *qp, int qstride, int64_t n, double circ[3], double imp[3]);
int remesh_compact(Ctx *c, GView w, const int vn[3], const int lo[3], const int hi[3], double thr,
                   const double x0[3], const double h[3], double dv, int neqpm, double **xp_out, double **qp_out,
                   DBuf &bx, DBuf &bq, int64_t *count);

// Poisson (k_poisson.hip)
// Dirichlet 7-point solve on a window of nodes [0, wn) of the views (window = whole view
// extent described by wn).  rhs: RHS view; bc: view holding Dirichlet values on the
// window faces (nullptr for zero BCs); out: view receiving the solution (interior; the
// faces are written with the BC values or zero).  h: node spacings.
// reuse (two solves of one window with the same RHS, the second with Dirichlet values):
// 1 keeps the work array after the x and y forward passes, 2 starts from the kept array with
// the transform of the folded Dirichlet terms taken off instead of running those passes
// (falls back to the full solve when the kept array does not match the window).
// form: the arithmetic of the passes -- SOLVE_CTX the context's precision (FP32 passes in the
// mixed form), SOLVE_FP64, SOLVE_FP32, SOLVE_INV64 (forward passes and diagonal in FP32, the
// inverse passes in FP64: the accuracy of an FP64 solve where second differences are taken).
enum SolveForm { SOLVE_CTX = -1, SOLVE_FP64 = 0, SOLVE_FP32 = 1, SOLVE_INV64 = 2 };
// face_nd (zero BCs only): instead of the solution, the one-sided normal derivative at every
// node of every face into that table (the layout of normal_derivs_table), from the spectrum.
// part: only the nodes on the planes x = xs[0], xs[1], y = ys[...], z = zs[...] (interior
// nodes; equal margins in x and y: xs[1] = nx + 1 - xs[0]) and the interior nodes 1 + a step
// (a >= 1) in all three axes are written; the rest of the interior keeps its old values.
struct SolvePartial {
    int xs[2], ys[2], zs[2];
    int step;
};
int poisson_dirichlet(Ctx *c, int ncomp, const int wn[3], const double h[3], GView rhs, const GView *bc,
                      GView out, int reuse = 0, int form = SOLVE_CTX, double *face_nd = nullptr,
                      const SolvePartial *part = nullptr);

// Complex FFT plan of length L (Stockham radices, device twiddle table exp(-2 pi i t / L)),
// cached per context.
struct FftShape {
    int L, nstages, radix[24];
    const double2 *tw;
};
int fft_shape(Ctx *c, int L, FftShape *out);

// Boundary (k_boundary.hip)
int normal_derivs(Ctx *c, int ncomp, const Grid &g, GView s0, double *sig, double *cxyz, double *wgt,
                  int64_t *nbound);
// The same in two steps: the face-node table (ncomp values per node of each face, faces XMIN,
// XMAX, YMIN, YMAX, ZMIN, ZMAX, each ncomp x N0 x N1 with (N0, N1) = (NY, NZ), (NX, NZ),
// (NX, NY), first index fastest; the context keeps its buffer), filled by the caller, then the
// per-cell sources from it.
double *normal_derivs_table(Ctx *c, int ncomp, const int n[3], int *st);
int normal_derivs_cells(Ctx *c, int ncomp, const Grid &g, const double *fd, double *sig, double *cxyz, double *wgt,
                        int64_t *nbound);
// levmax >= 0: the tree form with that many levels (pm.inp tree = 1), else the direct sums;
// ibctyp 2 point sources, 3 constant-source panels (needs `scratch`, a grid of sol's shape
// whose boundary nodes it may overwrite, for the direct form).
int boundary_sums(Ctx *c, int ncomp, const Grid &g, int64_t nbound, const double *sig, const double *cxyz,
                  const double *wgt, GView sol, bool fine, int levmax = -1, int ibctyp = 2,
                  const GView *scratch = nullptr);
// SOL(t) += sum_s a(c, s) / |x_t - x_s| over the face-cell centres (spectral / lattice /
// direct evaluation by the context's setting).
int point_sums(Ctx *c, const Grid &g, int64_t nbound, const double *cxyz, const double *a, GView sol, bool fine);
// Panel sources (k_panel.hip), direct or tree form.
int boundary_sums_panels(Ctx *c, const Grid &g, int64_t nb, const double *sig, const double *cxyz, const double *ds,
                         GView sol, bool fine, int levmax, GView scratch);
// Spectral evaluation of the same sums (k_bspec.hip); returns VPMX_ERR_UNSUP when the
// block's faces are too long for its in-LDS FFTs, so the caller can use the lattice kernel.
// slot: 0 the fine block, 1 the coarse grid (each keeps its own plan across calls).
int boundary_sums_spectral(Ctx *c, const Grid &g, int64_t nbound, const double *a, GView sol, int slot);
// Its transpose: abar(c, s) = sum over boundary nodes t of gbar(c, t) / |x_t - x_s| (the adjoint
// of the sources a); VPMX_ERR_UNSUP where the forward plan does not build.
int boundary_sums_spectral_adj(Ctx *c, const Grid &g, int64_t nbound, GView gb, double *abar, int slot);
void free_bspec(Ctx *c);
// Column-major C_b = op(A_b) op(B_b) for b < batch (FP64 on the matrix cores, k_dgemm.hip; the
// tile shape is a tuner choice, every shape gives the same bits).
int dgemm(Ctx *c, bool ta, bool tb, int M, int N, int K, const double *A, int lda, int64_t sa, const double *B,
          int ldb, int64_t sb, double *C, int ldc, int64_t sc, int batch);
// C_b = A_b op(B_b).
int dgemm_batched(Ctx *c, bool transb, int M, int N, int K, const double *A, int lda, int64_t sa, const double *B,
                  int ldb, int64_t sb, double *C, int ldc, int64_t sc, int batch);
// Tree form (k_tree.hip): sig (3 x nbound), cxyz (3 x nbound), ds (nbound) of normal_derivs.
// panels: tree_calc_3d_s (ibctyp = 3) instead of tree_calc_3d.
int boundary_sums_tree(Ctx *c, const Grid &g, int64_t nbound, const double *sig, const double *cxyz, const double *ds,
                       GView sol, int levmax, bool fine, bool panels);
int64_t tree_entries(Ctx *c, const int n[3], int levmax);
void free_tree(Ctx *c);

// Grid kernels (k_grid.hip)
int fill(Ctx *c, double *p, int64_t n, double v);
int zero_faces(Ctx *c, const int n[3], GView v, int ncomp = 3);  // the six faces of a grid (k_poisson.hip)
int curl_kernel(Ctx *c, const int n[3], const double h[3], GView psi, GView vel);
// curl and stretching in one pass (the velocity grid written too); same values as the two kernels
int curl_stretch_kernel(Ctx *c, const int n[3], const double h[3], GView psi, GView rhs, GView vel, GView out);
int stretch_kernel(Ctx *c, const int n[3], const double h[3], GView rhs, GView vel, GView out);
int diffuse_kernel(Ctx *c, const int n[3], const double h[3], double vis, GView rhs, GView out);
int coarse_sample(Ctx *c, int ncomp, const int nsamp[3], int nc, const int *cmap_dev, GView fine, GView coarse,
                  const int cn[3]);
int coarse_laplacian(Ctx *c, int ncomp, const int cn[3], const int lo[3], const int hi[3], const double hc[3],
                     GView samp, GView rhs, bool accumulate = false);
struct InterpArgs {
    int nbl[6];           // inner box NNbl_bl (1-based)
    double xbl[3], h[3];  // block origin and fine spacing
    double xc[3], hc[3];  // coarse origin, spacing
    int cn[3];            // coarse dims
    int fp_lo[3], fp_hi[3];  // footprint of the block on the coarse grid (1-based coarse index)
};
int face_interp(Ctx *c, int ncomp, const InterpArgs &a, GView coarse_sol, GView samp, GView sol_bl);
// Several yaps blocks, of equal or uneven sizes (all blocks' arrays back to back, block b's
// from MbBlock::off).
struct MbBlock {
    int64_t off;             // first double of the block's arrays in the blocks' buffers (3 x its nodes)
    int64_t f0;              // first of its inner-face nodes in the list over all blocks
    int nbl[6];              // inner box NNbl_bl (1-based)
    int n[3];                // block grid NNbl
    int fp_lo[3], fp_hi[3];  // sampled footprint on the coarse grid (1-based, inclusive)
    int bi, bj, bk;          // block position (1-based)
    double xbl[3];           // block origin
};
struct MbArgs {
    double h[3], xc[3], hc[3];
    int cn[3];
    int nbi, nbj, nbk, nblocks;
    int64_t sstride;  // doubles per block in the sample array (3 x coarse nodes)
    int64_t ntot;     // inner-face nodes of all blocks
};
int face_interp_blocks(Ctx *c, const MbArgs &a, const MbBlock *blk_dev, GView csol, const double *samp, double *sol,
                       double *tmp);
struct Box6 {
    int v[6];
};
// rhsscat_3d / solget_3d for one block: n block dims, nbl inner box, g0 global node of the
// inner box's first node (1-based), zmin min faces to zero (block not first on that axis).
int rhsscat_block(Ctx *c, const int n[3], const int nbl[6], const int g0[3], const int zmin[3], GView glob,
                  GView blk);
int solget_block(Ctx *c, const int nbl[6], const int g0[3], GView blk, GView glob);
int soa_to_aos(Ctx *c, int ncomp_src, int ncomp_dst, const int n[3], GView src, double fill_last, double *dst);
int aos_to_soa(Ctx *c, int ncomp, const int n[3], const double *src, GView dst);

// Adjoint stages (k_adjoint.hip): transposes of the curl, the fold of Dirichlet values, the
// boundary sums, the normal derivatives, the TSC face correction, the coarse Laplacian and
// sampling; probe interpolation and its transpose; the particles' gradients.
int curl_adj(Ctx *c, const int n[3], const double h[3], GView ub, GView pb);
int fold_adj(Ctx *c, const int n[3], const double h[3], GView pb, GView y, GView out);
int bsum_adj(Ctx *c, const Grid &g, int64_t ns, const double *cxyz, GView gb, double *abar);
// (writes every node of the grid g in s0b)
int nderiv_adj(Ctx *c, const Grid &g, int64_t nb, const double *abar, const double *ds, GView s0b);
int face_interp_adj(Ctx *c, const InterpArgs &a, GView fb, GView phib, GView cb, GView sb);
int coarse_laplacian_adj(Ctx *c, const int cn[3], const int lo[3], const int hi[3], const double hc[3], GView rb,
                         GView sb);
int coarse_sample_adj(Ctx *c, const int ns[3], int nc, const int *cmap, GView coarse, GView fine);
int axpy_view(Ctx *c, const int n[3], GView in, GView out);
// dq = (1/DV) P^T wb, dx = its position derivative (acc: added to dq, dx instead)
int grad_particles(Ctx *c, const Grid &g, const double *xp, const double *qp, int qs, int64_t n, GView wb, double idv,
                   double *dq, double *dx, bool acc = false);
// M2P^T: ubar_p = au*us_p, gbar_p = ag*gs_p scattered to ubg / sbg (vol_p GP weight), xacc += dx
int m2p_adj(Ctx *c, const Grid &g, const double *xp, const double *qp, int qs, int64_t n, const double *us, double au,
            const double *gs, double ag, GView vel, GView str, GView ubg, GView sbg, double *xacc);
// stretching^T: rb = S_R^T sb (written), ub += S_u^T sb
int stretch_adj(Ctx *c, const int n[3], const double h[3], GView R, GView u, GView sb, GView rb, GView ub);
// remesh^T: node particles' strength adjoints added to their nodes
int scatter_nodes(Ctx *c, const Grid &g, const double *xp, const double *qb, int64_t n, GView out);
int axpy(Ctx *c, int64_t n, double a, const double *x, double *y);
int probe_interp(Ctx *c, const Grid &g, const double *pts, int64_t np, GView v, double *u);
int probe_scatter(Ctx *c, const Grid &g, const double *pts, int64_t np, const double *ub, GView out);

}  // namespace vpmx
