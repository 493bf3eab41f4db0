// Particle-side kernels: cell sort, gather-form projection (P2M), interpolation (M2P),
// particle updates, reductions and the remesh compaction.
//
// P2M follows project_particles_3D + project_vol3d (pmproject.f90:43-153) but is
// reorganised as a deterministic gather: particles are counting-sorted by cell (ties
// kept in particle order), and every node sums the particles of the 4x4x4 cells whose
// stencil covers it, so no floating-point atomics are used and the result does not
// depend on scheduling.
#include <algorithm>
#include <cstdio>
#include <mutex>

#include <hip/hip_ext.h>

#include "tune.h"
#include "vpmx_internal.h"
#include "vpmx_ctx.h"

namespace vpmx {

// Projection functions of projection_fun (pmproject.f90:391-431).
__device__ __forceinline__ double proj_fun(int itype, double x) {
#pragma clang fp contract(off)
    double a = fabs(x), r = 0.0;
    if (itype == 4) {
        if (a < 1.0) r = 1.0 - 2.5 * (a * a) + 1.5 * (a * a * a);
        else if (a <= 2.0) r = 0.5 * ((2.0 - a) * (2.0 - a)) * (1.0 - a);
    } else if (itype == 3) {
        if (a < 0.5) r = 0.5 * ((a + 1.5) * (a + 1.5)) - 1.5 * ((a + 0.5) * (a + 0.5));
        else if (a <= 1.5) r = 0.5 * ((-a + 1.5) * (-a + 1.5));
    } else if (itype == 2) {
        r = (a > 1.0) ? 0.0 : 1.0 - a;
    } else if (itype == 5) {
        if (a <= 0.5) r = 1.0 - a * a;
        else if (a <= 1.5) r = 0.5 * (1.0 - a) * (2.0 - a);
    } else if (itype == 6) {
        if (a <= 1.0) r = 1.0 - a * a;
        else if (a <= 2.0) r = 0.5 * (1.0 - a) * (2.0 - a);
    }
    return r;
}

// Node offset of the stencil relative to the particle's cell: nodes inode-ips .. inode+ipf.
__host__ __device__ inline int stencil_ips(int itype) { return itype == 2 ? 0 : 1; }
__host__ __device__ inline int stencil_ipf(int itype) { return itype == 2 ? 1 : 2; }

// ------------------------------------------------------------------ scans
// Three-phase exclusive scan of int32 counts (n entries) into out (n+1 entries).
// Exclusive prefix sum out[0..n) of in[0..n), out[n] = total, in three passes over tiles of
// SCAN_TILE inputs: the tile sums, their scan (one work-group), then every tile's scan plus
// its offset.  Reads the input twice and writes the output once; in == out is allowed.
constexpr int SCAN_THREADS = 256, SCAN_PER = 16, SCAN_TILE = SCAN_THREADS * SCAN_PER;

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}
// Exclusive scan of one value per thread over the work-group (at most 16 waves); *total gets
// the sum.  Contains work-group barriers: every thread must call it.
__device__ __forceinline__ int block_excl_scan(int v, int *total) {
    __shared__ int ws[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int inc = wave_incl_scan(v, lane);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < nw; q++) {
        const int x = ws[q];
        off += q < w ? x : 0;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const int *in, int64_t n, int *bsum) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const int64_t i = base + q * SCAN_THREADS + threadIdx.x;
        s += i < n ? in[i] : 0;
    }
    int tot;
    (void)block_excl_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}
// bsum[0..nb) -> exclusive scan in place, bsum[nb] = total; one work-group of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan_tiles(int *bsum, int64_t nb) {
    const int64_t per = (nb + 1023) / 1024, lo = threadIdx.x * per, hi = lo + per < nb ? lo + per : nb;
    int s = 0;
    for (int64_t i = lo; i < hi; i++) s += bsum[i];
    int tot;
    int run = block_excl_scan(s, &tot);
    for (int64_t i = lo; i < hi; i++) {
        const int v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 0) bsum[nb] = tot;
}
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_out(const int *in, int *out, int64_t n, const int *bsum,
                                                           int64_t nb) {
    __shared__ int tile[SCAN_TILE + SCAN_TILE / 16];  // one pad slot per 16: thread rows are conflict-free
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const int i = q * SCAN_THREADS + t;
        const int64_t g = base + i;
        tile[i + (i >> 4)] = g < n ? in[g] : 0;
    }
    __syncthreads();
    int v[SCAN_PER], s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_PER; r++) {
        v[r] = tile[t * 17 + r];
        s += v[r];
    }
    int tot;
    int run = block_excl_scan(s, &tot) + bsum[blockIdx.x];
#pragma unroll
    for (int r = 0; r < SCAN_PER; r++) {
        tile[t * 17 + r] = run;
        run += v[r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const int i = q * SCAN_THREADS + t;
        const int64_t g = base + i;
        if (g < n) out[g] = tile[i + (i >> 4)];
    }
    if (blockIdx.x == 0 && t == 0) out[n] = bsum[nb];
}

__global__ void k_zero_int(int *p, int64_t n);

// Particle count of a launch: n, or the count the device holds at *dn (at most n, the
// capacity the grid was sized for) in the device-count mode of captured steps.
__device__ __forceinline__ int64_t pcount(int64_t n, const int *dn) {
    if (!dn) return n;
    const int64_t d = *dn;
    return d < n ? d : n;
}

// Arrays of up to SCAN_SMALL entries in one work-group of 1024 threads (16 consecutive entries
// per thread): one launch instead of three, which is what the small grids of batched clouds
// need.
constexpr int64_t SCAN_SMALL = 1024 * 16;
__global__ void __launch_bounds__(1024) k_scan_small(const int *in, int *out, int64_t n) {
    const int64_t lo = (int64_t)threadIdx.x * 16;
    int v[16], s = 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        v[r] = lo + r < n ? in[lo + r] : 0;
        s += v[r];
    }
    int tot;
    int run = block_excl_scan(s, &tot);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        if (lo + r < n) out[lo + r] = run;
        run += v[r];
    }
    if (threadIdx.x == 0) out[n] = tot;
}

int exclusive_scan(Ctx *c, const int *in, int *out, int64_t n) {
    if (n <= 0) {
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, out, (int64_t)1);
        VPMX_HIP(c, hipGetLastError());
        return VPMX_OK;
    }
    if (n <= SCAN_SMALL) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, c->stream, in, out, n);
        VPMX_HIP(c, hipGetLastError());
        return VPMX_OK;
    }
    int st = VPMX_OK;
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    int *bsum = (int *)ensure(c, c->scan_tmp, sizeof(int) * (size_t)(nb + 2), &st);
    if (st) return st;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, c->stream, in, n, bsum);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, c->stream, bsum, nb);
    hipLaunchKernelGGL(k_scan_out, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, c->stream, in, out, n, bsum, nb);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// ------------------------------------------------------------------ cell sort
__global__ void k_zero_int(int *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Cell of a particle: inode = int((x - X0)/DX) + 1 (pmproject.f90:59-61), kept when
// inode is in [-1, n+1] in every direction (any node of its stencil can be on the grid).
__global__ void k_cell_keys(Grid g, const double *xp, int64_t n, const int *dn, int cnx, int cny, int cnz, int *keys,
                            int *count) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pcount(n, dn)) return;
    int ci[3];
    bool ok = true;
    for (int d = 0; d < 3; d++) {
        double t = (xp[3 * p + d] - g.x0[d]) / g.h[d];
        if (!(t > -3.0 && t < (double)g.n[d] + 3.0)) { ok = false; ci[d] = 0; continue; }
        int inode = (int)t + 1;
        ci[d] = inode + 1;
        if (ci[d] < 0 || ci[d] > g.n[d] + 2) ok = false;
    }
    int key = -1;
    if (ok) {
        key = ci[0] + cnx * (ci[1] + cny * ci[2]);
        atomicAdd(&count[key], 1);
    }
    keys[p] = key;
}
// Each particle takes a slot of its cell by counting the cell's count back down, so the counts
// are all zero again when the scatter has run (the next sort needs no clearing pass).
__global__ void k_cell_scatter(const int *keys, int64_t n, const int *dn, const int *start, int *count, int *perm) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pcount(n, dn)) return;
    int key = keys[p];
    if (key < 0) return;
    int pos = start[key] + atomicSub(&count[key], 1) - 1;
    perm[pos] = (int)p;
}
// Order the particles of each cell by particle index (the scatter above is unordered): one
// thread per sorted slot, the slot that opens a cell of two or more particles sorts it.
__global__ void k_cell_order(const int *start, const int *keys, int64_t ncells, int *perm) {
    const int64_t sl = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= (int64_t)start[ncells]) return;
    const int cidx = keys[perm[sl]];
    const int s = start[cidx];
    const int cnt = start[cidx + 1] - s;
    if (s != sl || cnt < 2) return;
    for (int a = 1; a < cnt; a++) {
        int v = perm[s + a];
        int b = a - 1;
        while (b >= 0 && perm[s + b] > v) {
            perm[s + b + 1] = perm[s + b];
            b--;
        }
        perm[s + b + 1] = v;
    }
}

int sort_particles(Ctx *c, const Grid &g, const double *xp, int64_t n, SortState &s) {
    int st = VPMX_OK;
    c->sort_n = n;
    s.cn[0] = g.n[0] + 3; s.cn[1] = g.n[1] + 3; s.cn[2] = g.n[2] + 3;
    s.ncells = (int64_t)s.cn[0] * s.cn[1] * s.cn[2];
    int *keys = (int *)ensure(c, s.keys, sizeof(int) * (size_t)(n > 0 ? n : 1), &st); if (st) return st;
    int *perm = (int *)ensure(c, s.perm, sizeof(int) * (size_t)(n > 0 ? n : 1), &st); if (st) return st;
    const void *was = s.count.p;
    int *count = (int *)ensure(c, s.count, sizeof(int) * (size_t)s.ncells, &st); if (st) return st;
    int *start = (int *)ensure(c, s.start, sizeof(int) * (size_t)(s.ncells + 1), &st); if (st) return st;
    unsigned gp = (unsigned)((n + 255) / 256);
    // the counts are zero after every completed sort (the scatter counts them back down); a new
    // buffer, or a sort that stopped half way, clears them first
    if (count != was || !s.count_zero) {
        const int64_t words = (int64_t)(s.count.bytes / sizeof(int));
        hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, c->stream, count, words);
    }
    s.count_zero = false;
    if (n > 0)
        hipLaunchKernelGGL(k_cell_keys, dim3(gp), dim3(256), 0, c->stream, g, xp, n, c->dn, s.cn[0], s.cn[1], s.cn[2],
                           keys, count);
    VPMX_TRY(exclusive_scan(c, count, start, s.ncells));
    if (n > 0) {
        hipLaunchKernelGGL(k_cell_scatter, dim3(gp), dim3(256), 0, c->stream, keys, n, c->dn, start, count, perm);
        hipLaunchKernelGGL(k_cell_order, dim3(gp), dim3(256), 0, c->stream, start, keys, s.ncells, perm);
    }
    VPMX_HIP(c, hipGetLastError());
    s.count_zero = true;
    return VPMX_OK;
}

// ------------------------------------------------------------------ P2M gather
// One thread per view node.  View node (a,b,c) is grid node (i,j,k) = (a-wlo0+1, ...)
// (1-based); nodes outside the grid window are written 0 (the block halo of rhsscat_3d).
template <int NC>
__global__ void __launch_bounds__(256) k_p2m_gather(Grid g, const int *start, const int *perm, int cnx, int cny,
                                                     const double *xp, const double *qp, int qstride, int itype,
                                                     double div, GView out, int vnx, int vny, int vnz, int wlo0,
                                                     int wlo1, int wlo2) {
#pragma clang fp contract(off)
    int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nv = (int64_t)vnx * vny * vnz;
    if (tid >= nv) return;
    int a = (int)(tid % vnx);
    int b = (int)((tid / vnx) % vny);
    int cc = (int)(tid / ((int64_t)vnx * vny));
    int i = a - wlo0 + 1, j = b - wlo1 + 1, k = cc - wlo2 + 1;
    double acc[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) acc[q] = 0.0;
    if (i >= 1 && i <= g.n[0] && j >= 1 && j <= g.n[1] && k >= 1 && k <= g.n[2]) {
        const int ips = stencil_ips(itype), ipf = stencil_ipf(itype);
        const double X = g.x0[0] + 0.0, Y = g.x0[1], Z = g.x0[2];
        // node coordinate offsets (i-1)*DX as in x = (XP - XMIN - (i-1)*DX)/DX
        const double ox = (double)(i - 1) * g.h[0], oy = (double)(j - 1) * g.h[1], oz = (double)(k - 1) * g.h[2];
        for (int ck = k - ipf; ck <= k + ips; ck++) {
            for (int cj = j - ipf; cj <= j + ips; cj++) {
                int64_t row = (int64_t)cnx * ((cj + 1) + (int64_t)cny * (ck + 1));
                int p0 = start[row + (i - ipf + 1)];
                int p1 = start[row + (i + ips + 1) + 1];
                for (int s = p0; s < p1; s++) {
                    int p = perm[s];
                    double px = xp[3 * (int64_t)p], py = xp[3 * (int64_t)p + 1], pz = xp[3 * (int64_t)p + 2];
                    double fx = proj_fun(itype, (px - X - ox) / g.h[0]);
                    double fy = proj_fun(itype, (py - Y - oy) / g.h[1]);
                    double fz = proj_fun(itype, (pz - Z - oz) / g.h[2]);
                    double f = fx * fy * fz;
                    const double *q = qp + (int64_t)qstride * p;
#pragma unroll
                    for (int r = 0; r < NC; r++) acc[r] = acc[r] + f * q[r];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < NC; r++) acc[r] = acc[r] / div;
    }
    int64_t o = out.at(a, b, cc);
#pragma unroll
    for (int r = 0; r < NC; r++) out.p[r * out.cs + o] = acc[r];
}

// M4' (itype 4) weight, the expression of proj_fun.
__device__ __forceinline__ double m4p(double x) {
#pragma clang fp contract(off)
    const double a = fabs(x);
    double r = 0.0;
    if (a < 1.0) r = 1.0 - 2.5 * (a * a) + 1.5 * (a * a * a);
    else if (a <= 2.0) r = 0.5 * ((2.0 - a) * (2.0 - a)) * (1.0 - a);
    return r;
}

template <typename T>
__device__ __forceinline__ T m4p_t(T x) {
#pragma clang fp contract(off)
    const T a = fabs(x);
    T r = T(0);
    if (a < T(1)) r = T(1) - T(2.5) * (a * a) + T(1.5) * (a * a * a);
    else if (a <= T(2)) r = T(0.5) * ((T(2) - a) * (T(2) - a)) * (T(1) - a);
    return r;
}

// Particles per cell row (the start table is too large for the caches; the row counts stay
// in L2 and let the projection skip the start loads of empty rows).
__global__ void k_row_counts(const int *__restrict__ start, int cnx, int64_t nrows, int *__restrict__ rowcnt) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) rowcnt[r] = start[(r + 1) * cnx] - start[r * cnx];
}

// M4' projection on node tiles.  A work-group takes 16 x 4 x 4 nodes (one per thread); the
// particles that reach them lie in 19 x 7 x 7 cells (the particles of cell c are visited by
// nodes c-1..c+2), 49 cell rows whose particles are contiguous in cell order.  The rows'
// particles are staged in LDS, each with the M4' weights of the 4 x 4 x 4 nodes that visit
// it (W((x - x0)/h - (i - 1)) for node i), formed once per tile, in chunks of whole rows that
// fit; every thread then sums its node's 16 rows, the four cells of each, rows z-outer,
// y-inner, cells and particles in cell order: deterministic, no atomics.  A row with more particles than a chunk holds is read straight from memory in the
// same order.  Tiles without particles are written by k_p2m_zero_tiles.
// Round of the one-work-group-per-tile launches (xcd_slice): 256 work-groups per XCD.
constexpr int64_t XCD_ROUND = 2048;

constexpr int PT_X = 16, PT_Y = 4, PT_Z = 4;
constexpr int PC_X = PT_X + 3, PC_Y = PT_Y + 3, PC_Z = PT_Z + 3;
constexpr int PT_ROWS = PC_Y * PC_Z;
// The staged particles in LDS, one array per field (neighbouring lanes read neighbouring
// particles, so a field read is a unit-stride access across the wave): w[4 d + q][p] the
// weight of the q-th node that visits particle p along axis d (nodes c_d - 1 .. c_d + 2 for
// the particle's cell c_d in the cell order), q[c][p] the strengths.
template <class T, int MAXP>
struct P2mStage {
    T w[12][MAXP];
    T q[3][MAXP];
};

// Cell-ordered particle records for the tiled projection: for sorted particle s (particle
// perm[s]) R[16 s + 4 d + q] = W(t_d - (c_d + q - 2)), the M4' weight of the q-th node that
// visits it along axis d (node c_d - 1 + q, c_d its cell as the sort computed it,
// int((x - x0)/h) + 1, t_d = (x - x0)/h), and R[16 s + 12 + c] its strengths: 16-word records,
// written and read with vector accesses.  Formed once per projection instead of once per tile
// that stages the particle.
template <class T>
__global__ void k_p2m_sorted(const int *__restrict__ perm, const int *__restrict__ total, int64_t n, Grid g,
                             const double *__restrict__ xp, const double *__restrict__ qp, int qstride,
                             T *__restrict__ R) {
#pragma clang fp contract(off)
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || s >= (int64_t)*total) return;
    const int64_t p = perm[s];
    T r[16];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double x = xp[3 * p + d];
        const int cell = (int)((x - g.x0[d]) / g.h[d]) + 1;
        const double t = (x - g.x0[d]) * (1.0 / g.h[d]);
#pragma unroll
        for (int q = 0; q < 4; q++) r[4 * d + q] = m4p_t<T>((T)(t - (double)(cell + q - 2)));
        r[12 + d] = (T)qp[(int64_t)qstride * p + d];
    }
    r[15] = T(0);
    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int NV = 16 * (int)sizeof(T) / (int)sizeof(V);
    V *dst = (V *)(R + 16 * s);
#pragma unroll
    for (int v = 0; v < NV; v++) {
        V x;
        __builtin_memcpy(&x, r + v * (int)(sizeof(V) / sizeof(T)), sizeof(V));
        dst[v] = x;
    }
}

// Tiles whose 49 cell rows hold any particle (a superset of the tiles with work); thread per
// tile, so the row-count loads of many tiles are in flight together.
__global__ void k_p2m_tile_flags(int ntiles, int ntx, int nty, int vnx, int vny, int vnz, Grid g, int cny,
                                 const int *__restrict__ rowcnt, int wlo1, int wlo2, int *__restrict__ flag) {
    const int tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= ntiles) return;
    const int ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int J0 = ty * PT_Y - wlo1 + 1, K0 = tz * PT_Z - wlo2 + 1;
    int any = 0;
    for (int r = 0; r < PT_ROWS; r++) {
        const int cj = J0 - 2 + r % PC_Y, ck = K0 - 2 + r / PC_Y;
        if (cj >= -1 && cj <= g.n[1] + 1 && ck >= -1 && ck <= g.n[2] + 1)
            any |= rowcnt[(cj + 1) + (int64_t)cny * (ck + 1)] > 0;
    }
    flag[tile] = any;
}

// Zeros for the nodes of the view in tiles without particles (one thread per node).
__global__ void k_p2m_zero_tiles(int ntx, int nty, int vnx, int vny, int vnz, const int *__restrict__ flag, GView out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)vnx * vny * vnz) return;
    const int va = (int)(t % vnx), vb = (int)((t / vnx) % vny), vc = (int)(t / ((int64_t)vnx * vny));
    const int tile = va / PT_X + ntx * (vb / PT_Y + nty * (vc / PT_Z));
    if (flag[tile]) return;
    const int64_t o = out.at(va, vb, vc);
#pragma unroll
    for (int r = 0; r < 3; r++) out.p[r * out.cs + o] = 0.0;
}

// The same for the tiles that held particles in the previous projection and hold none now
// (the other empty tiles are still zero): a work-group walks ZT_TILES tiles, one thread per
// node of a tile (PT_X x PT_Y x PT_Z = 256), and skips the tiles with nothing to clear.
constexpr int ZT_TILES = 16;
static_assert(PT_X * PT_Y * PT_Z == 256, "one thread per node of a tile");
__global__ void __launch_bounds__(256) k_p2m_zero_changed(int ntiles, int ntx, int nty, int vnx, int vny, int vnz,
                                                          const int *__restrict__ flag, const int *__restrict__ prev,
                                                          GView out) {
    const int xa = threadIdx.x % PT_X, ya = (threadIdx.x / PT_X) % PT_Y, za = threadIdx.x / (PT_X * PT_Y);
    for (int u = 0; u < ZT_TILES; u++) {
        const int tile = blockIdx.x * ZT_TILES + u;
        if (tile >= ntiles) return;
        if (flag[tile] || !prev[tile]) continue;
        const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
        const int va = tx * PT_X + xa, vb = ty * PT_Y + ya, vc = tz * PT_Z + za;
        if (va >= vnx || vb >= vny || vc >= vnz) continue;
        const int64_t o = out.at(va, vb, vc);
#pragma unroll
        for (int r = 0; r < 3; r++) out.p[r * out.cs + o] = 0.0;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_p2m_tile(Grid g, const int *__restrict__ start,
                                                  const int *__restrict__ rowcnt, int cnx, int cny,
                                                  const T *__restrict__ R, double div, GView out, int vnx,
                                                  int vny, int vnz, int wlo0,
                                                  int wlo1, int wlo2, const int *__restrict__ list,
                                                  const int *__restrict__ nlist) {
#pragma clang fp contract(off)
    // particles per chunk (measured on the 273-node grid: 256 FP32 / 192 FP64 keep 5-6
    // work-groups per CU, enough to cover the load latencies of the row tables and staging)
    constexpr int MAXP = sizeof(T) == 4 ? 256 : 192;
    __shared__ P2mStage<T, MAXP> S;
    __shared__ int cst[PT_ROWS][PC_X + 1];  // sorted index of the first particle of each cell, and the row's end
    __shared__ int rbase[PT_ROWS + 1];      // exclusive prefix of the rows' counts
    __shared__ int rowok[PT_ROWS];
    __shared__ int any;
    const int tid = threadIdx.x;
    const int ntx = (vnx + PT_X - 1) / PT_X, nty = (vny + PT_Y - 1) / PT_Y;
    const int64_t slot = xcd_slice(blockIdx.x, gridDim.x, XCD_ROUND);
    if (slot >= *nlist) return;
    const int tile = list[slot];
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    const int lx = tid & (PT_X - 1), ly = (tid / PT_X) & (PT_Y - 1), lz = tid / (PT_X * PT_Y);
    const int va = tx * PT_X + lx, vb = ty * PT_Y + ly, vc = tz * PT_Z + lz;
    // 1-based grid nodes: this thread's, and the tile's first
    const int i = va - wlo0 + 1, j = vb - wlo1 + 1, k = vc - wlo2 + 1;
    const int I0 = tx * PT_X - wlo0 + 1, J0 = ty * PT_Y - wlo1 + 1, K0 = tz * PT_Z - wlo2 + 1;
    if (tid == 0) any = 0;
    __syncthreads();
    if (tid < PT_ROWS) {
        const int cj = J0 - 2 + tid % PC_Y, ck = K0 - 2 + tid / PC_Y;
        int ok = cj >= -1 && cj <= g.n[1] + 1 && ck >= -1 && ck <= g.n[2] + 1;
        if (ok) ok = rowcnt[(cj + 1) + (int64_t)cny * (ck + 1)] > 0;
        rowok[tid] = ok;
        if (ok) any = 1;
    }
    __syncthreads();
    T acc[3] = {T(0), T(0), T(0)};
    if (any) {
        for (int e = tid; e < PT_ROWS * (PC_X + 1); e += blockDim.x) {
            const int r = e / (PC_X + 1), cidx = e - r * (PC_X + 1);
            int v = 0;
            if (rowok[r]) {
                const int cj = J0 - 2 + r % PC_Y, ck = K0 - 2 + r / PC_Y;
                const int64_t row = (int64_t)cnx * ((cj + 1) + (int64_t)cny * (ck + 1));
                const int cx = min(max(I0 - 2 + cidx, -1), g.n[0] + 2);
                v = start[row + cx + 1];
            }
            cst[r][cidx] = v;
        }
        __syncthreads();
        if (tid < 64) {  // the rows' counts, scanned across the first wave
            int v = tid < PT_ROWS ? cst[tid][PC_X] - cst[tid][0] : 0;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(v, d, 64);
                if (tid >= d) v += o;
            }
            if (tid < PT_ROWS) rbase[tid + 1] = v;
            if (tid == 0) rbase[0] = 0;
        }
        __syncthreads();
        int rb = 0;
        while (rb < PT_ROWS) {
            int re = rb;
            while (re < PT_ROWS && rbase[re + 1] - rbase[rb] <= MAXP) re++;
            const bool direct = re == rb;  // one row larger than a chunk
            if (direct) re = rb + 1;
            const int cnt = rbase[re] - rbase[rb];
            if (!direct && cnt > 0) {
                for (int e = tid; e < cnt; e += blockDim.x) {
                    // the row of staged element e (binary search over the chunk's rows)
                    int lo = rb, hi = re - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (rbase[mid] - rbase[rb] <= e) lo = mid;
                        else hi = mid - 1;
                    }
                    const int64_t sidx = cst[lo][0] + (e - (rbase[lo] - rbase[rb]));
                    using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
                    constexpr int NV = 16 * (int)sizeof(T) / (int)sizeof(V), PV = (int)(sizeof(V) / sizeof(T));
                    T r[16];
                    const V *src = (const V *)(R + 16 * sidx);
#pragma unroll
                    for (int v = 0; v < NV; v++) {
                        const V x = src[v];
                        __builtin_memcpy(r + v * PV, &x, sizeof(V));
                    }
#pragma unroll
                    for (int q = 0; q < 12; q++) S.w[q][e] = r[q];
#pragma unroll
                    for (int c = 0; c < 3; c++) S.q[c][e] = r[12 + c];
                }
            }
            __syncthreads();
            // this node's rows of the chunk: z-outer, y-inner; in row (dy, dz)
            // and cell dc of its 4 x 4 x 4 neighbourhood the node is the (3 - dc, 3 - dy, 3 - dz)-th
            // visitor of the particles
#pragma unroll
            for (int dz = 0; dz < 4; dz++) {
#pragma unroll
                for (int dy = 0; dy < 4; dy++) {
                    const int r = (ly + dy) + PC_Y * (lz + dz);
                    if (r < rb || r >= re) continue;
                    if (!direct) {
                        const int off = rbase[r] - rbase[rb] - cst[r][0];
#pragma unroll
                        for (int dc = 0; dc < 4; dc++) {
                            const int p0 = cst[r][lx + dc] + off, p1 = cst[r][lx + dc + 1] + off;
                            for (int p = p0; p < p1; p++) {
                                const T f = (S.w[3 - dc][p] * S.w[4 + 3 - dy][p]) * S.w[8 + 3 - dz][p];
                                acc[0] = acc[0] + f * S.q[0][p];
                                acc[1] = acc[1] + f * S.q[1][p];
                                acc[2] = acc[2] + f * S.q[2][p];
                            }
                        }
                    } else {
#pragma unroll
                        for (int dc = 0; dc < 4; dc++) {
                            for (int64_t sp = cst[r][lx + dc]; sp < cst[r][lx + dc + 1]; sp++) {
                                const T *w = R + 16 * sp;
                                const T f = (w[3 - dc] * w[7 - dy]) * w[11 - dz];
                                acc[0] = acc[0] + f * w[12];
                                acc[1] = acc[1] + f * w[13];
                                acc[2] = acc[2] + f * w[14];
                            }
                        }
                    }
                }
            }
            __syncthreads();  // the next chunk reuses the LDS
            rb = re;
        }
    }
    if (va >= vnx || vb >= vny || vc >= vnz) return;
    const bool in = i >= 1 && i <= g.n[0] && j >= 1 && j <= g.n[1] && k >= 1 && k <= g.n[2];
    const int64_t o = out.at(va, vb, vc);
#pragma unroll
    for (int r = 0; r < 3; r++) out.p[r * out.cs + o] = in ? (double)acc[r] / div : 0.0;
}

__global__ void k_m2p_tile_list(int ntiles, const int *__restrict__ flag, const int *__restrict__ pos,
                                int *__restrict__ list);

int p2m_gather(Ctx *c, const Grid &g, const SortState &s, const double *xp, const double *qp, int qstride, int ncomp,
               int itype, double div, GView out, const int vn[3], const int wlo[3], int kclass, bool f32,
               bool zero_empty, bool keep_zeros) {
    c->tile_map = false;
    if ((!keep_zeros && out.p == c->zt_ptr) || (keep_zeros && !(ncomp == 3 && itype == 4)))
        c->zt_valid = false;  // another writer of that buffer, or a path that writes every node
    int64_t nv = (int64_t)vn[0] * vn[1] * vn[2];
    unsigned grid = (unsigned)((nv + 255) / 256);
    const int *start = (const int *)s.start.p, *perm = (const int *)s.perm.p;
    hipEvent_t e0, e1;
    if (ncomp == 3 && itype == 4) {
        // node tiles with the particles' records staged in LDS
        int st = VPMX_OK;
        const int64_t n = c->sort_n;
        // buffers first (a growing one drains the stream), then the path, timed as a whole
        // (records, row counts, tile list, zeros, tiles); the cell-ordered particle records
        // use the space of the cell-ordered copy, which this path does not use
        const size_t ws = f32 ? sizeof(float) : sizeof(double);
        void *Rv = ensure(c, c->sort.xs, ws * 16 * (size_t)(n > 0 ? n : 1), &st);
        if (st) return st;
        const int64_t nrows = (int64_t)s.cn[1] * s.cn[2];
        int *rowcnt = (int *)ensure(c, c->sort.rowcnt, sizeof(int) * (size_t)nrows, &st);
        if (st) return st;
        const int ntx0 = (vn[0] + PT_X - 1) / PT_X, nty0 = (vn[1] + PT_Y - 1) / PT_Y;
        const int64_t tiles0 = (int64_t)ntx0 * nty0 * ((vn[2] + PT_Z - 1) / PT_Z);
        if (!ensure(c, c->tile_flag, sizeof(int) * (size_t)tiles0, &st)) return st;
        if (!ensure(c, c->tile_pos, sizeof(int) * (size_t)(tiles0 + 1), &st)) return st;
        if (!ensure(c, c->tile_list, sizeof(int) * (size_t)tiles0, &st)) return st;
        if (keep_zeros && !ensure(c, c->zt_prev, sizeof(int) * (size_t)tiles0, &st)) return st;
        const bool timed_all = kclass >= 0 && kernel_events(c, kclass, &e0, &e1);
        if (timed_all) VPMX_HIP(c, hipEventRecord(e0, c->stream));
        if (n > 0) {
            const dim3 gs((unsigned)((n + 255) / 256));
            if (f32)
                hipLaunchKernelGGL(k_p2m_sorted<float>, gs, dim3(256), 0, c->stream, perm, start + s.ncells, n, g,
                                   xp, qp, qstride, (float *)Rv);
            else
                hipLaunchKernelGGL(k_p2m_sorted<double>, gs, dim3(256), 0, c->stream, perm, start + s.ncells, n,
                                   g, xp, qp, qstride, (double *)Rv);
        }
        hipLaunchKernelGGL(k_row_counts, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, c->stream, start,
                           s.cn[0], nrows, rowcnt);
        const int ntx = (vn[0] + PT_X - 1) / PT_X, nty = (vn[1] + PT_Y - 1) / PT_Y;
        const int64_t tiles = (int64_t)ntx * nty * ((vn[2] + PT_Z - 1) / PT_Z);
        int *flag = (int *)ensure(c, c->tile_flag, sizeof(int) * (size_t)tiles, &st); if (st) return st;
        int *pos = (int *)ensure(c, c->tile_pos, sizeof(int) * (size_t)(tiles + 1), &st); if (st) return st;
        int *list = (int *)ensure(c, c->tile_list, sizeof(int) * (size_t)tiles, &st); if (st) return st;
        const unsigned gt = (unsigned)((tiles + 255) / 256);
        hipLaunchKernelGGL(k_p2m_tile_flags, dim3(gt), dim3(256), 0, c->stream, (int)tiles, ntx, nty, vn[0], vn[1],
                           vn[2], g, s.cn[1], rowcnt, wlo[1], wlo[2], flag);
        VPMX_TRY(exclusive_scan(c, flag, pos, tiles));
        hipLaunchKernelGGL(k_m2p_tile_list, dim3(gt), dim3(256), 0, c->stream, (int)tiles, flag, pos, list);
        if (zero_empty) {
            // buffers are all in place here, so alloc_gen is final for this call
            const bool same = keep_zeros && c->zt_valid && out.p == c->zt_ptr && c->zt_gen == c->alloc_gen &&
                              std::equal(vn, vn + 3, c->zt_vn) && std::equal(wlo, wlo + 3, c->zt_wlo);
            if (same)
                hipLaunchKernelGGL(k_p2m_zero_changed, dim3((unsigned)((tiles + ZT_TILES - 1) / ZT_TILES)), dim3(256),
                                   0, c->stream, (int)tiles, ntx, nty, vn[0], vn[1], vn[2], flag,
                                   (const int *)c->zt_prev.p, out);
            else
                hipLaunchKernelGGL(k_p2m_zero_tiles, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, c->stream, ntx,
                                   nty, vn[0], vn[1], vn[2], flag, out);
            if (keep_zeros) {
                VPMX_HIP(c, hipMemcpyAsync(c->zt_prev.p, flag, sizeof(int) * (size_t)tiles, hipMemcpyDeviceToDevice,
                                           c->stream));
                c->zt_valid = true;
                c->zt_ptr = out.p;
                std::copy(vn, vn + 3, c->zt_vn);
                std::copy(wlo, wlo + 3, c->zt_wlo);
                c->zt_gen = c->alloc_gen;
            }
        } else {
            c->tile_map = true;
            for (int d = 0; d < 3; d++) c->tile_map_vn[d] = vn[d];
            c->tile_map_ntx = ntx;
            c->tile_map_nty = nty;
        }
        const void *fn = f32 ? (const void *)k_p2m_tile<float> : (const void *)k_p2m_tile<double>;
        int cnx = s.cn[0], cny = s.cn[1], vx = vn[0], vy = vn[1], vz = vn[2], w0 = wlo[0], w1 = wlo[1], w2 = wlo[2];
        const int *lst = list, *nl = pos + tiles;
        void *args[] = {(void *)&g, (void *)&start, (void *)&rowcnt, &cnx, &cny, &Rv, &div,
                        &out, &vx, &vy, &vz, &w0, &w1, &w2, (void *)&lst, (void *)&nl};
        VPMX_HIP(c, hipLaunchKernel(fn, dim3((unsigned)tiles), dim3(256), args, 0, c->stream));
        VPMX_HIP(c, hipGetLastError());
        if (timed_all) VPMX_HIP(c, hipEventRecord(e1, c->stream));
        return VPMX_OK;
    }
    if (ncomp == 3 && kclass >= 0 && kernel_events(c, kclass, &e0, &e1)) {
        hipExtLaunchKernelGGL(k_p2m_gather<3>, dim3(grid), dim3(256), 0, c->stream, e0, e1, 0, g, start, perm, s.cn[0],
                              s.cn[1], xp, qp, qstride, itype, div, out, vn[0], vn[1], vn[2], wlo[0], wlo[1], wlo[2]);
    } else if (ncomp == 3)
        hipLaunchKernelGGL(k_p2m_gather<3>, dim3(grid), dim3(256), 0, c->stream, g, start, perm, s.cn[0], s.cn[1], xp,
                           qp, qstride, itype, div, out, vn[0], vn[1], vn[2], wlo[0], wlo[1], wlo[2]);
    else if (ncomp == 4)
        hipLaunchKernelGGL(k_p2m_gather<4>, dim3(grid), dim3(256), 0, c->stream, g, start, perm, s.cn[0], s.cn[1], xp,
                           qp, qstride, itype, div, out, vn[0], vn[1], vn[2], wlo[0], wlo[1], wlo[2]);
    else
        return fail(c, VPMX_ERR_UNSUP, "p2m_gather: ncomp must be 3 or 4");
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// ------------------------------------------------------------------ M2P
// back_to_particles_3D itype 1 (vpm_time.f90:30-107): UP = sum W vel, GP = (sum W SOL)*QP(neqpm+1).
template <int ITYPEB>
__global__ void __launch_bounds__(256) k_m2p(Grid g, const double *xp, const double *qp, int qstride, int itype,
                                              GView vel, GView sol, double *up, double *gp, int64_t n, const int *dn) {
#pragma clang fp contract(off)
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pcount(n, dn)) return;
    const int ips = stencil_ips(itype), ipf = stencil_ipf(itype);
    double px = xp[3 * p], py = xp[3 * p + 1], pz = xp[3 * p + 2];
    double u0 = 0, u1 = 0, u2 = 0, g0 = 0, g1 = 0, g2 = 0;
    double tx = (px - g.x0[0]) / g.h[0], ty = (py - g.x0[1]) / g.h[1], tz = (pz - g.x0[2]) / g.h[2];
    bool inside = (tx > -3.0 && tx < g.n[0] + 3.0) && (ty > -3.0 && ty < g.n[1] + 3.0) && (tz > -3.0 && tz < g.n[2] + 3.0);
    if (inside) {
        int in = (int)tx + 1, jn = (int)ty + 1, kn = (int)tz + 1;
        double fx[4], fy[4], fz[4];
        for (int q = 0; q < 4; q++) {
            fx[q] = proj_fun(itype, (px - g.x0[0] - (double)(in - ips + q - 1) * g.h[0]) / g.h[0]);
            fy[q] = proj_fun(itype, (py - g.x0[1] - (double)(jn - ips + q - 1) * g.h[1]) / g.h[1]);
            fz[q] = proj_fun(itype, (pz - g.x0[2] - (double)(kn - ips + q - 1) * g.h[2]) / g.h[2]);
        }
        int w = ips + ipf + 1;
        for (int c = 0; c < w; c++) {
            int k = kn - ips + c;
            for (int b = 0; b < w; b++) {
                int j = jn - ips + b;
                for (int a = 0; a < w; a++) {
                    int i = in - ips + a;
                    double f = fx[a] * fy[b] * fz[c];
                    if (i < 1 || i > g.n[0] || j < 1 || j > g.n[1] || k < 1 || k > g.n[2]) continue;
                    int64_t ov = vel.at(i - 1, j - 1, k - 1), os = sol.at(i - 1, j - 1, k - 1);
                    if (ITYPEB == 1) {
                        u0 = u0 + f * vel.p[ov];
                        u1 = u1 + f * vel.p[vel.cs + ov];
                        u2 = u2 + f * vel.p[2 * vel.cs + ov];
                    }
                    g0 = g0 + f * sol.p[os];
                    g1 = g1 + f * sol.p[sol.cs + os];
                    g2 = g2 + f * sol.p[2 * sol.cs + os];
                }
            }
        }
    }
    double vol = qp[(int64_t)qstride * p + (qstride - 1)];
    up[3 * p] = u0; up[3 * p + 1] = u1; up[3 * p + 2] = u2;
    gp[3 * p] = g0 * vol; gp[3 * p + 1] = g1 * vol; gp[3 * p + 2] = g2 * vol;
}

// back_to_particles_3D itype 1 by cell tiles: a work-group owns M2P_TX x M2P_TY x M2P_TZ
// cells, stages the velocity and stretching values of the (TX+3)(TY+3)(TZ+3) nodes its
// particles can reach in LDS (node-major, 0 outside the grid), then each thread takes one
// of the tile's particles (cell order) and sums its stencil from LDS in the order of k_m2p.
constexpr int M2P_TX = 8, M2P_TY = 8, M2P_TZ = 4;
constexpr int M2P_NX = M2P_TX + 3, M2P_NY = M2P_TY + 3, M2P_NZ = M2P_TZ + 3;
constexpr int M2P_ROWS = M2P_TY * M2P_TZ;
static_assert(M2P_ROWS <= 64, "the row scan runs in one wave");

// Tiles holding particles: flag per tile (from each particle's cell key; every writer of a
// flag writes 1), then (after an exclusive scan) their list.
__global__ void k_m2p_flag_particles(const int *__restrict__ keys, int64_t n, const int *dn, int cnx, int cny,
                                     int ntx, int nty, int *__restrict__ flag) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pcount(n, dn)) return;
    const int key = keys[p];
    if (key < 0) return;
    const int cx = key % cnx, cy = (key / cnx) % cny, cz = key / (cnx * cny);
    flag[cx / M2P_TX + ntx * (cy / M2P_TY + nty * (cz / M2P_TZ))] = 1;
}
__global__ void k_m2p_tile_list(int ntiles, const int *__restrict__ flag, const int *__restrict__ pos,
                                int *__restrict__ list) {
    const int tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile < ntiles && flag[tile]) list[pos[tile]] = tile;
}

// T: arithmetic of the node values, weights and sums (double: FP64 form; float: mixed form,
// positions and offsets still in FP64, UP and GP written in FP64).
template <class T>
__global__ void __launch_bounds__(256) k_m2p_tile(Grid g, const int *__restrict__ start, const int *__restrict__ perm,
                                                  int cnx, int cny, int cnz, const int *__restrict__ list,
                                                  const int *__restrict__ nlist, const double *__restrict__ xp,
                                                  const double *__restrict__ qp, int qstride, int itype, GView vel,
                                                  GView sol, double *__restrict__ up, double *__restrict__ gp) {
#pragma clang fp contract(off)
    __shared__ T nodes[M2P_NX * M2P_NY * M2P_NZ * 6];
    __shared__ int rs[M2P_ROWS + 1], rp[M2P_ROWS];
    const int ntx = (cnx + M2P_TX - 1) / M2P_TX, nty = (cny + M2P_TY - 1) / M2P_TY;
    const int ntot = *nlist;
    for (int li = (int)xcd_slice(blockIdx.x, gridDim.x, gridDim.x); li < ntot; li += gridDim.x) {
    const int tile = list[li];
    const int tx = tile % ntx, ty = (tile / ntx) % nty, tz = tile / (ntx * nty);
    // cell indices (stored index = cell + 1) of the tile: [cx0, cx0 + TX) clipped to the grid
    const int cx0 = tx * M2P_TX, cy0 = ty * M2P_TY, cz0 = tz * M2P_TZ;
    const int cx1 = min(cx0 + M2P_TX, cnx);
    const int tid = threadIdx.x;
    if (tid < 64) {  // the rows' first particles and counts, the counts scanned across the wave
        int cnt = 0, p0 = 0;
        if (tid < M2P_ROWS) {
            const int cyy = cy0 + tid % M2P_TY, czz = cz0 + tid / M2P_TY;
            if (cyy < cny && czz < cnz) {
                const int64_t row = (int64_t)cnx * (cyy + (int64_t)cny * czz);
                p0 = start[row + cx0];
                cnt = start[row + cx1] - p0;
            }
        }
        int v = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(v, d, 64);
            if (tid >= d) v += o;
        }
        if (tid < M2P_ROWS) {
            rp[tid] = p0;
            rs[tid + 1] = v;
        }
        if (tid == 0) rs[0] = 0;
    }
    __syncthreads();
    const int total = rs[M2P_ROWS];
    // node region: a particle of stored cell cs (inode = cs - 1, the sort's cell) reaches the
    // nodes inode-1 .. inode+2 = cs-2 .. cs+1, so the tile's cells [cx0, cx0+TX) reach global
    // nodes (1-based) cx0-2 .. cx0+TX: region index a <-> node cx0 - 2 + a, a in [0, TX+3)
    const int i0 = cx0 - 2, j0 = cy0 - 2, k0 = cz0 - 2;
    for (int e = tid; e < M2P_NX * M2P_NY * M2P_NZ; e += blockDim.x) {
        const int a = e % M2P_NX, b = (e / M2P_NX) % M2P_NY, cc = e / (M2P_NX * M2P_NY);
        const int i = i0 + a, j = j0 + b, k = k0 + cc;
        const bool in = i >= 1 && i <= g.n[0] && j >= 1 && j <= g.n[1] && k >= 1 && k <= g.n[2];
        double v[6] = {0, 0, 0, 0, 0, 0};
        if (in) {
            const int64_t ov = vel.at(i - 1, j - 1, k - 1), os = sol.at(i - 1, j - 1, k - 1);
            v[0] = vel.p[ov]; v[1] = vel.p[vel.cs + ov]; v[2] = vel.p[2 * vel.cs + ov];
            v[3] = sol.p[os]; v[4] = sol.p[sol.cs + os]; v[5] = sol.p[2 * sol.cs + os];
        }
#pragma unroll
        for (int q = 0; q < 6; q++) nodes[6 * e + q] = (T)v[q];
    }
    __syncthreads();
    const int ips = stencil_ips(itype), ipf = stencil_ipf(itype);
    const int w = ips + ipf + 1;
    for (int t = tid; t < total; t += blockDim.x) {
        int r = 0;
        while (rs[r + 1] <= t) r++;
        const int sidx = rp[r] + (t - rs[r]);
        const int64_t p = perm[sidx];
        const double px = xp[3 * p], py = xp[3 * p + 1], pz = xp[3 * p + 2];
        T u0 = 0, u1 = 0, u2 = 0, g0 = 0, g1 = 0, g2 = 0;
        const double txx = (px - g.x0[0]) / g.h[0], tyy = (py - g.x0[1]) / g.h[1], tzz = (pz - g.x0[2]) / g.h[2];
        const int in = (int)txx + 1, jn = (int)tyy + 1, kn = (int)tzz + 1;
        T fx[4], fy[4], fz[4];
        for (int q = 0; q < 4; q++) {
            fx[q] = (T)proj_fun(itype, (px - g.x0[0] - (double)(in - ips + q - 1) * g.h[0]) / g.h[0]);
            fy[q] = (T)proj_fun(itype, (py - g.x0[1] - (double)(jn - ips + q - 1) * g.h[1]) / g.h[1]);
            fz[q] = (T)proj_fun(itype, (pz - g.x0[2] - (double)(kn - ips + q - 1) * g.h[2]) / g.h[2]);
        }
        for (int c = 0; c < w; c++) {
            const int k = kn - ips + c;
            for (int b = 0; b < w; b++) {
                const int j = jn - ips + b;
                for (int a = 0; a < w; a++) {
                    const int i = in - ips + a;
                    const T f = fx[a] * fy[b] * fz[c];
                    const T *nd = nodes + 6 * ((i - i0) + M2P_NX * ((j - j0) + M2P_NY * (k - k0)));
                    u0 = u0 + f * nd[0];
                    u1 = u1 + f * nd[1];
                    u2 = u2 + f * nd[2];
                    g0 = g0 + f * nd[3];
                    g1 = g1 + f * nd[4];
                    g2 = g2 + f * nd[5];
                }
            }
        }
        const double vol = qp[(int64_t)qstride * p + (qstride - 1)];
        up[3 * p] = (double)u0; up[3 * p + 1] = (double)u1; up[3 * p + 2] = (double)u2;
        gp[3 * p] = (double)g0 * vol; gp[3 * p + 1] = (double)g1 * vol; gp[3 * p + 2] = (double)g2 * vol;
    }
    __syncthreads();  // the next tile reuses the LDS
    }
}

// Particles outside the cell grid reach no node: UP = GP = 0.
__global__ void k_m2p_outside(const int *__restrict__ keys, int64_t n, const int *dn, double *up, double *gp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pcount(n, dn) || keys[p] >= 0) return;
    for (int q = 0; q < 3; q++) up[3 * p + q] = 0.0, gp[3 * p + q] = 0.0;
}

int m2p_gather(Ctx *c, const Grid &g, const SortState &s, const double *xp, const double *qp, int qstride, int itype,
               GView vel, GView sol, double *up, double *gp, int64_t n, int itypeb) {
    c->tile_map = false;  // the interpolation tiles reuse tile_flag
    if (n <= 0) return VPMX_OK;
    hipEvent_t e0, e1;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (itypeb == 1 && c->sort_n == n && s.cn[0] == g.n[0] + 3 && s.cn[1] == g.n[1] + 3 &&
        s.cn[2] == g.n[2] + 3) {
        const int ntx = (s.cn[0] + M2P_TX - 1) / M2P_TX, nty = (s.cn[1] + M2P_TY - 1) / M2P_TY,
                  ntz = (s.cn[2] + M2P_TZ - 1) / M2P_TZ;
        const int tiles = ntx * nty * ntz;
        const int *start = (const int *)s.start.p, *perm = (const int *)s.perm.p, *keys = (const int *)s.keys.p;
        int st = VPMX_OK;
        int *flag = (int *)ensure(c, c->tile_flag, sizeof(int) * (size_t)tiles, &st); if (st) return st;
        int *pos = (int *)ensure(c, c->tile_pos, sizeof(int) * (size_t)(tiles + 1), &st); if (st) return st;
        int *list = (int *)ensure(c, c->tile_list, sizeof(int) * (size_t)tiles, &st); if (st) return st;
        const unsigned gt = (unsigned)((tiles + 255) / 256);
        // timed as a whole (tile list and the tiles)
        const bool timed = kernel_events(c, K_M2P, &e0, &e1);
        if (timed) VPMX_HIP(c, hipEventRecord(e0, c->stream));
        hipLaunchKernelGGL(k_m2p_outside, grid, dim3(256), 0, c->stream, keys, n, c->dn, up, gp);
        hipLaunchKernelGGL(k_zero_int, dim3(gt), dim3(256), 0, c->stream, flag, (int64_t)tiles);
        hipLaunchKernelGGL(k_m2p_flag_particles, grid, dim3(256), 0, c->stream, keys, n, c->dn, s.cn[0], s.cn[1], ntx,
                           nty, flag);
        VPMX_TRY(exclusive_scan(c, flag, pos, tiles));
        hipLaunchKernelGGL(k_m2p_tile_list, dim3(gt), dim3(256), 0, c->stream, tiles, flag, pos, list);
        // persistent: the work-groups per CU the tuner picks, at most what one CU holds (LDS-bound:
        // 40 KB of FP64 node values, 20 KB of FP32 in the mixed form); the kernel's results do
        // not depend on the grid, so the tuner's trials are real launches
        const bool f32 = c->precision == 1;
        const void *fn = f32 ? (const void *)k_m2p_tile<float> : (const void *)k_m2p_tile<double>;
        int gn = (int)s.cn[0], gny = (int)s.cn[1], gnz = (int)s.cn[2];
        const int *nl = pos + tiles;
        void *args[] = {(void *)&g, (void *)&start, (void *)&perm, &gn, &gny, &gnz, (void *)&list, (void *)&nl,
                        (void *)&xp, (void *)&qp, &qstride, &itype, &vel, &sol, &up, &gp};
        char key[128];
        std::snprintf(key, sizeof key, "interpolation tiles %s cells=%dx%dx%d", f32 ? "FP32" : "FP64", gn, gny, gnz);
        const char *names[1] = {"node tiles in LDS"};
        Ctx::Tuned ch;
        auto trial = [&](int, int pc) -> int {
            VPMX_HIP(c, hipLaunchKernel(fn, dim3((unsigned)(pc * c->num_cu)), dim3(256), args, 0, c->stream));
            return VPMX_OK;
        };
        VPMX_TRY(tune(c, key, 1, &fn, names, 0, 256, trial, &ch));
        const unsigned wg = (unsigned)(ch.per_cu * c->num_cu);
        VPMX_HIP(c, hipLaunchKernel(fn, dim3(wg), dim3(256), args, 0, c->stream));
        VPMX_HIP(c, hipGetLastError());
        if (timed) VPMX_HIP(c, hipEventRecord(e1, c->stream));
        return VPMX_OK;
    }
    if (itypeb == 2)
        hipLaunchKernelGGL(k_m2p<2>, grid, dim3(256), 0, c->stream, g, xp, qp, qstride, itype, vel, sol, up, gp, n,
                           c->dn);
    else if (kernel_events(c, K_M2P, &e0, &e1))
        hipExtLaunchKernelGGL(k_m2p<1>, grid, dim3(256), 0, c->stream, e0, e1, 0, g, xp, qp, qstride, itype, vel, sol,
                              up, gp, n, c->dn);
    else
        hipLaunchKernelGGL(k_m2p<1>, grid, dim3(256), 0, c->stream, g, xp, qp, qstride, itype, vel, sol, up, gp, n,
                           c->dn);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// ------------------------------------------------------------------ updates
// x = x0 + (u + uinf)*dt ; q(1:3) = q0(1:3) - g*dt  (test.f90:104-117 with FACDEF = 1)
__global__ void k_euler(double *xp, double *qp, int qs, const double *up, const double *gp, int64_t n, const int *dn,
                        double dt, double ux, double uy, double uz, const double *x0, const double *q0) {
#pragma clang fp contract(off)
    // one thread per array element (positions, then strengths), so neighbouring threads touch
    // neighbouring doubles
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = pcount(n, dn), nx = 3 * np;
    if (e < nx) {
        const int d = (int)(e % 3);
        const double ui = d == 0 ? ux : (d == 1 ? uy : uz);
        xp[e] = x0[e] + (up[e] + ui) * dt;
    } else if (e - nx < (int64_t)qs * np) {
        const int64_t f = e - nx, p = f / qs;
        const int d = (int)(f - p * qs);
        qp[f] = d < 3 ? q0[f] - gp[3 * p + d] * dt : q0[f];
    }
}
int euler_update(Ctx *c, double *xp, double *qp, int qs, const double *up, const double *gp, int64_t n, double dt,
                 const double uinf[3], const double *x0, const double *q0) {
    if (n <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_euler, dim3((unsigned)(((3 + qs) * n + 255) / 256)), dim3(256), 0, c->stream, xp, qp, qs, up, gp, n, c->dn,
                       dt, uinf[0], uinf[1], uinf[2], x0, q0);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// ------------------------------------------------------------------ reductions
// Deterministic two-pass reductions: fixed per-block tree, then one block over partials.
template <int NV, int MODE>  // MODE 0: sum, MODE 1: bbox (min 0..2, max 3..5)
__global__ void __launch_bounds__(256) k_reduce_particles(const double *xp, const double *qp, int qs, int64_t n,
                                                           double *part) {
#pragma clang fp contract(off)
    __shared__ double s[NV][256];
    int t = threadIdx.x;
    double v[NV];
    for (int q = 0; q < NV; q++) v[q] = MODE == 1 ? (q < 3 ? 1e308 : -1e308) : 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + t; p < n; p += (int64_t)gridDim.x * 256) {
        if (MODE == 1) {
            for (int d = 0; d < 3; d++) {
                double x = xp[3 * p + d];
                v[d] = fmin(v[d], x);
                v[3 + d] = fmax(v[3 + d], x);
            }
        } else {
            const double *x = xp + 3 * p, *q = qp + (int64_t)qs * p;
            double w0 = -q[0], w1 = -q[1], w2 = -q[2];
            v[0] += w0; v[1] += w1; v[2] += w2;
            v[3] += x[1] * w2 - x[2] * w1;
            v[4] += x[2] * w0 - x[0] * w2;
            v[5] += x[0] * w1 - x[1] * w0;
        }
    }
    for (int q = 0; q < NV; q++) s[q][t] = v[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < NV; q++) {
                if (MODE == 1) s[q][t] = q < 3 ? fmin(s[q][t], s[q][t + off]) : fmax(s[q][t], s[q][t + off]);
                else s[q][t] = s[q][t] + s[q][t + off];
            }
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < NV; q++) part[(int64_t)blockIdx.x * NV + q] = s[q][0];
}
// The partials in a fixed order: thread t folds partials t, t + 256, ... in turn, then a
// fixed tree over the 256 threads (the same bits for the same partials, whatever the launch).
template <int NV, int MODE>
__global__ void __launch_bounds__(256) k_reduce_final(const double *part, int nb, double *out, double hs = 1.0) {
#pragma clang fp contract(off)
    __shared__ double s[NV][256];
    const int t = threadIdx.x;
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = MODE == 1 ? (q < 3 ? 1e308 : -1e308) : 0.0;
    for (int b = t; b < nb; b += 256)
#pragma unroll
        for (int q = 0; q < NV; q++) {
            const double x = part[(int64_t)b * NV + q];
            if (MODE == 1) v[q] = q < 3 ? fmin(v[q], x) : fmax(v[q], x);
            else v[q] = v[q] + x;
        }
#pragma unroll
    for (int q = 0; q < NV; q++) s[q][t] = v[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            for (int q = 0; q < NV; q++) {
                if (MODE == 1) s[q][t] = q < 3 ? fmin(s[q][t], s[q][t + off]) : fmax(s[q][t], s[q][t + off]);
                else s[q][t] = s[q][t] + s[q][t + off];
            }
        __syncthreads();
    }
    if (t < NV) out[t] = (MODE == 0 && t >= 3) ? hs * s[t][0] : s[t][0];
}

static int reduce6(Ctx *c, int mode, const double *xp, const double *qp, int qs, int64_t n, double out[6]) {
    int st = VPMX_OK;
    int nb = (int)((n + 255) / 256);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    double *part = (double *)ensure(c, c->red_tmp, sizeof(double) * (size_t)(6 * nb + 6), &st);
    if (st) return st;
    if (mode == 1) {
        hipLaunchKernelGGL((k_reduce_particles<6, 1>), dim3(nb), dim3(256), 0, c->stream, xp, qp, qs, n, part);
        hipLaunchKernelGGL((k_reduce_final<6, 1>), dim3(1), dim3(256), 0, c->stream, part, nb, part + 6 * nb);
    } else {
        hipLaunchKernelGGL((k_reduce_particles<6, 0>), dim3(nb), dim3(256), 0, c->stream, xp, qp, qs, n, part);
        hipLaunchKernelGGL((k_reduce_final<6, 0>), dim3(1), dim3(256), 0, c->stream, part, nb, part + 6 * nb);
    }
    VPMX_HIP(c, hipGetLastError());
    VPMX_HIP(c, hipMemcpyAsync(out, part + 6 * nb, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    return VPMX_OK;
}
// ------------------------------------------------------------------ allocation-free diagnostics
// The partial sums of diagnostics_noalloc live in this fixed device array instead of a buffer
// the library allocates: RED_MAXB blocks of 6 sums (48 KiB + 64 B).  Each GPU holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives for the device (scratch_init), never through the array's name.
constexpr int RED_MAXB = 1024;
static __device__ __align__(256) double scratch[6 * RED_MAXB + 8];
constexpr int SCRATCH_MAXDEV = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[SCRATCH_MAXDEV] = {};

int scratch_init(Ctx *c) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device < 0 || c->device >= SCRATCH_MAXDEV)
        return fail(c, VPMX_ERR_UNSUP, "allocation-free entry points: device index above 63");
    if (!g_scratch_ptr[c->device]) {
        // (the context's device is current: vpmx_create set it)
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, VPMX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[c->device] = (double *)p;
    }
    c->scratch = g_scratch_ptr[c->device];
    return VPMX_OK;
}

int diagnostics_noalloc(Ctx *c, const double *xp, const double *qp, int qs, int64_t n, double *out) {
    if (!c->scratch) return fail(c, VPMX_ERR_STATE, "allocation-free diagnostics: the device array is not set up");
    if (n < 0 || n > 0x7fffffffLL) return fail(c, VPMX_ERR_ARG, "allocation-free diagnostics: n must be 0..2^31-1");
    if (qs < 3) return fail(c, VPMX_ERR_ARG, "allocation-free diagnostics: qstride must be at least 3");
    if (!out || (n > 0 && (!xp || !qp))) return fail(c, VPMX_ERR_ARG, "allocation-free diagnostics: null pointer");
    // the block count of reduce6, so both entry points sum in the same order
    int nb = (int)((n + 255) / 256);
    if (nb > RED_MAXB) nb = RED_MAXB;
    if (nb < 1) nb = 1;
    double *part = c->scratch;  // every block writes its partials: nothing to clear
    hipLaunchKernelGGL((k_reduce_particles<6, 0>), dim3(nb), dim3(256), 0, c->stream, xp, qp, qs, n, part);
    hipLaunchKernelGGL((k_reduce_final<6, 0>), dim3(1), dim3(256), 0, c->stream, part, nb, out, 0.5);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

int particle_bbox(Ctx *c, const double *xp, int64_t n, double mn[3], double mx[3]) {
    double o[6];
    VPMX_TRY(reduce6(c, 1, xp, nullptr, 0, n, o));
    for (int d = 0; d < 3; d++) { mn[d] = o[d]; mx[d] = o[3 + d]; }
    return VPMX_OK;
}
int diagnostics(Ctx *c, const double *xp, const double *qp, int qs, int64_t n, double circ[3], double imp[3]) {
    double o[6];
    VPMX_TRY(reduce6(c, 0, xp, qp, qs, n, o));
    for (int d = 0; d < 3; d++) { circ[d] = o[d]; imp[d] = 0.5 * o[3 + d]; }
    return VPMX_OK;
}

// ------------------------------------------------------------------ remesh
// remesh_particles_3d (vpm_remesh.f90:137-195), ncell_rem=1: nodes k->j->i over [lo,hi],
// kept when sqrt(w1^2+w2^2+w3^2) >= thr; the kept nodes keep that order (prefix sum).
// tmap (optional): node tiles of the projection holding values (ntx x nty per plane); nodes
// of the other tiles were not written and are zero.
__global__ void k_remesh_flag(GView w, int lx, int ly, int lz, int nx, int ny, int nz, double thr, int *flag,
                              const int *__restrict__ tmap, int ntx, int nty) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nt = (int64_t)nx * ny * nz;
    if (t >= nt) return;
    int a = (int)(t % nx), b = (int)((t / nx) % ny), cc = (int)(t / ((int64_t)nx * ny));
    const int i = lx - 1 + a, j = ly - 1 + b, k = lz - 1 + cc;
    if (tmap && !tmap[i / PT_X + ntx * (j / PT_Y + nty * (k / PT_Z))]) {
        flag[t] = 0;
        return;
    }
    int64_t o = w.at(i, j, k);
    double r0 = w.p[o], r1 = w.p[w.cs + o], r2 = w.p[2 * w.cs + o];
    double wmag = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    flag[t] = (wmag < thr) ? 0 : 1;
}
__global__ void k_remesh_write(GView w, int lx, int ly, int lz, int nx, int ny, int nz, const int *flag,
                               const int *pos, double x0, double y0, double z0, double hx, double hy, double hz,
                               double dv, int neqpm, double *xo, double *qo) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nt = (int64_t)nx * ny * nz;
    if (t >= nt || !flag[t]) return;
    int a = (int)(t % nx), b = (int)((t / nx) % ny), cc = (int)(t / ((int64_t)nx * ny));
    int i = lx + a, j = ly + b, k = lz + cc;
    int64_t o = w.at(i - 1, j - 1, k - 1);
    int64_t p = pos[t];
    xo[3 * p] = x0 + hx * (double)(i - 1);
    xo[3 * p + 1] = y0 + hy * (double)(j - 1);
    xo[3 * p + 2] = z0 + hz * (double)(k - 1);
    int nr = neqpm + 1;
    for (int q = 0; q < neqpm; q++) qo[(int64_t)nr * p + q] = w.p[q * w.cs + o] * dv;
    qo[(int64_t)nr * p + neqpm] = dv;
}

__global__ void k_copy_count(const int *src, int *dst) {
    if (threadIdx.x == 0) *dst = *src;
}

int remesh_compact(Ctx *c, GView w, const int vn[3], const int lo[3], const int hi[3], double thr, const double x0[3],
                   const double h[3], double dv, int neqpm, double **xp_out, double **qp_out, DBuf &bx, DBuf &bq,
                   int64_t *count) {
    int st = VPMX_OK;
    int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) { *count = 0; return VPMX_OK; }
    int64_t nt = (int64_t)nx * ny * nz;
    int *flag = (int *)ensure(c, c->rm_flag, sizeof(int) * (size_t)nt, &st); if (st) return st;
    int *pos = (int *)ensure(c, c->rm_pos, sizeof(int) * (size_t)(nt + 1), &st); if (st) return st;
    unsigned gr = (unsigned)((nt + 255) / 256);
    // the tile map of the projection that filled w, if it left empty tiles unwritten
    const bool tm = c->tile_map && c->tile_map_vn[0] == vn[0] && c->tile_map_vn[1] == vn[1] && c->tile_map_vn[2] == vn[2];
    hipLaunchKernelGGL(k_remesh_flag, dim3(gr), dim3(256), 0, c->stream, w, lo[0], lo[1], lo[2], nx, ny, nz, thr, flag,
                       tm ? (const int *)c->tile_flag.p : nullptr, c->tile_map_ntx, c->tile_map_nty);
    c->tile_map = false;
    VPMX_TRY(exclusive_scan(c, flag, pos, nt));
    int nr = neqpm + 1;
    if (c->dn) {
        // device-count mode: the kept nodes go straight into the resident arrays (their
        // capacity holds every node of the window: checked when the mode was set up) and the
        // count to the device, with no drain; the projection that read the particles precedes
        // this on the stream
        if (nt > c->cap) return fail(c, VPMX_ERR_STATE, "remesh window exceeds the particle capacity");
        double *xo = (double *)c->xp.p, *qo = (double *)c->qp.p;
        hipLaunchKernelGGL(k_remesh_write, dim3(gr), dim3(256), 0, c->stream, w, lo[0], lo[1], lo[2], nx, ny, nz, flag,
                           pos, x0[0], x0[1], x0[2], h[0], h[1], h[2], dv, neqpm, xo, qo);
        hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(64), 0, c->stream, pos + nt, c->dn);
        VPMX_HIP(c, hipGetLastError());
        *xp_out = xo;
        *qp_out = qo;
        *count = -1;
        return VPMX_OK;
    }
    int total = 0;
    VPMX_HIP(c, hipMemcpyAsync(&total, pos + nt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    double *xo = (double *)ensure(c, bx, sizeof(double) * 3 * (size_t)(total > 0 ? total : 1), &st); if (st) return st;
    double *qo = (double *)ensure(c, bq, sizeof(double) * nr * (size_t)(total > 0 ? total : 1), &st); if (st) return st;
    hipLaunchKernelGGL(k_remesh_write, dim3(gr), dim3(256), 0, c->stream, w, lo[0], lo[1], lo[2], nx, ny, nz, flag, pos,
                       x0[0], x0[1], x0[2], h[0], h[1], h[2], dv, neqpm, xo, qo);
    VPMX_HIP(c, hipGetLastError());
    *xp_out = xo;
    *qp_out = qo;
    *count = total;
    return VPMX_OK;
}

}  // namespace vpmx
