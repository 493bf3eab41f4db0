// Infinite-domain boundary values: normal-derivative face sources and their FP64
// direct Green's-function sums onto the boundary nodes.
//
//   normal_derivs : calc_normalderiv_3d (pinfdomain.f90:294-547)
//   boundary_sums : Bounds3d -> calc_boundinf_3d (pmbound.f90:80-157, 415-515), ibctyp=2
//
// Sum: SOL(c,t) += sum_s -sigma_c(s) * ds(s) / (4 pi |x_t - x_s|)  over every face-cell
// source s, for every boundary node t (X faces full, Y faces without the x-edges, Z
// faces without x- and y-edges).  Targets are spread over threads (2 per thread),
// sources stream through LDS in tiles, and the source range is split over a second
// grid dimension whose partial sums are added in a fixed order afterwards, so the
// result is deterministic.  1/r is v_rsq_f64 refined by one third-order step.
#include <hip/hip_ext.h>

#include "bnodes.h"
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

// Sources of face `f` (0 XMIN, 1 XMAX, 2 YMIN, 3 YMAX, 4 ZMIN, 5 ZMAX), in the reference's
// order: k outer / j inner on X faces, k outer / i inner on Y faces, j outer / i inner on Z.
// Face-node offsets of the node-derivative table: face f holds ncomp x N0 x N1 values (N0, N1
// its in-plane node counts: x faces (NY, NZ), y faces (NX, NZ), z faces (NX, NY)).
struct FaceNodes {
    int64_t off[7];
};
__host__ __device__ inline void face_dims(const int n[3], int f, int *N0, int *N1) {
    const int axis = f / 2;
    *N0 = axis == 0 ? n[1] : n[0];
    *N1 = axis == 2 ? n[1] : n[2];
}

// The one-sided 5-point normal derivative (1/h_n)(25/12 f0 - 4 f1 + 3 f2 - 4/3 f3 + 1/4 f4) at
// every node of every face, once per node (a node is a corner of up to four face cells):
// thread per (face, node), the in-plane index with the faster memory stride first.
__global__ void __launch_bounds__(256) k_face_node_derivs(Grid g, GView s0, int ncomp, FaceNodes fo, double *fd) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nn = fo.off[6] / ncomp;
    if (t >= nn) return;
    int f = 0;
    while (f < 5 && t >= fo.off[f + 1] / ncomp) f++;
    const int64_t r = t - fo.off[f] / ncomp;
    int N0, N1;
    face_dims(g.n, f, &N0, &N1);
    const int a = (int)(r % N0) + 1, b = (int)(r / N0) + 1;
    const int axis = f / 2, hi = f & 1, sg = hi ? -1 : 1, pl = hi ? g.n[axis] : 1;
    const double a1 = 25.0 / 12.0, a2 = -4.0, a3 = 3.0, a4 = -4.0 / 3.0, a5 = 0.25;
    const double hn = g.h[axis];
    for (int c = 0; c < ncomp; c++) {
        const double *sp = s0.p + c * s0.cs;
        double v[5];
        for (int st = 0; st < 5; st++) {
            const int q = pl + sg * st;
            const int ii = axis == 0 ? q : a, jj = axis == 1 ? q : (axis == 0 ? a : b), kk = axis == 2 ? q : b;
            v[st] = sp[s0.at(ii - 1, jj - 1, kk - 1)];
        }
        fd[fo.off[f] + (int64_t)c * N0 * N1 + r] = (1.0 / hn) * ((((a1 * v[0] + a2 * v[1]) + a3 * v[2]) + a4 * v[3]) + a5 * v[4]);
    }
}

__global__ void k_normal_derivs(Grid g, const double *__restrict__ fd, FaceNodes fo, int ncomp, int64_t nb, double *sig,
                                double *cxyz, double *a, double *ds) {
#pragma clang fp contract(off)
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t fX = (int64_t)(NY - 1) * (NZ - 1), fY = (int64_t)(NX - 1) * (NZ - 1), fZ = (int64_t)(NX - 1) * (NY - 1);
    const double DX = g.h[0], DY = g.h[1], DZ = g.h[2];
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    int face, i0, i1;  // face and the two in-plane 1-based cell indices
    int64_t r;
    if (s < 2 * fX) { face = (int)(s / fX); r = s % fX; i0 = (int)(r % (NY - 1)) + 1; i1 = (int)(r / (NY - 1)) + 1; }
    else if (s < 2 * (fX + fY)) { int64_t u = s - 2 * fX; face = 2 + (int)(u / fY); r = u % fY; i0 = (int)(r % (NX - 1)) + 1; i1 = (int)(r / (NX - 1)) + 1; }
    else { int64_t u = s - 2 * (fX + fY); face = 4 + (int)(u / fZ); r = u % fZ; i0 = (int)(r % (NX - 1)) + 1; i1 = (int)(r / (NX - 1)) + 1; }
    const int axis = face / 2, hi = face & 1;
    const int pl = hi ? g.n[axis] : 1;  // plane node index
    // corner order of psi1..psi4: X faces (j,k),(j1,k),(j1,k1),(j,k1); Y faces (i,k),(i1,k),(i1,k1),(i,k1);
    // Z faces (i,j),(i1,j),(i1,j1),(i,j1); their derivatives from the face-node table
    const int du[4] = {0, 1, 1, 0}, dv[4] = {0, 0, 1, 1};
    int N0, N1;
    face_dims(g.n, face, &N0, &N1);
    for (int c = 0; c < ncomp; c++) {
        const double *fc = fd + fo.off[face] + (int64_t)c * N0 * N1;
        double p[4];
        for (int q = 0; q < 4; q++) p[q] = fc[(i0 + du[q] - 1) + (int64_t)(i1 + dv[q] - 1) * N0];
        double sv = 0.25 * (((p[0] + p[1]) + p[2]) + p[3]);
        sig[c * nb + s] = sv;
    }
    // centroid = 0.25*(((s1+s2)+s3)+s4) with the reference's corner order per face
    auto X = [&](int i) { return g.x0[0] + (double)(i - 1) * DX; };
    auto Y = [&](int j) { return g.x0[1] + (double)(j - 1) * DY; };
    auto Z = [&](int k) { return g.x0[2] + (double)(k - 1) * DZ; };
    auto C4 = [](double p1, double p2, double p3, double p4) { return 0.25 * (((p1 + p2) + p3) + p4); };
    double cx, cy, cz, dsv;
    if (axis == 0) {
        double x = X(pl);
        cx = C4(x, x, x, x);
        int j = i0, k = i1;
        cy = hi ? C4(Y(j + 1), Y(j + 1), Y(j), Y(j)) : C4(Y(j), Y(j), Y(j + 1), Y(j + 1));
        cz = C4(Z(k), Z(k + 1), Z(k + 1), Z(k));
        dsv = DY * DZ;
    } else if (axis == 1) {
        double y = Y(pl);
        cy = C4(y, y, y, y);
        int i = i0, k = i1;
        cx = hi ? C4(X(i + 1), X(i + 1), X(i), X(i)) : C4(X(i), X(i), X(i + 1), X(i + 1));
        cz = C4(Z(k), Z(k + 1), Z(k + 1), Z(k));
        dsv = DX * DZ;
    } else {
        double z = Z(pl);
        cz = C4(z, z, z, z);
        int i = i0, j = i1;
        cx = hi ? C4(X(i + 1), X(i + 1), X(i), X(i)) : C4(X(i), X(i), X(i + 1), X(i + 1));
        cy = C4(Y(j), Y(j + 1), Y(j + 1), Y(j));
        dsv = DX * DY;
    }
    cxyz[s] = cx;
    cxyz[nb + s] = cy;
    cxyz[2 * nb + s] = cz;
    ds[s] = dsv;
    for (int c = 0; c < ncomp; c++) a[c * nb + s] = -(sig[c * nb + s] * dsv) / PI4;
}

static FaceNodes face_nodes(int ncomp, const int n[3]) {
    FaceNodes fo;
    fo.off[0] = 0;
    for (int f = 0; f < 6; f++) {
        int N0, N1;
        face_dims(n, f, &N0, &N1);
        fo.off[f + 1] = fo.off[f] + (int64_t)ncomp * N0 * N1;
    }
    return fo;
}

double *normal_derivs_table(Ctx *c, int ncomp, const int n[3], int *st) {
    const FaceNodes fo = face_nodes(ncomp, n);
    return (double *)ensure(c, c->nd_face, sizeof(double) * (size_t)fo.off[6], st);
}

int normal_derivs_cells(Ctx *c, int ncomp, const Grid &g, const double *fd, double *sig, double *cxyz, double *wgt,
                        int64_t *nbound) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    if (NX < 6 || NY < 6 || NZ < 6) return fail(c, VPMX_ERR_ARG, "normal derivatives need >= 6 nodes per axis");
    const int64_t nb = 2 * ((int64_t)(NY - 1) * (NZ - 1) + (int64_t)(NX - 1) * (NZ - 1) + (int64_t)(NX - 1) * (NY - 1));
    const FaceNodes fo = face_nodes(ncomp, g.n);
    // wgt holds a (ncomp*nb) followed by ds (nb)
    hipLaunchKernelGGL(k_normal_derivs, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c->stream, g, fd, fo, ncomp,
                       nb, sig, cxyz, wgt, wgt + (int64_t)ncomp * nb);
    VPMX_HIP(c, hipGetLastError());
    *nbound = nb;
    return VPMX_OK;
}

int normal_derivs(Ctx *c, int ncomp, const Grid &g, GView s0, double *sig, double *cxyz, double *wgt, int64_t *nbound) {
    if (g.n[0] < 6 || g.n[1] < 6 || g.n[2] < 6) return fail(c, VPMX_ERR_ARG, "normal derivatives need >= 6 nodes per axis");
    // the derivative at each face node once, then each cell's four-corner average
    const FaceNodes fo = face_nodes(ncomp, g.n);
    int st = VPMX_OK;
    double *fd = normal_derivs_table(c, ncomp, g.n, &st);
    if (st) return st;
    const int64_t nn = fo.off[6] / ncomp;
    hipLaunchKernelGGL(k_face_node_derivs, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, g, s0, ncomp,
                       fo, fd);
    VPMX_HIP(c, hipGetLastError());
    return normal_derivs_cells(c, ncomp, g, fd, sig, cxyz, wgt, nbound);
}

__device__ __forceinline__ double rsqrt_f64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    double h = r2 * y;
    double e = __builtin_fma(-h, y, 1.0);
    double p = __builtin_fma(0.375, e, 0.5);
    return __builtin_fma(y * e, p, y);
}

constexpr int BS_THREADS = 256;
constexpr int BS_TPT = 2;
constexpr int BS_TILE = 256;

// Partial sums over one source chunk for BS_THREADS*BS_TPT targets.
// FINE=1 for the block grid, 0 for the coarse grid (separate kernel names in profiles).
template <int FINE>
__global__ void __launch_bounds__(BS_THREADS) k_boundsum(Grid g, int64_t nt, int64_t ns, int64_t chunk,
                                                          const double *__restrict__ cxyz, const double *__restrict__ a,
                                                          double *__restrict__ part) {
    __shared__ double sx[BS_TILE], sy[BS_TILE], sz[BS_TILE], s0[BS_TILE], s1[BS_TILE], s2[BS_TILE];
    const int64_t tb = (int64_t)blockIdx.x * BS_THREADS * BS_TPT;
    double X[BS_TPT], Y[BS_TPT], Z[BS_TPT], acc0[BS_TPT], acc1[BS_TPT], acc2[BS_TPT];
#pragma unroll
    for (int r = 0; r < BS_TPT; r++) {
        int64_t t = tb + threadIdx.x + r * BS_THREADS;
        int i = 1, j = 1, k = 1, pl = 0;
        if (t < nt) target_node(g, t, &i, &j, &k, &pl);
        X[r] = g.x0[0] + (double)(i - 1) * g.h[0];
        Y[r] = g.x0[1] + (double)(j - 1) * g.h[1];
        Z[r] = g.x0[2] + (double)(k - 1) * g.h[2];
        acc0[r] = acc1[r] = acc2[r] = 0.0;
    }
    const int64_t sb = (int64_t)blockIdx.y * chunk;
    const int64_t se = sb + chunk < ns ? sb + chunk : ns;
    for (int64_t t0 = sb; t0 < se; t0 += BS_TILE) {
        const int cnt = (int)(se - t0 < BS_TILE ? se - t0 : BS_TILE);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += BS_THREADS) {
            int64_t s = t0 + q;
            sx[q] = cxyz[s];
            sy[q] = cxyz[ns + s];
            sz[q] = cxyz[2 * ns + s];
            s0[q] = a[s];
            s1[q] = a[ns + s];
            s2[q] = a[2 * ns + s];
        }
        __syncthreads();
        for (int q = 0; q < cnt; q++) {
            const double px = sx[q], py = sy[q], pz = sz[q], w0 = s0[q], w1 = s1[q], w2 = s2[q];
#pragma unroll
            for (int r = 0; r < BS_TPT; r++) {
                double dx = px - X[r], dy = py - Y[r], dz = pz - Z[r];
                double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
                double inv = rsqrt_f64(r2);
                acc0[r] = __builtin_fma(w0, inv, acc0[r]);
                acc1[r] = __builtin_fma(w1, inv, acc1[r]);
                acc2[r] = __builtin_fma(w2, inv, acc2[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < BS_TPT; r++) {
        int64_t t = tb + threadIdx.x + r * BS_THREADS;
        if (t >= nt) continue;
        double *pp = part + (int64_t)blockIdx.y * 3 * nt;
        pp[t] = acc0[r];
        pp[nt + t] = acc1[r];
        pp[2 * nt + t] = acc2[r];
    }
}

// SOL(target) = SOL(target) + partial_0 + partial_1 + ... (fixed order)
__global__ void k_boundsum_reduce(Grid g, int64_t nt, int nsplit, const double *part, GView sol) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i, j, k, pl;
    target_node(g, t, &i, &j, &k, &pl);
    int64_t o = sol.at(i - 1, j - 1, k - 1);
    for (int c = 0; c < 3; c++) {
        double v = sol.p[c * sol.cs + o];
        for (int s = 0; s < nsplit; s++) v = v + part[(int64_t)s * 3 * nt + c * nt + t];
        sol.p[c * sol.cs + o] = v;
    }
}

// Exact reference arithmetic including the r <= 1e-5 self term; used when a source can
// come within 1e-5 of a boundary node (grid spacings below ~1.4e-5).
__global__ void k_boundsum_exact(Grid g, int64_t nt, int64_t ns, const double *cxyz, const double *sig,
                                 const double *ds, GView sol) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i, j, k, pl;
    target_node(g, t, &i, &j, &k, &pl);
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    const double lg = log((sqrt(2.0) - 1) / (sqrt(2.0) + 1));
    double X = g.x0[0] + (double)(i - 1) * g.h[0], Y = g.x0[1] + (double)(j - 1) * g.h[1],
           Z = g.x0[2] + (double)(k - 1) * g.h[2];
    double hs = pl == 2 ? (g.h[1] + g.h[2]) : (pl == 1 ? (g.h[0] + g.h[2]) : (g.h[0] + g.h[1]));
    int64_t o = sol.at(i - 1, j - 1, k - 1);
    double acc[3];
    for (int c = 0; c < 3; c++) acc[c] = sol.p[c * sol.cs + o];
    for (int64_t s = 0; s < ns; s++) {
        double XR = cxyz[s] - X, YR = cxyz[ns + s] - Y, ZR = cxyz[2 * ns + s] - Z;
        double r = sqrt(XR * XR + YR * YR + ZR * ZR);
        double DS = ds[s];
        for (int c = 0; c < 3; c++) {
            double sv = sig[c * ns + s];
            if (r > 1e-05) acc[c] = acc[c] - sv * DS / (PI4 * r);
            else acc[c] = acc[c] + sv * hs * lg / PI4;
        }
    }
    for (int c = 0; c < 3; c++) sol.p[c * sol.cs + o] = acc[c];
}

// ---------------------------------------------------------------- lattice form
// Sources (face-cell centres) and targets (face nodes) sit on regular lattices.  For a
// target line and a source line running along the same axis C, the distance of a pair
// depends only on the two line positions and on the index offset e = tc - sc along C:
//   r^2 = K(line_t, line_s) + ((2e - 1) hC / 2)^2
// (coordinates in half-cells: node i -> 2(i-1), cell j -> 2(j-1)+1).  Every pair's term
// is still summed directly; a wave computes each distinct 1/r of its line once per
// source line into an LDS table and reads it back for all pairs with that offset.
struct LatPair {
    int At, ht, As, hs;     // target / source face normal axes and hi flags
    int C, Bt, Sa;          // line axis, target line-index axis, source line-index axis
    int tlo[3], thi[3];     // target node ranges (1-based) on the target face
    int nlines;             // target lines
    int nsl;                // source lines
    int nsc;                // sources per source line (cells along C)
    int fs;                 // source face id 0..5 (partial-sum slice)
    int64_t soff;           // index of the source face's first source
    int64_t sstride_c;      // source index stride along C
    int64_t sstride_l;      // source index stride along the source line axis
    // resolved per pair on the host (no runtime-indexed arrays in the kernel)
    int blo, clo, chi;      // first target line index, target range along C
    int Pt, Ps, plane;      // plane coordinates (half-cells) and the target plane node index
    double hA, hS, hC, hB;  // spacings along At, As, C, Bt
};

constexpr int LW = 4;  // waves (target lines) per work-group

__device__ __forceinline__ int64_t target_index(const Grid &g, int face_axis, int hi, int i, int j, int k) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t fX = (int64_t)NY * NZ, fY = (int64_t)(NX - 2) * NZ, fZ = (int64_t)(NX - 2) * (NY - 2);
    if (face_axis == 0) return hi * fX + (int64_t)(k - 1) * NY + (j - 1);
    if (face_axis == 1) return 2 * fX + hi * fY + (int64_t)(k - 1) * (NX - 2) + (i - 2);
    return 2 * fX + 2 * fY + hi * fZ + (int64_t)(j - 2) * (NX - 2) + (i - 2);
}

template <int M>
__global__ void __launch_bounds__(64 * LW) k_bs_lattice(Grid g, const LatPair *__restrict__ pairs, int nsp,
                                                         int tabn, int nsplit, const double *__restrict__ a,
                                                         int64_t ns, double *__restrict__ part, int64_t nt) {
    extern __shared__ double lds[];
    const LatPair P = pairs[blockIdx.y];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *tab = lds + w * tabn;
    double *aw = lds + LW * tabn;  // aw[3][nsp]
    const int line = blockIdx.x * LW + w;
    const bool active = line < P.nlines;
    const int b = P.blo + line;
    const int clo = P.clo, chi = P.chi;
    const int emin = clo - nsp;
    const double hA = P.hA, hS = P.hS, hC = P.hC;
    const int Pt = P.Pt, Ps = P.Ps;
    const int split = blockIdx.z;
    const int l0 = (int)((int64_t)P.nsl * split / nsplit), l1 = (int)((int64_t)P.nsl * (split + 1) / nsplit);
    double acc[M][3];
#pragma unroll
    for (int m = 0; m < M; m++) acc[m][0] = acc[m][1] = acc[m][2] = 0.0;
    const int tc0 = clo + lane * M;
    for (int sl = l0; sl < l1; sl++) {
        const int aidx = sl + 1;  // source line index (cell) along the source line axis
        __syncthreads();
        // source weights of this line (zero-padded to nsp)
        for (int q = threadIdx.x; q < nsp; q += 64 * LW) {
            double v0 = 0.0, v1 = 0.0, v2 = 0.0;
            if (q < P.nsc) {
                int64_t s = P.soff + (int64_t)sl * P.sstride_l + (int64_t)q * P.sstride_c;
                v0 = a[s];
                v1 = a[ns + s];
                v2 = a[2 * ns + s];
            }
            aw[q] = v0;
            aw[nsp + q] = v1;
            aw[2 * nsp + q] = v2;
        }
        // K for (target line b, source line aidx)
        double d1, d2;
        if (P.At != P.As) {
            d1 = (double)(Pt - (2 * (aidx - 1) + 1)) * hA * 0.5;
            d2 = (double)(2 * (b - 1) - Ps) * hS * 0.5;
        } else {
            d1 = (double)(Pt - Ps) * hA * 0.5;
            d2 = (double)(2 * (b - 1) - (2 * (aidx - 1) + 1)) * P.hB * 0.5;
        }
        const double K = d1 * d1 + d2 * d2;
        if (active)
            for (int q = lane; q < tabn; q += 64) {
                const int e = emin + q;
                const double dc = (double)(2 * e - 1) * hC * 0.5;
                tab[q] = rsqrt_f64(K + dc * dc);
            }
        __syncthreads();
        if (!active) continue;
        // sc runs over 1..nsp in blocks of M; window r[q] = tab[tc0 - s0 - (M-1) + q - emin]
        for (int s0 = 1; s0 <= nsp; s0 += M) {
            const int base = tc0 - s0 - (M - 1) - emin;
            double r[2 * M - 1];
#pragma unroll
            for (int q = 0; q < 2 * M - 1; q++) r[q] = tab[base + q];
#pragma unroll
            for (int d = 0; d < M; d++) {
                const double w0 = aw[s0 - 1 + d], w1 = aw[nsp + s0 - 1 + d], w2 = aw[2 * nsp + s0 - 1 + d];
#pragma unroll
                for (int m = 0; m < M; m++) {
                    const double t = r[m - d + M - 1];
                    acc[m][0] = __builtin_fma(w0, t, acc[m][0]);
                    acc[m][1] = __builtin_fma(w1, t, acc[m][1]);
                    acc[m][2] = __builtin_fma(w2, t, acc[m][2]);
                }
            }
        }
    }
    if (!active) return;
    double *pp = part + (int64_t)(P.fs * nsplit + split) * 3 * nt;
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int tc = tc0 + m;
        if (tc > chi) continue;
        const int pl = P.plane;
        const int ii = P.At == 0 ? pl : (P.Bt == 0 ? b : tc);
        const int jj = P.At == 1 ? pl : (P.Bt == 1 ? b : tc);
        const int kk = P.At == 2 ? pl : (P.Bt == 2 ? b : tc);
        const int64_t t = target_index(g, P.At, P.ht, ii, jj, kk);
        pp[t] = acc[m][0];
        pp[nt + t] = acc[m][1];
        pp[2 * nt + t] = acc[m][2];
    }
}

// SOL(target) = SOL(target) + sum over (source face, split) partials in a fixed order.
__global__ void k_bs_lattice_reduce(Grid g, int64_t nt, int nslices, const double *part, GView sol) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i, j, k, pl;
    target_node(g, t, &i, &j, &k, &pl);
    int64_t o = sol.at(i - 1, j - 1, k - 1);
    for (int c = 0; c < 3; c++) {
        double v = sol.p[c * sol.cs + o];
        for (int s = 0; s < nslices; s++) v = v + part[(int64_t)s * 3 * nt + c * nt + t];
        sol.p[c * sol.cs + o] = v;
    }
}

static void make_pairs(const Grid &g, std::vector<LatPair> &out) {
    const int N[3] = {g.n[0], g.n[1], g.n[2]};
    // source face offsets in the reference order
    int64_t fsz[3] = {(int64_t)(N[1] - 1) * (N[2] - 1), (int64_t)(N[0] - 1) * (N[2] - 1),
                      (int64_t)(N[0] - 1) * (N[1] - 1)};
    int64_t off[6];
    int64_t o = 0;
    for (int f = 0; f < 6; f++) {
        off[f] = o;
        o += fsz[f / 2];
    }
    for (int ft = 0; ft < 6; ft++)
        for (int fs = 0; fs < 6; fs++) {
            LatPair P{};
            P.At = ft / 2; P.ht = ft & 1; P.As = fs / 2; P.hs = fs & 1;
            if (P.At != P.As) {
                P.C = 3 - P.At - P.As;
                P.Bt = P.As;
                P.Sa = P.At;
            } else {
                int u = P.At == 0 ? 1 : 0, v = P.At == 2 ? 1 : 2;
                P.C = v;
                P.Bt = u;
                P.Sa = u;
            }
            // target ranges: X faces full; Y faces x in [2,NX-1]; Z faces x in [2,NX-1], y in [2,NY-1]
            for (int d = 0; d < 3; d++) { P.tlo[d] = 1; P.thi[d] = N[d]; }
            if (P.At >= 1) { P.tlo[0] = 2; P.thi[0] = N[0] - 1; }
            if (P.At == 2) { P.tlo[1] = 2; P.thi[1] = N[1] - 1; }
            P.tlo[P.At] = P.thi[P.At] = P.ht ? N[P.At] : 1;
            P.nlines = P.thi[P.Bt] - P.tlo[P.Bt] + 1;
            P.nsl = N[P.Sa] - 1;
            P.nsc = N[P.C] - 1;
            P.fs = fs;
            P.soff = off[fs];
            // source in-plane axes p < q: index = (q_cell-1)*(N_p-1) + (p_cell-1)
            int p = P.As == 0 ? 1 : 0, q = P.As == 2 ? 1 : 2;
            int64_t sp = 1, sq = N[p] - 1;
            P.sstride_c = (P.C == p) ? sp : sq;
            P.sstride_l = (P.Sa == p) ? sp : sq;
            P.blo = P.tlo[P.Bt];
            P.clo = P.tlo[P.C];
            P.chi = P.thi[P.C];
            P.Pt = P.ht ? 2 * (N[P.At] - 1) : 0;
            P.Ps = P.hs ? 2 * (N[P.As] - 1) : 0;
            P.plane = P.ht ? N[P.At] : 1;
            P.hA = g.h[P.At];
            P.hS = g.h[P.As];
            P.hC = g.h[P.C];
            P.hB = g.h[P.Bt];
            out.push_back(P);
        }
}

template <int M>
static int launch_lattice(Ctx *c, const Grid &g, const LatPair *pairs_d, int maxlines, int nsp, int tabn, int nsplit,
                          const double *a, int64_t ns, double *part, int64_t nt, bool fine) {
    size_t lds = sizeof(double) * ((size_t)LW * tabn + 3 * (size_t)nsp);
    dim3 grid((unsigned)((maxlines + LW - 1) / LW), 36, (unsigned)nsplit);
    hipEvent_t e0, e1;
    if (fine && kernel_events(c, K_BOUNDSUM, &e0, &e1)) {
        hipExtLaunchKernelGGL(k_bs_lattice<M>, grid, dim3(64 * LW), (uint32_t)lds, c->stream, e0, e1, 0, g, pairs_d,
                              nsp, tabn, nsplit, a, ns, part, nt);
    } else {
        hipLaunchKernelGGL(k_bs_lattice<M>, grid, dim3(64 * LW), lds, c->stream, g, pairs_d, nsp, tabn, nsplit, a, ns,
                           part, nt);
    }
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

static int boundary_sums_lattice(Ctx *c, const Grid &g, int64_t ns, const double *a, GView sol, bool fine) {
    int st = VPMX_OK;
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    int64_t nt = 2 * ((int64_t)NY * NZ + (int64_t)(NX - 2) * NZ + (int64_t)(NX - 2) * (NY - 2));
    int nmax = NX > NY ? NX : NY;
    if (NZ > nmax) nmax = NZ;
    const int M = (nmax + 63) / 64;
    if (M > 8) return fail(c, VPMX_ERR_UNSUP, "lattice boundary sums: faces longer than 512 nodes");
    std::vector<LatPair> pairs;
    make_pairs(g, pairs);
    int maxlines = 0, maxsc = 0;
    for (auto &P : pairs) {
        if (P.nlines > maxlines) maxlines = P.nlines;
        if (P.nsc > maxsc) maxsc = P.nsc;
    }
    const int nsp = (maxsc + M - 1) / M * M;
    const int tabn = 64 * M + nsp;
    size_t lds = sizeof(double) * ((size_t)LW * tabn + 3 * (size_t)nsp);
    if (lds > 65536) return fail(c, VPMX_ERR_UNSUP, "lattice boundary sums: LDS table too large");
    // enough work-groups to fill the device several times over
    int64_t base_wg = (int64_t)((maxlines + LW - 1) / LW) * 36;
    int nsplit = (int)((4096 + base_wg - 1) / base_wg);
    if (nsplit < 1) nsplit = 1;
    if (nsplit > 8) nsplit = 8;
    LatPair *pd = (LatPair *)ensure(c, fine ? c->lat_pairs : c->lat_pairs_c, sizeof(LatPair) * pairs.size(), &st);
    if (st) return st;
    VPMX_HIP(c, hipMemcpyAsync(pd, pairs.data(), sizeof(LatPair) * pairs.size(), hipMemcpyHostToDevice, c->stream));
    double *part = (double *)ensure(c, c->bs_part, sizeof(double) * 3 * (size_t)nt * 6 * nsplit, &st);
    if (st) return st;
    if (fine) {
        // FP64 work of the launch: 3 FMA (6 flops) per source-target pair, plus 20 flops per
        // table entry (r^2, v_rsq_f64 and its refinement) -- every wave builds one table of
        // tabn entries per (target line, source line)
        double tables = 0.0;
        for (auto &P : pairs) tables += (double)P.nlines * P.nsl * tabn;
        c->k_work[K_BOUNDSUM] = 6.0 * (double)nt * (double)ns + 20.0 * tables;
        mark(c, ST_BOUNDS, 0);
    }
    switch (M) {
        case 1: st = launch_lattice<1>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        case 2: st = launch_lattice<2>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        case 3: st = launch_lattice<3>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        case 4: st = launch_lattice<4>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        case 5: st = launch_lattice<5>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        case 6: st = launch_lattice<6>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        case 7: st = launch_lattice<7>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
        default: st = launch_lattice<8>(c, g, pd, maxlines, nsp, tabn, nsplit, a, ns, part, nt, fine); break;
    }
    if (st) return st;
    if (fine) mark(c, ST_BOUNDS, 1);
    hipLaunchKernelGGL(k_bs_lattice_reduce, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, g, nt,
                       6 * nsplit, part, sol);
    VPMX_HIP(c, hipGetLastError());
    if (fine) c->last_pairs += (double)nt * (double)ns;
    return VPMX_OK;
}

int boundary_sums(Ctx *c, int ncomp, const Grid &g, int64_t ns, const double *sig, const double *cxyz,
                  const double *wgt, GView sol, bool fine, int levmax, int ibctyp, const GView *scratch) {
    if (ncomp != 3) return fail(c, VPMX_ERR_UNSUP, "boundary sums: 3 components");
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    int64_t nt = 2 * ((int64_t)NY * NZ + (int64_t)(NX - 2) * NZ + (int64_t)(NX - 2) * (NY - 2));
    const double *a = wgt, *ds = wgt + 3 * ns;
    if (ibctyp == 3) {  // panel sources: calc_boundinf_3d_s / calc_boundinf_3d_lev_s
        if (levmax < 0 && !scratch) return fail(c, VPMX_ERR_STATE, "panel sums need a scratch grid");
        return boundary_sums_panels(c, g, ns, sig, cxyz, ds, sol, fine, levmax, scratch ? *scratch : sol);
    }
    if (levmax >= 0) {  // Bounds3d_lev (infdomain_3d with itree = 1, pinfdomain.f90:128-150)
        if (fine) mark(c, ST_BOUNDS, 0);
        VPMX_TRY(boundary_sums_tree(c, g, ns, sig, cxyz, ds, sol, levmax, fine, false));
        if (fine) mark(c, ST_BOUNDS, 1);
        return VPMX_OK;
    }
    double hmin = g.h[0] < g.h[1] ? g.h[0] : g.h[1];
    if (g.h[2] < hmin) hmin = g.h[2];
    // nearest source-target distance is h_min/sqrt(2) (face-cell centroid vs node)
    if (hmin * 0.70710678 <= 1.5e-5) {
        hipLaunchKernelGGL(k_boundsum_exact, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, g, nt, ns, cxyz,
                           sig, ds, sol);
        VPMX_HIP(c, hipGetLastError());
        return VPMX_OK;
    }
    return point_sums(c, g, ns, cxyz, a, sol, fine);
}

int point_sums(Ctx *c, const Grid &g, int64_t ns, const double *cxyz, const double *a, GView sol, bool fine) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    int64_t nt = 2 * ((int64_t)NY * NZ + (int64_t)(NX - 2) * NZ + (int64_t)(NX - 2) * (NY - 2));
    if (fine && c->bs_mode == 2) {
        mark(c, ST_BOUNDS, 0);
        hipEvent_t e0, e1;
        const bool timed = kernel_events(c, K_BOUNDSUM, &e0, &e1);
        if (timed) VPMX_HIP(c, hipEventRecord(e0, c->stream));
        const int st = boundary_sums_spectral(c, g, ns, a, sol, 0);
        if (st == VPMX_OK) {
            if (timed) VPMX_HIP(c, hipEventRecord(e1, c->stream));
            mark(c, ST_BOUNDS, 1);
            c->last_pairs += (double)nt * (double)ns;
            return VPMX_OK;
        }
        if (st != VPMX_ERR_UNSUP) return st;
        if (timed) c->kev_pending[K_BOUNDSUM] -= 1;  // nothing ran: drop the interval
        mark(c, ST_BOUNDS, 1);  // faces too long for the in-LDS transforms: lattice kernel
    } else if (c->bs_mode == 2) {  // the coarse grid: its own plan
        const int st = boundary_sums_spectral(c, g, ns, a, sol, 1);
        if (st == VPMX_OK) return VPMX_OK;
        if (st != VPMX_ERR_UNSUP) return st;
    }
    if (c->bs_mode != 1) return boundary_sums_lattice(c, g, ns, a, sol, fine);
    int64_t tblocks = (nt + BS_THREADS * BS_TPT - 1) / (BS_THREADS * BS_TPT);
    // split sources so the launch has >= ~2048 work-groups, each with >= 1024 sources
    int64_t nsplit = (2048 + tblocks - 1) / tblocks;
    int64_t maxsplit = (ns + 1023) / 1024;
    if (nsplit > maxsplit) nsplit = maxsplit;
    if (nsplit < 1) nsplit = 1;
    if (nsplit > 64) nsplit = 64;
    int64_t chunk = (ns + nsplit - 1) / nsplit;
    chunk = (chunk + BS_TILE - 1) / BS_TILE * BS_TILE;
    nsplit = (ns + chunk - 1) / chunk;
    int st = VPMX_OK;
    double *part = (double *)ensure(c, c->bs_part, sizeof(double) * 3 * (size_t)nt * nsplit, &st);
    if (st) return st;
    if (fine) {
        mark(c, ST_BOUNDS, 0);
        hipEvent_t e0, e1;
        if (kernel_events(c, K_BOUNDSUM, &e0, &e1))
            hipExtLaunchKernelGGL(k_boundsum<1>, dim3((unsigned)tblocks, (unsigned)nsplit), dim3(BS_THREADS), 0,
                                  c->stream, e0, e1, 0, g, nt, ns, chunk, cxyz, a, part);
        else
            hipLaunchKernelGGL(k_boundsum<1>, dim3((unsigned)tblocks, (unsigned)nsplit), dim3(BS_THREADS), 0,
                               c->stream, g, nt, ns, chunk, cxyz, a, part);
        mark(c, ST_BOUNDS, 1);
    } else {
        hipLaunchKernelGGL(k_boundsum<0>, dim3((unsigned)tblocks, (unsigned)nsplit), dim3(BS_THREADS), 0, c->stream, g,
                           nt, ns, chunk, cxyz, a, part);
    }
    hipLaunchKernelGGL(k_boundsum_reduce, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, g, nt,
                       (int)nsplit, part, sol);
    VPMX_HIP(c, hipGetLastError());
    if (fine) c->last_pairs += (double)nt * (double)ns;
    return VPMX_OK;
}

}  // namespace vpmx
