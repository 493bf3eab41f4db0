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
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

// Sources of face `f` (0 XMIN, 1 XMAX, 2 YMIN, 3 YMAX, 4 ZMIN, 5 ZMAX), in the reference's
// order: k outer / j inner on X faces, k outer / i inner on Y faces, j outer / i inner on Z.
__global__ void k_normal_derivs(Grid g, GView s0, int ncomp, int64_t nb, double *sig, double *cxyz, double *a,
                                double *ds) {
#pragma clang fp contract(off)
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t fX = (int64_t)(NY - 1) * (NZ - 1), fY = (int64_t)(NX - 1) * (NZ - 1), fZ = (int64_t)(NX - 1) * (NY - 1);
    const double a1 = 25.0 / 12.0, a2 = -4.0, a3 = 3.0, a4 = -4.0 / 3.0, a5 = 0.25;
    const double DX = g.h[0], DY = g.h[1], DZ = g.h[2];
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    int face, i0, i1;  // face and the two in-plane 1-based cell indices
    int64_t r;
    if (s < 2 * fX) { face = (int)(s / fX); r = s % fX; i0 = (int)(r % (NY - 1)) + 1; i1 = (int)(r / (NY - 1)) + 1; }
    else if (s < 2 * (fX + fY)) { int64_t u = s - 2 * fX; face = 2 + (int)(u / fY); r = u % fY; i0 = (int)(r % (NX - 1)) + 1; i1 = (int)(r / (NX - 1)) + 1; }
    else { int64_t u = s - 2 * (fX + fY); face = 4 + (int)(u / fZ); r = u % fZ; i0 = (int)(r % (NX - 1)) + 1; i1 = (int)(r / (NX - 1)) + 1; }
    const int axis = face / 2, hi = face & 1;
    const int sg = hi ? -1 : 1;
    const int pl = hi ? g.n[axis] : 1;  // plane node index
    // node index of corner (du, dv) -> (i,j,k) 1-based
    auto node = [&](int du, int dv, int step, int *ii, int *jj, int *kk) {
        if (axis == 0) { *ii = pl + sg * step; *jj = i0 + du; *kk = i1 + dv; }
        else if (axis == 1) { *ii = i0 + du; *jj = pl + sg * step; *kk = i1 + dv; }
        else { *ii = i0 + du; *jj = i1 + dv; *kk = pl + sg * step; }
    };
    // corner order of psi1..psi4: X faces (j,k),(j1,k),(j1,k1),(j,k1); Y faces (i,k),(i1,k),(i1,k1),(i,k1);
    // Z faces (i,j),(i1,j),(i1,j1),(i,j1)
    const int du[4] = {0, 1, 1, 0}, dv[4] = {0, 0, 1, 1};
    const double hn = g.h[axis];
    for (int c = 0; c < ncomp; c++) {
        const double *sp = s0.p + c * s0.cs;
        double p[4];
        for (int q = 0; q < 4; q++) {
            int ii, jj, kk;
            double f[5];
            for (int st = 0; st < 5; st++) {
                node(du[q], dv[q], st, &ii, &jj, &kk);
                f[st] = sp[s0.at(ii - 1, jj - 1, kk - 1)];
            }
            p[q] = (1.0 / hn) * ((((a1 * f[0] + a2 * f[1]) + a3 * f[2]) + a4 * f[3]) + a5 * f[4]);
        }
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

int normal_derivs(Ctx *c, int ncomp, const Grid &g, GView s0, double *sig, double *cxyz, double *wgt, int64_t *nbound) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    if (NX < 6 || NY < 6 || NZ < 6) return fail(c, VPMX_ERR_ARG, "normal derivatives need >= 6 nodes per axis");
    int64_t nb = 2 * ((int64_t)(NY - 1) * (NZ - 1) + (int64_t)(NX - 1) * (NZ - 1) + (int64_t)(NX - 1) * (NY - 1));
    // wgt holds a (ncomp*nb) followed by ds (nb)
    hipLaunchKernelGGL(k_normal_derivs, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c->stream, g, s0, ncomp, nb,
                       sig, cxyz, wgt, wgt + (int64_t)ncomp * nb);
    VPMX_HIP(c, hipGetLastError());
    *nbound = nb;
    return VPMX_OK;
}

// Boundary node t (reference face order) -> 1-based (i,j,k) and plane code.
__device__ __forceinline__ void target_node(const Grid &g, int64_t t, int *i, int *j, int *k, int *pl) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t fX = (int64_t)NY * NZ, fY = (int64_t)(NX - 2) * NZ, fZ = (int64_t)(NX - 2) * (NY - 2);
    if (t < 2 * fX) {
        int f = (int)(t / fX); int64_t r = t % fX;
        *i = f ? NX : 1; *j = (int)(r % NY) + 1; *k = (int)(r / NY) + 1; *pl = 2;
    } else if (t < 2 * (fX + fY)) {
        int64_t u = t - 2 * fX; int f = (int)(u / fY); int64_t r = u % fY;
        *j = f ? NY : 1; *i = (int)(r % (NX - 2)) + 2; *k = (int)(r / (NX - 2)) + 1; *pl = 1;
    } else {
        int64_t u = t - 2 * (fX + fY); int f = (int)(u / fZ); int64_t r = u % fZ;
        *k = f ? NZ : 1; *i = (int)(r % (NX - 2)) + 2; *j = (int)(r / (NX - 2)) + 2; *pl = 3;
    }
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

int boundary_sums(Ctx *c, int ncomp, const Grid &g, int64_t ns, const double *sig, const double *cxyz,
                  const double *wgt, GView sol, bool fine) {
    if (ncomp != 3) return fail(c, VPMX_ERR_UNSUP, "boundary sums: 3 components");
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    int64_t nt = 2 * ((int64_t)NY * NZ + (int64_t)(NX - 2) * NZ + (int64_t)(NX - 2) * (NY - 2));
    const double *a = wgt, *ds = wgt + 3 * ns;
    double hmin = g.h[0] < g.h[1] ? g.h[0] : g.h[1];
    if (g.h[2] < hmin) hmin = g.h[2];
    // nearest source-target distance is h_min/sqrt(2) (face-cell centroid vs node)
    if (hmin * 0.70710678 <= 1.5e-5) {
        hipLaunchKernelGGL(k_boundsum_exact, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, g, nt, ns, cxyz,
                           sig, ds, sol);
        VPMX_HIP(c, hipGetLastError());
        return VPMX_OK;
    }
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
        hipLaunchKernelGGL(k_boundsum<1>, dim3((unsigned)tblocks, (unsigned)nsplit), dim3(BS_THREADS), 0, c->stream, g,
                           nt, ns, chunk, cxyz, a, part);
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
