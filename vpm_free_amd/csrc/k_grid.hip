// Grid kernels: fills, curl and stretching stencils, the yaps coarse-grid glue, and
// layout conversions.
//
//   curl_kernel     u = curl(psi) by central differences   (vpm_gcalc.f90:17-40)
//   stretch_kernel  SOL(m) = -sum_j d_j(RHS_j u_m)          (vpm_gcalc.f90:66-132)
//   coarse_sample   fine solution every ncoarse nodes       (yaps3d.f90:104-138)
//   coarse_laplacian 7-point Laplacian of the sample        (yaps3d.f90:1990-2016)
//   face_interp     TSC coarse correction on inner faces    (yaps3d.f90:340-395, 1914-1983)
#include <cstdio>

#include "tune.h"
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

__global__ void k_fill(double *p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
int fill(Ctx *c, double *p, int64_t n, double v) {
    if (n <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n, v);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// Node (i,j,k) 0-based.  Interior 1..n-2 gets the curl, the rest 0 (velvr*_pm = 0 before).
// One thread per node on a (x, y, z) launch (64 x 4 nodes per work-group, one z plane), so no
// index divisions; the differences are scaled by 1/(2h) (within an ulp of the reference's
// division by 2h).
__global__ void __launch_bounds__(256) k_curl(int nx, int ny, int nz, double ix2, double iy2, double iz2, GView psi,
                                               GView vel) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6), k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    double u = 0.0, v = 0.0, w = 0.0;
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
        const double *p0 = psi.p, *p1 = psi.p + psi.cs, *p2 = psi.p + 2 * psi.cs;
        const int64_t o = psi.at(i, j, k);
        const int64_t xp = o + 1, xm = o - 1, yp = o + psi.py, ym = o - psi.py, zp = o + psi.pz, zm = o - psi.pz;
        double dpsidx1 = (p1[xp] - p1[xm]) * ix2, dpsidx2 = (p2[xp] - p2[xm]) * ix2;
        double dpsidy0 = (p0[yp] - p0[ym]) * iy2, dpsidy2 = (p2[yp] - p2[ym]) * iy2;
        double dpsidz0 = (p0[zp] - p0[zm]) * iz2, dpsidz1 = (p1[zp] - p1[zm]) * iz2;
        u = dpsidy2 - dpsidz1;
        v = -(dpsidx2 - dpsidz0);
        w = dpsidx1 - dpsidy0;
    }
    const int64_t o = vel.at(i, j, k);
    vel.p[o] = u;
    vel.p[vel.cs + o] = v;
    vel.p[2 * vel.cs + o] = w;
}

static dim3 node_grid(const int n[3]) {
    return dim3((unsigned)((n[0] + 63) / 64), (unsigned)((n[1] + 3) / 4), (unsigned)n[2]);
}

int curl_kernel(Ctx *c, const int n[3], const double h[3], GView psi, GView vel) {
    if ((int64_t)n[0] * n[1] * n[2] <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_curl, node_grid(n), dim3(256), 0, c->stream, n[0], n[1], n[2], 1.0 / (2 * h[0]),
                       1.0 / (2 * h[1]), 1.0 / (2 * h[2]), psi, vel);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// Interior 2..n-3 (0-based) gets the stretching term, the rest 0 (SOL_pm = 0 before).  Same
// launch shape and scaling as k_curl.
__global__ void __launch_bounds__(256) k_stretch(int nx, int ny, int nz, double ix2, double iy2, double iz2,
                                                  GView rhs, GView vel, GView out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6), k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    double r[3] = {0.0, 0.0, 0.0};
    if (i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3 && k >= 2 && k <= nz - 3) {
        const double *R0 = rhs.p, *R1 = rhs.p + rhs.cs, *R2 = rhs.p + 2 * rhs.cs;
        const int64_t ro = rhs.at(i, j, k), vo = vel.at(i, j, k);
        const double a0p = R0[ro + 1], a0m = R0[ro - 1], a1p = R1[ro + rhs.py], a1m = R1[ro - rhs.py];
        const double a2p = R2[ro + rhs.pz], a2m = R2[ro - rhs.pz];
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double *u = vel.p + m * vel.cs + vo;
            double wdudx = -(a0p * u[1] - (a0m)*u[-1]) * ix2;
            double wdvdy = -(a1p * u[vel.py] - (a1m)*u[-vel.py]) * iy2;
            double wdwdz = -(a2p * u[vel.pz] - (a2m)*u[-vel.pz]) * iz2;
            r[m] = wdudx + wdvdy + wdwdz;
        }
    }
    const int64_t o = out.at(i, j, k);
    out.p[o] = r[0];
    out.p[out.cs + o] = r[1];
    out.p[2 * out.cs + o] = r[2];
}
int stretch_kernel(Ctx *c, const int n[3], const double h[3], GView rhs, GView vel, GView out) {
    if ((int64_t)n[0] * n[1] * n[2] <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_stretch, node_grid(n), dim3(256), 0, c->stream, n[0], n[1], n[2], 1.0 / (2 * h[0]),
                       1.0 / (2 * h[1]), 1.0 / (2 * h[2]), rhs, vel, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// curl and stretching in one pass (calc_velocity_serial_3d, vpm_gcalc.f90:1-132): a work-group
// owns a TX x TY column tile (256 threads) and marches through CS_KC planes of z, keeping the
// velocity of three planes (with a one-node halo in x and y) in LDS; each plane's velocity is
// formed once per tile (halo nodes by the neighbouring tiles too) with k_curl's expressions,
// and the stretching term reads its velocity neighbours from LDS with k_stretch's expressions
// -- the same values bit for bit, without writing the velocity grid and reading it back.  The
// tile shape is a tuner choice (the halo's share against the x extent of the rows).
constexpr int CS_KC = 32;
// Tiles go out XCD by XCD in contiguous slices of each round of R work-groups (xcd_slice; R
// about the work-groups resident at once), so the tiles that share halo columns, and the
// planes each of them reads three times, meet in one L2.
template <int TX, int TY>
__global__ void __launch_bounds__(256) k_curl_stretch(int nx, int ny, int nz, double ix2, double iy2, double iz2,
                                                       GView psi, GView rhs, GView vel, GView out, int R) {
#pragma clang fp contract(off)
    static_assert(TX * TY == 256, "one thread per column of the tile");
    constexpr int CS_HX = TX + 2, CS_HY = TY + 2;
    __shared__ double V[3][3][CS_HY][CS_HX];  // [plane mod 3][component][y][x]
    const int64_t gx = gridDim.x, gy = gridDim.y;
    const int64_t lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int64_t slot = xcd_slice(lin, gx * gy * gridDim.z, R);
    const int i0 = (int)(slot % gx) * TX, j0 = (int)((slot / gx) % gy) * TY, kb = (int)(slot / (gx * gy)) * CS_KC;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const double *p0 = psi.p, *p1 = psi.p + psi.cs, *p2 = psi.p + 2 * psi.cs;
    auto fill = [&](int kk) {  // velocity of plane kk on the tile and its halo
        const int sl = (kk + 3) % 3;
        for (int q = threadIdx.x; q < CS_HX * CS_HY; q += 256) {
            const int hx = q % CS_HX, hy = q / CS_HX, i = i0 - 1 + hx, j = j0 - 1 + hy;
            double u = 0.0, v = 0.0, w = 0.0;
            if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && kk >= 1 && kk <= nz - 2) {
                const int64_t o = psi.at(i, j, kk);
                const int64_t xp = o + 1, xm = o - 1, yp = o + psi.py, ym = o - psi.py, zp = o + psi.pz, zm = o - psi.pz;
                double dpsidx1 = (p1[xp] - p1[xm]) * ix2, dpsidx2 = (p2[xp] - p2[xm]) * ix2;
                double dpsidy0 = (p0[yp] - p0[ym]) * iy2, dpsidy2 = (p2[yp] - p2[ym]) * iy2;
                double dpsidz0 = (p0[zp] - p0[zm]) * iz2, dpsidz1 = (p1[zp] - p1[zm]) * iz2;
                u = dpsidy2 - dpsidz1;
                v = -(dpsidx2 - dpsidz0);
                w = dpsidx1 - dpsidy0;
            }
            V[sl][0][hy][hx] = u;
            V[sl][1][hy][hx] = v;
            V[sl][2][hy][hx] = w;
        }
    };
    fill(kb - 1);
    fill(kb);
    const int kend = kb + CS_KC < nz ? kb + CS_KC : nz;
    const int i = i0 + tx, j = j0 + ty;
    for (int k = kb; k < kend; k++) {
        fill(k + 1);
        __syncthreads();
        const int sl = k % 3, sp = (k + 1) % 3, sm = (k + 2) % 3;
        if (i < nx && j < ny) {
            const int64_t ov = vel.at(i, j, k);
            for (int c = 0; c < 3; c++) vel.p[c * vel.cs + ov] = V[sl][c][ty + 1][tx + 1];
            double r[3] = {0.0, 0.0, 0.0};
            if (i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3 && k >= 2 && k <= nz - 3) {
                const double *R0 = rhs.p, *R1 = rhs.p + rhs.cs, *R2 = rhs.p + 2 * rhs.cs;
                const int64_t ro = rhs.at(i, j, k);
                const double a0p = R0[ro + 1], a0m = R0[ro - 1], a1p = R1[ro + rhs.py], a1m = R1[ro - rhs.py];
                const double a2p = R2[ro + rhs.pz], a2m = R2[ro - rhs.pz];
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    double wdudx = -(a0p * V[sl][m][ty + 1][tx + 2] - (a0m)*V[sl][m][ty + 1][tx]) * ix2;
                    double wdvdy = -(a1p * V[sl][m][ty + 2][tx + 1] - (a1m)*V[sl][m][ty][tx + 1]) * iy2;
                    double wdwdz = -(a2p * V[sp][m][ty + 1][tx + 1] - (a2m)*V[sm][m][ty + 1][tx + 1]) * iz2;
                    r[m] = wdudx + wdvdy + wdwdz;
                }
            }
            const int64_t oo = out.at(i, j, k);
            out.p[oo] = r[0];
            out.p[out.cs + oo] = r[1];
            out.p[2 * out.cs + oo] = r[2];
        }
        __syncthreads();  // the next iteration refills the slot of plane k - 1
    }
}
// The same with each stream-function value read from memory once per tile: every thread owns
// up to NS columns of the tile's region with a two-node halo ((TX + 4) x (TY + 4)) and keeps
// psi_0 and psi_1 of its columns for three planes in registers (the z differences), while
// each plane's three components go to one LDS plane the x and y differences read; the
// stretching term keeps RHS_2 of its column for three planes the same way.  The expressions
// and their operands are k_curl's and k_stretch's, so the values are the same bit for bit.
template <int TX, int TY>
__global__ void __launch_bounds__(256) k_curl_stretch_w(int nx, int ny, int nz, double ix2, double iy2, double iz2,
                                                         GView psi, GView rhs, GView vel, GView out, int R) {
#pragma clang fp contract(off)
    static_assert(TX * TY == 256, "one thread per column of the tile");
    constexpr int CS_HX = TX + 2, CS_HY = TY + 2, RX = TX + 4, RY = TY + 4;
    constexpr int NS = (RX * RY + 255) / 256;
    __shared__ double V[3][3][CS_HY][CS_HX];  // velocity: [plane mod 3][component][y][x]
    __shared__ double P[3][RY][RX];           // stream function of one plane: [component][y][x]
    const int64_t gx = gridDim.x, gy = gridDim.y;
    const int64_t lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int64_t slot = xcd_slice(lin, gx * gy * gridDim.z, R);
    const int i0 = (int)(slot % gx) * TX, j0 = (int)((slot / gx) % gy) * TY, kb = (int)(slot / (gx * gy)) * CS_KC;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const double *p0 = psi.p, *p1 = psi.p + psi.cs, *p2 = psi.p + 2 * psi.cs;
    // the thread's region columns: (hx, hy) <-> node (i0 - 2 + hx, j0 - 2 + hy)
    int64_t co[NS];
    bool cin[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int q = tid + 256 * s, hx = q % RX, hy = q / RX, i = i0 - 2 + hx, j = j0 - 2 + hy;
        cin[s] = q < RX * RY && i >= 0 && i < nx && j >= 0 && j < ny;
        co[s] = cin[s] ? psi.at(i, j, 0) : 0;
    }
    auto ld = [&](const double *pc, int s, int k) -> double {
        return (cin[s] && k >= 0 && k < nz) ? pc[co[s] + (int64_t)k * psi.pz] : 0.0;
    };
    // w0/w1[s][t]: psi_0 / psi_1 of column s at planes kk - 1 + t while plane kk is formed
    double w0[NS][3], w1[NS][3];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        w0[s][1] = ld(p0, s, kb - 2);
        w1[s][1] = ld(p1, s, kb - 2);
        w0[s][2] = ld(p0, s, kb - 1);
        w1[s][2] = ld(p1, s, kb - 1);
    }
    const int kend = kb + CS_KC < nz ? kb + CS_KC : nz;
    const int i = i0 + tx, j = j0 + ty;
    const bool tin = i < nx && j < ny;
    const bool sint = i >= 2 && i <= nx - 3 && j >= 2 && j <= ny - 3;
    const double *R0 = rhs.p, *R1 = rhs.p + rhs.cs, *R2 = rhs.p + 2 * rhs.cs;
    const int64_t rc = tin ? rhs.at(i, j, 0) : 0;
    auto ldr2 = [&](int k) -> double { return (tin && k >= 0 && k < nz) ? R2[rc + (int64_t)k * rhs.pz] : 0.0; };
    // RHS_2 of the thread's column at planes k - 1, k, k + 1 of the stretching plane k = kk - 1
    // (after the shift at the top of iteration kk; the first iteration has k = kb - 2)
    double r2[3] = {0.0, ldr2(kb - 3), ldr2(kb - 2)};
    for (int kk = kb - 1; kk <= kend; kk++) {
        // plane kk into LDS, psi_0 / psi_1 of plane kk + 1 into the windows
#pragma unroll
        for (int s = 0; s < NS; s++) {
            w0[s][0] = w0[s][1];
            w1[s][0] = w1[s][1];
            w0[s][1] = w0[s][2];
            w1[s][1] = w1[s][2];
            w0[s][2] = ld(p0, s, kk + 1);
            w1[s][2] = ld(p1, s, kk + 1);
            const int q = tid + 256 * s;
            if (q < RX * RY) {
                const int hx = q % RX, hy = q / RX;
                P[0][hy][hx] = w0[s][1];
                P[1][hy][hx] = w1[s][1];
                P[2][hy][hx] = ld(p2, s, kk);
            }
        }
        __syncthreads();
        // velocity of plane kk on the tile and its one-node halo
        const int sl = (kk + 3) % 3;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int q = tid + 256 * s, hx = q % RX, hy = q / RX;
            if (q >= RX * RY || hx < 1 || hx > TX + 2 || hy < 1 || hy > TY + 2) continue;
            const int gi = i0 - 2 + hx, gj = j0 - 2 + hy;
            double u = 0.0, v = 0.0, w = 0.0;
            if (gi >= 1 && gi <= nx - 2 && gj >= 1 && gj <= ny - 2 && kk >= 1 && kk <= nz - 2) {
                double dpsidx1 = (P[1][hy][hx + 1] - P[1][hy][hx - 1]) * ix2;
                double dpsidx2 = (P[2][hy][hx + 1] - P[2][hy][hx - 1]) * ix2;
                double dpsidy0 = (P[0][hy + 1][hx] - P[0][hy - 1][hx]) * iy2;
                double dpsidy2 = (P[2][hy + 1][hx] - P[2][hy - 1][hx]) * iy2;
                double dpsidz0 = (w0[s][2] - w0[s][0]) * iz2, dpsidz1 = (w1[s][2] - w1[s][0]) * iz2;
                u = dpsidy2 - dpsidz1;
                v = -(dpsidx2 - dpsidz0);
                w = dpsidx1 - dpsidy0;
            }
            V[sl][0][hy - 1][hx - 1] = u;
            V[sl][1][hy - 1][hx - 1] = v;
            V[sl][2][hy - 1][hx - 1] = w;
        }
        __syncthreads();
        // stretching of plane k = kk - 1 (its velocity neighbours kk - 2 .. kk are in the ring)
        const int k = kk - 1;
        r2[0] = r2[1];
        r2[1] = r2[2];
        r2[2] = ldr2(k + 1);
        if (k >= kb && k < kend && tin) {
            const int s0 = k % 3, sp = (k + 1) % 3, sm = (k + 2) % 3;
            const int64_t ov = vel.at(i, j, k);
            for (int c = 0; c < 3; c++) vel.p[c * vel.cs + ov] = V[s0][c][ty + 1][tx + 1];
            double r[3] = {0.0, 0.0, 0.0};
            if (sint && k >= 2 && k <= nz - 3) {
                const int64_t ro = rc + (int64_t)k * rhs.pz;
                const double a0p = R0[ro + 1], a0m = R0[ro - 1], a1p = R1[ro + rhs.py], a1m = R1[ro - rhs.py];
                const double a2p = r2[2], a2m = r2[0];
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    double wdudx = -(a0p * V[s0][m][ty + 1][tx + 2] - (a0m)*V[s0][m][ty + 1][tx]) * ix2;
                    double wdvdy = -(a1p * V[s0][m][ty + 2][tx + 1] - (a1m)*V[s0][m][ty][tx + 1]) * iy2;
                    double wdwdz = -(a2p * V[sp][m][ty + 1][tx + 1] - (a2m)*V[sm][m][ty + 1][tx + 1]) * iz2;
                    r[m] = wdudx + wdvdy + wdwdz;
                }
            }
            const int64_t oo = out.at(i, j, k);
            out.p[oo] = r[0];
            out.p[out.cs + oo] = r[1];
            out.p[2 * out.cs + oo] = r[2];
        }
    }
}

int curl_stretch_kernel(Ctx *c, const int n[3], const double h[3], GView psi, GView rhs, GView vel, GView out) {
    if ((int64_t)n[0] * n[1] * n[2] <= 0) return VPMX_OK;
    static const int shp[5][2] = {{64, 4}, {32, 8}, {128, 2}, {64, 4}, {32, 8}};
    const void *fns[5] = {(const void *)k_curl_stretch<64, 4>, (const void *)k_curl_stretch<32, 8>,
                          (const void *)k_curl_stretch<128, 2>, (const void *)k_curl_stretch_w<64, 4>,
                          (const void *)k_curl_stretch_w<32, 8>};
    const char *names[5] = {"64 x 4 column tiles", "32 x 8 column tiles", "128 x 2 column tiles",
                            "64 x 4 column tiles, stream function once per tile",
                            "32 x 8 column tiles, stream function once per tile"};
    double ix2 = 1.0 / (2 * h[0]), iy2 = 1.0 / (2 * h[1]), iz2 = 1.0 / (2 * h[2]);
    int nx = n[0], ny = n[1], nz = n[2];
    int R = c->num_cu * 4 / 8 * 8;  // about the work-groups resident at once (4 per CU)
    if (R < 8) R = 8;
    void *args[] = {&nx, &ny, &nz, &ix2, &iy2, &iz2, &psi, &rhs, &vel, &out, &R};
    auto grid_of = [&](int fi) {
        return dim3((unsigned)((n[0] + shp[fi][0] - 1) / shp[fi][0]), (unsigned)((n[1] + shp[fi][1] - 1) / shp[fi][1]),
                    (unsigned)((n[2] + CS_KC - 1) / CS_KC));
    };
    // trials write the real outputs: every variant writes the same values to every node
    auto trial = [&](int fi, int) -> int {
        VPMX_HIP(c, hipLaunchKernel(fns[fi], grid_of(fi), dim3(256), args, 0, c->stream));
        return VPMX_OK;
    };
    char key[128];
    std::snprintf(key, sizeof key, "curl and stretching %dx%dx%d", n[0], n[1], n[2]);
    Ctx::Tuned ch;
    VPMX_TRY(tune(c, key, 5, fns, names, 0, 256, trial, &ch, false));
    VPMX_HIP(c, hipLaunchKernel(fns[ch.fi], grid_of(ch.fi), dim3(256), args, 0, c->stream));
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// diffuse_vort_3d (vpm_gcalc.f90:138-210), neqpm = 3: out = -NI * lap(RHS) per component on the
// interior nodes (7-point second differences, the reference's term order), 0 elsewhere.
__global__ void __launch_bounds__(256) k_diffuse(int nx, int ny, int nz, double dx2, double dy2, double dz2,
                                                 double vis, GView rhs, GView out) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    double r[3] = {0.0, 0.0, 0.0};
    if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2) {
        const int64_t c0 = rhs.at(i, j, k);
        const int64_t xp = rhs.at(i + 1, j, k), xm = rhs.at(i - 1, j, k);
        const int64_t yp = rhs.at(i, j + 1, k), ym = rhs.at(i, j - 1, k);
        const int64_t zp = rhs.at(i, j, k + 1), zm = rhs.at(i, j, k - 1);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double *R = rhs.p + m * rhs.cs;
            const double dwxdx = (R[xp] - 2 * R[c0] + R[xm]) / dx2;
            const double dwydy = (R[yp] - 2 * R[c0] + R[ym]) / dy2;
            const double dwzdz = (R[zp] - 2 * R[c0] + R[zm]) / dz2;
            r[m] = -vis * (dwxdx + dwydy + dwzdz);
        }
    }
    const int64_t o = out.at(i, j, k);
    for (int m = 0; m < 3; m++) out.p[m * out.cs + o] = r[m];
}

int diffuse_kernel(Ctx *c, const int n[3], const double h[3], double vis, GView rhs, GView out) {
    int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_diffuse, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, n[0], n[1], n[2],
                       h[0] * h[0], h[1] * h[1], h[2] * h[2], vis, rhs, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// coarse(c, cmap_x[a], cmap_y[b], cmap_z[cc]) = fine(c, a*nc, b*nc, cc*nc)  (0-based)
__global__ void k_coarse_sample(int ncomp, int sx, int sy, int sz, int nc, const int *cmap, GView fine, GView coarse) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)sx * sy * sz) return;
    int a = (int)(t % sx), b = (int)((t / sx) % sy), cc = (int)(t / ((int64_t)sx * sy));
    int ci = cmap[a] - 1, cj = cmap[sx + b] - 1, ck = cmap[sx + sy + cc] - 1;
    int64_t of = fine.at(a * nc, b * nc, cc * nc), oc = coarse.at(ci, cj, ck);
    for (int q = 0; q < ncomp; q++) coarse.p[q * coarse.cs + oc] = fine.p[q * fine.cs + of];
}
int coarse_sample(Ctx *c, int ncomp, const int ns[3], int nc, const int *cmap_dev, GView fine, GView coarse,
                  const int cn[3]) {
    (void)cn;
    int64_t nt = (int64_t)ns[0] * ns[1] * ns[2];
    hipLaunchKernelGGL(k_coarse_sample, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, ncomp, ns[0], ns[1],
                       ns[2], nc, cmap_dev, fine, coarse);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// rhs(i,j,k) for lo < i,j,k < hi (1-based, footprint interior); 0 elsewhere.  accumulate:
// rhs += the Laplacian on the footprint interior, other nodes untouched (the sum over
// blocks of yaps3d.f90:190-195).
__global__ void k_coarse_laplacian(int ncomp, int nx, int ny, int nz, int lx, int ly, int lz, int hx, int hy, int hz,
                                   double c0, double c1, double c2, GView s, GView rhs, int accumulate) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    int i = (int)(t % nx) + 1, j = (int)((t / nx) % ny) + 1, k = (int)(t / ((int64_t)nx * ny)) + 1;
    int64_t o = rhs.at(i - 1, j - 1, k - 1);
    bool in = i > lx && i < hx && j > ly && j < hy && k > lz && k < hz;
    for (int q = 0; q < ncomp; q++) {
        double v = 0.0;
        if (in) {
            const double *S = s.p + q * s.cs;
            double sc = S[s.at(i - 1, j - 1, k - 1)];
            double ax = ((S[s.at(i, j - 1, k - 1)] - 2 * sc) + S[s.at(i - 2, j - 1, k - 1)]) / c0;
            double ay = ((S[s.at(i - 1, j, k - 1)] - 2 * sc) + S[s.at(i - 1, j - 2, k - 1)]) / c1;
            double az = ((S[s.at(i - 1, j - 1, k)] - 2 * sc) + S[s.at(i - 1, j - 1, k - 2)]) / c2;
            v = 0.0 + ((ax + ay) + az);
        }
        if (!accumulate) rhs.p[q * rhs.cs + o] = v;
        else if (in) rhs.p[q * rhs.cs + o] = rhs.p[q * rhs.cs + o] + v;
    }
}
int coarse_laplacian(Ctx *c, int ncomp, const int cn[3], const int lo[3], const int hi[3], const double hc[3],
                     GView samp, GView rhs, bool accumulate) {
    int64_t nt = (int64_t)cn[0] * cn[1] * cn[2];
    hipLaunchKernelGGL(k_coarse_laplacian, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, ncomp, cn[0],
                       cn[1], cn[2], lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], hc[0] * hc[0], hc[1] * hc[1],
                       hc[2] * hc[2], samp, rhs, accumulate ? 1 : 0);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

__device__ __forceinline__ double tsc(double x) {
#pragma clang fp contract(off)
    double a = fabs(x);
    if (a > 1.5) return 0.0;
    if (a >= 0.5) return 0.5 * ((-a + 1.5) * (-a + 1.5));
    return 0.5 * ((a + 1.5) * (a + 1.5)) - 1.5 * ((a + 0.5) * (a + 0.5));
}

// interp_stencil for every node of the six faces of the inner box, in face order.
__global__ void __launch_bounds__(256) k_face_interp(int ncomp, InterpArgs a, int64_t nface, GView csol, GView samp,
                                                      GView sol) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nface) return;
    const int XS = a.nbl[0], YS = a.nbl[1], ZS = a.nbl[2], XF = a.nbl[3], YF = a.nbl[4], ZF = a.nbl[5];
    const int64_t f12 = (int64_t)(YF - YS + 1) * (ZF - ZS + 1), f34 = (int64_t)(XF - XS - 1) * (ZF - ZS + 1);
    int i, j, k;
    if (t < 2 * f12) {
        int f = (int)(t / f12); int64_t r = t % f12;
        i = f ? XF : XS; j = YS + (int)(r % (YF - YS + 1)); k = ZS + (int)(r / (YF - YS + 1));
    } else if (t < 2 * (f12 + f34)) {
        int64_t u = t - 2 * f12; int f = (int)(u / f34); int64_t r = u % f34;
        j = f ? YF : YS; i = XS + 1 + (int)(r % (XF - XS - 1)); k = ZS + (int)(r / (XF - XS - 1));
    } else {
        int64_t u = t - 2 * (f12 + f34); int64_t f56 = (int64_t)(XF - XS - 1) * (YF - YS - 1);
        int f = (int)(u / f56); int64_t r = u % f56;
        k = f ? ZF : ZS; i = XS + 1 + (int)(r % (XF - XS - 1)); j = YS + 1 + (int)(r / (XF - XS - 1));
    }
    double X0 = a.xbl[0] + (double)(i - 1) * a.h[0];
    double X1 = a.xbl[1] + (double)(j - 1) * a.h[1];
    double X2 = a.xbl[2] + (double)(k - 1) * a.h[2];
    int in = (int)round((X0 - a.xc[0]) / a.hc[0]) + 1;
    int jn = (int)round((X1 - a.xc[1]) / a.hc[1]) + 1;
    int kn = (int)round((X2 - a.xc[2]) / a.hc[2]) + 1;
    // map_nodes of the block: coarse nodes inside its sampled footprint
    auto inmap = [&](int p, int q, int r) {
        return p >= a.fp_lo[0] && p <= a.fp_hi[0] && q >= a.fp_lo[1] && q <= a.fp_hi[1] && r >= a.fp_lo[2] &&
               r <= a.fp_hi[2];
    };
    int nnb = (!inmap(in - 1, jn - 1, kn - 1) || !inmap(in + 1, jn + 1, kn + 1)) ? 1 : 0;
    int64_t o = sol.at(i - 1, j - 1, k - 1);
    double acc[3];
    for (int q = 0; q < ncomp; q++) acc[q] = sol.p[q * sol.cs + o];
    for (int kc = kn - 1; kc <= kn + 1; kc++)
        for (int jc = jn - 1; jc <= jn + 1; jc++)
            for (int ic = in - 1; ic <= in + 1; ic++) {
                double xc = (X0 - a.xc[0] - (double)(ic - 1) * a.hc[0]) / a.hc[0];
                double yc = (X1 - a.xc[1] - (double)(jc - 1) * a.hc[1]) / a.hc[1];
                double zc = (X2 - a.xc[2] - (double)(kc - 1) * a.hc[2]) / a.hc[2];
                double fw = tsc(xc) * tsc(yc) * tsc(zc);
                bool ok = ic >= 1 && ic <= a.cn[0] && jc >= 1 && jc <= a.cn[1] && kc >= 1 && kc <= a.cn[2];
                int64_t oc = ok ? csol.at(ic - 1, jc - 1, kc - 1) : 0;
                for (int q = 0; q < ncomp; q++) {
                    double cs = ok ? csol.p[q * csol.cs + oc] : 0.0;
                    double sm = ok ? samp.p[q * samp.cs + oc] : 0.0;
                    double add_sample = sm + 0.0;
                    double add = nnb ? sm + 0.0 : 0.0;
                    acc[q] = acc[q] + fw * ((cs - add_sample) + add);
                }
            }
    for (int q = 0; q < ncomp; q++) sol.p[q * sol.cs + o] = acc[q];
}
int face_interp(Ctx *c, int ncomp, const InterpArgs &a, GView csol, GView samp, GView sol) {
    int64_t nx = a.nbl[3] - a.nbl[0] + 1, ny = a.nbl[4] - a.nbl[1] + 1, nz = a.nbl[5] - a.nbl[2] + 1;
    int64_t nface = 2 * (ny * nz + (nx - 2) * nz + (nx - 2) * (ny - 2));
    hipLaunchKernelGGL(k_face_interp, dim3((unsigned)((nface + 255) / 256)), dim3(256), 0, c->stream, ncomp, a, nface,
                       csol, samp, sol);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// ---------------------------------------------------------------- several yaps blocks
// Node t of the six faces of an inner box [XS..XF] x [YS..YF] x [ZS..ZF] (1-based), in the
// order of yaps3d.f90:320-395: x faces full, y faces without the x edges, z faces without
// both.
__device__ __forceinline__ void inner_face_node(const int nbl[6], int64_t t, int *i, int *j, int *k) {
    const int XS = nbl[0], YS = nbl[1], ZS = nbl[2], XF = nbl[3], YF = nbl[4], ZF = nbl[5];
    const int64_t f12 = (int64_t)(YF - YS + 1) * (ZF - ZS + 1), f34 = (int64_t)(XF - XS - 1) * (ZF - ZS + 1);
    if (t < 2 * f12) {
        int f = (int)(t / f12); int64_t r = t % f12;
        *i = f ? XF : XS; *j = YS + (int)(r % (YF - YS + 1)); *k = ZS + (int)(r / (YF - YS + 1));
    } else if (t < 2 * (f12 + f34)) {
        int64_t u = t - 2 * f12; int f = (int)(u / f34); int64_t r = u % f34;
        *j = f ? YF : YS; *i = XS + 1 + (int)(r % (XF - XS - 1)); *k = ZS + (int)(r / (XF - XS - 1));
    } else {
        int64_t u = t - 2 * (f12 + f34); int64_t f56 = (int64_t)(XF - XS - 1) * (YF - YS - 1);
        int f = (int)(u / f56); int64_t r = u % f56;
        *k = f ? ZF : ZS; *i = XS + 1 + (int)(r % (XF - XS - 1)); *j = YS + 1 + (int)(r / (XF - XS - 1));
    }
}

// The block of face node t of the list over all blocks (blocks in order, f0 ascending).
__device__ __forceinline__ int face_block(const MbBlock *__restrict__ blk, int nblocks, int64_t t) {
    int b = nblocks - 1;
    while (b > 0 && t < blk[b].f0) b--;
    return b;
}

// interp_stencil (yaps3d.f90:1914-1983) for the face nodes of every block in one launch.
// Block b's arrays: fine + blk[b].off (3 components of its nodes, SoA), samp + b * sstride
// (3 x coarse grid).  fine holds every block's infinite-domain solution as it is before
// this step; the new face values go to out (read nothing that this launch writes).
__global__ void __launch_bounds__(256) k_face_interp_mb(MbArgs a, const MbBlock *__restrict__ blk, GView csol,
                                                         const double *__restrict__ samp,
                                                         const double *__restrict__ fine, double *__restrict__ out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntot) return;
    const int b = face_block(blk, a.nblocks, t);
    const MbBlock &B = blk[b];
    const int64_t tl = t - B.f0;
    int i, j, k;
    inner_face_node(B.nbl, tl, &i, &j, &k);
    const double X0 = B.xbl[0] + (double)(i - 1) * a.h[0];
    const double X1 = B.xbl[1] + (double)(j - 1) * a.h[1];
    const double X2 = B.xbl[2] + (double)(k - 1) * a.h[2];
    const int in = (int)round((X0 - a.xc[0]) / a.hc[0]) + 1;
    const int jn = (int)round((X1 - a.xc[1]) / a.hc[1]) + 1;
    const int kn = (int)round((X2 - a.xc[2]) / a.hc[2]) + 1;
    // nnb: the block's sampled footprint lacks a corner of the 3x3x3 coarse stencil
    uint64_t nnb = 0;
    for (int bc = 0; bc < a.nblocks; bc++) {
        const MbBlock &C = blk[bc];
        const bool lo = in - 1 >= C.fp_lo[0] && in - 1 <= C.fp_hi[0] && jn - 1 >= C.fp_lo[1] && jn - 1 <= C.fp_hi[1] &&
                        kn - 1 >= C.fp_lo[2] && kn - 1 <= C.fp_hi[2];
        const bool hi = in + 1 >= C.fp_lo[0] && in + 1 <= C.fp_hi[0] && jn + 1 >= C.fp_lo[1] && jn + 1 <= C.fp_hi[1] &&
                        kn + 1 >= C.fp_lo[2] && kn + 1 <= C.fp_hi[2];
        if (!lo || !hi) nnb |= (uint64_t)1 << bc;
    }
    const int64_t B3 = (int64_t)B.n[0] * B.n[1] * B.n[2];
    const int64_t o = ((int64_t)(k - 1) * B.n[1] + (j - 1)) * B.n[0] + (i - 1);
    const double *own = fine + B.off;
    double acc[3] = {own[o], own[B3 + o], own[2 * B3 + o]};
    for (int kc = kn - 1; kc <= kn + 1; kc++)
        for (int jc = jn - 1; jc <= jn + 1; jc++)
            for (int ic = in - 1; ic <= in + 1; ic++) {
                const double xc = (X0 - a.xc[0] - (double)(ic - 1) * a.hc[0]) / a.hc[0];
                const double yc = (X1 - a.xc[1] - (double)(jc - 1) * a.hc[1]) / a.hc[1];
                const double zc = (X2 - a.xc[2] - (double)(kc - 1) * a.hc[2]) / a.hc[2];
                const double fw = tsc(xc) * tsc(yc) * tsc(zc);
                const bool ok = ic >= 1 && ic <= a.cn[0] && jc >= 1 && jc <= a.cn[1] && kc >= 1 && kc <= a.cn[2];
                const int64_t oc = ok ? csol.at(ic - 1, jc - 1, kc - 1) : 0;
                for (int q = 0; q < 3; q++) {
                    double add_sample = 0.0, add = 0.0;
                    for (int bc = 0; bc < a.nblocks; bc++) {
                        const double sm = ok ? samp[(int64_t)bc * a.sstride + q * csol.cs + oc] : 0.0;
                        add_sample = sm + add_sample;
                        if ((nnb >> bc) & 1) add = sm + add;
                    }
                    const double cs = ok ? csol.p[q * csol.cs + oc] : 0.0;
                    acc[q] = acc[q] + fw * ((cs - add_sample) + add);
                }
            }
    // neighbours whose footprint covers the stencil: their fine solution at the node (BBound)
    for (int kb = B.bk - 1; kb <= B.bk + 1; kb++)
        for (int jb = B.bj - 1; jb <= B.bj + 1; jb++)
            for (int ib = B.bi - 1; ib <= B.bi + 1; ib++) {
                if (ib < 1 || ib > a.nbi || jb < 1 || jb > a.nbj || kb < 1 || kb > a.nbk) continue;
                const int bc = (kb - 1) * a.nbi * a.nbj + (jb - 1) * a.nbi + ib - 1;
                if (((nnb >> bc) & 1) || bc == b) continue;
                const MbBlock &C = blk[bc];
                const int ii = i + (int)round((B.xbl[0] - C.xbl[0]) / a.h[0]);
                const int jj = j + (int)round((B.xbl[1] - C.xbl[1]) / a.h[1]);
                const int kk = k + (int)round((B.xbl[2] - C.xbl[2]) / a.h[2]);
                if (ii < 1 || ii > C.n[0] || jj < 1 || jj > C.n[1] || kk < 1 || kk > C.n[2]) continue;
                const double *fc = fine + C.off;
                const int64_t C3 = (int64_t)C.n[0] * C.n[1] * C.n[2];
                const int64_t on = ((int64_t)(kk - 1) * C.n[1] + (jj - 1)) * C.n[0] + (ii - 1);
                for (int q = 0; q < 3; q++) acc[q] = acc[q] + fc[q * C3 + on];
            }
    for (int q = 0; q < 3; q++) out[3 * t + q] = acc[q];
}

__global__ void k_face_store_mb(MbArgs a, const MbBlock *__restrict__ blk, const double *__restrict__ in_vals,
                                double *__restrict__ sol) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ntot) return;
    const MbBlock &B = blk[face_block(blk, a.nblocks, t)];
    int i, j, k;
    inner_face_node(B.nbl, t - B.f0, &i, &j, &k);
    const int64_t B3 = (int64_t)B.n[0] * B.n[1] * B.n[2];
    const int64_t o = ((int64_t)(k - 1) * B.n[1] + (j - 1)) * B.n[0] + (i - 1);
    double *sb = sol + B.off;
    for (int q = 0; q < 3; q++) sb[q * B3 + o] = in_vals[3 * t + q];
}

int face_interp_blocks(Ctx *c, const MbArgs &a, const MbBlock *blk_dev, GView csol, const double *samp, double *sol,
                       double *tmp) {
    const int64_t nt = a.ntot;
    if (nt <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_face_interp_mb, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, a, blk_dev, csol,
                       samp, sol, tmp);
    hipLaunchKernelGGL(k_face_store_mb, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, a, blk_dev, tmp,
                       sol);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// rhsscat_3d (vpm_mpi.f90:199-238): block = global window (global node g0 + block inner index
// offset) on the inner box, 0 elsewhere and on the min faces listed in zmin.
__global__ void k_rhsscat(int nx, int ny, int nz, Box6 nbl_, int gx, int gy, int gz, int zx, int zy, int zz,
                          GView glob, GView blk) {
    const int *nbl = nbl_.v;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx) + 1, j = (int)((t / nx) % ny) + 1, k = (int)(t / ((int64_t)nx * ny)) + 1;
    const bool in = i >= nbl[0] && i <= nbl[3] && j >= nbl[1] && j <= nbl[4] && k >= nbl[2] && k <= nbl[5];
    const bool zero = !in || (zx && i == nbl[0]) || (zy && j == nbl[1]) || (zz && k == nbl[2]);
    const int64_t o = blk.at(i - 1, j - 1, k - 1);
    const int64_t og = zero ? 0 : glob.at(gx + (i - nbl[0]) - 1, gy + (j - nbl[1]) - 1, gz + (k - nbl[2]) - 1);
    for (int q = 0; q < 3; q++) blk.p[q * blk.cs + o] = zero ? 0.0 : glob.p[q * glob.cs + og];
}

// solget_3d (vpm_mpi.f90:240-328) for one block: its inner box into the global grid.
__global__ void k_solget(int wx, int wy, int wz, Box6 nbl_, int gx, int gy, int gz, GView blk, GView glob) {
    const int *nbl = nbl_.v;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)wx * wy * wz) return;
    const int a = (int)(t % wx), b = (int)((t / wx) % wy), cc = (int)(t / ((int64_t)wx * wy));
    const int64_t o = blk.at(nbl[0] - 1 + a, nbl[1] - 1 + b, nbl[2] - 1 + cc);
    const int64_t og = glob.at(gx - 1 + a, gy - 1 + b, gz - 1 + cc);
    for (int q = 0; q < 3; q++) glob.p[q * glob.cs + og] = blk.p[q * blk.cs + o];
}

int rhsscat_block(Ctx *c, const int n[3], const int nbl[6], const int g0[3], const int zmin[3], GView glob,
                  GView blk) {
    Box6 bx;
    for (int d = 0; d < 6; d++) bx.v[d] = nbl[d];
    const int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_rhsscat, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, n[0], n[1], n[2], bx,
                       g0[0], g0[1], g0[2], zmin[0], zmin[1], zmin[2], glob, blk);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

int solget_block(Ctx *c, const int nbl[6], const int g0[3], GView blk, GView glob) {
    Box6 bx;
    for (int d = 0; d < 6; d++) bx.v[d] = nbl[d];
    const int w[3] = {nbl[3] - nbl[0] + 1, nbl[4] - nbl[1] + 1, nbl[5] - nbl[2] + 1};
    const int64_t nt = (int64_t)w[0] * w[1] * w[2];
    hipLaunchKernelGGL(k_solget, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, w[0], w[1], w[2], bx,
                       g0[0], g0[1], g0[2], blk, glob);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// dst[c + ncd*(i + nx*(j + ny*k))] = src(c,i,j,k) for c < ncs; extra component = fill_last.
__global__ void k_soa_to_aos(int ncs, int ncd, int nx, int ny, int nz, GView src, double fl, double *dst) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    int64_t o = src.at(i, j, k);
    for (int q = 0; q < ncd; q++) dst[(int64_t)ncd * t + q] = q < ncs ? src.p[q * src.cs + o] : fl;
}
int soa_to_aos(Ctx *c, int ncs, int ncd, const int n[3], GView src, double fl, double *dst) {
    int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, ncs, ncd, n[0], n[1],
                       n[2], src, fl, dst);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
__global__ void k_aos_to_soa(int nc, int nx, int ny, int nz, const double *src, GView dst) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    int64_t o = dst.at(i, j, k);
    for (int q = 0; q < nc; q++) dst.p[q * dst.cs + o] = src[(int64_t)nc * t + q];
}
int aos_to_soa(Ctx *c, int nc, const int n[3], const double *src, GView dst) {
    int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, nc, n[0], n[1], n[2],
                       src, dst);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

}  // namespace vpmx
