// Adjoint of one vpm(WhatToDo=2) velocity evaluation (vpm.f90:230-313) for one yaps block:
// the gradient of J = 1/2 sum_k |u(p_k)|^2 -- u the velocity grid interpolated at probe
// points p_k -- with respect to the particle strengths QP(1:3) and positions XP.
//
// Forward (every stage linear in the strengths):
//   w = P(x) q / DV                  projection (pmproject.f90:43-153)
//   s0 = D0_B(w)                     zero-BC solve on the block (pmsolve.f90:367-432)
//   a = A s0                         face-cell sources: 4-corner 5-point normal derivatives,
//                                    times -area/(4 pi) (pinfdomain.f90:294-547)
//   g = BS a                         boundary sums on the block faces (pmbound.f90:415-515)
//   phi = D_B(w; g)                  Dirichlet solve with those face values
//   smp = Samp phi, r_c = Lc smp     coarse samples and footprint Laplacian (yaps3d.f90:104-196)
//   psi_c = Pinf_C(r_c)              the coarse infinite-domain solve (same chain)
//   f = phi|faces + T (psi_c - smp)  TSC face correction (yaps3d.f90:1914-1983)
//   psi = D_I(w; f)                  inner-box solve
//   u = curl psi                     central differences (vpm_gcalc.f90:1-60)
//   u(p_k) = M u                     interpolation at the probes
// Adjoint: every stage transposed.  The sine-transform solve D0 with zero boundary values is
// symmetric; a solve with face values g is D(w; g) = D0(w - F g) inside and g on the faces (F
// folds the face values into the neighbouring interior nodes, times 1/h^2), so its transpose
// takes an adjoint y on the grid to D0(y_inside) for w and y_faces - F^T D0(y_inside) for g.
// The transposed boundary sums are the same 1/r sums from face nodes to face-cell centres;
// the transposed stencils (normal derivatives, TSC, coarse Laplacian, curl) are gathers or
// scatters of the forward weights.  FP64 throughout; the scatters use FP64 atomics (the
// gradient's last bits depend on their order).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

// ---------------------------------------------------------------- curl^T
// Forward (on nodes 1..n-2 of each axis, 0-based; 0 elsewhere):
//   u_x = (psi_z(j+1) - psi_z(j-1))/(2 dy) - (psi_y(k+1) - psi_y(k-1))/(2 dz), and cyclic.
// Transpose: psibar_z(j) += (ubar_x(j-1) - ubar_x(j+1))/(2 dy) ... with ubar zero outside the
// nodes where u is formed.
__global__ void k_curl_adj(int nx, int ny, int nz, double hx, double hy, double hz, GView ub, GView pb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    auto U = [&](int c, int a, int b, int d) -> double {
        if (a < 1 || a > nx - 2 || b < 1 || b > ny - 2 || d < 1 || d > nz - 2) return 0.0;
        return ub.p[c * ub.cs + ub.at(a, b, d)];
    };
    // u_x: +dpsi_z/dy - dpsi_y/dz ; u_y: +dpsi_x/dz - dpsi_z/dx ; u_z: +dpsi_y/dx - dpsi_x/dy
    const double px = (U(1, i, j, k - 1) - U(1, i, j, k + 1)) / (2.0 * hz) - (U(2, i, j - 1, k) - U(2, i, j + 1, k)) / (2.0 * hy);
    const double py = (U(2, i - 1, j, k) - U(2, i + 1, j, k)) / (2.0 * hx) - (U(0, i, j, k - 1) - U(0, i, j, k + 1)) / (2.0 * hz);
    const double pz = (U(0, i, j - 1, k) - U(0, i, j + 1, k)) / (2.0 * hy) - (U(1, i - 1, j, k) - U(1, i + 1, j, k)) / (2.0 * hx);
    const int64_t o = pb.at(i, j, k);
    pb.p[o] = px;
    pb.p[pb.cs + o] = py;
    pb.p[2 * pb.cs + o] = pz;
}

// ---------------------------------------------------------------- F^T (face values)
// out = psibar on the faces of an n-node window minus y(inner neighbour along the normal)/h^2
// for the non-edge face nodes; 0 inside.
__global__ void k_fold_adj(int nx, int ny, int nz, double ihx, double ihy, double ihz, GView pb, GView y, GView out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    const bool fx = i == 0 || i == nx - 1, fy = j == 0 || j == ny - 1, fz = k == 0 || k == nz - 1;
    const int64_t o = out.at(i, j, k);
    for (int c = 0; c < 3; c++) {
        double v = 0.0;
        if (fx || fy || fz) {
            v = pb.p[c * pb.cs + pb.at(i, j, k)];
            const double *Y = y.p + c * y.cs;
            if (fx && !fy && !fz) v = v - Y[y.at(i == 0 ? 1 : nx - 2, j, k)] * ihx;
            if (fy && !fx && !fz) v = v - Y[y.at(i, j == 0 ? 1 : ny - 2, k)] * ihy;
            if (fz && !fx && !fy) v = v - Y[y.at(i, j, k == 0 ? 1 : nz - 2)] * ihz;
        }
        out.p[c * out.cs + o] = v;
    }
}

// ---------------------------------------------------------------- BS^T
// abar(c, s) = sum over every face node t of gbar(c, t) / |x_t - x_s|, sources s at the face
// cells' centroids (cxyz of the forward); the face nodes are read from the grid array gbar.
constexpr int BSA_TILE = 256;
__global__ void __launch_bounds__(256) k_bsum_adj(Grid g, int64_t ns, const double *__restrict__ cxyz, GView gb,
                                                  int64_t nt, double *__restrict__ abar) {
    __shared__ double tx[BSA_TILE], ty[BSA_TILE], tz[BSA_TILE], t0[BSA_TILE], t1[BSA_TILE], t2[BSA_TILE];
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t fyz = (int64_t)NY * NZ, fxz = (int64_t)(NX - 2) * NZ, fxy = (int64_t)(NX - 2) * (NY - 2);
    double sx = 0, sy = 0, sz = 0;
    if (s < ns) {
        sx = cxyz[s];
        sy = cxyz[ns + s];
        sz = cxyz[2 * ns + s];
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t base = 0; base < nt; base += BSA_TILE) {
        __syncthreads();
        const int64_t tt = base + threadIdx.x;
        if (tt < nt) {
            // face node tt: X faces full, Y faces without the x-edges, Z faces without both
            int i, j, k;
            if (tt < 2 * fyz) {
                const int f = (int)(tt / fyz);
                const int64_t r = tt % fyz;
                i = f ? NX - 1 : 0; j = (int)(r % NY); k = (int)(r / NY);
            } else if (tt < 2 * (fyz + fxz)) {
                const int64_t u = tt - 2 * fyz;
                const int f = (int)(u / fxz);
                const int64_t r = u % fxz;
                j = f ? NY - 1 : 0; i = 1 + (int)(r % (NX - 2)); k = (int)(r / (NX - 2));
            } else {
                const int64_t u = tt - 2 * (fyz + fxz);
                const int f = (int)(u / fxy);
                const int64_t r = u % fxy;
                k = f ? NZ - 1 : 0; i = 1 + (int)(r % (NX - 2)); j = 1 + (int)(r / (NX - 2));
            }
            tx[threadIdx.x] = g.x0[0] + (double)i * g.h[0];
            ty[threadIdx.x] = g.x0[1] + (double)j * g.h[1];
            tz[threadIdx.x] = g.x0[2] + (double)k * g.h[2];
            const int64_t o = gb.at(i, j, k);
            t0[threadIdx.x] = gb.p[o];
            t1[threadIdx.x] = gb.p[gb.cs + o];
            t2[threadIdx.x] = gb.p[2 * gb.cs + o];
        } else {
            tx[threadIdx.x] = ty[threadIdx.x] = tz[threadIdx.x] = 1e30;
            t0[threadIdx.x] = t1[threadIdx.x] = t2[threadIdx.x] = 0.0;
        }
        __syncthreads();
        const int m = (int)((nt - base) < BSA_TILE ? (nt - base) : BSA_TILE);
        for (int q = 0; q < m; q++) {
            const double dx = tx[q] - sx, dy = ty[q] - sy, dz = tz[q] - sz;
            const double ir = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            acc[0] += t0[q] * ir;
            acc[1] += t1[q] * ir;
            acc[2] += t2[q] * ir;
        }
    }
    if (s < ns)
        for (int c = 0; c < 3; c++) abar[c * ns + s] = acc[c];
}

// ---------------------------------------------------------------- A^T
// Forward: a(c, s) = -(ds/4pi) * 1/4 sum_corners (1/h_n) sum_st coef_st s0(corner, layer st),
// s a face cell, its corners nodes of the face plane and st = 0..4 the layers inward.
// Transpose, as a gather: node t of layer st of face f gets coef_st times the sum of
// -(ds/4pi)/4/h_n abar(c, s) over the (up to four) cells s of f it is a corner of, summed over
// the faces whose five layers hold it (near edges, several) in face order -- every node of the
// grid written (zeros outside the layers), no atomics, the same sum on every run.
__global__ void __launch_bounds__(256) k_nderiv_adj(Grid g, int64_t nb, const double *__restrict__ abar,
                                                    const double *__restrict__ ds, GView s0b) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)NX * NY * NZ) return;
    const int crd[3] = {(int)(t % NX) + 1, (int)((t / NX) % NY) + 1, (int)(t / ((int64_t)NX * NY)) + 1};
    const int64_t fX = (int64_t)(NY - 1) * (NZ - 1), fY = (int64_t)(NX - 1) * (NZ - 1), fZ = (int64_t)(NX - 1) * (NY - 1);
    const double coef[5] = {25.0 / 12.0, -4.0, 3.0, -4.0 / 3.0, 0.25};
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    const int du[4] = {0, 1, 1, 0}, dv[4] = {0, 0, 1, 1};
    double acc[3] = {0.0, 0.0, 0.0};
    for (int face = 0; face < 6; face++) {
        const int axis = face / 2, hi = face & 1;
        const int st = hi ? g.n[axis] - crd[axis] : crd[axis] - 1;
        if (st < 0 || st > 4) continue;
        // the face's in-plane node coordinates (the forward's i0 + du, i1 + dv) and counts
        const int a = axis == 0 ? crd[1] : crd[0], b = axis == 2 ? crd[1] : crd[2];
        const int N0 = axis == 0 ? NY : NX, N1 = axis == 2 ? NY : NZ;
        const int64_t base = axis == 0 ? (int64_t)face * fX
                             : axis == 1 ? 2 * fX + (int64_t)(face - 2) * fY
                                         : 2 * (fX + fY) + (int64_t)(face - 4) * fZ;
        for (int q = 0; q < 4; q++) {
            const int i0 = a - du[q], i1 = b - dv[q];
            if (i0 < 1 || i0 > N0 - 1 || i1 < 1 || i1 > N1 - 1) continue;
            const int64_t s = base + (i0 - 1) + (int64_t)(i1 - 1) * (N0 - 1);
            if (s >= nb) continue;
            const double w = -(ds[s] / PI4) * 0.25 / g.h[axis];
            for (int c = 0; c < 3; c++) acc[c] += w * coef[st] * abar[c * nb + s];
        }
    }
    const int64_t o = s0b.at(crd[0] - 1, crd[1] - 1, crd[2] - 1);
    for (int c = 0; c < 3; c++) s0b.p[c * s0b.cs + o] = acc[c];
}

// ---------------------------------------------------------------- T^T
// Forward (one block): f(t) += sum_27 w (psi_c(k) - smp(k) + [nnb] smp(k)) at the inner-box
// face nodes t.  Transpose: psibar_c(k) += w fbar(t), smpbar(k) += w fbar(t) ([nnb] - 1), and
// phibar(t) += fbar(t) (the face value carries the fine solution).
__device__ __forceinline__ double tsc_w(double x) {
    double a = fabs(x);
    if (a > 1.5) return 0.0;
    if (a >= 0.5) return 0.5 * ((-a + 1.5) * (-a + 1.5));
    return 0.5 * ((a + 1.5) * (a + 1.5)) - 1.5 * ((a + 0.5) * (a + 0.5));
}
// Gather form: one wave per coarse node k sums w(t, k) fbar(t) over the inner-box face nodes t
// whose stencil (in(t) +- 1 on every axis) holds k -- a window of about 3 ncoarse fine nodes per
// in-plane axis on each face within one coarse cell of k.  The lanes split each face's window
// and a fixed butterfly adds their sums, so no atomics are needed and the result does not
// depend on scheduling.
constexpr int FIA_WAVES = 4;
__global__ void __launch_bounds__(64 * FIA_WAVES) k_face_interp_adj(InterpArgs a, GView fb, GView cb, GView sb) {
    const int64_t t = (int64_t)blockIdx.x * FIA_WAVES + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= (int64_t)a.cn[0] * a.cn[1] * a.cn[2]) return;  // whole waves
    const int cc[3] = {(int)(t % a.cn[0]) + 1, (int)((t / a.cn[0]) % a.cn[1]) + 1,
                       (int)(t / ((int64_t)a.cn[0] * a.cn[1])) + 1};
    auto inmap = [&](int p, int q, int r) {
        return p >= a.fp_lo[0] && p <= a.fp_hi[0] && q >= a.fp_lo[1] && q <= a.fp_hi[1] && r >= a.fp_lo[2] &&
               r <= a.fp_hi[2];
    };
    auto cnode = [&](int d, int i) {  // nearest coarse node (1-based) of fine node i on axis d
        const double X = a.xbl[d] + (double)(i - 1) * a.h[d];
        return (int)round((X - a.xc[d]) / a.hc[d]) + 1;
    };
    double acc_c[3] = {0, 0, 0}, acc_s[3] = {0, 0, 0};
    for (int f = 0; f < 6; f++) {
        const int ax = f >> 1, P = (f & 1) ? a.nbl[ax + 3] : a.nbl[ax];
        if (abs(cnode(ax, P) - cc[ax]) > 1) continue;
        const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
        // face node ranges: X faces whole, Y faces without the x-edges, Z faces without x and y
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = a.nbl[d];
            hi[d] = a.nbl[d + 3];
        }
        if (ax >= 1) { lo[0]++; hi[0]--; }
        if (ax == 2) { lo[1]++; hi[1]--; }
        int r0[3], r1[3];
        for (int d : {u, v}) {
            const double f0 = (a.xc[d] - a.xbl[d] + ((double)cc[d] - 2.5) * a.hc[d]) / a.h[d] + 1.0;
            const double f1 = (a.xc[d] - a.xbl[d] + ((double)cc[d] + 0.5) * a.hc[d]) / a.h[d] + 1.0;
            r0[d] = max(lo[d], (int)floor(f0) - 1);
            r1[d] = min(hi[d], (int)ceil(f1) + 1);
        }
        const int nu = r1[u] - r0[u] + 1, nv = r1[v] - r0[v] + 1;
        if (nu <= 0 || nv <= 0) continue;
        for (int q0 = lane; q0 < nu * nv; q0 += 64) {
            int n3[3];
            n3[ax] = P;
            n3[u] = r0[u] + q0 % nu;
            n3[v] = r0[v] + q0 / nu;
            const int in = cnode(0, n3[0]), jn = cnode(1, n3[1]), kn = cnode(2, n3[2]);
            if (abs(in - cc[0]) > 1 || abs(jn - cc[1]) > 1 || abs(kn - cc[2]) > 1) continue;
            const double X0 = a.xbl[0] + (double)(n3[0] - 1) * a.h[0];
            const double X1 = a.xbl[1] + (double)(n3[1] - 1) * a.h[1];
            const double X2 = a.xbl[2] + (double)(n3[2] - 1) * a.h[2];
            const double xc = (X0 - a.xc[0] - (double)(cc[0] - 1) * a.hc[0]) / a.hc[0];
            const double yc = (X1 - a.xc[1] - (double)(cc[1] - 1) * a.hc[1]) / a.hc[1];
            const double zc = (X2 - a.xc[2] - (double)(cc[2] - 1) * a.hc[2]) / a.hc[2];
            const double fw = tsc_w(xc) * tsc_w(yc) * tsc_w(zc);
            const int nnb = (!inmap(in - 1, jn - 1, kn - 1) || !inmap(in + 1, jn + 1, kn + 1)) ? 1 : 0;
            const int64_t o = fb.at(n3[0] - 1, n3[1] - 1, n3[2] - 1);
            for (int q = 0; q < 3; q++) {
                const double fv = fb.p[q * fb.cs + o];
                acc_c[q] += fw * fv;
                acc_s[q] += fw * fv * (double)(nnb - 1);
            }
        }
    }
    for (int q = 0; q < 3; q++)
        for (int m = 32; m >= 1; m >>= 1) {
            acc_c[q] += __shfl_xor(acc_c[q], m, 64);
            acc_s[q] += __shfl_xor(acc_s[q], m, 64);
        }
    if (lane != 0) return;
    const int64_t oc = cb.at(cc[0] - 1, cc[1] - 1, cc[2] - 1), os = sb.at(cc[0] - 1, cc[1] - 1, cc[2] - 1);
    for (int q = 0; q < 3; q++) {
        cb.p[q * cb.cs + oc] += acc_c[q];
        sb.p[q * sb.cs + os] += acc_s[q];
    }
}
// phibar(t) += fbar(t) on the inner-box faces (the face value carries the fine solution).
__global__ void k_face_carry_adj(int XS, int YS, int ZS, int nx, int ny, int nz, GView fb, GView phib) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    if (!(i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1)) return;
    const int64_t o = fb.at(XS - 1 + i, YS - 1 + j, ZS - 1 + k), op = phib.at(XS - 1 + i, YS - 1 + j, ZS - 1 + k);
    for (int q = 0; q < 3; q++) phib.p[q * phib.cs + op] += fb.p[q * fb.cs + o];
}

// ---------------------------------------------------------------- Lc^T, Samp^T
// Forward: r(i) = sum_axes (s(i+e) - 2 s(i) + s(i-e))/h^2 on the footprint interior lo < i < hi.
// Transpose (gather): sbar(m) += sum over interior i in {m, m +- e} of rbar(i) * weight.
__global__ void k_coarse_laplacian_adj(int nx, int ny, int nz, int lx, int ly, int lz, int hx, int hy, int hz,
                                       double c0, double c1, double c2, GView rb, GView sb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx) + 1, j = (int)((t / nx) % ny) + 1, k = (int)(t / ((int64_t)nx * ny)) + 1;
    auto in = [&](int p, int q, int r) { return p > lx && p < hx && q > ly && q < hy && r > lz && r < hz; };
    auto R = [&](int c, int p, int q, int r) { return in(p, q, r) ? rb.p[c * rb.cs + rb.at(p - 1, q - 1, r - 1)] : 0.0; };
    const int64_t o = sb.at(i - 1, j - 1, k - 1);
    for (int c = 0; c < 3; c++) {
        double v = -2.0 * R(c, i, j, k) * (1.0 / c0 + 1.0 / c1 + 1.0 / c2);
        v += (R(c, i + 1, j, k) + R(c, i - 1, j, k)) / c0;
        v += (R(c, i, j + 1, k) + R(c, i, j - 1, k)) / c1;
        v += (R(c, i, j, k + 1) + R(c, i, j, k - 1)) / c2;
        sb.p[c * sb.cs + o] += v;
    }
}
// fine(a*nc, b*nc, cc*nc) += coarse(cmap[a], cmap[b], cmap[cc])
__global__ void k_coarse_sample_adj(int sx, int sy, int sz, int nc, const int *cmap, GView coarse, GView fine) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)sx * sy * sz) return;
    const int a = (int)(t % sx), b = (int)((t / sx) % sy), cc = (int)(t / ((int64_t)sx * sy));
    const int ci = cmap[a] - 1, cj = cmap[sx + b] - 1, ck = cmap[sx + sy + cc] - 1;
    const int64_t of = fine.at(a * nc, b * nc, cc * nc), oc = coarse.at(ci, cj, ck);
    for (int q = 0; q < 3; q++) fine.p[q * fine.cs + of] += coarse.p[q * coarse.cs + oc];
}

// out += in (three components over an n-node view)
__global__ void k_axpy_view(int nx, int ny, int nz, GView in, GView out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    for (int c = 0; c < 3; c++) out.p[c * out.cs + out.at(i, j, k)] += in.p[c * in.cs + in.at(i, j, k)];
}

// ---------------------------------------------------------------- position gradient
// dJ/dx_p(d) = (1/DV) sum_c q_p(c) sum_nodes dW/dx_d(x_p) wbar(c, node), the M4' weights of the
// projection differentiated (W(x) = prod_d w((x_d - X_d)/h_d), w' of the M4' kernel).
__device__ __forceinline__ double m4_w(double x) {
    const double a = fabs(x);
    if (a < 1.0) return 1.0 - 2.5 * a * a + 1.5 * a * a * a;
    if (a <= 2.0) return 0.5 * (2.0 - a) * (2.0 - a) * (1.0 - a);
    return 0.0;
}
__device__ __forceinline__ double m4_dw(double x) {
    const double a = fabs(x), sg = x < 0.0 ? -1.0 : 1.0;
    if (a < 1.0) return sg * (-5.0 * a + 4.5 * a * a);
    if (a <= 2.0) return sg * (-(2.0 - a) * (1.0 - a) - 0.5 * (2.0 - a) * (2.0 - a));
    return 0.0;
}
__global__ void k_grad_x(Grid g, const double *xp, const double *qp, int qs, int64_t n, GView wb, double idv,
                         double *dq, double *dx, int acc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double t[3];
    int c0[3];
    for (int d = 0; d < 3; d++) {
        t[d] = (xp[3 * p + d] - g.x0[d]) / g.h[d];
        c0[d] = (int)floor(t[d]);  // node c0 is the 0-based node at or below the particle
    }
    double gq[3] = {0, 0, 0}, gx[3] = {0, 0, 0};
    for (int kk = c0[2] - 1; kk <= c0[2] + 2; kk++) {
        if (kk < 0 || kk >= g.n[2]) continue;
        const double wz = m4_w(t[2] - kk), dz = m4_dw(t[2] - kk) / g.h[2];
        for (int jj = c0[1] - 1; jj <= c0[1] + 2; jj++) {
            if (jj < 0 || jj >= g.n[1]) continue;
            const double wy = m4_w(t[1] - jj), dy = m4_dw(t[1] - jj) / g.h[1];
            for (int ii = c0[0] - 1; ii <= c0[0] + 2; ii++) {
                if (ii < 0 || ii >= g.n[0]) continue;
                const double wx = m4_w(t[0] - ii), dxw = m4_dw(t[0] - ii) / g.h[0];
                const int64_t o = wb.at(ii, jj, kk);
                double s = 0.0;
                for (int c = 0; c < 3; c++) {
                    const double v = wb.p[c * wb.cs + o];
                    gq[c] += wx * wy * wz * v;
                    s += qp[(int64_t)qs * p + c] * v;
                }
                gx[0] += dxw * wy * wz * s;
                gx[1] += wx * dy * wz * s;
                gx[2] += wx * wy * dz * s;
            }
        }
    }
    for (int d = 0; d < 3; d++) {
        dq[3 * p + d] = (acc ? dq[3 * p + d] : 0.0) + gq[d] * idv;
        dx[3 * p + d] = (acc ? dx[3 * p + d] : 0.0) + gx[d] * idv;
    }
}

// ---------------------------------------------------------------- M2P^T
// Forward (back_to_particles_3D, vpm_time.f90:30-107): UP_p = sum W u, GP_p = vol_p sum W s over
// the 4 x 4 x 4 nodes around x_p, W the M4' weights.  With upstream adjoints
// ubar_p = au * us_p and gbar_p = ag * gs_p: the grid adjoints ubar_grid = sum_p W ubar_p and
// sbar_grid = sum_p W vol_p gbar_p are projections of those particle values (the forward's
// sorted gather P2M), and the position adjoint xbar_p += sum grad W (u . ubar_p + vol_p s .
// gbar_p) is a gather (this kernel, with the scaled particle values written for the P2M).
__global__ void k_m2p_adj_x(Grid g, const double *xp, const double *qp, int qs, int64_t n, const double *us,
                            double au, const double *gs, double ag, GView vel, GView str, double *xacc, double *ubp,
                            double *sbp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double t[3], ub[3], gb[3];
    int c0[3];
    const double vol = qp[(int64_t)qs * p + (qs - 1)];
    for (int d = 0; d < 3; d++) {
        t[d] = (xp[3 * p + d] - g.x0[d]) / g.h[d];
        c0[d] = (int)floor(t[d]);
        ub[d] = au * us[3 * p + d];
        gb[d] = ag * gs[3 * p + d] * vol;
        ubp[3 * p + d] = ub[d];
        sbp[3 * p + d] = gb[d];
    }
    double gx[3] = {0, 0, 0};
    for (int kk = c0[2] - 1; kk <= c0[2] + 2; kk++) {
        if (kk < 0 || kk >= g.n[2]) continue;
        const double wz = m4_w(t[2] - kk), dz = m4_dw(t[2] - kk) / g.h[2];
        for (int jj = c0[1] - 1; jj <= c0[1] + 2; jj++) {
            if (jj < 0 || jj >= g.n[1]) continue;
            const double wy = m4_w(t[1] - jj), dy = m4_dw(t[1] - jj) / g.h[1];
            for (int ii = c0[0] - 1; ii <= c0[0] + 2; ii++) {
                if (ii < 0 || ii >= g.n[0]) continue;
                const double wx = m4_w(t[0] - ii), dxw = m4_dw(t[0] - ii) / g.h[0];
                const int64_t ov = vel.at(ii, jj, kk), os = str.at(ii, jj, kk);
                double sdot = 0.0;
                for (int c = 0; c < 3; c++)
                    sdot += vel.p[c * vel.cs + ov] * ub[c] + str.p[c * str.cs + os] * gb[c];
                gx[0] += dxw * wy * wz * sdot;
                gx[1] += wx * dy * wz * sdot;
                gx[2] += wx * wy * dz * sdot;
            }
        }
    }
    for (int d = 0; d < 3; d++) xacc[3 * p + d] += gx[d];
}
// The scatter form of the same transpose (FP64 atomics), kept for checking the gather form.
__global__ void k_m2p_adj(Grid g, const double *xp, const double *qp, int qs, int64_t n, const double *us, double au,
                          const double *gs, double ag, GView vel, GView str, GView ubg, GView sbg, double *xacc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double t[3], ub[3], gb[3];
    int c0[3];
    const double vol = qp[(int64_t)qs * p + (qs - 1)];
    for (int d = 0; d < 3; d++) {
        t[d] = (xp[3 * p + d] - g.x0[d]) / g.h[d];
        c0[d] = (int)floor(t[d]);
        ub[d] = au * us[3 * p + d];
        gb[d] = ag * gs[3 * p + d] * vol;
    }
    double gx[3] = {0, 0, 0};
    for (int kk = c0[2] - 1; kk <= c0[2] + 2; kk++) {
        if (kk < 0 || kk >= g.n[2]) continue;
        const double wz = m4_w(t[2] - kk), dz = m4_dw(t[2] - kk) / g.h[2];
        for (int jj = c0[1] - 1; jj <= c0[1] + 2; jj++) {
            if (jj < 0 || jj >= g.n[1]) continue;
            const double wy = m4_w(t[1] - jj), dy = m4_dw(t[1] - jj) / g.h[1];
            for (int ii = c0[0] - 1; ii <= c0[0] + 2; ii++) {
                if (ii < 0 || ii >= g.n[0]) continue;
                const double wx = m4_w(t[0] - ii), dxw = m4_dw(t[0] - ii) / g.h[0];
                const double w = wx * wy * wz;
                const int64_t ov = vel.at(ii, jj, kk), os = str.at(ii, jj, kk);
                double sdot = 0.0;
                for (int c = 0; c < 3; c++)
                    sdot += vel.p[c * vel.cs + ov] * ub[c] + str.p[c * str.cs + os] * gb[c];
                gx[0] += dxw * wy * wz * sdot;
                gx[1] += wx * dy * wz * sdot;
                gx[2] += wx * wy * dz * sdot;
                const int64_t ou = ubg.at(ii, jj, kk), osb = sbg.at(ii, jj, kk);
                for (int c = 0; c < 3; c++) {
                    atomicAdd(&ubg.p[c * ubg.cs + ou], w * ub[c]);
                    atomicAdd(&sbg.p[c * sbg.cs + osb], w * gb[c]);
                }
            }
        }
    }
    for (int d = 0; d < 3; d++) xacc[3 * p + d] += gx[d];
}

// ---------------------------------------------------------------- stretching^T
// Forward (interior nodes 2..n-3, 0 elsewhere): s_m = -sum_j (R_j(+e_j) u_m(+e_j) - R_j(-e_j)
// u_m(-e_j)) / (2 h_j).  With D_j sbar(p) = (sbar(p + e_j) - sbar(p - e_j)) / (2 h_j), sbar taken
// as 0 off the interior: Rbar_j(p) = sum_m u_m(p) D_j sbar_m(p) (written to rb) and
// ubar_m(p) += sum_j R_j(p) D_j sbar_m(p).
__global__ void k_stretch_adj(int nx, int ny, int nz, double ix2, double iy2, double iz2, GView R, GView u, GView sb,
                              GView rb, GView ub) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((int64_t)nx * ny));
    auto S = [&](int m, int a, int b, int d) -> double {
        if (a < 2 || a > nx - 3 || b < 2 || b > ny - 3 || d < 2 || d > nz - 3) return 0.0;
        return sb.p[m * sb.cs + sb.at(a, b, d)];
    };
    const int64_t oR = R.at(i, j, k), ou = u.at(i, j, k), orb = rb.at(i, j, k), oub = ub.at(i, j, k);
    double D[3][3];  // D[j][m]
    for (int m = 0; m < 3; m++) {
        D[0][m] = (S(m, i + 1, j, k) - S(m, i - 1, j, k)) * ix2;
        D[1][m] = (S(m, i, j + 1, k) - S(m, i, j - 1, k)) * iy2;
        D[2][m] = (S(m, i, j, k + 1) - S(m, i, j, k - 1)) * iz2;
    }
    double uv[3], Rv[3];
    for (int c = 0; c < 3; c++) {
        uv[c] = u.p[c * u.cs + ou];
        Rv[c] = R.p[c * R.cs + oR];
    }
    for (int jj = 0; jj < 3; jj++) rb.p[jj * rb.cs + orb] = uv[0] * D[jj][0] + uv[1] * D[jj][1] + uv[2] * D[jj][2];
    for (int m = 0; m < 3; m++) ub.p[m * ub.cs + oub] += Rv[0] * D[0][m] + Rv[1] * D[1][m] + Rv[2] * D[2][m];
}

// ---------------------------------------------------------------- remesh^T
// Particles that sit on grid nodes (the output of a remesh): their strength adjoints (3 per
// particle) added to those nodes of out.
__global__ void k_scatter_nodes(Grid g, const double *xp, const double *qb, int64_t n, GView out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int idx[3];
    for (int d = 0; d < 3; d++) idx[d] = (int)llround((xp[3 * p + d] - g.x0[d]) / g.h[d]);
    if (idx[0] < 0 || idx[0] >= g.n[0] || idx[1] < 0 || idx[1] >= g.n[1] || idx[2] < 0 || idx[2] >= g.n[2]) return;
    const int64_t o = out.at(idx[0], idx[1], idx[2]);
    for (int c = 0; c < 3; c++) atomicAdd(&out.p[c * out.cs + o], qb[3 * p + c]);
}

// y += a x over n doubles
__global__ void k_axpy(int64_t n, double a, const double *x, double *y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

}  // namespace vpmx

namespace vpmx {

// ---------------------------------------------------------------- probes
// u(p) = sum W(p - x_node) vel(node) (M4', the 4 x 4 x 4 nodes around p) and its transpose,
// the probe values scattered onto the grid with the same weights.
__global__ void k_probe_interp(Grid g, const double *pts, int64_t np, GView v, double *u) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double t[3];
    int c0[3];
    for (int d = 0; d < 3; d++) {
        t[d] = (pts[3 * p + d] - g.x0[d]) / g.h[d];
        c0[d] = (int)floor(t[d]);
    }
    double acc[3] = {0, 0, 0};
    for (int kk = c0[2] - 1; kk <= c0[2] + 2; kk++) {
        if (kk < 0 || kk >= g.n[2]) continue;
        for (int jj = c0[1] - 1; jj <= c0[1] + 2; jj++) {
            if (jj < 0 || jj >= g.n[1]) continue;
            for (int ii = c0[0] - 1; ii <= c0[0] + 2; ii++) {
                if (ii < 0 || ii >= g.n[0]) continue;
                const double w = m4_w(t[0] - ii) * m4_w(t[1] - jj) * m4_w(t[2] - kk);
                const int64_t o = v.at(ii, jj, kk);
                for (int c = 0; c < 3; c++) acc[c] += w * v.p[c * v.cs + o];
            }
        }
    }
    for (int c = 0; c < 3; c++) u[3 * p + c] = acc[c];
}
__global__ void k_probe_scatter(Grid g, const double *pts, int64_t np, const double *ub, GView out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    double t[3];
    int c0[3];
    for (int d = 0; d < 3; d++) {
        t[d] = (pts[3 * p + d] - g.x0[d]) / g.h[d];
        c0[d] = (int)floor(t[d]);
    }
    for (int kk = c0[2] - 1; kk <= c0[2] + 2; kk++) {
        if (kk < 0 || kk >= g.n[2]) continue;
        for (int jj = c0[1] - 1; jj <= c0[1] + 2; jj++) {
            if (jj < 0 || jj >= g.n[1]) continue;
            for (int ii = c0[0] - 1; ii <= c0[0] + 2; ii++) {
                if (ii < 0 || ii >= g.n[0]) continue;
                const double w = m4_w(t[0] - ii) * m4_w(t[1] - jj) * m4_w(t[2] - kk);
                const int64_t o = out.at(ii, jj, kk);
                for (int c = 0; c < 3; c++) atomicAdd(&out.p[c * out.cs + o], w * ub[3 * p + c]);
            }
        }
    }
}

// ---------------------------------------------------------------- launchers
static unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

int curl_adj(Ctx *c, const int n[3], const double h[3], GView ub, GView pb) {
    const int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_curl_adj, dim3(blocks_of(nt)), dim3(256), 0, c->stream, n[0], n[1], n[2], h[0], h[1], h[2],
                       ub, pb);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int fold_adj(Ctx *c, const int n[3], const double h[3], GView pb, GView y, GView out) {
    const int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_fold_adj, dim3(blocks_of(nt)), dim3(256), 0, c->stream, n[0], n[1], n[2], 1.0 / (h[0] * h[0]),
                       1.0 / (h[1] * h[1]), 1.0 / (h[2] * h[2]), pb, y, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int bsum_adj(Ctx *c, const Grid &g, int64_t ns, const double *cxyz, GView gb, double *abar) {
    const int64_t nt = 2 * ((int64_t)g.n[1] * g.n[2] + (int64_t)(g.n[0] - 2) * g.n[2] +
                            (int64_t)(g.n[0] - 2) * (g.n[1] - 2));
    hipLaunchKernelGGL(k_bsum_adj, dim3(blocks_of(ns)), dim3(256), 0, c->stream, g, ns, cxyz, gb, nt, abar);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int nderiv_adj(Ctx *c, const Grid &g, int64_t nb, const double *abar, const double *ds, GView s0b) {
    const int64_t nt = (int64_t)g.n[0] * g.n[1] * g.n[2];
    hipLaunchKernelGGL(k_nderiv_adj, dim3(blocks_of(nt)), dim3(256), 0, c->stream, g, nb, abar, ds, s0b);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int face_interp_adj(Ctx *c, const InterpArgs &a, GView fb, GView phib, GView cb, GView sb) {
    const int nx = a.nbl[3] - a.nbl[0] + 1, ny = a.nbl[4] - a.nbl[1] + 1, nz = a.nbl[5] - a.nbl[2] + 1;
    const int64_t nc = (int64_t)a.cn[0] * a.cn[1] * a.cn[2], nw = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_face_interp_adj, dim3((unsigned)((nc + FIA_WAVES - 1) / FIA_WAVES)), dim3(64 * FIA_WAVES), 0,
                       c->stream, a, fb, cb, sb);
    VPMX_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_face_carry_adj, dim3(blocks_of(nw)), dim3(256), 0, c->stream, a.nbl[0], a.nbl[1], a.nbl[2], nx,
                       ny, nz, fb, phib);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int coarse_laplacian_adj(Ctx *c, const int cn[3], const int lo[3], const int hi[3], const double hc[3], GView rb,
                         GView sb) {
    const int64_t nt = (int64_t)cn[0] * cn[1] * cn[2];
    hipLaunchKernelGGL(k_coarse_laplacian_adj, dim3(blocks_of(nt)), dim3(256), 0, c->stream, cn[0], cn[1], cn[2], lo[0],
                       lo[1], lo[2], hi[0], hi[1], hi[2], hc[0] * hc[0], hc[1] * hc[1], hc[2] * hc[2], rb, sb);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int coarse_sample_adj(Ctx *c, const int ns[3], int nc, const int *cmap, GView coarse, GView fine) {
    const int64_t nt = (int64_t)ns[0] * ns[1] * ns[2];
    hipLaunchKernelGGL(k_coarse_sample_adj, dim3(blocks_of(nt)), dim3(256), 0, c->stream, ns[0], ns[1], ns[2], nc, cmap,
                       coarse, fine);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int axpy_view(Ctx *c, const int n[3], GView in, GView out) {
    const int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_axpy_view, dim3(blocks_of(nt)), dim3(256), 0, c->stream, n[0], n[1], n[2], in, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int grad_particles(Ctx *c, const Grid &g, const double *xp, const double *qp, int qs, int64_t n, GView wb, double idv,
                   double *dq, double *dx, bool acc) {
    if (n <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_grad_x, dim3(blocks_of(n)), dim3(256), 0, c->stream, g, xp, qp, qs, n, wb, idv, dq, dx,
                       acc ? 1 : 0);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int m2p_adj(Ctx *c, const Grid &g, const double *xp, const double *qp, int qs, int64_t n, const double *us, double au,
            const double *gs, double ag, GView vel, GView str, GView ubg, GView sbg, double *xacc) {
    const int64_t G = (int64_t)g.n[0] * g.n[1] * g.n[2];
    VPMX_TRY(fill(c, ubg.p, 3 * G, 0.0));
    VPMX_TRY(fill(c, sbg.p, 3 * G, 0.0));
    if (n <= 0) return VPMX_OK;
    if (c->adj_m2p_scatter) {
        hipLaunchKernelGGL(k_m2p_adj, dim3(blocks_of(n)), dim3(256), 0, c->stream, g, xp, qp, qs, n, us, au, gs, ag,
                           vel, str, ubg, sbg, xacc);
        VPMX_HIP(c, hipGetLastError());
        return VPMX_OK;
    }
    int st = VPMX_OK;
    double *ubp = (double *)ensure(c, c->adj_pu, sizeof(double) * 3 * (size_t)n, &st); if (st) return st;
    double *sbp = (double *)ensure(c, c->adj_ps, sizeof(double) * 3 * (size_t)n, &st); if (st) return st;
    hipLaunchKernelGGL(k_m2p_adj_x, dim3(blocks_of(n)), dim3(256), 0, c->stream, g, xp, qp, qs, n, us, au, gs, ag, vel,
                       str, xacc, ubp, sbp);
    VPMX_HIP(c, hipGetLastError());
    VPMX_TRY(sort_particles(c, g, xp, n, c->sort));
    const int vn[3] = {g.n[0], g.n[1], g.n[2]}, wlo[3] = {0, 0, 0};
    VPMX_TRY(p2m_gather(c, g, c->sort, xp, ubp, 3, 3, 4, 1.0, ubg, vn, wlo, -1, false, false, false));
    VPMX_TRY(p2m_gather(c, g, c->sort, xp, sbp, 3, 3, 4, 1.0, sbg, vn, wlo, -1, false, false, false));
    return VPMX_OK;
}
int stretch_adj(Ctx *c, const int n[3], const double h[3], GView R, GView u, GView sb, GView rb, GView ub) {
    const int64_t nt = (int64_t)n[0] * n[1] * n[2];
    hipLaunchKernelGGL(k_stretch_adj, dim3(blocks_of(nt)), dim3(256), 0, c->stream, n[0], n[1], n[2],
                       1.0 / (2 * h[0]), 1.0 / (2 * h[1]), 1.0 / (2 * h[2]), R, u, sb, rb, ub);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int scatter_nodes(Ctx *c, const Grid &g, const double *xp, const double *qb, int64_t n, GView out) {
    if (n <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_scatter_nodes, dim3(blocks_of(n)), dim3(256), 0, c->stream, g, xp, qb, n, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int axpy(Ctx *c, int64_t n, double a, const double *x, double *y) {
    if (n <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_axpy, dim3(blocks_of(n)), dim3(256), 0, c->stream, n, a, x, y);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int probe_interp(Ctx *c, const Grid &g, const double *pts, int64_t np, GView v, double *u) {
    if (np <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_probe_interp, dim3(blocks_of(np)), dim3(256), 0, c->stream, g, pts, np, v, u);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}
int probe_scatter(Ctx *c, const Grid &g, const double *pts, int64_t np, const double *ub, GView out) {
    if (np <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_probe_scatter, dim3(blocks_of(np)), dim3(256), 0, c->stream, g, pts, np, ub, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

}  // namespace vpmx
