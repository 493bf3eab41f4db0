// Dirichlet 7-point Poisson solve by sine transforms (replaces the MKL Poisson Library
// call of solve_eq_3d / solve_eq_0_3d, pmsolve.f90:284-432).
//
// The system is the one MKL's d_Helmholtz_3D(q=0,'DDDDDD') solves: for every interior
// node, sum_axes (u_{-} - 2u + u_{+})/h^2 = RHS, u given on the window faces.  Face
// values are folded into the right-hand side, then the interior problem is diagonalised
// by DST-I along x, y and z:  u_hat = rhs_hat / -(lx + ly + lz),
// l_a(k) = 4 sin^2(pi k / (2(m_a+1))) / h_a^2.
//
// Each pass transforms a batch of lines per work-group in LDS.  A DST-I of length m is
// read off the complex FFT of length L = 2(m+1) of the odd extension; two real lines
// share one complex FFT (real and imaginary parts).  The FFT is a mixed-radix Stockham
// autosort (any factorisation of L; radices 2, 3, 4 specialised, others generic).
// The z pass does forward transform, diagonal scaling and inverse transform without
// leaving LDS, so one solve costs 5 sweeps over the interior.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <type_traits>

#include <hip/hip_ext.h>

#include "fft_lds.h"
#include "tune.h"
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

enum { DST_XF = 0, DST_Y = 1, DST_Z = 2, DST_XI = 3 };

struct DstArgs {
    int m;              // line length
    int L;              // FFT length 2(m+1)
    int npairs;         // line pairs per work-group
    int nstages;
    int radix[24];
    int64_t nlines;     // lines per component
    int mx, my, mz;     // interior extents
    int px;             // row pitch of the work array (elements, >= mx)
    int64_t wcs;        // component stride of the work array
    double *w;          // work array (mx,my,mz) per component
    double *wo;         // where the one-wave kernels store their results (normally w)
    const double2 *tw;  // exp(-2 pi i t / L), t < L
    // fold / store
    GView rhs, bc, out;
    int has_bc;
    double ih2[3];
    double scale;
    const double *lam[3];  // eigenvalues (z pass)
    double norm;
    const double *sn;      // sin(j pi / N), j <= N/2 (half-length passes)
    int ncomp;             // components (one-wave kernels: all in one grid)
    int order;             // one-wave kernels' batch order: 0 the two halves of a batch pair on
                           // work-groups w and w+8 (one XCD) at once, 1 one after the other on
                           // one work-group (the second half's loads find the lines in L2)
    // one-wave z pass with zsum set: forward transform and diagonal only (the spectrum stored),
    // and per line sum_k zw[k-1] u_k (low z face) and sum_k (-1)^(k+1) zw[k-1] u_k (high) into
    // zsum[side][comp][ky][kx]
    double *zsum;
    const double *zw;
};

// Mixed-radix Stockham FFT of npairs sequences of length L held in LDS; returns the
// buffer that holds the result.
__device__ double2 *fft_lds(double2 *A, double2 *B, const DstArgs &a) {
    const int L = a.L;
    int Ns = 1;
    for (int s = 0; s < a.nstages; s++) {
        const int R = a.radix[s];
        const int nbf = L / R;
        const int total = a.npairs * nbf;
        const int twstep = L / (Ns * R);
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            const int pair = t / nbf, j = t - pair * nbf;
            const int k0 = j % Ns;
            const double2 *src = A + pair * L;
            double2 *dst = B + pair * L;
            const int idxD = (j / Ns) * Ns * R + k0;
            if (R == 4) {
                double2 v0 = src[j], v1 = src[j + nbf], v2 = src[j + 2 * nbf], v3 = src[j + 3 * nbf];
                if (k0) {
                    v1 = cmul(v1, a.tw[k0 * twstep]);
                    v2 = cmul(v2, a.tw[2 * k0 * twstep]);
                    v3 = cmul(v3, a.tw[3 * k0 * twstep]);
                }
                double2 s02 = cadd(v0, v2), d02 = csub(v0, v2), s13 = cadd(v1, v3), d13 = csub(v1, v3);
                // -i * d13
                double2 md = make_double2(d13.y, -d13.x);
                dst[idxD] = cadd(s02, s13);
                dst[idxD + Ns] = cadd(d02, md);
                dst[idxD + 2 * Ns] = csub(s02, s13);
                dst[idxD + 3 * Ns] = csub(d02, md);
            } else if (R == 2) {
                double2 v0 = src[j], v1 = src[j + nbf];
                if (k0) v1 = cmul(v1, a.tw[k0 * twstep]);
                dst[idxD] = cadd(v0, v1);
                dst[idxD + Ns] = csub(v0, v1);
            } else if (R == 3) {
                double2 v0 = src[j], v1 = src[j + nbf], v2 = src[j + 2 * nbf];
                if (k0) {
                    v1 = cmul(v1, a.tw[k0 * twstep]);
                    v2 = cmul(v2, a.tw[2 * k0 * twstep]);
                }
                const double c3 = -0.5, s3 = 0.86602540378443864676;
                double2 t1 = cadd(v1, v2), t2 = csub(v1, v2);
                double2 base = make_double2(v0.x + c3 * t1.x, v0.y + c3 * t1.y);
                // -i*s3*t2 = (s3*t2.y, -s3*t2.x)
                double2 rot = make_double2(s3 * t2.y, -s3 * t2.x);
                dst[idxD] = cadd(v0, t1);
                dst[idxD + Ns] = cadd(base, rot);
                dst[idxD + 2 * Ns] = csub(base, rot);
            } else {
                const int Lr = L / R;
                for (int k = 0; k < R; k++) {
                    double2 acc = make_double2(0.0, 0.0);
                    for (int r = 0; r < R; r++) {
                        double2 v = src[j + r * nbf];
                        int ti = (r * k0 * twstep + ((r * k) % R) * Lr) % L;
                        acc = cadd(acc, cmul(v, a.tw[ti]));
                    }
                    dst[idxD + k * Ns] = acc;
                }
            }
        }
        __syncthreads();
        double2 *tmp = A;
        A = B;
        B = tmp;
        Ns *= R;
    }
    return A;
}

// Address of point pt (0-based interior index) of line gl in the work array.
__device__ __forceinline__ int64_t line_addr(int mode, const DstArgs &a, int64_t gl, int pt) {
    if (mode == DST_XF || mode == DST_XI) return gl * a.mx + pt;  // lines are rows (j,k)
    if (mode == DST_Y) {                                          // lines (i,k), stride mx
        int64_t i = gl % a.mx, k = gl / a.mx;
        return i + (int64_t)a.mx * ((int64_t)pt + (int64_t)a.my * k);
    }
    return gl + (int64_t)a.mx * a.my * pt;  // z: lines (i,j), stride mx*my
}

template <int MODE>
__global__ void __launch_bounds__(256) k_dst_pass(DstArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    const int L = a.L, m = a.m, nl = 2 * a.npairs;
    double2 *A = lds, *B = lds + a.npairs * L;
    const int comp = blockIdx.y;
    const int64_t l0 = (int64_t)blockIdx.x * nl;
    double *w = a.w + comp * a.wcs;
    // clear the two zero entries of each odd extension
    for (int t = threadIdx.x; t < a.npairs; t += blockDim.x) {
        A[t * L] = make_double2(0.0, 0.0);
        A[t * L + m + 1] = make_double2(0.0, 0.0);
    }
    // load
    const int ne = nl * m;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        int line, pt;
        if (MODE == DST_XF || MODE == DST_XI) { line = e / m; pt = e - line * m; }
        else { pt = e / nl; line = e - pt * nl; }
        int64_t gl = l0 + line;
        double v = 0.0;
        if (gl < a.nlines) {
            if (MODE == DST_XF) {
                int j = (int)(gl % a.my), k = (int)(gl / a.my);
                int i = pt;
                // window node (i+1, j+1, k+1)
                int64_t o = a.rhs.at(i + 1, j + 1, k + 1);
                v = a.rhs.p[comp * a.rhs.cs + o];
                if (a.has_bc) {
                    const double *bp = a.bc.p + comp * a.bc.cs;
                    if (i == 0) v -= bp[a.bc.at(0, j + 1, k + 1)] * a.ih2[0];
                    if (i == a.mx - 1) v -= bp[a.bc.at(a.mx + 1, j + 1, k + 1)] * a.ih2[0];
                    if (j == 0) v -= bp[a.bc.at(i + 1, 0, k + 1)] * a.ih2[1];
                    if (j == a.my - 1) v -= bp[a.bc.at(i + 1, a.my + 1, k + 1)] * a.ih2[1];
                    if (k == 0) v -= bp[a.bc.at(i + 1, j + 1, 0)] * a.ih2[2];
                    if (k == a.mz - 1) v -= bp[a.bc.at(i + 1, j + 1, a.mz + 1)] * a.ih2[2];
                }
            } else {
                v = w[line_addr(MODE, a, gl, pt)];
            }
        }
        double *Ad = (double *)A;
        int pair = line >> 1, ri = line & 1;
        Ad[2 * (pair * L + pt + 1) + ri] = v;
        Ad[2 * (pair * L + L - pt - 1) + ri] = -v;
    }
    __syncthreads();
    double2 *R = fft_lds(A, B, a);
    if (MODE == DST_Z) {
        // S_k = (-Im Z_k / 2, Re Z_k / 2); scale by norm / -(lx+ly+lz); odd extension again
        double2 *C = (R == A) ? B : A;
        for (int t = threadIdx.x; t < a.npairs * (m + 2); t += blockDim.x) {
            int pair = t / (m + 2), k = t - pair * (m + 2);
            if (k == 0 || k == m + 1) {
                C[pair * L + k] = make_double2(0.0, 0.0);
                continue;
            }
            double2 z = R[pair * L + k];
            double sa = -0.5 * z.y, sb = 0.5 * z.x;
            int64_t ga = l0 + 2 * pair, gb = ga + 1;
            double lz = a.lam[2][k - 1];
            double fa = 0.0, fb = 0.0;
            if (ga < a.nlines) {
                int ix = (int)(ga % a.mx), iy = (int)(ga / a.mx);
                fa = a.norm / (-(a.lam[0][ix] + a.lam[1][iy] + lz));
            }
            if (gb < a.nlines) {
                int ix = (int)(gb % a.mx), iy = (int)(gb / a.mx);
                fb = a.norm / (-(a.lam[0][ix] + a.lam[1][iy] + lz));
            }
            double2 v = make_double2(sa * fa, sb * fb);
            C[pair * L + k] = v;
            C[pair * L + L - k] = make_double2(-v.x, -v.y);
        }
        __syncthreads();
        double2 *O = (C == A) ? B : A;
        R = fft_lds(C, O, a);
    }
    // store
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        int line, pt;
        if (MODE == DST_XF || MODE == DST_XI) { line = e / m; pt = e - line * m; }
        else { pt = e / nl; line = e - pt * nl; }
        int64_t gl = l0 + line;
        if (gl >= a.nlines) continue;
        int pair = line >> 1;
        double2 z = R[pair * L + pt + 1];
        double v = (line & 1) ? 0.5 * z.x : -0.5 * z.y;
        if (MODE == DST_XI) {
            int j = (int)(gl % a.my), k = (int)(gl / a.my);
            a.out.p[comp * a.out.cs + a.out.at(pt + 1, j + 1, k + 1)] = v * a.scale;
        } else {
            w[line_addr(MODE, a, gl, pt)] = v * a.scale;
        }
    }
}

// ------------------------------------------------------------------ single-buffer passes
// Same transform as k_dst_pass with one LDS buffer: every stage loads each butterfly's R
// inputs into registers, synchronises, and writes the outputs to their Stockham
// positions.  Lines are mapped to threads without 64-bit divisions, and the number of
// lines per work-group is a power of two.
template <int MODE, int MAXR>
__global__ void __launch_bounds__(DST_THREADS) k_dst_pass2(DstArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    __shared__ double lamxy[16];
    const int L = a.L, m = a.m, np = a.npairs, nl = 2 * np;
    double2 *A = lds;
    double2 *twl = lds + np * L;  // twiddle table exp(-2 pi i t / L) in LDS
    const int comp = blockIdx.y;
    const int64_t l0 = (int64_t)blockIdx.x * nl;
    double *w = a.w + comp * a.wcs;
    const int tid = threadIdx.x;
    for (int t = tid; t < L; t += DST_THREADS) twl[t] = a.tw[t];
    for (int t = tid; t < np; t += DST_THREADS) {
        A[t * L] = make_double2(0.0, 0.0);
        A[t * L + m + 1] = make_double2(0.0, 0.0);
    }
    if (MODE == DST_Z && tid < nl) {
        const int64_t gl = l0 + tid;
        double v = 0.0;
        if (gl < a.nlines) {
            const int ix = (int)(gl % a.mx), iy = (int)(gl / a.mx);
            v = a.lam[0][ix] + a.lam[1][iy];
        }
        lamxy[tid] = v;
    }
    double *Ad = (double *)A;
    if (MODE == DST_XF || MODE == DST_XI) {
        for (int line = 0; line < nl; line++) {
            const int64_t gl = l0 + line;
            const int pair = line >> 1, ri = line & 1;
            if (gl >= a.nlines) {
                for (int pt = tid; pt < m; pt += DST_THREADS) {
                    Ad[2 * (pair * L + pt + 1) + ri] = 0.0;
                    Ad[2 * (pair * L + L - pt - 1) + ri] = -0.0;
                }
                continue;
            }
            const int j = (int)(gl % a.my), k = (int)(gl / a.my);
            if (MODE == DST_XF) {
                const double *rp = a.rhs.p + comp * a.rhs.cs;
                const double *bp = a.bc.p + comp * a.bc.cs;
                for (int pt = tid; pt < m; pt += DST_THREADS) {
                    double v = rp[a.rhs.at(pt + 1, j + 1, k + 1)];
                    if (a.has_bc) {
                        if (pt == 0) v -= bp[a.bc.at(0, j + 1, k + 1)] * a.ih2[0];
                        if (pt == a.mx - 1) v -= bp[a.bc.at(a.mx + 1, j + 1, k + 1)] * a.ih2[0];
                        if (j == 0) v -= bp[a.bc.at(pt + 1, 0, k + 1)] * a.ih2[1];
                        if (j == a.my - 1) v -= bp[a.bc.at(pt + 1, a.my + 1, k + 1)] * a.ih2[1];
                        if (k == 0) v -= bp[a.bc.at(pt + 1, j + 1, 0)] * a.ih2[2];
                        if (k == a.mz - 1) v -= bp[a.bc.at(pt + 1, j + 1, a.mz + 1)] * a.ih2[2];
                    }
                    Ad[2 * (pair * L + pt + 1) + ri] = v;
                    Ad[2 * (pair * L + L - pt - 1) + ri] = -v;
                }
            } else {
                const double *src = w + gl * a.mx;
                for (int pt = tid; pt < m; pt += DST_THREADS) {
                    const double v = src[pt];
                    Ad[2 * (pair * L + pt + 1) + ri] = v;
                    Ad[2 * (pair * L + L - pt - 1) + ri] = -v;
                }
            }
        }
    } else {
        const int line = tid & (nl - 1);
        const int pair = line >> 1, ri = line & 1;
        const int64_t gl = l0 + line;
        const bool ok = gl < a.nlines;
        int64_t base = 0, stride;
        if (MODE == DST_Y) {
            stride = a.mx;
            if (ok) base = (gl % a.mx) + (int64_t)a.mx * a.my * (gl / a.mx);
        } else {
            stride = (int64_t)a.mx * a.my;
            base = gl;
        }
        const int step = DST_THREADS / nl;
        for (int pt = tid / nl; pt < m; pt += step) {
            const double v = ok ? w[base + pt * stride] : 0.0;
            Ad[2 * (pair * L + pt + 1) + ri] = v;
            Ad[2 * (pair * L + L - pt - 1) + ri] = -v;
        }
    }
    __syncthreads();
    fft_reg<MAXR>(A, twl, a);
    if (MODE == DST_Z) {
        // S_k = (-Im Z_k / 2, Re Z_k / 2) scaled by norm / -(lx+ly+lz), odd extension again (in place:
        // entry k is read and rewritten by one thread, entries L-k are not read)
        for (int t = tid; t < np * (m + 2); t += DST_THREADS) {
            const int pair = t / (m + 2), k = t - pair * (m + 2);
            if (k == 0 || k == m + 1) {
                A[pair * L + k] = make_double2(0.0, 0.0);
                continue;
            }
            const double2 z = A[pair * L + k];
            const double lz = a.lam[2][k - 1];
            const double fa = a.norm / (-(lamxy[2 * pair] + lz)), fb = a.norm / (-(lamxy[2 * pair + 1] + lz));
            const double2 v = make_double2(-0.5 * z.y * fa, 0.5 * z.x * fb);
            A[pair * L + k] = v;
            A[pair * L + L - k] = make_double2(-v.x, -v.y);
        }
        __syncthreads();
        fft_reg<MAXR>(A, twl, a);
    }
    if (MODE == DST_XF || MODE == DST_XI) {
        for (int line = 0; line < nl; line++) {
            const int64_t gl = l0 + line;
            if (gl >= a.nlines) break;
            const int pair = line >> 1;
            const int j = (int)(gl % a.my), k = (int)(gl / a.my);
            for (int pt = tid; pt < m; pt += DST_THREADS) {
                const double2 z = A[pair * L + pt + 1];
                const double v = (line & 1) ? 0.5 * z.x : -0.5 * z.y;
                if (MODE == DST_XI) a.out.p[comp * a.out.cs + a.out.at(pt + 1, j + 1, k + 1)] = v * a.scale;
                else w[gl * a.mx + pt] = v * a.scale;
            }
        }
    } else {
        const int line = tid & (nl - 1);
        const int pair = line >> 1;
        const int64_t gl = l0 + line;
        if (gl < a.nlines) {
            int64_t base, stride;
            if (MODE == DST_Y) {
                stride = a.mx;
                base = (gl % a.mx) + (int64_t)a.mx * a.my * (gl / a.mx);
            } else {
                stride = (int64_t)a.mx * a.my;
                base = gl;
            }
            const int step = DST_THREADS / nl;
            for (int pt = tid / nl; pt < m; pt += step) {
                const double2 z = A[pair * L + pt + 1];
                const double v = (line & 1) ? 0.5 * z.x : -0.5 * z.y;
                w[base + pt * stride] = v * a.scale;
            }
        }
    }
}

// ------------------------------------------------------------------ half-length passes
// DST-I of length m through a complex FFT of length N = m + 1 instead of 2(m + 1) (the
// sine-transform algorithm of Press et al., Numerical Recipes, sec. 12.3): with
//   y_0 = 0,  y_j = sin(j pi / N) (x_j + x_{N-j}) + (x_j - x_{N-j}) / 2,   Y = FFT_N(y),
// the transform F_k = sum_j x_j sin(pi j k / N) is
//   F_{2k} = -Im Y_k,   F_{2k+1} = F_{2k-1} + Re Y_k,   F_1 = Re Y_0 / 2,
// (Y with the exp(+) sign; our FFT has exp(-), which flips Im).  Two real lines share one
// complex FFT, so a line costs N * 8 bytes of LDS: 16 lines of a 289-node block fit in
// 42 KB, and the strided y and z passes read and write 128-byte runs (16 neighbouring
// lines).  The odd outputs are a prefix sum, done by one wave per line pair.

// y from x in place; slot j of pair p holds (x_j of line 2p, x_j of line 2p+1).
// With fold (the x forward pass), slots 1 and N-1 first lose the x-end Dirichlet terms
// fold[2l] * ih2 and fold[2l+1] * ih2 of each line l (fold in LDS).
__device__ __forceinline__ void dst_pre(double2 *A, const double *__restrict__ sn, int N, int np,
                                        const double *fold = nullptr, double ih2 = 0.0) {
    const int H = N / 2;  // pairs (j, N-j), j = 1..H (j = N/2 maps to itself)
    for (int t = threadIdx.x; t < np * H; t += DST_THREADS) {
        const int p = t / H, j = t - p * H + 1;
        double2 *B = A + p * N;
        double2 x = B[j], y = B[N - j];
        if (fold && j == 1) {
            const double *fa = fold + 4 * p, *fb = fa + 2;  // lines 2p and 2p+1
            if (N == 2) {  // one interior point: both x ends fold into slot 1
                x = make_double2((x.x - fa[0] * ih2) - fa[1] * ih2, (x.y - fb[0] * ih2) - fb[1] * ih2);
                y = x;
            } else {
                x = make_double2(x.x - fa[0] * ih2, x.y - fb[0] * ih2);
                y = make_double2(y.x - fa[1] * ih2, y.y - fb[1] * ih2);
            }
        }
        const double sj = sn[j];
        const double2 sp = make_double2(sj * (x.x + y.x), sj * (x.y + y.y));
        const double2 am = make_double2(0.5 * (x.x - y.x), 0.5 * (x.y - y.y));
        B[j] = cadd(sp, am);
        B[N - j] = csub(sp, am);
    }
    for (int p = threadIdx.x; p < np; p += DST_THREADS) A[p * N] = make_double2(0.0, 0.0);
}

constexpr int DST_PPT = (DST_EMAX / 2 + 8 + DST_THREADS - 1) / DST_THREADS;

// FFT output -> F_j in slot j (j = 1..N-1) of every pair.  Ends with a barrier.
__device__ __forceinline__ void dst_post(double2 *A, int N, int np) {
    const int H = N / 2 + 1;  // k = 0..N/2
    double2 ev[DST_PPT], od[DST_PPT];
    int base[DST_PPT], kk[DST_PPT];
#pragma unroll
    for (int u = 0; u < DST_PPT; u++) {
        const int t = threadIdx.x + u * DST_THREADS;
        base[u] = -1;
        kk[u] = 0;
        ev[u] = od[u] = make_double2(0.0, 0.0);
        if (t < np * H) {
            const int p = t / H, k = t - p * H;
            const double2 z = A[p * N + k], zc = A[p * N + (k == 0 ? 0 : N - k)];
            // line 2p: Y_k = (z + conj zc) / 2; line 2p+1: Y_k = (z - conj zc) / 2i
            ev[u] = make_double2(-0.5 * (z.y - zc.y), 0.5 * (z.x - zc.x));  // F_{2k} = -Im Y_k (exp(+))
            od[u] = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y + zc.y));   // Re Y_k
            if (k == 0) od[u] = make_double2(0.5 * od[u].x, 0.5 * od[u].y);
            base[u] = p * N;
            kk[u] = k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < DST_PPT; u++) {
        if (base[u] < 0) continue;
        const int k = kk[u];
        if (k >= 1 && 2 * k <= N - 1) A[base[u] + 2 * k] = ev[u];
        if (2 * k + 1 <= N - 1) A[base[u] + 2 * k + 1] = od[u];
    }
    __syncthreads();
    // odd slots: running sum over k, one wave per pair: each lane sums a contiguous chunk,
    // the chunk totals are scanned across the wave
    const int K = N / 2, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int C = (K + 63) / 64;
    for (int p = wave; p < np; p += DST_THREADS / 64) {
        double2 *B = A + p * N;
        const int k0 = lane * C < K ? lane * C : K, k1 = k0 + C < K ? k0 + C : K;
        double2 s = make_double2(0.0, 0.0);
        for (int k = k0; k < k1; k++) s = cadd(s, B[2 * k + 1]);
        double2 inc = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const double ox = __shfl_up(inc.x, d, 64), oy = __shfl_up(inc.y, d, 64);
            if (lane >= d) inc = make_double2(inc.x + ox, inc.y + oy);
        }
        double2 run = make_double2(__shfl_up(inc.x, 1, 64), __shfl_up(inc.y, 1, 64));
        if (lane == 0) run = make_double2(0.0, 0.0);
        for (int k = k0; k < k1; k++) {
            run = cadd(run, B[2 * k + 1]);
            B[2 * k + 1] = run;
        }
    }
    __syncthreads();
}

// Values of one batch of lines held by a thread between the prefetch and the LDS write.
constexpr int DST_PF = (2 * DST_EMAX + DST_THREADS - 1) / DST_THREADS;

// Element u of this thread in a batch of nl lines (nl a power of two): x passes give each
// line DST_THREADS / nl neighbouring threads, which walk its points (the line index is
// fixed per thread); y and z passes walk columns (neighbouring threads take neighbouring
// lines, 8*nl contiguous bytes per point).  Both cover m <= DST_PF * DST_THREADS / nl.
template <int MODE, int NT = DST_THREADS>
__device__ __forceinline__ void dst_elem(int m, int nl, int u, int *line, int *pt) {
    if (MODE == DST_XF || MODE == DST_XI) {
        const int tpl = NT / nl;
        *line = threadIdx.x / tpl;
        *pt = (threadIdx.x & (tpl - 1)) + u * tpl;
    } else {
        *line = threadIdx.x & (nl - 1);
        *pt = threadIdx.x / nl + u * (NT / nl);
    }
}

// Global offset (within a component) of point pt of line gl of the work array.
// Rows of the work array have pitch px >= mx; y lines are (i, k) and z lines (i, j) with i
// running over the pitch, so a batch of neighbouring lines starts on a row-aligned x.
template <int MODE>
__device__ __forceinline__ int64_t dst_addr(const DstArgs &a, int gl, int pt) {
    if (MODE == DST_XF || MODE == DST_XI) return (int64_t)gl * a.px + pt;
    if (MODE == DST_Y) return (int64_t)(gl % a.px) + (int64_t)a.px * a.my * (gl / a.px) + (int64_t)pt * a.px;
    return (int64_t)gl + (int64_t)pt * a.px * a.my;
}

// Line gl exists (padding lines of the y and z passes, x >= mx, are skipped).
template <int MODE>
__device__ __forceinline__ bool dst_line_ok(const DstArgs &a, int gl) {
    if (gl >= a.nlines) return false;
    if (MODE == DST_Y || MODE == DST_Z) return gl % a.px < a.mx;
    return true;
}

// Issue the loads of batch b into registers (no wait).  For the x forward pass, thread
// l < nl also fetches the Dirichlet values beyond both x ends of line l: fx = (x0, x1).
template <int MODE, class T>
__device__ __forceinline__ void dst_prefetch(const DstArgs &a, const T *w, int comp, int b, int nl,
                                             T (&v)[DST_PF], double (&fx)[2]) {
#pragma unroll
    for (int u = 0; u < DST_PF; u++) {
        int line, pt;
        dst_elem<MODE>(a.m, nl, u, &line, &pt);
        const int gl = b * nl + line;
        v[u] = (T)0;
        if (line < nl && pt < a.m && dst_line_ok<MODE>(a, gl)) {
            if (MODE == DST_XF) {
                const int j = gl % a.my, k = gl / a.my;
                v[u] = (T)a.rhs.p[comp * a.rhs.cs + a.rhs.at(pt + 1, j + 1, k + 1)];
            } else {
                v[u] = w[dst_addr<MODE>(a, gl, pt)];
            }
        }
    }
    if (MODE == DST_XF) {
        fx[0] = fx[1] = 0.0;
        const int gl = b * nl + (int)threadIdx.x;
        if (a.has_bc && (int)threadIdx.x < nl && gl < a.nlines) {
            const int j = gl % a.my, k = gl / a.my;
            const double *bp = a.bc.p + comp * a.bc.cs;
            fx[0] = bp[a.bc.at(0, j + 1, k + 1)];
            fx[1] = bp[a.bc.at(a.mx + 1, j + 1, k + 1)];
        }
    }
}

// Persistent, software-pipelined passes: each work-group walks batches b, b + gridDim.x,
// ...; the loads of the next batch are in flight while the current one is transformed.
template <int MODE, int MAXR>
__global__ void __launch_bounds__(DST_THREADS) k_dst_pass3(DstArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    __shared__ double lamxy[16];
    __shared__ double foldx[32];
    const int N = a.L, m = a.m, np = a.npairs, nl = 2 * np;
    const int nbatch = (int)((a.nlines + nl - 1) / nl);
    double2 *A = lds;
    double2 *twl = lds + np * N;  // exp(-2 pi i t / N)
    double *Ad = (double *)A;
    const int comp = blockIdx.y;
    double *w = a.w + comp * a.wcs;
    const int tid = threadIdx.x;
    for (int t = tid; t < N; t += DST_THREADS) twl[t] = a.tw[t];
    double v[DST_PF], fx[2];
    int b = blockIdx.x;
    if (b < nbatch) dst_prefetch<MODE>(a, w, comp, b, nl, v, fx);
    for (; b < nbatch; b += gridDim.x) {
        // registers -> LDS slot pt+1 (real part: even line, imaginary part: odd line)
#pragma unroll
        for (int u = 0; u < DST_PF; u++) {
            int line, pt;
            dst_elem<MODE>(m, nl, u, &line, &pt);
            if (line >= nl || pt >= m) continue;
            double val = v[u];
            if (MODE == DST_XF && a.has_bc) {
                const int gl = b * nl + line;
                if (gl < a.nlines) {
                    const int j = gl % a.my, k = gl / a.my;
                    const double *bp = a.bc.p + comp * a.bc.cs;
                    if (j == 0) val -= bp[a.bc.at(pt + 1, 0, k + 1)] * a.ih2[1];
                    if (j == a.my - 1) val -= bp[a.bc.at(pt + 1, a.my + 1, k + 1)] * a.ih2[1];
                    if (k == 0) val -= bp[a.bc.at(pt + 1, j + 1, 0)] * a.ih2[2];
                    if (k == a.mz - 1) val -= bp[a.bc.at(pt + 1, j + 1, a.mz + 1)] * a.ih2[2];
                }
            }
            Ad[2 * ((line >> 1) * N + pt + 1) + (line & 1)] = val;
        }
        if (MODE == DST_XF && tid < nl) {
            foldx[2 * tid] = fx[0];
            foldx[2 * tid + 1] = fx[1];
        }
        if (MODE == DST_Z && tid < nl) {
            const int gl = b * nl + tid;
            double lv = 0.0;
            if (gl < a.nlines) lv = a.lam[0][gl % a.mx] + a.lam[1][gl / a.mx];
            lamxy[tid] = lv;
        }
        __syncthreads();
        if (b + (int)gridDim.x < nbatch) dst_prefetch<MODE>(a, w, comp, b + gridDim.x, nl, v, fx);
        if (MODE == DST_XF && a.has_bc) dst_pre(A, a.sn, N, np, foldx, a.ih2[0]);
        else dst_pre(A, a.sn, N, np);
        __syncthreads();
        fft_reg<MAXR>(A, twl, a);
        dst_post(A, N, np);
        if (MODE == DST_Z) {
            for (int t = tid; t < np * m; t += DST_THREADS) {
                const int p = t / m, j = t - p * m + 1;
                const double lz = a.lam[2][j - 1];
                const double fa = a.norm / (-(lamxy[2 * p] + lz)), fb = a.norm / (-(lamxy[2 * p + 1] + lz));
                const double2 z = A[p * N + j];
                A[p * N + j] = make_double2(z.x * fa, z.y * fb);
            }
            __syncthreads();
            dst_pre(A, a.sn, N, np);
            __syncthreads();
            fft_reg<MAXR>(A, twl, a);
            dst_post(A, N, np);
        }
        // LDS slot pt+1 -> global
#pragma unroll
        for (int u = 0; u < DST_PF; u++) {
            int line, pt;
            dst_elem<MODE>(m, nl, u, &line, &pt);
            const int gl = b * nl + line;
            if (line >= nl || pt >= m || gl >= a.nlines) continue;
            const double val = Ad[2 * ((line >> 1) * N + pt + 1) + (line & 1)] * a.scale;
            if (MODE == DST_XI) {
                const int j = gl % a.my, k = gl / a.my;
                a.out.p[comp * a.out.cs + a.out.at(pt + 1, j + 1, k + 1)] = val;
            } else {
                w[dst_addr<MODE>(a, gl, pt)] = val;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ one wave per line pair
// The passes of lines up to WAVE_LMAX - 1 points: 16 lines (8 pairs) per batch, wave w
// transforms pair w alone (pre-processing, FFT, odd-output scan, and for z the diagonal
// and the inverse), so a batch needs three work-group barriers -- after the LDS fill,
// before the stores, after the stores.  Batches are pipelined as in k_dst_pass3.

template <class C, class Buf>
__device__ __forceinline__ void wave_pre(Buf B, const double *__restrict__ sn, int N, int lane, const double *fold,
                                         double ih2) {
    using T = typename Cx<C>::R;
    const int H = N / 2;
    for (int j = 1 + lane; j <= H; j += 64) {
        C x = B[j], y = B[N - j];
        if (fold && j == 1) {  // fold = (x0, x1) of line 2p, then of line 2p+1
            const T f0 = (T)(fold[0] * ih2), f1 = (T)(fold[1] * ih2), f2 = (T)(fold[2] * ih2), f3 = (T)(fold[3] * ih2);
            if (N == 2) {
                x = Cx<C>::mk((x.x - f0) - f1, (x.y - f2) - f3);
                y = x;
            } else {
                x = Cx<C>::mk(x.x - f0, x.y - f2);
                y = Cx<C>::mk(y.x - f1, y.y - f3);
            }
        }
        const T sj = (T)sn[j], h = (T)0.5;
        const C sp = Cx<C>::mk(sj * (x.x + y.x), sj * (x.y + y.y));
        const C am = Cx<C>::mk(h * (x.x - y.x), h * (x.y - y.y));
        B[j] = cadd(sp, am);
        B[N - j] = csub(sp, am);
    }
    if (lane == 0) B[0] = Cx<C>::mk(0, 0);
    __builtin_amdgcn_wave_barrier();
}

template <class C, class Buf>
__device__ __forceinline__ void wave_post(Buf B, int N, int lane) {
    using T = typename Cx<C>::R;
    constexpr int PPL = (WAVE_LMAX / 2 + 1 + 63) / 64;
    const int H = N / 2 + 1;
    const T h = (T)0.5;
    C ev[PPL], od[PPL];
#pragma unroll
    for (int i = 0; i < PPL; i++) {
        const int k = lane + 64 * i;
        ev[i] = od[i] = Cx<C>::mk(0, 0);
        if (k < H) {
            const C z = B[k], zc = B[k == 0 ? 0 : N - k];
            ev[i] = Cx<C>::mk(-h * (z.y - zc.y), h * (z.x - zc.x));
            od[i] = Cx<C>::mk(h * (z.x + zc.x), h * (z.y + zc.y));
            if (k == 0) od[i] = Cx<C>::mk(h * od[i].x, h * od[i].y);
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < PPL; i++) {
        const int k = lane + 64 * i;
        if (k >= H) continue;
        if (k >= 1 && 2 * k <= N - 1) B[2 * k] = ev[i];
        if (2 * k + 1 <= N - 1) B[2 * k + 1] = od[i];
    }
    __builtin_amdgcn_wave_barrier();
    const int K = N / 2, Cn = (K + 63) / 64;
    const int k0 = lane * Cn < K ? lane * Cn : K, k1 = k0 + Cn < K ? k0 + Cn : K;
    C s = Cx<C>::mk(0, 0);
    for (int k = k0; k < k1; k++) s = cadd(s, B[2 * k + 1]);
    C inc = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T ox = __shfl_up(inc.x, d, 64), oy = __shfl_up(inc.y, d, 64);
        if (lane >= d) inc = Cx<C>::mk(inc.x + ox, inc.y + oy);
    }
    C run = Cx<C>::mk(__shfl_up(inc.x, 1, 64), __shfl_up(inc.y, 1, 64));
    if (lane == 0) run = Cx<C>::mk(0, 0);
    for (int k = k0; k < k1; k++) {
        run = cadd(run, B[2 * k + 1]);
        B[2 * k + 1] = run;
    }
    __builtin_amdgcn_wave_barrier();
}

// Odd-extension form (FP32): DST-I of length m from the complex FFT of length L = 2(m+1)
// of the odd extension (0, x_1..x_m, 0, -x_m..-x_1); with two lines packed as real and
// imaginary parts, F_k = -Im Z_k / 2 for the first and Re Z_k / 2 for the second.  No
// sine weights and no running sums, so in FP32 its error stays at the FFT's own (about
// 1e-7 relative, ten times below the half-length form's).  The kernel writes the mirror
// image in its LDS fill and reads F_k in its drain; wave_pre_odd serves the x forward pass
// with Dirichlet values, whose folds change points after the fill.
template <class C, class Buf>
__device__ __forceinline__ void wave_pre_odd(Buf B, int m, int lane, const double *fold, double ih2) {
    using T = typename Cx<C>::R;
    const int L = 2 * (m + 1);
    if (fold && lane == 0) {  // x ends: fold = (x0, x1) of line 2p, then of line 2p+1
        const T f0 = (T)(fold[0] * ih2), f1 = (T)(fold[1] * ih2), f2 = (T)(fold[2] * ih2), f3 = (T)(fold[3] * ih2);
        C x = B[1];
        x = Cx<C>::mk(x.x - f0, x.y - f2);
        if (m == 1) {
            x = Cx<C>::mk(x.x - f1, x.y - f3);
            B[1] = x;
        } else {
            B[1] = x;
            const C y = B[m];
            B[m] = Cx<C>::mk(y.x - f1, y.y - f3);
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int j = 1 + lane; j <= m; j += 64) {
        const C x = B[j];
        B[L - j] = Cx<C>::mk(-x.x, -x.y);
    }
    if (lane == 0) {
        B[0] = Cx<C>::mk(0, 0);
        B[m + 1] = Cx<C>::mk(0, 0);
    }
    __builtin_amdgcn_wave_barrier();
}


// Addressing in the one-wave kernels (NL = 16 or 32 lines per batch, 32 threads per line).
// dst_elem gives every thread one line (x passes: 32 threads per line; y and z passes: line
// tid mod NL) and the points pt0 + 32u of it, so per batch a thread forms its line's offsets
// once: element u is read at src + u * src_step, written at dst + u * dst_step, and sits in
// LDS at a fixed stride.  NL = 32 in FP32 gives every point of a y or z batch a whole
// 128-byte row segment (16 lines leave the other half to another batch, which must find it in
// L2).
constexpr int WV_NL = 16, WV_PS = DST_THREADS / WV_NL;
static_assert(WV_PS == 32, "32 threads per line");
struct WaveAddr {
    int64_t src, src_step;  // work array, or the RHS view (x forward)
    int64_t dst, dst_step;  // work array, or the output view (x inverse)
    bool ok;                // the line exists
};

template <int MODE>
__device__ __forceinline__ WaveAddr wave_addr(const DstArgs &a, int comp, int gl, int pt0) {
    WaveAddr w;
    w.ok = dst_line_ok<MODE>(a, gl);
    if (MODE == DST_XF || MODE == DST_XI) {
        const int j = gl % a.my, k = gl / a.my;
        const int64_t wo = (int64_t)gl * a.px + pt0;
        w.src = MODE == DST_XF ? comp * a.rhs.cs + a.rhs.at(pt0 + 1, j + 1, k + 1) : wo;
        w.dst = MODE == DST_XI ? comp * a.out.cs + a.out.at(pt0 + 1, j + 1, k + 1) : wo;
        w.src_step = w.dst_step = WV_PS;
    } else {
        w.src = w.dst = dst_addr<MODE>(a, gl, pt0);
        w.src_step = w.dst_step = MODE == DST_Y ? (int64_t)WV_PS * a.px : (int64_t)WV_PS * a.px * a.my;
    }
    return w;
}

// Loads of batch b into registers (no wait), as dst_prefetch.  Every load is issued (one
// whose element does not exist reads element 0 instead), so no branch separates the loads and
// none has to wait for another; `ok` says whether the line exists, and the fill keeps only
// the points below m.  PT: double for the x forward pass (the RHS), converted at the fill.
template <int MODE, int NL, class T, class PT>
__device__ __forceinline__ void wave_prefetch(const DstArgs &a, const T *w, int comp, int b, int line, int pt0,
                                              int m, PT (&v)[DST_PF], bool &ok, double (&fx)[2]) {
    const WaveAddr ad = wave_addr<MODE>(a, comp, b * NL + line, pt0);
    ok = ad.ok;
#pragma unroll
    for (int u = 0; u < DST_PF; u++) {
        const int64_t o = (ad.ok && pt0 + u * WV_PS < m) ? ad.src + u * ad.src_step : 0;
        if constexpr (MODE == DST_XF) v[u] = a.rhs.p[o];
        else v[u] = w[o];
    }
    if (MODE == DST_XF) {
        fx[0] = fx[1] = 0.0;
        const int gl = b * NL + (int)threadIdx.x;
        if (a.has_bc && (int)threadIdx.x < NL && gl < a.nlines) {
            const int j = gl % a.my, k = gl / a.my;
            const double *bp = a.bc.p + comp * a.bc.cs;
            fx[0] = bp[a.bc.at(0, j + 1, k + 1)];
            fx[1] = bp[a.bc.at(a.mx + 1, j + 1, k + 1)];
        }
    }
}

// -n / d for d > 0 from the hardware reciprocal and Newton steps (within an ulp or two of
// the quotient, at a fraction of the cost of the IEEE division sequence).
__device__ __forceinline__ double neg_quot(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    double q = n * r;
    q = fma(r, fma(-d, q, n), q);
    return -q;
}

template <int FIX>
__host__ __device__ constexpr int fix_len() {
    if constexpr (FIX >= 0) return FixedFft<FIX>::L;
    else return 0;
}

// Padding of the wave kernels' LDS sequences (PadBuf shift, 0: none).  Measured: the y passes
// with compile-time plans gain 10-16% (their Stockham writes conflicted most), the FP32 z and
// x forward passes of the 289^3 block 5-8% (576 = 4.4.4.9); the FP64 x passes, the FP32 x
// inverse (register-capped: it spills, -66%), the 544-point FP32 z pass (-47%) and the
// runtime plans lose more to the extra registers and address arithmetic than the conflicts
// cost them.
__host__ __device__ constexpr int wave_pad_shift(bool dbl, int mode, int fix) {
    // the 8.8.9 plan: one slot per 8 elements keeps its radix-8 stores (strides 8 and 64
    // elements for groups of 8 lanes) on distinct banks in every pass
    if (fix == 6 || fix == 7 || fix == 8 || fix == 10) return 3;
    // the 16.4.9 and 16.17 plans: their radix-16 stage writes 16 consecutive elements per lane
    // (lanes 16 elements apart), which one slot per 16 elements spreads over the banks
    if (fix == 13 || fix == 14) return 4;
    return fix < 0 ? 0 : (mode == DST_Y || (mode == DST_Z && (dbl || fix == 0)) || (!dbl && fix == 0 && mode == DST_XF)) ? (dbl ? 3 : 4) : 0;
}
// Radix-17 roots as constants in the 272- and 544-point plans (measured per pass: faster where
// the extra registers cost no occupancy).
__host__ __device__ constexpr bool const_roots(bool dbl, int mode, int fix) {
    return dbl ? (fix == 3 && mode != DST_Z) : ((fix == 5 || fix == 8) && (mode == DST_Z || mode == DST_XI));
}

// T: arithmetic and work-array type (double: plain FP64 form, half-length transforms;
// float: the mixed form, odd-extension transforms, whose RHS, Dirichlet values and output
// stay FP64 and whose inner passes run in FP32).  P.L is the FFT length (m+1 or 2(m+1)).
template <int MODE, int MAXR, class T, int FIX, int NL>
__device__ __forceinline__ void dst_wave_body(const DstArgs &a, const WavePlan &P) {
#pragma clang fp contract(off)
    using C = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    extern __shared__ double2 lds_raw[];
    C *lds = (C *)lds_raw;
    __shared__ double lamxy[NL];
    __shared__ double foldx[2 * NL];
    __shared__ double lzs[MODE == DST_Z ? WAVE_LMAX : 1];  // lambda_z, read once per kernel
    constexpr int nl = NL, NT = NL * WV_PS;
    constexpr bool ODD = sizeof(T) == 4;
    // N: FFT length (2(m+1) in the odd-extension form, m+1 in the half-length form); both
    // constants with a compile-time plan
    const int N = FIX >= 0 ? fix_len<FIX>() : P.L;
    const int m = FIX >= 0 ? (ODD ? N / 2 - 1 : N - 1) : a.m;
    const int nbatch = (int)((a.nlines + nl - 1) / nl);
    // NL / 2 padded sequences (one per wave), then the twiddles
    constexpr int S = wave_pad_shift(sizeof(T) == 8, MODE, FIX);
    constexpr bool CROOTS = const_roots(sizeof(T) == 8, MODE, FIX);
    const int Np = pad_len<S>(N);
    C *A = lds;
    C *twl = lds + (NL / 2) * Np;  // exp(-2 pi i t / N)
    T *Ad = (T *)A;
    // LDS scalar of point pt (element pt + 1) of line `line` in a batch: lines 2p and 2p+1 are
    // the real and imaginary parts of sequence p
    auto slot_of = [&](int line, int pt) { return 2 * ((line >> 1) * Np + pad_pos<S>(pt + 1)) + (line & 1); };
    // this thread's line and first point (dst_elem), and the LDS stride of its elements
    const int tline = (MODE == DST_XF || MODE == DST_XI) ? (int)threadIdx.x / WV_PS : (int)threadIdx.x & (NL - 1);
    const int tpt0 = (MODE == DST_XF || MODE == DST_XI) ? (int)threadIdx.x & (WV_PS - 1) : (int)threadIdx.x / NL;
    constexpr int SSTEP = 2 * (WV_PS + (S ? WV_PS >> S : 0));
    const int slot0 = slot_of(tline, tpt0);
    // odd-extension form: the mirror element L - (pt + 1) of point pt, stepping down by SSTEP
    const int mslot0 = 2 * ((tline >> 1) * Np + pad_pos<S>(N - 1 - tpt0)) + (tline & 1);
    // (every component's batches in one persistent grid: global batch gb is batch
    // gb mod nbatch of component gb / nbatch)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool zf = MODE == DST_Z && a.zsum != nullptr;  // forward transform and diagonal only
    // (S = 0: the plain pointer, so those kernels compile exactly as unpadded code)
    using Buf = typename std::conditional<S == 0, C *, PadBuf<C, S>>::type;
    const Buf B{A + wave * Np};
    for (int t = tid; t < N; t += NT) twl[t] = Cx<C>::mk((T)a.tw[t].x, (T)a.tw[t].y);
    if (MODE == DST_Z)
        for (int t = tid; t < m; t += NT) lzs[t] = a.lam[2][t];
    using PT = typename std::conditional<MODE == DST_XF, double, T>::type;
    PT v[DST_PF];
    bool vok = false;
    double fx[2];
    // batch of iteration `it`: the two halves of a batch pair (neighbouring lines, one
    // 128-byte row segment in FP32) go to work-groups w and w+8, which share an XCD (and its
    // L2); gridDim.x is a multiple of 16
    const int G = gridDim.x, w8 = blockIdx.x & 7, slot = blockIdx.x >> 3, S2 = G >> 4;
    auto batch_of = [&](int it) {
        return a.order ? (it >> 1) * 2 * G + 2 * (int)blockIdx.x + (it & 1)
                       : 2 * (it * (G >> 1) + w8 * S2 + (slot >> 1)) + (slot & 1);
    };
    const int nbt = nbatch * a.ncomp;
    int it = 0, gb = batch_of(0);
    int comp = gb / nbatch, b = gb - comp * nbatch;
    if (gb < nbt) wave_prefetch<MODE, NL, T>(a, (T *)a.w + comp * a.wcs, comp, b, tline, tpt0, m, v, vok, fx);
    for (; gb < nbt; ++it) {
        T *w = (T *)a.wo + comp * a.wcs;
#pragma unroll
        for (int u = 0; u < DST_PF; u++)
            if (tpt0 + u * WV_PS < m) {
                const T x = vok ? (T)v[u] : (T)0;
                Ad[slot0 + u * SSTEP] = x;
                // odd-extension form: the mirror image -x at element L - (pt + 1) as well
                if constexpr (ODD) Ad[mslot0 - u * SSTEP] = -x;
            }
        if (MODE == DST_XF && a.has_bc) {
            // lines on the y or z faces of the window (a thread's line is fixed in x passes):
            // fold their Dirichlet neighbours; the x ends are folded in wave_pre
            int line, pt;
            dst_elem<MODE, NT>(m, nl, 0, &line, &pt);
            const int gl = b * nl + line;
            const int j = gl % a.my, k = gl / a.my;
            if (dst_line_ok<MODE>(a, gl) && (j == 0 || j == a.my - 1 || k == 0 || k == a.mz - 1)) {
                const double *bp = a.bc.p + comp * a.bc.cs;
#pragma unroll 1
                for (int u = 0; u < DST_PF; u++) {
                    dst_elem<MODE, NT>(m, nl, u, &line, &pt);
                    if (pt >= m) continue;
                    double f = 0.0;
                    if (j == 0) f += bp[a.bc.at(pt + 1, 0, k + 1)] * a.ih2[1];
                    if (j == a.my - 1) f += bp[a.bc.at(pt + 1, a.my + 1, k + 1)] * a.ih2[1];
                    if (k == 0) f += bp[a.bc.at(pt + 1, j + 1, 0)] * a.ih2[2];
                    if (k == a.mz - 1) f += bp[a.bc.at(pt + 1, j + 1, a.mz + 1)] * a.ih2[2];
                    T &slot = Ad[slot_of(line, pt)];
                    slot = (T)((double)slot - f);
                }
            }
        }
        if (MODE == DST_XF && tid < nl) {
            foldx[2 * tid] = fx[0];
            foldx[2 * tid + 1] = fx[1];
        }
        if (MODE == DST_Z && tid < nl) {
            const int gl = b * nl + tid;
            double lv = 0.0;
            if (dst_line_ok<MODE>(a, gl)) lv = a.lam[0][gl % a.px] + a.lam[1][gl / a.px];
            lamxy[tid] = lv;
        }
        __syncthreads();
        const int gbn = batch_of(it + 1);
        const int compn = gbn / nbatch, bn = gbn - compn * nbatch;
        if (gbn < nbt) wave_prefetch<MODE, NL, T>(a, (T *)a.w + compn * a.wcs, compn, bn, tline, tpt0, m, v, vok, fx);
        const double *fold = (MODE == DST_XF && a.has_bc) ? foldx + 4 * wave : nullptr;
        if constexpr (ODD) {
            // the fill wrote the odd extension; with Dirichlet values the folds change
            // points of it, so wave_pre_odd folds the x ends and writes the mirror again
            if (MODE == DST_XF && a.has_bc) {
                wave_pre_odd<C>(B, m, lane, fold, a.ih2[0]);
            } else {
                if (lane == 0) B[0] = B[m + 1] = Cx<C>::mk(0, 0);
                __builtin_amdgcn_wave_barrier();
            }
        } else {
            wave_pre<C>(B, a.sn, N, lane, fold, a.ih2[0]);
        }
        if constexpr (FIX >= 0) FixedFft<FIX>::template run<C, decltype(B), CROOTS>(B, twl, lane);
            else wave_fft<MAXR>(B, twl, P, lane);
        if (!ODD) wave_post<C>(B, N, lane);
        if (MODE == DST_Z) {
            const double la = lamxy[2 * wave], lb = lamxy[2 * wave + 1];
            double so[2] = {0.0, 0.0}, se[2] = {0.0, 0.0};  // z-face sums over odd / even modes
            for (int j = 1 + lane; j <= m; j += 64) {
                const double lz = lzs[j - 1];
                const C z = B[j];
                C g;
                if constexpr (ODD) {
                    // odd-extension form: the transform's outputs F = (-Im Z/2, Re Z/2), the
                    // diagonal (factors from the FP32 reciprocal, 1 ulp) and the odd extension
                    // of the result for the inverse transform, in one pass
                    const float nz = (float)a.norm, h = 0.5f;
                    const float fa = -nz * __builtin_amdgcn_rcpf((float)(la + lz));
                    const float fb = -nz * __builtin_amdgcn_rcpf((float)(lb + lz));
                    g = Cx<C>::mk((-h * z.y) * fa, (h * z.x) * fb);
                    B[j] = g;
                    if (!zf) B[N - j] = Cx<C>::mk(-g.x, -g.y);
                } else {
                    g = Cx<C>::mk(z.x * neg_quot(a.norm, la + lz), z.y * neg_quot(a.norm, lb + lz));
                    B[j] = g;
                }
                if (zf) {
                    const double wj = a.zw[j - 1];
                    if (j & 1) {
                        so[0] += wj * (double)g.x;
                        so[1] += wj * (double)g.y;
                    } else {
                        se[0] += wj * (double)g.x;
                        se[1] += wj * (double)g.y;
                    }
                }
            }
            if (zf) {
                // the line pair's sums over the wave: low face odd + even modes, high face odd - even
#pragma unroll
                for (int t = 0; t < 2; t++)
                    for (int d = 1; d < 64; d <<= 1) {
                        so[t] += __shfl_xor(so[t], d, 64);
                        se[t] += __shfl_xor(se[t], d, 64);
                    }
                if (lane < 2) {
                    const int gl = b * nl + 2 * wave + lane;
                    if (dst_line_ok<MODE>(a, gl)) {
                        const int kx = gl % a.px, ky = gl / a.px;
                        const double o = lane ? so[1] : so[0], e = lane ? se[1] : se[0];
                        a.zsum[((int64_t)comp * a.my + ky) * a.mx + kx] = o + e;
                        a.zsum[((int64_t)(a.ncomp + comp) * a.my + ky) * a.mx + kx] = o - e;
                    }
                }
            } else {
                if constexpr (ODD) {
                    if (lane == 0) B[0] = B[m + 1] = Cx<C>::mk(0, 0);
                    __builtin_amdgcn_wave_barrier();
                } else {
                    __builtin_amdgcn_wave_barrier();
                    wave_pre<C>(B, a.sn, N, lane, nullptr, 0.0);
                }
                if constexpr (FIX >= 0) FixedFft<FIX>::template run<C, decltype(B), CROOTS>(B, twl, lane);
                else wave_fft<MAXR>(B, twl, P, lane);
                if (!ODD) wave_post<C>(B, N, lane);
            }
        }
        __syncthreads();
        const WaveAddr ad = wave_addr<MODE>(a, comp, b * nl + tline, tpt0);
        // odd-extension form: point k of line 2p is -Im Z_k / 2, of line 2p+1 Re Z_k / 2 (the
        // other part of the same complex slot); a forward-only z pass stores the diagonal's
        // outputs as they are (line 2p in the real part)
        const T hs = ODD ? ((tline & 1) ? (T)0.5 : (T)-0.5) : (T)1;
#pragma unroll
        for (int u = 0; u < DST_PF; u++) {
            if (!ad.ok || tpt0 + u * WV_PS >= m) continue;
            const T val = (ODD && !zf) ? (hs * Ad[(slot0 + u * SSTEP) ^ 1]) * (T)a.scale
                                       : Ad[slot0 + u * SSTEP] * (T)a.scale;
            if (MODE == DST_XI) a.out.p[ad.dst + u * ad.dst_step] = (double)val;
            else w[ad.dst + u * ad.dst_step] = val;
        }
        __syncthreads();
        gb = gbn;
        comp = compn;
        b = bn;
    }
}

template <int MODE, int MAXR, class T, int FIX, int NL = WV_NL>
__global__ void __launch_bounds__(NL * WV_PS) k_dst_wave(DstArgs a, WavePlan P) {
    dst_wave_body<MODE, MAXR, T, FIX, NL>(a, P);
}
// The same with the registers capped for W waves per SIMD (W = 4: 128 VGPRs, two 8-wave
// work-groups per CU; W = 6: 80 VGPRs, three).  Measured on the 289- and 273-node blocks:
// the FP64 x and y passes of the 289-node block, the FP64 x inverse and the FP32 z pass of
// the 273-node box gain from four waves despite a few spilled registers, the FP32 x inverse
// pass of the 289-node block from six; the other passes lose more to the spills than the
// occupancy brings (the FP32 y pass at six waves: 8 % slower), or already fit.
template <int W, int MODE, int MAXR, class T, int FIX, int NL = WV_NL>
__global__ void __launch_bounds__(NL * WV_PS) __attribute__((amdgpu_waves_per_eu(W)))
k_dst_wave_cap(DstArgs a, WavePlan P) {
    dst_wave_body<MODE, MAXR, T, FIX, NL>(a, P);
}
__host__ __device__ constexpr int wave_cap(bool dbl, int mode, int fix) {
    if (dbl) return ((fix == 2 && (mode == DST_XF || mode == DST_Y)) || (fix == 3 && mode == DST_XI)) ? 4 : 0;
    if (fix == 0 && mode == DST_XI) return 6;
    return (fix == 5 && mode == DST_Z) ? 4 : 0;
}

// ---------------------------------------------------------------- mixed z pass
// The z pass of the mixed form's accurate solves: forward transform and diagonal from the
// FP32 work array in FP32 (odd-extension form), inverse transform in FP64 (half-length
// form), result stored FP64 into a.wo (same pitch and component stride).  A rounding error
// made before the diagonal is divided by the eigenvalue of its mode, so FP32 forward passes
// leave the solution's high modes accurate; one made after it is white noise on the solution,
// which the second differences of the curl and stretching stencils amplify by 1/h^2 -- so
// everything after the diagonal runs, and is stored, in FP64.  Each wave's LDS region holds
// either form's sequence (the FP32 one of 2(m+1) points and the FP64 one of m+1 points take
// the same bytes); between the two transforms a wave holds the diagonal's outputs in
// registers.
template <int MAXR, int FIX32, int FIX64>
__device__ __forceinline__ void dst_zmix_body(const DstArgs &a, const WavePlan &P32, const WavePlan &P64,
                                              const double2 *__restrict__ tw64) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds_raw[];
    __shared__ double lamxy[16];
    __shared__ double lzs[WAVE_LMAX];
    constexpr int nl = WV_NL;
    const int N32 = FIX32 >= 0 ? fix_len<FIX32>() : P32.L;  // 2(m+1)
    const int N64 = FIX64 >= 0 ? fix_len<FIX64>() : P64.L;  // m+1
    const int m = N64 - 1;
    const int nbatch = (int)((a.nlines + nl - 1) / nl);
    constexpr int S32 = wave_pad_shift(false, DST_Z, FIX32), S64 = wave_pad_shift(true, DST_Z, FIX64);
    constexpr bool CR32 = const_roots(false, DST_Z, FIX32), CR64 = const_roots(true, DST_Z, FIX64);
    const int Np32 = pad_len<S32>(N32), Np64 = pad_len<S64>(N64);
    const int Wq = (Np32 + 1) / 2 > Np64 ? (Np32 + 1) / 2 : Np64;  // double2 slots per wave
    float *Af = (float *)lds_raw;
    double *Ad = (double *)lds_raw;
    float2 *tw32l = (float2 *)(lds_raw + 8 * Wq);
    double2 *tw64l = lds_raw + 8 * Wq + N32 / 2;
    // scalar slot of element e of line `line` in a batch (lines 2p and 2p+1: real and
    // imaginary parts of wave p's sequence)
    auto slot32 = [&](int line, int e) { return 2 * ((line >> 1) * 2 * Wq + pad_pos<S32>(e)) + (line & 1); };
    auto slot64 = [&](int line, int e) { return 2 * ((line >> 1) * Wq + pad_pos<S64>(e)) + (line & 1); };
    const int tline = (int)threadIdx.x & (WV_NL - 1), tpt0 = (int)threadIdx.x / WV_NL;
    constexpr int SSTEP32 = 2 * (WV_PS + (S32 ? WV_PS >> S32 : 0));
    constexpr int SSTEP64 = 2 * (WV_PS + (S64 ? WV_PS >> S64 : 0));
    const int s32 = slot32(tline, tpt0 + 1), ms32 = slot32(tline, N32 - 1 - tpt0), s64 = slot64(tline, tpt0 + 1);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    using Buf32 = typename std::conditional<S32 == 0, float2 *, PadBuf<float2, S32>>::type;
    using Buf64 = typename std::conditional<S64 == 0, double2 *, PadBuf<double2, S64>>::type;
    const Buf32 B32{(float2 *)(lds_raw + wave * Wq)};
    const Buf64 B64{lds_raw + wave * Wq};
    for (int t = tid; t < N32; t += DST_THREADS) tw32l[t] = make_float2((float)a.tw[t].x, (float)a.tw[t].y);
    for (int t = tid; t < N64; t += DST_THREADS) tw64l[t] = tw64[t];
    for (int t = tid; t < m; t += DST_THREADS) lzs[t] = a.lam[2][t];
    float v[DST_PF];
    bool vok = false;
    double fx[2];
    const int G = gridDim.x, w8 = blockIdx.x & 7, slot = blockIdx.x >> 3, S2 = G >> 4;
    auto batch_of = [&](int it) {
        return a.order ? (it >> 1) * 2 * G + 2 * (int)blockIdx.x + (it & 1)
                       : 2 * (it * (G >> 1) + w8 * S2 + (slot >> 1)) + (slot & 1);
    };
    const int nbt = nbatch * a.ncomp;
    int it = 0, gb = batch_of(0);
    int comp = gb / nbatch, b = gb - comp * nbatch;
    if (gb < nbt) wave_prefetch<DST_Z, WV_NL, float>(a, (const float *)a.w + comp * a.wcs, comp, b, tline, tpt0, m, v, vok, fx);
    for (; gb < nbt; ++it) {
#pragma unroll
        for (int u = 0; u < DST_PF; u++)
            if (tpt0 + u * WV_PS < m) {
                const float x = vok ? v[u] : 0.f;
                Af[s32 + u * SSTEP32] = x;
                Af[ms32 - u * SSTEP32] = -x;  // odd extension
            }
        if (tid < nl) {
            const int gl = b * nl + tid;
            double lv = 0.0;
            if (dst_line_ok<DST_Z>(a, gl)) lv = a.lam[0][gl % a.px] + a.lam[1][gl / a.px];
            lamxy[tid] = lv;
        }
        __syncthreads();
        const int gbn = batch_of(it + 1);
        const int compn = gbn / nbatch, bn = gbn - compn * nbatch;
        if (gbn < nbt)
            wave_prefetch<DST_Z, WV_NL, float>(a, (const float *)a.w + compn * a.wcs, compn, bn, tline, tpt0, m, v, vok, fx);
        // forward transform in FP32
        if (lane == 0) B32[0] = B32[m + 1] = make_float2(0.f, 0.f);
        __builtin_amdgcn_wave_barrier();
        if constexpr (FIX32 >= 0) FixedFft<FIX32>::template run<float2, Buf32, CR32>(B32, tw32l, lane);
        else wave_fft<MAXR>(B32, tw32l, P32, lane);
        // F = (-Im Z/2, Re Z/2) times -norm / lambda in FP64 (the FP32 values convert exactly),
        // held in registers while the region changes form
        const double la = lamxy[2 * wave], lb = lamxy[2 * wave + 1];
        constexpr int GPL = (WAVE_LMAX + 63) / 64;
        double2 g[GPL];
#pragma unroll
        for (int i = 0; i < GPL; i++) {
            const int j = 1 + lane + 64 * i;
            g[i] = make_double2(0.0, 0.0);
            if (j <= m) {
                const float2 z = B32[j];
                const double lz = lzs[j - 1];
                g[i] = make_double2((double)(-0.5f * z.y) * neg_quot(a.norm, la + lz),
                                    (double)(0.5f * z.x) * neg_quot(a.norm, lb + lz));
            }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < GPL; i++) {
            const int j = 1 + lane + 64 * i;
            if (j <= m) B64[j] = g[i];
        }
        __builtin_amdgcn_wave_barrier();
        // inverse transform in FP64 (half-length form)
        wave_pre<double2>(B64, a.sn, N64, lane, nullptr, 0.0);
        if constexpr (FIX64 >= 0) FixedFft<FIX64>::template run<double2, Buf64, CR64>(B64, tw64l, lane);
        else wave_fft<MAXR>(B64, tw64l, P64, lane);
        wave_post<double2>(B64, N64, lane);
        __syncthreads();
        const WaveAddr ad = wave_addr<DST_Z>(a, comp, b * nl + tline, tpt0);
        double *wo = a.wo + comp * a.wcs;
#pragma unroll
        for (int u = 0; u < DST_PF; u++) {
            if (!ad.ok || tpt0 + u * WV_PS >= m) continue;
            wo[ad.dst + u * ad.dst_step] = Ad[s64 + u * SSTEP64] * a.scale;
        }
        __syncthreads();
        gb = gbn;
        comp = compn;
        b = bn;
    }
}

template <int MAXR, int FIX32, int FIX64>
__global__ void __launch_bounds__(DST_THREADS) k_dst_zmix(DstArgs a, WavePlan P32, WavePlan P64, const double2 *tw64) {
    dst_zmix_body<MAXR, FIX32, FIX64>(a, P32, P64, tw64);
}

// Zero the six faces of a window (zero Dirichlet values of solve_eq_0_3d).
__global__ void k_zero_faces(GView out, int nx, int ny, int nz, int ncomp) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t fxy = (int64_t)nx * ny, fxz = (int64_t)nx * nz, fyz = (int64_t)ny * nz;
    int64_t tot = 2 * (fxy + fxz + fyz);
    if (t >= tot) return;
    int i, j, k;
    if (t < 2 * fyz) { int f = (int)(t / fyz); int64_t r = t % fyz; i = f ? nx - 1 : 0; j = (int)(r % ny); k = (int)(r / ny); }
    else if (t < 2 * (fyz + fxz)) { int64_t u = t - 2 * fyz; int f = (int)(u / fxz); int64_t r = u % fxz; j = f ? ny - 1 : 0; i = (int)(r % nx); k = (int)(r / nx); }
    else { int64_t u = t - 2 * (fyz + fxz); int f = (int)(u / fxy); int64_t r = u % fxy; k = f ? nz - 1 : 0; i = (int)(r % nx); j = (int)(r / nx); }
    int64_t o = out.at(i, j, k);
    for (int c = 0; c < ncomp; c++) out.p[c * out.cs + o] = 0.0;
}

// ------------------------------------------------------------------ host side
struct FftPlan {
    int L = 0, nstages = 0, radix[24];
    double2 *tw = nullptr;
};

static int get_plan(Ctx *c, int L, FftPlan **out) {
    auto it = c->fft_plans.find(L);
    if (it != c->fft_plans.end()) { *out = (FftPlan *)it->second; return VPMX_OK; }
    FftPlan *p = new FftPlan();
    p->L = L;
    int r = L;
    while (r % 4 == 0) { p->radix[p->nstages++] = 4; r /= 4; }
    while (r % 2 == 0) { p->radix[p->nstages++] = 2; r /= 2; }
    for (int q = 3; r > 1;) {
        if (r % q == 0) {
            if (p->nstages >= 24) { delete p; return fail(c, VPMX_ERR_UNSUP, "FFT length has too many factors"); }
            p->radix[p->nstages++] = q; r /= q;
        } else q += 2;
    }
    for (int s = 0; s < p->nstages; s++)
        if (p->radix[s] > 64) {
            int R = p->radix[s];
            delete p;
            return fail(c, VPMX_ERR_UNSUP, "DST length 2(m+1)=" + std::to_string(L) + " has a prime factor " +
                                               std::to_string(R) + " > 64");
        }
    std::vector<double2> h(L);
    for (int t = 0; t < L; t++) {
        long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)L;
        h[t] = make_double2((double)cosl(ang), (double)(-sinl(ang)));
    }
    VPMX_HIP(c, hipMalloc(&p->tw, sizeof(double2) * L));
    VPMX_HIP(c, hipMemcpy(p->tw, h.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
    c->fft_plans[L] = p;
    *out = p;
    return VPMX_OK;
}

int fft_shape(Ctx *c, int L, FftShape *out) {
    FftPlan *p = nullptr;
    VPMX_TRY(get_plan(c, L, &p));
    out->L = p->L;
    out->nstages = p->nstages;
    for (int s = 0; s < p->nstages; s++) out->radix[s] = p->radix[s];
    out->tw = p->tw;
    return VPMX_OK;
}

void free_fft_plans(Ctx *c) {
    for (auto &kv : c->fft_plans) {
        FftPlan *p = (FftPlan *)kv.second;
        if (p->tw) (void)hipFree(p->tw);
        delete p;
    }
    c->fft_plans.clear();
    for (auto &kv : c->eig_cache) (void)hipFree(kv.second);
    c->eig_cache.clear();
    for (auto &kv : c->sin_cache) (void)hipFree(kv.second);
    c->sin_cache.clear();
    for (auto &kv : c->sinm_cache) (void)hipFree(kv.second);
    c->sinm_cache.clear();
    for (auto &kv : c->sinx_cache) (void)hipFree(kv.second);
    c->sinx_cache.clear();
    for (auto &kv : c->fdw_cache) (void)hipFree(kv.second);
    c->fdw_cache.clear();
    for (auto &kv : c->sinr_cache) (void)hipFree(kv.second);
    c->sinr_cache.clear();
    for (auto &kv : c->sinl_cache) (void)hipFree(kv.second);
    c->sinl_cache.clear();
}

// Eigenvalues 4 sin^2(pi k / (2(m+1))) / h^2 of the 1-D Dirichlet second difference,
// uploaded once per (m, h) and kept for the context's lifetime.
static int eigen(Ctx *c, int m, double h, DBuf &, const double **out) {
    auto key = std::make_pair(m, h);
    auto it = c->eig_cache.find(key);
    if (it != c->eig_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    std::vector<double> v(m);
    for (int k = 1; k <= m; k++) {
        double s = std::sin(M_PI * (double)k / (2.0 * (double)(m + 1)));
        v[k - 1] = 4.0 * s * s / (h * h);
    }
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * m));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    c->eig_cache[key] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

// Largest radix of the current plan, rounded to the radix class of k_dst_pass2 (0: the
// plan needs the two-buffer kernel).
static int radix_class(const DstArgs &a) {
    int mx = 0;
    for (int s = 0; s < a.nstages; s++) mx = a.radix[s] > mx ? a.radix[s] : mx;
    if (a.L > DST_EMAX || 2 * a.L > DST_LDS_MAX) return 0;  // lines + twiddle table must fit
    if (mx <= 4) return 4;
    if (mx <= 7) return 7;
    if (mx <= 17) return 17;
    if (mx <= 37) return 37;
    return 0;
}

template <int MODE>
static int launch_pass(Ctx *c, DstArgs a, int ncomp) {
    const int rc = c->dst_kernel == 1 ? 0 : radix_class(a);
    if (rc == 0) {
        int np = (int)(65536 / (sizeof(double2) * 2 * (size_t)a.L));
        a.npairs = np > 8 ? 8 : (np < 1 ? 1 : np);
        size_t lds = sizeof(double2) * 2 * (size_t)a.npairs * a.L;
        unsigned nblk = (unsigned)((a.nlines + 2 * a.npairs - 1) / (2 * a.npairs));
        hipLaunchKernelGGL(k_dst_pass<MODE>, dim3(nblk, ncomp), dim3(256), lds, c->stream, a);
    } else {
        // power-of-two line pairs per work-group, at most DST_EMAX complex entries in LDS
        int np = 1;
        while (np < 8 && 2 * np * a.L <= DST_EMAX && (2 * np + 1) * a.L <= DST_LDS_MAX) np *= 2;
        a.npairs = np;
        size_t lds = sizeof(double2) * (size_t)(np + 1) * a.L;
        unsigned nblk = (unsigned)((a.nlines + 2 * np - 1) / (2 * np));
        dim3 grid(nblk, ncomp);
        switch (rc) {
            case 4: hipLaunchKernelGGL((k_dst_pass2<MODE, 4>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
            case 7: hipLaunchKernelGGL((k_dst_pass2<MODE, 7>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
            case 17: hipLaunchKernelGGL((k_dst_pass2<MODE, 17>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
            default: hipLaunchKernelGGL((k_dst_pass2<MODE, 37>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
        }
    }
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

static void setup_axis(DstArgs &a, FftPlan *p, int m) {
    a.m = m;
    a.L = p->L;
    a.nstages = p->nstages;
    for (int s = 0; s < p->nstages; s++) a.radix[s] = p->radix[s];
    a.tw = p->tw;
    a.npairs = 1;
}

// The DST-I matrix sin(pi (t+1)(k+1) / (m+1)), t, k < m (row t), uploaded once per m.
static int sine_matrix(Ctx *c, int m, const double **out) {
    auto it = c->sinm_cache.find(m);
    if (it != c->sinm_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    std::vector<double> v((size_t)m * m);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < m; t++)
        for (int k = 0; k < m; k++) {
            // reduce (t+1)(k+1) mod 2(m+1) first so the argument stays below 2 pi
            const long long q = ((long long)(t + 1) * (k + 1)) % (2LL * (m + 1));
            v[(size_t)t * m + k] = (double)sinl(PI * (long double)q / (long double)(m + 1));
        }
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * v.size()));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    c->sinm_cache[m] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

// sin(j pi / N), j = 0..N/2, uploaded once per N.
static int sine_table(Ctx *c, int N, const double **out) {
    auto it = c->sin_cache.find(N);
    if (it != c->sin_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    std::vector<double> v(N / 2 + 1);
    for (int j = 0; j <= N / 2; j++)
        v[j] = (double)sinl(3.141592653589793238462643383279502884L * (long double)j / (long double)N);
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * v.size()));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    c->sin_cache[N] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

template <int MODE>
static int launch_pass3(Ctx *c, DstArgs a, int ncomp, int rc) {
    int np = 1;
    while (np < 8 && 2 * np * a.L <= DST_EMAX && (2 * np + 1) * a.L <= DST_LDS_MAX) np *= 2;
    a.npairs = np;
    const size_t lds = sizeof(double2) * (size_t)(np + 1) * a.L;
    // persistent grid: the work-groups the device holds at once, batches looped inside
    const int64_t nbatch = (a.nlines + 2 * np - 1) / (2 * np);
    const int per_cu = 1;
    int64_t resident = (int64_t)c->num_cu * per_cu / (ncomp > 0 ? ncomp : 1);
    if (resident < 1) resident = 1;
    dim3 grid((unsigned)(nbatch < resident ? nbatch : resident), ncomp);
    switch (rc) {
        case 4: hipLaunchKernelGGL((k_dst_pass3<MODE, 4>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
        case 7: hipLaunchKernelGGL((k_dst_pass3<MODE, 7>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
        case 17: hipLaunchKernelGGL((k_dst_pass3<MODE, 17>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
        default: hipLaunchKernelGGL((k_dst_pass3<MODE, 37>), grid, dim3(DST_THREADS), lds, c->stream, a); break;
    }
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// One pass along an axis of interior length m: the half-length kernel (FFT of m+1) when its
// plan fits, else the odd-extension kernels (FFT of 2(m+1)).
struct AxisPlan {
    int m;
    FftPlan *full, *half;
    const double *sn;
};

template <int MODE, class T>
static int launch_wave(Ctx *c, DstArgs a, int ncomp, int rc, const FftPlan *fp) {
    WavePlan P{};
    P.L = fp->L;
    P.nstages = fp->nstages;
    int ns = 1;
    for (int s = 0; s < fp->nstages; s++) {
        P.R[s] = fp->radix[s];
        P.Ns[s] = ns;
        P.mg[s] = ns > 1 ? (unsigned)((1ull << 32) / (unsigned long long)ns + 1ull) : 0u;
        ns *= fp->radix[s];
    }
    a.npairs = 8;
    // compile-time plans for the lengths of the common block sizes
    auto same = [&](const int *rr, int n) {
        if (fp->nstages != n) return false;
        for (int q = 0; q < n; q++)
            if (fp->radix[q] != rr[q]) return false;
        return true;
    };
    static const int r576[] = {4, 4, 4, 3, 3}, r288[] = {4, 4, 2, 3, 3}, r272[] = {4, 4, 17}, r544[] = {4, 4, 2, 17};
    int fix = -1;
    if (sizeof(T) == 4 && fp->L == 576 && same(r576, 5)) fix = 0;
    if (sizeof(T) == 8 && fp->L == 288 && same(r288, 5)) fix = 2;
    if (sizeof(T) == 8 && fp->L == 272 && same(r272, 3)) fix = 3;
    if (sizeof(T) == 4 && fp->L == 544 && same(r544, 4)) fix = 5;
    static const int r160[] = {4, 4, 2, 5}, r144[] = {4, 4, 3, 3};
    if (sizeof(T) == 8 && fp->L == 160 && same(r160, 4)) fix = 11;
    if (sizeof(T) == 8 && fp->L == 144 && same(r144, 4)) fix = 12;
    // kernel variants: with a compile-time plan the registers capped for four waves per SIMD
    // (six for the FP32 x inverse of the 576-point plan) or not, each in both batch orders;
    // FP32 compile-time plans also with 32-line batches (1024 threads) -- the tuner times all
    constexpr bool dbl = sizeof(T) == 8;
    // (each length's plan is instantiated only for the form that uses it)
    constexpr int c2 = wave_cap(dbl, MODE, 2), c3 = wave_cap(dbl, MODE, 3), c0 = wave_cap(dbl, MODE, 0),
                  c5 = wave_cap(dbl, MODE, 5);
    constexpr int k2 = c2 ? c2 : 4, k3 = c3 ? c3 : 4, k0 = c0 ? c0 : 4, k5 = c5 ? c5 : 4;
    struct Var {
        const void *fn;
        const char *name;
        int nl, order, fix;
    };
    Var vars[16];
    int nv = 0;
    auto add = [&](const void *fn, const char *name, int nl, int order, int vfix) {
        vars[nv++] = {fn, name, nl, order, vfix};
    };
    const void *unc = nullptr, *cap = nullptr, *unc32 = nullptr, *cap32 = nullptr;
    if constexpr (dbl) {
        if (fix == 2) {
            unc = (const void *)k_dst_wave<MODE, 4, T, 2>;
            cap = (const void *)k_dst_wave_cap<k2, MODE, 4, T, 2>;
        } else if (fix == 3) {
            unc = (const void *)k_dst_wave<MODE, 17, T, 3>;
            cap = (const void *)k_dst_wave_cap<k3, MODE, 17, T, 3>;
        } else if (fix == 11) {
            unc = (const void *)k_dst_wave<MODE, 9, T, 11>;
            cap = (const void *)k_dst_wave_cap<4, MODE, 9, T, 11>;
        } else if (fix == 12) {
            unc = (const void *)k_dst_wave<MODE, 9, T, 12>;
            cap = (const void *)k_dst_wave_cap<4, MODE, 9, T, 12>;
        }
    } else {
        if (fix == 0) {
            unc = (const void *)k_dst_wave<MODE, 4, T, 0>;
            cap = (const void *)k_dst_wave_cap<k0, MODE, 4, T, 0>;
            unc32 = (const void *)k_dst_wave<MODE, 4, T, 0, 32>;
            cap32 = (const void *)k_dst_wave_cap<k0, MODE, 4, T, 0, 32>;
        } else if (fix == 5) {
            unc = (const void *)k_dst_wave<MODE, 17, T, 5>;
            cap = (const void *)k_dst_wave_cap<k5, MODE, 17, T, 5>;
            unc32 = (const void *)k_dst_wave<MODE, 17, T, 5, 32>;
            cap32 = (const void *)k_dst_wave_cap<k5, MODE, 17, T, 5, 32>;
        }
    }
    if (unc) {
        add(unc, "uncapped", 16, 0, fix);
        add(cap, "capped", 16, 0, fix);
        add(unc, "uncapped, batch halves in turn", 16, 1, fix);
        add(cap, "capped, batch halves in turn", 16, 1, fix);
        if (unc32) {
            add(unc32, "uncapped, 32-line batches", 32, 0, fix);
            add(cap32, "capped, 32-line batches", 32, 0, fix);
        }
        if constexpr (dbl) {
            if (fix == 3) {  // the same plan with the radix-17 outputs split over four lanes
                add((const void *)k_dst_wave<MODE, 17, T, 9>, "radix-17 split", 16, 0, 9);
                add((const void *)k_dst_wave_cap<k3, MODE, 17, T, 9>, "radix-17 split, capped", 16, 0, 9);
                add((const void *)k_dst_wave<MODE, 17, T, 9>, "radix-17 split, batch halves in turn", 16, 1, 9);
                add((const void *)k_dst_wave_cap<k3, MODE, 17, T, 9>, "radix-17 split, capped, batch halves in turn",
                    16, 1, 9);
                // and the two-stage 16.17 plan, radix-17 split
                add((const void *)k_dst_wave<MODE, 17, T, 14>, "16.17 plan, radix-17 split", 16, 0, 14);
                add((const void *)k_dst_wave<MODE, 17, T, 14>, "16.17 plan, radix-17 split, batch halves in turn", 16,
                    1, 14);
            }
            if (fix == 2) {  // the three-stage 8.4.9 plan of the same length
                add((const void *)k_dst_wave<MODE, 9, T, 7>, "8.4.9 plan", 16, 0, 7);
                add((const void *)k_dst_wave_cap<k2, MODE, 9, T, 7>, "8.4.9 plan, capped", 16, 0, 7);
                add((const void *)k_dst_wave<MODE, 9, T, 7>, "8.4.9 plan, batch halves in turn", 16, 1, 7);
                add((const void *)k_dst_wave_cap<k2, MODE, 9, T, 7>, "8.4.9 plan, capped, batch halves in turn", 16,
                    1, 7);
            }
        }
        if constexpr (!dbl) {
            if (fix == 5) {  // the 17.8.4 plan with the radix-17 outputs split over two lanes
                add((const void *)k_dst_wave<MODE, 17, T, 10>, "17.8.4 plan, radix-17 split", 16, 0, 10);
                add((const void *)k_dst_wave_cap<k5, MODE, 17, T, 10>, "17.8.4 plan, radix-17 split, capped", 16, 0,
                    10);
            }
            if (fix == 5) {  // the three-stage 17.8.4 plan of the same length
                add((const void *)k_dst_wave<MODE, 17, T, 8>, "17.8.4 plan", 16, 0, 8);
                add((const void *)k_dst_wave_cap<k5, MODE, 17, T, 8>, "17.8.4 plan, capped", 16, 0, 8);
                add((const void *)k_dst_wave<MODE, 17, T, 8>, "17.8.4 plan, batch halves in turn", 16, 1, 8);
                add((const void *)k_dst_wave_cap<k5, MODE, 17, T, 8>, "17.8.4 plan, capped, batch halves in turn",
                    16, 1, 8);
            }
            if (fix == 0) {  // the three-stage 8.8.9 plan of the same length
                add((const void *)k_dst_wave<MODE, 9, T, 6>, "8.8.9 plan", 16, 0, 6);
                add((const void *)k_dst_wave_cap<k0, MODE, 9, T, 6>, "8.8.9 plan, capped", 16, 0, 6);
                add((const void *)k_dst_wave<MODE, 9, T, 6>, "8.8.9 plan, batch halves in turn", 16, 1, 6);
                add((const void *)k_dst_wave_cap<k0, MODE, 9, T, 6>, "8.8.9 plan, capped, batch halves in turn", 16,
                    1, 6);
                add((const void *)k_dst_wave<MODE, 9, T, 13>, "16.4.9 plan", 16, 0, 13);
                add((const void *)k_dst_wave_cap<k0, MODE, 9, T, 13>, "16.4.9 plan, capped", 16, 0, 13);
                add((const void *)k_dst_wave<MODE, 9, T, 13>, "16.4.9 plan, batch halves in turn", 16, 1, 13);
                add((const void *)k_dst_wave_cap<k0, MODE, 9, T, 13>, "16.4.9 plan, capped, batch halves in turn",
                    16, 1, 13);
            }
        }
    } else {
        fix = -1;
        const void *f = rc == 4 ? (const void *)k_dst_wave<MODE, 4, T, -1>
                      : rc == 7 ? (const void *)k_dst_wave<MODE, 7, T, -1>
                      : rc == 17 ? (const void *)k_dst_wave<MODE, 17, T, -1> : (const void *)k_dst_wave<MODE, 37, T, -1>;
        add(f, "uncapped", 16, 0, -1);
        add(f, "uncapped, batch halves in turn", 16, 1, -1);
    }
    // NL / 2 padded sequences and the twiddles (the kernel's LDS layout)
    using C = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
    auto np_of = [&](int vfix) {
        const int S = wave_pad_shift(sizeof(T) == 8, MODE, vfix);
        return S == 0 ? pad_len<0>(a.L) : S == 3 ? pad_len<3>(a.L) : pad_len<4>(a.L);
    };
    const void *fns[16];
    const char *names[16];
    size_t ldsv[16];
    int thr[16];
    {
        int q = 0;
        for (int r = 0; r < nv; r++) {
            // (static LDS of the kernel: lambda_z and the per-line tables, about 9 KB)
            const size_t l = sizeof(C) * (size_t)((vars[r].nl / 2) * np_of(vars[r].fix) + a.L);
            if (vars[r].nl > 16 && (int64_t)(l + 12288) > c->max_lds) continue;  // does not fit a work-group
            vars[q] = vars[r];
            fns[q] = vars[q].fn;
            names[q] = vars[q].name;
            ldsv[q] = l;
            thr[q] = vars[q].nl * WV_PS;
            q++;
        }
        nv = q;
    }
    a.ncomp = ncomp;
    // persistent grid: every component's batches in one grid, per_cu resident work-groups per
    // CU, a multiple of 16 (batch pairs on one XCD), no more than the batches need
    auto grid_of = [&](int per_cu, int nl) {
        const int64_t nbatch = (a.nlines + nl - 1) / nl;
        int64_t resident = (int64_t)c->num_cu * per_cu / 16 * 16;
        if (resident < 16) resident = 16;
        const int64_t need = (nbatch * (int64_t)ncomp + 15) / 16 * 16;
        return dim3((unsigned)(need < resident ? need : resident));
    };
    // the tuner's trials run on scratch arrays (the x inverse pass writes the output, which its
    // real launch then overwrites)
    static const char *mname[4] = {"x forward", "y", "z", "x inverse"};
    char key[160];
    std::snprintf(key, sizeof key, "sine transforms %s pass %s L=%d m=%d lines=%lld comps=%d%s", mname[MODE],
                  dbl ? "FP64" : "FP32", a.L, a.m, (long long)a.nlines, ncomp,
                  a.zsum ? " (forward and diagonal, z-face sums)" : "");
    Ctx::Tuned ch;
    {
        int st = VPMX_OK;
        const size_t wb = (sizeof(T) * (size_t)a.wcs * ncomp + 255) / 256 * 256;
        const size_t zb = a.zsum ? sizeof(double) * 2 * (size_t)ncomp * a.mx * a.my : 0;
        auto trial = [&](int fi, int pc) -> int {
            if (!c->tune_scratch.p || c->tune_scratch.bytes < wb + zb) {
                void *p = ensure(c, c->tune_scratch, wb + zb, &st);
                if (!p) return st;
                VPMX_TRY(fill(c, (double *)p, (int64_t)(c->tune_scratch.bytes / sizeof(double)), 0.0));
            }
            DstArgs t = a;
            t.w = t.wo = (double *)c->tune_scratch.p;
            if (a.zsum) t.zsum = (double *)((char *)c->tune_scratch.p + wb);
            t.order = vars[fi].order;
            WavePlan Pt = P;
            void *targs[] = {&t, &Pt};
            VPMX_HIP(c, hipLaunchKernel(fns[fi], grid_of(pc, vars[fi].nl), dim3(thr[fi]), targs, ldsv[fi], c->stream));
            return VPMX_OK;
        };
        VPMX_TRY(tune_v(c, key, nv, fns, names, ldsv, thr, trial, &ch));
    }
    const Var &V = vars[ch.fi];
    a.order = V.order;
    dim3 grid = grid_of(ch.per_cu, V.nl);
    hipEvent_t e0, e1;
    void *args[] = {&a, &P};
    constexpr int kc = MODE == DST_Y ? K_DST : K_DSTZ;
    if ((MODE == DST_Y || MODE == DST_Z) && a.m == c->dst_timed_m && !a.zsum && kernel_events(c, kc, &e0, &e1)) {
        // the fine block's y and z passes: timed by their own dispatch events; work = bytes
        // moved (read + write of every interior point)
        c->k_work[kc] = 2.0 * sizeof(T) * (double)a.mx * (MODE == DST_Y ? a.mz : a.my) * a.m * ncomp;
        VPMX_HIP(c, hipExtLaunchKernel(V.fn, grid, dim3(thr[ch.fi]), args, ldsv[ch.fi], c->stream, e0, e1, 0));
    } else {
        VPMX_HIP(c, hipLaunchKernel(V.fn, grid, dim3(thr[ch.fi]), args, ldsv[ch.fi], c->stream));
    }
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

static WavePlan wave_plan(const FftPlan *fp) {
    WavePlan P{};
    P.L = fp->L;
    P.nstages = fp->nstages;
    int ns = 1;
    for (int s = 0; s < fp->nstages; s++) {
        P.R[s] = fp->radix[s];
        P.Ns[s] = ns;
        P.mg[s] = ns > 1 ? (unsigned)((1ull << 32) / (unsigned long long)ns + 1ull) : 0u;
        ns *= fp->radix[s];
    }
    return P;
}

// The mixed z pass (k_dst_zmix) on the FP32 work array a.w, FP64 results into wo.
static int launch_zmix(Ctx *c, DstArgs a, int ncomp, const AxisPlan &X, double *wo) {
    const FftPlan *f32 = X.full, *f64 = X.half;
    const WavePlan P32 = wave_plan(f32), P64 = wave_plan(f64);
    setup_axis(a, (FftPlan *)f32, X.m);
    const int rc32 = radix_class(a);
    DstArgs ah = a;
    setup_axis(ah, (FftPlan *)f64, X.m);
    const int rc64 = radix_class(ah);
    const int rc = rc32 > rc64 ? rc32 : rc64;
    a.sn = X.sn;
    a.wo = wo;
    a.npairs = 8;
    a.ncomp = ncomp;
    auto same = [](const FftPlan *fp, const int *rr, int n) {
        if (fp->nstages != n) return false;
        for (int q = 0; q < n; q++)
            if (fp->radix[q] != rr[q]) return false;
        return true;
    };
    static const int r576[] = {4, 4, 4, 3, 3}, r288[] = {4, 4, 2, 3, 3}, r272[] = {4, 4, 17}, r544[] = {4, 4, 2, 17};
    // kernels (compile-time plans of both forms, or runtime plans), each in both batch orders
    struct ZK {
        const void *fn;
        int fix32, fix64;
        const char *name;
    };
    ZK zk[4];
    int nk = 1;
    if (f32->L == 576 && same(f32, r576, 5) && f64->L == 288 && same(f64, r288, 5)) {
        zk[0] = {(const void *)k_dst_zmix<4, 0, 2>, 0, 2, "FP32 forward, FP64 inverse"};
    } else if (f32->L == 544 && same(f32, r544, 4) && f64->L == 272 && same(f64, r272, 3)) {
        zk[0] = {(const void *)k_dst_zmix<17, 5, 3>, 5, 3, "FP32 forward, FP64 inverse"};
        zk[1] = {(const void *)k_dst_zmix<17, 10, 9>, 10, 9, "FP32 forward (17.8.4), FP64 inverse, radix-17 split"};
        zk[2] = {(const void *)k_dst_zmix<17, 5, 9>, 5, 9, "FP32 forward, FP64 inverse with radix-17 split"};
        zk[3] = {(const void *)k_dst_zmix<17, 5, 14>, 5, 14, "FP32 forward, FP64 inverse 16.17 with radix-17 split"};
        nk = 4;
    } else {
        zk[0] = {rc == 4 ? (const void *)k_dst_zmix<4, -1, -1>
                 : rc == 7 ? (const void *)k_dst_zmix<7, -1, -1>
                 : rc == 17 ? (const void *)k_dst_zmix<17, -1, -1> : (const void *)k_dst_zmix<37, -1, -1>,
                 -1, -1, "FP32 forward, FP64 inverse"};
    }
    // the kernel's LDS layout: eight regions of Wq double2 slots, then both twiddle tables
    auto lds_of = [&](int fix32, int fix64) {
        const int S32 = wave_pad_shift(false, DST_Z, fix32), S64 = wave_pad_shift(true, DST_Z, fix64);
        const int Np32 = S32 == 0 ? pad_len<0>(f32->L) : S32 == 3 ? pad_len<3>(f32->L) : pad_len<4>(f32->L);
        const int Np64 = S64 == 0 ? pad_len<0>(f64->L) : S64 == 3 ? pad_len<3>(f64->L) : pad_len<4>(f64->L);
        const int Wq = (Np32 + 1) / 2 > Np64 ? (Np32 + 1) / 2 : Np64;
        return sizeof(double2) * (size_t)(8 * Wq + f32->L / 2 + f64->L);
    };
    const void *fns[8];
    const char *names[8];
    size_t ldsv[8];
    int thr[8];
    char nbuf[8][96];
    for (int q = 0; q < nk; q++)
        for (int o = 0; o < 2; o++) {
            const int v = 2 * q + o;
            fns[v] = zk[q].fn;
            std::snprintf(nbuf[v], sizeof nbuf[v], "%s%s", zk[q].name, o ? ", batch halves in turn" : "");
            names[v] = nbuf[v];
            ldsv[v] = lds_of(zk[q].fix32, zk[q].fix64);
            thr[v] = DST_THREADS;
        }
    const int64_t nbatch = (a.nlines + 15) / 16;
    auto grid_of = [&](int per_cu) {
        int64_t resident = (int64_t)c->num_cu * per_cu / 16 * 16;
        if (resident < 16) resident = 16;
        const int64_t need = (nbatch * (int64_t)ncomp + 15) / 16 * 16;
        return dim3((unsigned)(need < resident ? need : resident));
    };
    char key[160];
    std::snprintf(key, sizeof key, "sine transforms z pass FP32/FP64 L=%d/%d m=%d lines=%lld comps=%d", f32->L, f64->L,
                  a.m, (long long)a.nlines, ncomp);
    Ctx::Tuned ch;
    {
        int st = VPMX_OK;
        // trials: FP32 input and FP64 output in scratch
        const size_t wb = (sizeof(float) + sizeof(double)) * (size_t)a.wcs * ncomp;
        auto trial = [&](int fi, int pc) -> int {
            if (!c->tune_scratch.p || c->tune_scratch.bytes < wb) {
                void *p = ensure(c, c->tune_scratch, wb, &st);
                if (!p) return st;
                VPMX_TRY(fill(c, (double *)p, (int64_t)(c->tune_scratch.bytes / sizeof(double)), 0.0));
            }
            DstArgs t = a;
            t.order = fi & 1;
            t.w = (double *)c->tune_scratch.p;
            t.wo = (double *)((char *)c->tune_scratch.p + sizeof(float) * (size_t)a.wcs * ncomp);
            WavePlan p32 = P32, p64 = P64;
            const double2 *t64 = f64->tw;
            void *targs[] = {&t, &p32, &p64, &t64};
            VPMX_HIP(c, hipLaunchKernel(fns[fi], grid_of(pc), dim3(DST_THREADS), targs, ldsv[fi], c->stream));
            return VPMX_OK;
        };
        VPMX_TRY(tune_v(c, key, 2 * nk, fns, names, ldsv, thr, trial, &ch));
    }
    a.order = ch.fi & 1;
    const void *fn = fns[ch.fi];
    const size_t lds = ldsv[ch.fi];
    dim3 grid = grid_of(ch.per_cu);
    const double2 *tw64 = f64->tw;
    WavePlan p32 = P32, p64 = P64;
    void *args[] = {&a, &p32, &p64, &tw64};
    hipEvent_t e0, e1;
    if (a.m == c->dst_timed_m && kernel_events(c, K_DSTZ, &e0, &e1)) {
        // FP32 read + FP64 write of every interior point
        c->k_work[K_DSTZ] = (4.0 + 8.0) * (double)a.mx * a.my * a.m * ncomp;
        VPMX_HIP(c, hipExtLaunchKernel(fn, grid, dim3(DST_THREADS), args, lds, c->stream, e0, e1, 0));
    } else {
        VPMX_HIP(c, hipLaunchKernel(fn, grid, dim3(DST_THREADS), args, lds, c->stream));
    }
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// The one-wave-per-pair kernel: FP64 lines of up to WAVE_LMAX - 1 points (FFT of m+1),
// FP32 lines of up to WAVE_LMAX_F / 2 - 1 points (FFT of 2(m+1)), radices up to 37.
static int wave_class(const AxisPlan &P, bool f32) {
    const FftPlan *fp = f32 ? P.full : P.half;
    if (fp->L > (f32 ? WAVE_LMAX_F : WAVE_LMAX) || fp->nstages > 12) return 0;
    DstArgs a{};
    setup_axis(a, (FftPlan *)fp, P.m);
    return radix_class(a);
}

template <int MODE>
static int run_pass(Ctx *c, DstArgs a, int ncomp, const AxisPlan &P, bool f32) {
    if (f32) {
        setup_axis(a, P.full, P.m);
        return launch_wave<MODE, float>(c, a, ncomp, wave_class(P, true), P.full);
    }
    const int wc = wave_class(P, false);
    if (c->dst_kernel == 4 && wc) {
        setup_axis(a, P.half, P.m);
        a.sn = P.sn;
        return launch_wave<MODE, double>(c, a, ncomp, wc, P.half);
    }
    if (c->dst_kernel >= 3) {
        setup_axis(a, P.half, P.m);
        a.sn = P.sn;
        const int rc = radix_class(a);
        if (rc) return launch_pass3<MODE>(c, a, ncomp, rc);
    }
    setup_axis(a, P.full, P.m);
    return launch_pass<MODE>(c, a, ncomp);
}

// ---------------------------------------------------------------- forward-pass reuse
// The x and y forward passes are linear, and the second solve of pmesh_inf differs from the
// first only by the folded Dirichlet terms F (on the layers next to the six faces), so after
// those passes its work array is the first solve's minus
//   C(kx,ky,k) = sum_{i,j} F(i,j,k) S_x(i,kx) S_y(j,ky),   S(i,k) = sin(pi i k / (m+1)),
// which reduces to 1-D sine transforms of the x and y faces' values and a 2-D one of the z
// faces' values:
//   C = ih2x (S_x(1,kx) X0(ky,k) + S_x(mx,kx) X1(ky,k)) + ih2y (S_y(1,ky) Y0(kx,k) + S_y(my,ky) Y1(kx,k))
//     + ih2z ([k = 1] Z0(kx,ky) + [k = mz] Z1(kx,ky)).
// FP64 tables; one pass forms W2 = W1 - C in the work array's precision.
// Face values into dense matrices, job q = 6 comp + face: x faces (face 0, 1) [k][j],
// y faces (2, 3) [k][i], z faces (4, 5) [i][j] (the first index is the row).
__global__ void k_fold_gather(GView bc, int mx, int my, int mz, double *ax, double *ay, double *az) {
    const int q = blockIdx.y, comp = q / 6, f = q - 6 * comp;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double *b = bc.p + comp * bc.cs;
    const int side = f & 1;
    if (f < 2) {
        if (e >= (int64_t)mz * my) return;
        const int k = (int)(e / my), j = (int)(e - (int64_t)k * my);
        ax[(int64_t)(2 * comp + side) * mz * my + e] = b[bc.at(side ? mx + 1 : 0, j + 1, k + 1)];
    } else if (f < 4) {
        if (e >= (int64_t)mz * mx) return;
        const int k = (int)(e / mx), i = (int)(e - (int64_t)k * mx);
        ay[(int64_t)(2 * comp + side) * mz * mx + e] = b[bc.at(i + 1, side ? my + 1 : 0, k + 1)];
    } else {
        if (e >= (int64_t)mx * my) return;
        const int i = (int)(e / my), j = (int)(e - (int64_t)i * my);
        az[(int64_t)(2 * comp + side) * mx * my + e] = b[bc.at(i + 1, j + 1, side ? mz + 1 : 0)];
    }
}

// Tables per (comp, side) pair p = 2 comp + side: X[p] [k][ky] (mz x my), Y[p] [k][kx]
// (mz x mx), Z[p] [ky][kx] (my x mx).  A work-group takes `span` (<= FC_SPAN) rows ky of one plane
// (k, comp), a thread one kx: the kx-dependent factors are loaded once, the ky-dependent ones
// are uniform (scalar loads), and FC_UNROLL row loads are in flight per thread.
constexpr int FC_SPAN = 48, FC_UNROLL = 8;
template <class T>
__global__ void __launch_bounds__(320) k_fold_correct(const T *w1, T *w2, int px, int mx,
                                                      int my, int mz, const double *__restrict__ X,
                                                      const double *__restrict__ Y, const double *__restrict__ Z,
                                                      double ihx, double ihy, double ihz, const double *__restrict__ Sx,
                                                      const double *__restrict__ Sy, int span) {
    const int kx = threadIdx.x, k = blockIdx.y % mz, comp = blockIdx.y / mz;
    const int ky0 = blockIdx.x * span, ky1 = ky0 + span < my ? ky0 + span : my;
    const double *X0 = X + (int64_t)(2 * comp) * mz * my + (int64_t)k * my, *X1 = X0 + (int64_t)mz * my;
    const double *Y0 = Y + (int64_t)(2 * comp) * mz * mx + (int64_t)k * mx, *Y1 = Y0 + (int64_t)mz * mx;
    const double *Z0 = Z + (int64_t)(2 * comp) * my * mx, *Z1 = Z0 + (int64_t)my * mx;
    const double *Sy1 = Sy + (int64_t)(my - 1) * my;
    // the rows' factors, staged once per work-group
    __shared__ double rf[4][FC_SPAN];
    for (int t = threadIdx.x; t < 4 * FC_SPAN; t += blockDim.x) {
        const int q = t / FC_SPAN, y = ky0 + (t - q * FC_SPAN);  // rows past the span are not read
        double v = 0.0;
        if (y < ky1) v = q == 0 ? X0[y] : (q == 1 ? X1[y] : (q == 2 ? Sy[y] : Sy1[y]));
        rf[q][t - q * FC_SPAN] = v;
    }
    __syncthreads();
    if (kx >= px) return;
    // the row padding (kx >= mx) is copied along, so every row is written whole
    const bool in = kx < mx;
    const double ax0 = in ? ihx * Sx[kx] : 0.0, ax1 = in ? ihx * Sx[(int64_t)(mx - 1) * mx + kx] : 0.0;
    const double by0 = in ? ihy * Y0[kx] : 0.0, by1 = in ? ihy * Y1[kx] : 0.0;
    const int64_t base = ((int64_t)comp * mz + k) * my * px + kx;
    for (int ky = ky0; ky < ky1; ky += FC_UNROLL) {
        T v[FC_UNROLL];
#pragma unroll
        for (int r = 0; r < FC_UNROLL; r++) {
            // every load issued (a row past the span reads the last row again): a conditional
            // load would wait before the next one is issued
            const int y = ky + r < ky1 ? ky + r : ky1 - 1;
            v[r] = w1[base + (int64_t)y * px];
        }
#pragma unroll
        for (int r = 0; r < FC_UNROLL; r++) {
            const int y = ky + r, yl = y - ky0;
            if (y >= ky1) break;
            double cr = (ax0 * rf[0][yl] + ax1 * rf[1][yl]) + (by0 * rf[2][yl] + by1 * rf[3][yl]);
            if (in && k == 0) cr += ihz * Z0[(int64_t)y * mx + kx];
            if (in && k == mz - 1) cr += ihz * Z1[(int64_t)y * mx + kx];
            w2[base + (int64_t)y * px] = (T)((double)v[r] - cr);
        }
    }
}

// W (the work array, holding the kept transforms) -= C from the Dirichlet values in bc.  The face transforms
// are products with the DST-I matrices (row-major out[a][k] = sum_t in[a][t] S[t][k], i.e.
// column-major out^T = S^T in^T).
static int fold_correct(Ctx *c, const DstArgs &a, int ncomp, bool f32, const GView &bc) {
    int st = VPMX_OK;
    const int mx = a.mx, my = a.my, mz = a.mz, np = 2 * ncomp;
    const double *Sx = nullptr, *Sy = nullptr;
    VPMX_TRY(sine_matrix(c, mx, &Sx));
    VPMX_TRY(sine_matrix(c, my, &Sy));
    const int64_t fx = (int64_t)mz * my, fy = (int64_t)mz * mx, fz = (int64_t)mx * my;
    // gathered faces (fx + fy + fz per pair), then X, Y, T, Z (fx + fy + 2 fz per pair)
    double *tab = (double *)ensure(c, c->fold_tab, sizeof(double) * (size_t)(np * (2 * fx + 2 * fy + 3 * fz)), &st);
    if (st) return st;
    double *ax = tab, *ay = ax + np * fx, *az = ay + np * fy;         // gathered faces
    double *X = az + np * fz, *Y = X + np * fx, *T = Y + np * fy, *Z = T + np * fz;  // transforms
    const int64_t fmax = std::max(fx, std::max(fy, fz));
    hipLaunchKernelGGL(k_fold_gather, dim3((unsigned)((fmax + 255) / 256), (unsigned)(6 * ncomp)), dim3(256), 0,
                       c->stream, bc, mx, my, mz, ax, ay, az);
    VPMX_HIP(c, hipGetLastError());
    // X[p][k][ky] = sum_j ax[p][k][j] Sy[j][ky];  Y[p][k][kx] = sum_i ay[p][k][i] Sx[i][kx]
    VPMX_TRY(dgemm_batched(c, false, my, np * mz, my, Sy, my, 0, ax, my, 0, X, my, 0, 1));
    VPMX_TRY(dgemm_batched(c, false, mx, np * mz, mx, Sx, mx, 0, ay, mx, 0, Y, mx, 0, 1));
    // T[p][i][ky] = sum_j az[p][i][j] Sy[j][ky];  Z[p][ky][kx] = sum_i T[p][i][ky] Sx[i][kx]
    VPMX_TRY(dgemm_batched(c, false, my, np * mx, my, Sy, my, 0, az, my, 0, T, my, 0, 1));
    VPMX_TRY(dgemm_batched(c, true, mx, my, mx, Sx, mx, 0, T, my, fz, Z, mx, fz, np));
    // one thread per kx: rows of up to 320 (the one-wave kernels' limit is below that)
    if (a.px > 320) return fail(c, VPMX_ERR_UNSUP, "Poisson: forward-pass reuse needs rows of at most 320 points");
    // small grids (the coarse problem): short spans, so the grid still fills the device
    const int span = my <= 96 ? 8 : FC_SPAN;
    const dim3 g((unsigned)((my + span - 1) / span), (unsigned)(mz * ncomp));
    const dim3 blk((unsigned)((a.px + 63) / 64 * 64));
    if (f32)
        hipLaunchKernelGGL(k_fold_correct<float>, g, blk, 0, c->stream, (const float *)a.w, (float *)a.w, a.px,
                           mx, my, mz, X, Y, Z, a.ih2[0], a.ih2[1], a.ih2[2], Sx, Sy, span);
    else
        hipLaunchKernelGGL(k_fold_correct<double>, g, blk, 0, c->stream, (const double *)a.w, a.w, a.px, mx,
                           my, mz, X, Y, Z, a.ih2[0], a.ih2[1], a.ih2[2], Sx, Sy, span);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

// ---------------------------------------------------------------- normal derivatives from the spectrum
// The zero-BC solve of pmesh is read only through the one-sided normal derivative
// (1/h)(25/12 f0 - 4 f1 + 3 f2 - 4/3 f3 + 1/4 f4) at every face node (f_s the value s nodes
// inside, f0 = 0 on the face; pinfdomain.f90:294-547).  The solution is a sine series in x and y
// once the z pass has run (work array in (kx, ky, z)), so on an x face the derivative is
// sum_kx W_x(kx) (series in y) with W(k) = (1/h)(-4 sin(pi k/(m+1)) + 3 sin(2 pi k/(m+1))
// - 4/3 sin(3 pi k/(m+1)) + 1/4 sin(4 pi k/(m+1))) on the low face and (-1)^(k+1) W(k) on the
// high one; likewise on the y faces, and on the z faces the stencil acts on the four z planes
// next to each face.  One read of the work array gives the three face sums (FP64); products
// with the sine matrices (FP64, matrix cores) take them to the face nodes.  This replaces the
// y and x inverse passes of that solve.

constexpr int FR_COLS = 5;  // kx columns per lane: rows of at most 320 points

// x and y face sums of one (z plane, component): waves take rows ky = wave + 4 j (so a wave's
// rows share their parity), R rows at a time with all their loads in flight; lanes take columns
// kx = lane + 64 q (the parity of kx is the lane's, so the low and high x sums are the even-lane
// sum plus / minus the odd-lane sum).  xp[b][k][ky], yp[b][k][kx] with b = component (low
// face) or ncomp + component (high face) and k = z + 1; the columns k = 0 and mz + 1 (the nodes
// on the z faces) are zero.
template <typename T, int R>
__global__ void __launch_bounds__(256) k_face_reduce_xy(const T *__restrict__ w, int64_t wcs, int px, int mx, int my,
                                                        int mz, int ncomp, const double *__restrict__ wx,
                                                        const double *__restrict__ wy, double *__restrict__ xp,
                                                        double *__restrict__ yp) {
    __shared__ double part[4][FR_COLS * 64];
    const int z = blockIdx.x, comp = blockIdx.y;
    const int NZ = mz + 2;
    const T *__restrict__ pl = w + comp * wcs + (int64_t)z * px * my;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    double wxl[FR_COLS], ys[FR_COLS];
#pragma unroll
    for (int q = 0; q < FR_COLS; q++) {
        const int kx = lane + 64 * q;
        wxl[q] = kx < mx ? wx[kx] : 0.0;
        ys[q] = 0.0;
    }
    double *__restrict__ xlo = xp + ((int64_t)comp * NZ + z + 1) * my;
    double *__restrict__ xhi = xp + ((int64_t)(ncomp + comp) * NZ + z + 1) * my;
    for (int ky0 = wave; ky0 < my; ky0 += 4 * R) {
        T v[R][FR_COLS];
        // every load in bounds (indexes clamped), values outside the rows masked afterwards, so
        // no load waits behind a branch
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int ky = ky0 + 4 * r;
            const T *__restrict__ row = pl + (int64_t)min(ky, my - 1) * px;
#pragma unroll
            for (int q = 0; q < FR_COLS; q++) v[r][q] = row[min(lane + 64 * q, mx - 1)];
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int q = 0; q < FR_COLS; q++)
                if (ky0 + 4 * r >= my || lane + 64 * q >= mx) v[r][q] = (T)0;
        double s[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int ky = ky0 + 4 * r;
            const double wyk = wy[min(ky, my - 1)];  // (the masked rows add zeros)
            s[r] = 0.0;
#pragma unroll
            for (int q = 0; q < FR_COLS; q++) {
                s[r] += wxl[q] * (double)v[r][q];
                ys[q] += wyk * (double)v[r][q];
            }
        }
#pragma unroll
        for (int d = 2; d < 64; d <<= 1)
#pragma unroll
            for (int r = 0; r < R; r++) s[r] += __shfl_xor(s[r], d, 64);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int ky = ky0 + 4 * r;
            const double o = __shfl_xor(s[r], 1, 64);
            if (lane == 0 && ky < my) {
                xlo[ky] = s[r] + o;
                xhi[ky] = s[r] - o;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < FR_COLS; q++) part[wave][lane + 64 * q] = ys[q];
    __syncthreads();
    // waves 0 and 2 hold the even rows, 1 and 3 the odd ones
    for (int kx = tid; kx < mx; kx += 256) {
        const double e = part[0][kx] + part[2][kx], o = part[1][kx] + part[3][kx];
        yp[((int64_t)comp * NZ + z + 1) * mx + kx] = e + o;
        yp[((int64_t)(ncomp + comp) * NZ + z + 1) * mx + kx] = e - o;
    }
    for (int side = 0; side < 2; side++) {
        if (z != (side ? mz - 1 : 0)) continue;
        const int k = side ? NZ - 1 : 0;
        for (int b = comp; b < 2 * ncomp; b += ncomp) {
            for (int t = tid; t < my; t += 256) xp[((int64_t)b * NZ + k) * my + t] = 0.0;
            for (int t = tid; t < mx; t += 256) yp[((int64_t)b * NZ + k) * mx + t] = 0.0;
        }
    }
}

// z face sums: the stencil on the four planes next to each z face, zp[b][ky][kx].
template <typename T>
__global__ void __launch_bounds__(256) k_face_reduce_z(const T *__restrict__ w, int64_t wcs, int px, int mx, int my,
                                                       int mz, int ncomp, double ihz, double *__restrict__ zp) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int comp = blockIdx.y;
    if (t >= (int64_t)mx * my) return;
    const int kx = (int)(t % mx), ky = (int)(t / mx);
    const T *__restrict__ p = w + comp * wcs + (int64_t)ky * px + kx;
    const int64_t pz = (int64_t)px * my;
    const double a2 = -4.0, a3 = 3.0, a4 = -4.0 / 3.0, a5 = 0.25;
    for (int side = 0; side < 2; side++) {
        const int z0 = side ? mz - 1 : 0, dz = side ? -1 : 1;
        const double v1 = (double)p[z0 * pz], v2 = (double)p[(z0 + dz) * pz], v3 = (double)p[(z0 + 2 * dz) * pz],
                     v4 = (double)p[(z0 + 3 * dz) * pz];
        zp[((int64_t)(side * ncomp + comp) * my + ky) * mx + kx] = ihz * (((a2 * v1 + a3 * v2) + a4 * v3) + a5 * v4);
    }
}

// W(k), k = 1..m (index k - 1), uploaded once per (m, h).
static int face_weights(Ctx *c, int m, double h, const double **out) {
    auto key = std::make_pair(m, h);
    auto it = c->fdw_cache.find(key);
    if (it != c->fdw_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    const long double PI = 3.141592653589793238462643383279502884L;
    const long double a[4] = {-4.0L, 3.0L, -4.0L / 3.0L, 0.25L};
    std::vector<double> v(m);
    for (int k = 1; k <= m; k++) {
        long double s = 0.0L;
        for (int q = 1; q <= 4; q++) {
            const long long r = ((long long)k * q) % (2LL * (m + 1));
            s += a[q - 1] * sinl(PI * (long double)r / (long double)(m + 1));
        }
        v[k - 1] = (double)(s / (long double)h);
    }
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * m));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    c->fdw_cache[key] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

// The DST-I matrix with a zero row above and below, sin(pi r (k+1) / (m+1)) for r = 0..m+1,
// k < m, column-major (m+2) x m (row r = node r of the line), uploaded once per m.
static int sine_matrix_ext(Ctx *c, int m, const double **out) {
    auto it = c->sinx_cache.find(m);
    if (it != c->sinx_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    const int N = m + 2;
    std::vector<double> v((size_t)N * m);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < m; k++)
        for (int r = 0; r < N; r++) {
            const long long q = ((long long)r * (k + 1)) % (2LL * (m + 1));
            v[(size_t)k * N + r] = q % (m + 1) == 0 ? 0.0 : (double)sinl(PI * (long double)q / (long double)(m + 1));
        }
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * v.size()));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    c->sinx_cache[m] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

// Work buffers of the spectral normal derivatives (the context keeps them): x / y face sums,
// z face sums, the z faces' partial products, and the x / y faces' partial products when the z
// pass left z in sine space.
struct FaceBufs {
    double *xp, *yp, *zp, *v, *tx, *ty;
};
static int face_bufs(Ctx *c, int mx, int my, int mz, int ncomp, FaceBufs *fb) {
    int st = VPMX_OK;
    const int np = 2 * ncomp, NX = mx + 2, NY = my + 2, NZ = mz + 2;
    const int64_t nxp = (int64_t)np * NZ * my, nyp = (int64_t)np * NZ * mx, nzp = (int64_t)np * my * mx,
                  nv = (int64_t)np * my * NX, ntx = (int64_t)np * NZ * NY, nty = (int64_t)np * NZ * NX;
    double *p = (double *)ensure(c, c->fd_spec, sizeof(double) * (size_t)(nxp + nyp + nzp + nv + ntx + nty), &st);
    if (st) return st;
    fb->xp = p;
    fb->yp = fb->xp + nxp;
    fb->zp = fb->yp + nyp;
    fb->v = fb->zp + nzp;
    fb->tx = fb->v + nv;
    fb->ty = fb->tx + ntx;
    return VPMX_OK;
}

// The face-node table fd (faces XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX; each ncomp x N0 x N1 with
// (N0, N1) = (NY, NZ), (NX, NZ), (NX, NY), the first index fastest) from the work array after
// the z pass (a.w, FP32 when f32).  zspec: the z pass stopped after the diagonal (z still in
// sine space) and left the z faces' sums in fb.zp; else the z faces are summed here from the
// planes next to them.  Every entry is written (zero on the face edges).
static int face_derivs_spectral(Ctx *c, const DstArgs &a, int ncomp, bool f32, const double h[3], double *fd,
                                bool zspec, const FaceBufs &fb) {
    const int mx = a.mx, my = a.my, mz = a.mz, np = 2 * ncomp;
    const int NX = mx + 2, NY = my + 2, NZ = mz + 2;
    if (mx > 64 * FR_COLS) return fail(c, VPMX_ERR_UNSUP, "spectral normal derivatives: rows of at most 320 points");
    if (mx < 4 || my < 4 || mz < 4) return fail(c, VPMX_ERR_ARG, "normal derivatives need >= 6 nodes per axis");
    const double *wx = nullptr, *wy = nullptr, *SX = nullptr, *SY = nullptr, *SZ = nullptr;
    VPMX_TRY(face_weights(c, mx, h[0], &wx));
    VPMX_TRY(face_weights(c, my, h[1], &wy));
    VPMX_TRY(sine_matrix_ext(c, mx, &SX));
    VPMX_TRY(sine_matrix_ext(c, my, &SY));
    if (zspec) VPMX_TRY(sine_matrix_ext(c, mz, &SZ));
    double *xp = fb.xp, *yp = fb.yp, *zp = fb.zp, *v = fb.v;
    // rows in flight per wave: two candidates per precision, same bits (the row order of every
    // sum is fixed)
    const void *fns[2];
    if (f32) {
        fns[0] = (const void *)k_face_reduce_xy<float, 4>;
        fns[1] = (const void *)k_face_reduce_xy<float, 8>;
    } else {
        fns[0] = (const void *)k_face_reduce_xy<double, 2>;
        fns[1] = (const void *)k_face_reduce_xy<double, 4>;
    }
    const char *names[2] = {f32 ? "4 rows in flight" : "2 rows in flight", f32 ? "8 rows in flight" : "4 rows in flight"};
    const void *wp = a.w;
    int64_t wcs = a.wcs;
    int px = a.px, ncv = ncomp, mxv = mx, myv = my, mzv = mz;
    auto trial = [&](int var, int) -> int {
        void *kargs[] = {(void *)&wp, &wcs, &px, &mxv, &myv, &mzv, &ncv, &wx, &wy, &xp, &yp};
        VPMX_HIP(c, hipLaunchKernel(fns[var], dim3((unsigned)mz, (unsigned)ncomp), dim3(256), kargs, 0, c->stream));
        return VPMX_OK;
    };
    char key[128];
    std::snprintf(key, sizeof key, "face sums %s m=%d,%d,%d", f32 ? "FP32" : "FP64", mx, my, mz);
    Ctx::Tuned ch;
    VPMX_TRY(tune(c, key, 2, fns, names, 0, 256, trial, &ch, false));
    VPMX_TRY(trial(ch.fi, 0));
    if (!zspec) {
        const dim3 gz((unsigned)(((int64_t)mx * my + 255) / 256), (unsigned)ncomp);
        if (f32)
            hipLaunchKernelGGL(k_face_reduce_z<float>, gz, dim3(256), 0, c->stream, (const float *)a.w, a.wcs, a.px, mx,
                               my, mz, ncomp, 1.0 / h[2], zp);
        else
            hipLaunchKernelGGL(k_face_reduce_z<double>, gz, dim3(256), 0, c->stream, (const double *)a.w, a.wcs, a.px,
                               mx, my, mz, ncomp, 1.0 / h[2], zp);
        VPMX_HIP(c, hipGetLastError());
    }
    const int64_t oy = (int64_t)np * NY * NZ, oz = oy + (int64_t)np * NX * NZ;
    if (zspec) {
        // x faces: tx[b][kz][j] = sum_ky S_y(j, ky) xp[b][kz][ky], fd[b][k][j] = sum_kz tx[b][kz][j] S_z(k, kz)
        VPMX_TRY(dgemm(c, false, false, NY, np * NZ, my, SY, NY, 0, xp, my, 0, fb.tx, NY, 0, 1));
        VPMX_TRY(dgemm(c, false, true, NY, NZ, mz, fb.tx + NY, NY, (int64_t)NY * NZ, SZ, NZ, 0, fd, NY,
                       (int64_t)NY * NZ, np));
        // y faces likewise with S_x
        VPMX_TRY(dgemm(c, false, false, NX, np * NZ, mx, SX, NX, 0, yp, mx, 0, fb.ty, NX, 0, 1));
        VPMX_TRY(dgemm(c, false, true, NX, NZ, mz, fb.ty + NX, NX, (int64_t)NX * NZ, SZ, NZ, 0, fd + oy, NX,
                       (int64_t)NX * NZ, np));
    } else {
        // x faces: fd[b][k][j] = sum_ky S_y(j, ky) xp[b][k][ky]; y faces likewise with S_x
        VPMX_TRY(dgemm(c, false, false, NY, np * NZ, my, SY, NY, 0, xp, my, 0, fd, NY, 0, 1));
        VPMX_TRY(dgemm(c, false, false, NX, np * NZ, mx, SX, NX, 0, yp, mx, 0, fd + oy, NX, 0, 1));
    }
    // z faces: v[b][ky][i] = sum_kx S_x(i, kx) zp[b][ky][kx], fd[b][j][i] = sum_ky v[b][ky][i] S_y(j, ky)
    VPMX_TRY(dgemm(c, false, false, NX, np * my, mx, SX, NX, 0, zp, mx, 0, v, NX, 0, 1));
    VPMX_TRY(dgemm(c, false, true, NX, NY, my, v, NX, (int64_t)NX * my, SY, NY, 0, fd + oz, NX, (int64_t)NX * NY, np));
    return VPMX_OK;
}

// ---------------------------------------------------------------- selected values of a solve
// The second fine solve of a single yaps block is read only on the inner box's faces (the TSC
// correction adds to them, the final solve takes them as Dirichlet values) and at the coarse
// sampling nodes (every ncoarse-th node); the final solve overwrites the rest of the inner box
// and nothing reads the margin (yaps3d.f90:104-138,254-261,340-395).  From the work array after
// the z pass (kx, ky, z): the x-face values are sum_kx sin(pi kx p / (mx+1)) (series in y) --
// face sums with sine weights, whose high-face weights are (-1)^(kx+1) times the low ones when
// the margins are equal -- then S_y; the y faces likewise; the z faces and the sample planes are
// planes of the work array, taken through S_x and S_y (full rows for the faces, the sample rows
// for the samples).  FP64 products on the matrix cores replace the y and x inverse passes.

// sin(pi p k / (m+1)), k = 1..m (index k - 1), uploaded once per (m, p).
static int sine_row(Ctx *c, int m, int p, const double **out) {
    auto key = std::make_pair(m, p);
    auto it = c->sinr_cache.find(key);
    if (it != c->sinr_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    const long double PI = 3.141592653589793238462643383279502884L;
    std::vector<double> v(m);
    for (int k = 1; k <= m; k++) {
        const long long q = ((long long)p * k) % (2LL * (m + 1));
        v[k - 1] = q % (m + 1) == 0 ? 0.0 : (double)sinl(PI * (long double)q / (long double)(m + 1));
    }
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * m));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * m, hipMemcpyHostToDevice));
    c->sinr_cache[key] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

// Rows p = a s (a = 1..m/s) of the DST-I matrix, sin(pi p (k+1) / (m+1)), column-major
// (m/s) x m, uploaded once per (m, s).
static int sine_rows(Ctx *c, int m, int step, const double **out) {
    auto key = std::make_pair(m, step);
    auto it = c->sinl_cache.find(key);
    if (it != c->sinl_cache.end()) {
        *out = (const double *)it->second;
        return VPMX_OK;
    }
    const int nr = m / step;
    const long double PI = 3.141592653589793238462643383279502884L;
    std::vector<double> v((size_t)nr * m);
    for (int k = 0; k < m; k++)
        for (int r = 0; r < nr; r++) {
            const long long q = ((long long)(r + 1) * step * (k + 1)) % (2LL * (m + 1));
            v[(size_t)k * nr + r] = q % (m + 1) == 0 ? 0.0 : (double)sinl(PI * (long double)q / (long double)(m + 1));
        }
    void *d = nullptr;
    VPMX_HIP(c, hipMalloc(&d, sizeof(double) * v.size()));
    VPMX_HIP(c, hipMemcpy(d, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    c->sinl_cache[key] = d;
    *out = (const double *)d;
    return VPMX_OK;
}

// Planes z = z0 + t dz (t < np) of the work array, FP64: out[t][comp][ky][kx].
template <typename T>
__global__ void __launch_bounds__(256) k_gather_planes(const T *__restrict__ w, int64_t wcs, int px, int mx, int my,
                                                       int ncomp, int z0, int dz, double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t pl = (int64_t)mx * my;
    if (e >= pl) return;
    const int b = blockIdx.y, t = b / ncomp, comp = b % ncomp;
    const int kx = (int)(e % mx), ky = (int)(e / mx);
    out[(int64_t)b * pl + e] = (double)w[comp * wcs + (int64_t)(z0 + t * dz) * px * my + (int64_t)ky * px + kx];
}

struct PartialTabs {
    const double *tx, *ty, *tz, *ps;  // x / y / z face values, sample values
    int xs[2], ys[2], zs[2];           // face nodes (1-based)
    int step, nlx, nly, nlz;           // sample step and interior sample counts
};

// Writes the selected nodes of `out` once each: the x faces, the y faces off the x faces, the
// z faces off both, the interior samples off every face.
__global__ void __launch_bounds__(256) k_partial_scatter(PartialTabs t, int nx, int ny, int nz, int ncomp, GView out) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int mx = nx - 2, my = ny - 2, mz = nz - 2;
    const int64_t nxf = (int64_t)2 * ncomp * my * mz, nyf = (int64_t)2 * ncomp * mx * mz, nzf = (int64_t)2 * ncomp * mx * my;
    const int64_t nsm = (int64_t)ncomp * t.nlx * t.nly * t.nlz;
    auto onx = [&](int i) { return i == t.xs[0] || i == t.xs[1]; };
    auto ony = [&](int j) { return j == t.ys[0] || j == t.ys[1]; };
    auto onz = [&](int k) { return k == t.zs[0] || k == t.zs[1]; };
    int i, j, k, comp;
    double v;
    if (e < nxf) {
        j = 2 + (int)(e % my); e /= my;
        k = 2 + (int)(e % mz); e /= mz;
        const int b = (int)e, side = b / ncomp;
        comp = b % ncomp;
        i = t.xs[side];
        v = t.tx[((int64_t)b * nz + (k - 1)) * ny + (j - 1)];
    } else if ((e -= nxf) < nyf) {
        i = 2 + (int)(e % mx); e /= mx;
        k = 2 + (int)(e % mz); e /= mz;
        const int b = (int)e, side = b / ncomp;
        comp = b % ncomp;
        j = t.ys[side];
        if (onx(i)) return;
        v = t.ty[((int64_t)b * nz + (k - 1)) * nx + (i - 1)];
    } else if ((e -= nyf) < nzf) {
        i = 2 + (int)(e % mx); e /= mx;
        j = 2 + (int)(e % my); e /= my;
        const int b = (int)e, side = b / ncomp;
        comp = b % ncomp;
        k = t.zs[side];
        if (onx(i) || ony(j)) return;
        v = t.tz[((int64_t)b * ny + (j - 1)) * nx + (i - 1)];
    } else if ((e -= nzf) < nsm) {
        const int ax = (int)(e % t.nlx); e /= t.nlx;
        const int ay = (int)(e % t.nly); e /= t.nly;
        const int b = (int)e;  // plane * ncomp + comp
        const int az = b / ncomp;
        comp = b % ncomp;
        i = 1 + (ax + 1) * t.step;
        j = 1 + (ay + 1) * t.step;
        k = 1 + (az + 1) * t.step;
        if (onx(i) || ony(j) || onz(k)) return;
        v = t.ps[((int64_t)b * t.nly + ay) * t.nlx + ax];
    } else {
        return;
    }
    out.p[comp * out.cs + out.at(i - 1, j - 1, k - 1)] = v;
}

static int partial_values(Ctx *c, const DstArgs &a, int ncomp, bool f32, const SolvePartial &sp, GView out) {
    int st = VPMX_OK;
    const int mx = a.mx, my = a.my, mz = a.mz, np = 2 * ncomp;
    const int NX = mx + 2, NY = my + 2, NZ = mz + 2;
    const int step = sp.step;
    if (mx > 64 * FR_COLS) return fail(c, VPMX_ERR_UNSUP, "selected solve values: rows of at most 320 points");
    for (int d = 0; d < 3; d++) {
        const int n = d == 0 ? NX : (d == 1 ? NY : NZ);
        const int *f = d == 0 ? sp.xs : (d == 1 ? sp.ys : sp.zs);
        if (f[0] < 2 || f[1] > n - 1 || f[0] >= f[1]) return fail(c, VPMX_ERR_ARG, "selected solve values: faces inside the grid");
        if (d < 2 && f[1] != n + 1 - f[0]) return fail(c, VPMX_ERR_ARG, "selected solve values: equal margins in x and y");
    }
    if (step < 1) return fail(c, VPMX_ERR_ARG, "selected solve values: sample step");
    const int nlx = mx / step, nly = my / step, nlz = mz / step;
    const double *wx = nullptr, *wy = nullptr, *SX = nullptr, *SY = nullptr, *LX = nullptr, *LY = nullptr;
    VPMX_TRY(sine_row(c, mx, sp.xs[0] - 1, &wx));
    VPMX_TRY(sine_row(c, my, sp.ys[0] - 1, &wy));
    VPMX_TRY(sine_matrix_ext(c, mx, &SX));
    VPMX_TRY(sine_matrix_ext(c, my, &SY));
    if (nlx > 0 && nly > 0 && nlz > 0) {
        VPMX_TRY(sine_rows(c, mx, step, &LX));
        VPMX_TRY(sine_rows(c, my, step, &LY));
    }
    // face sums, gathered planes, and the products' outputs
    const int64_t nxp = (int64_t)np * NZ * my, nyp = (int64_t)np * NZ * mx, nzc = (int64_t)np * my * mx;
    const int64_t nsc = (int64_t)ncomp * nlz * my * mx, ntx = (int64_t)np * NZ * NY, nty = (int64_t)np * NZ * NX;
    const int64_t nv = (int64_t)np * my * NX, ntz = (int64_t)np * NY * NX, nt5 = (int64_t)ncomp * nlz * my * nlx;
    const int64_t nps = (int64_t)ncomp * nlz * nly * nlx;
    double *xp = (double *)ensure(c, c->fp_spec,
                                  sizeof(double) * (size_t)(nxp + nyp + nzc + nsc + ntx + nty + nv + ntz + nt5 + nps), &st);
    if (st) return st;
    double *yp = xp + nxp, *zc = yp + nyp, *sc = zc + nzc, *tx = sc + nsc, *ty = tx + ntx, *v = ty + nty, *tz = v + nv;
    double *t5 = tz + ntz, *ps = t5 + nt5;
    // x and y face sums (the kernel of the normal derivatives, with sine weights; its tuning key)
    {
        const void *fns[2];
        if (f32) {
            fns[0] = (const void *)k_face_reduce_xy<float, 4>;
            fns[1] = (const void *)k_face_reduce_xy<float, 8>;
        } else {
            fns[0] = (const void *)k_face_reduce_xy<double, 2>;
            fns[1] = (const void *)k_face_reduce_xy<double, 4>;
        }
        const char *names[2] = {f32 ? "4 rows in flight" : "2 rows in flight", f32 ? "8 rows in flight" : "4 rows in flight"};
        const void *wp = a.w;
        int64_t wcs = a.wcs;
        int px = a.px, ncv = ncomp, mxv = mx, myv = my, mzv = mz;
        auto trial = [&](int var, int) -> int {
            void *kargs[] = {(void *)&wp, &wcs, &px, &mxv, &myv, &mzv, &ncv, &wx, &wy, &xp, &yp};
            VPMX_HIP(c, hipLaunchKernel(fns[var], dim3((unsigned)mz, (unsigned)ncomp), dim3(256), kargs, 0, c->stream));
            return VPMX_OK;
        };
        char key[128];
        std::snprintf(key, sizeof key, "face sums %s m=%d,%d,%d", f32 ? "FP32" : "FP64", mx, my, mz);
        Ctx::Tuned ch;
        VPMX_TRY(tune(c, key, 2, fns, names, 0, 256, trial, &ch, false));
        VPMX_TRY(trial(ch.fi, 0));
    }
    // the z-face planes and the interior sample planes
    const int64_t pl = (int64_t)mx * my;
    const unsigned gpl = (unsigned)((pl + 255) / 256);
    if (f32) {
        hipLaunchKernelGGL(k_gather_planes<float>, dim3(gpl, (unsigned)np), dim3(256), 0, c->stream, (const float *)a.w,
                           a.wcs, a.px, mx, my, ncomp, sp.zs[0] - 2, sp.zs[1] - sp.zs[0], zc);
        if (nlz > 0 && nlx > 0 && nly > 0)
            hipLaunchKernelGGL(k_gather_planes<float>, dim3(gpl, (unsigned)(ncomp * nlz)), dim3(256), 0, c->stream,
                               (const float *)a.w, a.wcs, a.px, mx, my, ncomp, step - 1, step, sc);
    } else {
        hipLaunchKernelGGL(k_gather_planes<double>, dim3(gpl, (unsigned)np), dim3(256), 0, c->stream, (const double *)a.w,
                           a.wcs, a.px, mx, my, ncomp, sp.zs[0] - 2, sp.zs[1] - sp.zs[0], zc);
        if (nlz > 0 && nlx > 0 && nly > 0)
            hipLaunchKernelGGL(k_gather_planes<double>, dim3(gpl, (unsigned)(ncomp * nlz)), dim3(256), 0, c->stream,
                               (const double *)a.w, a.wcs, a.px, mx, my, ncomp, step - 1, step, sc);
    }
    VPMX_HIP(c, hipGetLastError());
    // x faces tx[b][k][j] = sum_ky S_y(j, ky) xp[b][k][ky]; y faces ty[b][k][i]
    VPMX_TRY(dgemm(c, false, false, NY, np * NZ, my, SY, NY, 0, xp, my, 0, tx, NY, 0, 1));
    VPMX_TRY(dgemm(c, false, false, NX, np * NZ, mx, SX, NX, 0, yp, mx, 0, ty, NX, 0, 1));
    // z faces: v[b][ky][i] = sum_kx S_x(i, kx) zc[b][ky][kx]; tz[b][j][i] = sum_ky v[b][ky][i] S_y(j, ky)
    VPMX_TRY(dgemm(c, false, false, NX, np * my, mx, SX, NX, 0, zc, mx, 0, v, NX, 0, 1));
    VPMX_TRY(dgemm(c, false, true, NX, NY, my, v, NX, (int64_t)NX * my, SY, NY, 0, tz, NX, (int64_t)NX * NY, np));
    // samples: t5[b][ky][a] = sum_kx L_x(a, kx) sc[b][ky][kx]; ps[b][ay][ax] = sum_ky t5[b][ky][ax] L_y(ay, ky)
    if (nlx > 0 && nly > 0 && nlz > 0) {
        VPMX_TRY(dgemm(c, false, false, nlx, ncomp * nlz * my, mx, LX, nlx, 0, sc, mx, 0, t5, nlx, 0, 1));
        VPMX_TRY(dgemm(c, false, true, nlx, nly, my, t5, nlx, (int64_t)nlx * my, LY, nly, 0, ps, nlx,
                       (int64_t)nlx * nly, ncomp * nlz));
    }
    PartialTabs t;
    t.tx = tx; t.ty = ty; t.tz = tz; t.ps = ps;
    for (int q = 0; q < 2; q++) { t.xs[q] = sp.xs[q]; t.ys[q] = sp.ys[q]; t.zs[q] = sp.zs[q]; }
    t.step = step;
    t.nlx = nlx; t.nly = nly; t.nlz = nlz;
    const int64_t tot = (int64_t)np * (my * (int64_t)mz + mx * (int64_t)mz + mx * (int64_t)my) + nps;
    hipLaunchKernelGGL(k_partial_scatter, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, t, NX, NY, NZ,
                       ncomp, out);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

int poisson_dirichlet(Ctx *c, int ncomp, const int wn[3], const double h[3], GView rhs, const GView *bc, GView out,
                      int reuse, int form, double *face_nd, const SolvePartial *part) {
    int st = VPMX_OK;
    int m[3] = {wn[0] - 2, wn[1] - 2, wn[2] - 2};
    if (m[0] < 1 || m[1] < 1 || m[2] < 1) return fail(c, VPMX_ERR_ARG, "Poisson window needs >= 3 nodes per axis");
    if (bc && bc->p != out.p) return fail(c, VPMX_ERR_ARG, "Poisson: BC view must be the output view");
    for (int d = 0; d < 3; d++)
        if (2 * (int64_t)(m[d] + 1) * sizeof(double2) * 2 > 65536)
            return fail(c, VPMX_ERR_UNSUP, "Poisson: line length too long for LDS");
    double *w = nullptr;
    AxisPlan ax[3];
    for (int d = 0; d < 3; d++) {
        ax[d].m = m[d];
        VPMX_TRY(get_plan(c, 2 * (m[d] + 1), &ax[d].full));
        VPMX_TRY(get_plan(c, m[d] + 1, &ax[d].half));
        VPMX_TRY(sine_table(c, m[d] + 1, &ax[d].sn));
    }
    DstArgs a{};
    a.mx = m[0]; a.my = m[1]; a.mz = m[2];
    a.px = m[0];
    a.rhs = rhs;
    a.out = out;
    a.has_bc = bc ? 1 : 0;
    if (bc) a.bc = *bc;
    for (int d = 0; d < 3; d++) a.ih2[d] = 1.0 / (h[d] * h[d]);
    VPMX_TRY(eigen(c, m[0], h[0], c->lam[0], &a.lam[0]));
    VPMX_TRY(eigen(c, m[1], h[1], c->lam[1], &a.lam[1]));
    VPMX_TRY(eigen(c, m[2], h[2], c->lam[2], &a.lam[2]));
    a.norm = (2.0 / (m[0] + 1)) * (2.0 / (m[1] + 1)) * (2.0 / (m[2] + 1));
    a.scale = 1.0;
    // the FP32 passes run when every axis takes the one-wave kernel in both forms; SOLVE_CTX
    // follows the context's precision (FP32 passes in the mixed form)
    if (form == SOLVE_CTX) form = c->precision == 1 ? SOLVE_FP32 : SOLVE_FP64;
    const bool wave32 = wave_class(ax[0], true) && wave_class(ax[1], true) && wave_class(ax[2], true);
    const bool wave64 = wave_class(ax[0], false) && wave_class(ax[1], false) && wave_class(ax[2], false);
    const bool inv64 = form == SOLVE_INV64 && wave32 && wave64 && c->dst_kernel == 4;
    const bool f32 = (form == SOLVE_FP32 && wave32) || inv64;
    if (inv64) reuse = 0;  // (the fine solves' reuse keeps FP32 arrays)
    // one-wave kernels on every axis: rows of the work array padded to 128 bytes
    const bool wave_all = f32 || (c->dst_kernel == 4 && wave_class(ax[0], false) && wave_class(ax[1], false) &&
                                  wave_class(ax[2], false));
    if (wave_all) {
        const int q = f32 ? 32 : 16;
        a.px = (m[0] + q - 1) / q * q;
    }
    a.wcs = (int64_t)a.px * m[1] * m[2];
    w = (double *)ensure(c, c->dst_work, sizeof(double) * (size_t)a.wcs * ncomp, &st);
    if (st) return st;
    a.w = a.wo = w;
    const int key[6] = {m[0], m[1], m[2], a.px, ncomp, f32 ? 1 : 0};
    const size_t wbytes = (f32 ? sizeof(float) : sizeof(double)) * (size_t)a.wcs * ncomp;
    if (reuse == 2 && bc && std::equal(key, key + 6, c->save_key)) {
        // the kept x/y forward transforms of the RHS (still in the work array), less those of
        // the folded Dirichlet terms
        VPMX_TRY(fold_correct(c, a, ncomp, f32, *bc));
        c->save_key[5] = -1;
    } else {
        // x forward (with boundary folding)
        a.nlines = (int64_t)m[1] * m[2];
        VPMX_TRY(run_pass<DST_XF>(c, a, ncomp, ax[0], f32));
        // y forward
        a.nlines = (int64_t)a.px * m[2];
        VPMX_TRY(run_pass<DST_Y>(c, a, ncomp, ax[1], f32));
        c->save_key[5] = -1;
        // keep them: the rest of this solve runs in a second array (one-wave kernels only; the
        // correction pass takes a row per work-group)
        if (reuse == 1 && !bc && wave_all && a.px <= 320) {
            double *w2 = (double *)ensure(c, c->dst_save, wbytes, &st);
            if (st) return st;
            a.wo = w2;
            std::copy(key, key + 6, c->save_key);
        }
    }
    // z forward + diagonal + inverse (into wo), then the inverse passes in place there; for
    // the face derivatives alone, with the one-wave kernel, z forward + diagonal and the z
    // faces' sums
    a.nlines = (int64_t)a.px * m[1];
    FaceBufs fb{};
    const bool zspec = face_nd && !inv64 && (f32 || (c->dst_kernel == 4 && wave_class(ax[2], false)));
    if (face_nd) {
        VPMX_TRY(face_bufs(c, m[0], m[1], m[2], ncomp, &fb));
        if (zspec) {
            a.zsum = fb.zp;
            VPMX_TRY(face_weights(c, m[2], h[2], &a.zw));
        }
    }
    if (inv64) {
        // FP32 forward transforms and diagonal; the inverse passes in FP64
        double *w64 = (double *)ensure(c, c->dst_work64, sizeof(double) * (size_t)a.wcs * ncomp, &st);
        if (st) return st;
        VPMX_TRY(launch_zmix(c, a, ncomp, ax[2], w64));
        a.w = a.wo = w64;
    } else {
        VPMX_TRY(run_pass<DST_Z>(c, a, ncomp, ax[2], f32));
        a.w = a.wo;
    }
    if (face_nd) {
        if (inv64 || bc) return fail(c, VPMX_ERR_ARG, "Poisson: face derivatives of a zero-BC solve only");
        a.zsum = nullptr;
        return face_derivs_spectral(c, a, ncomp, f32, h, face_nd, zspec, fb);
    }
    if (part) {
        if (inv64) return fail(c, VPMX_ERR_ARG, "Poisson: selected values of the FP32/FP64 passes only");
        return partial_values(c, a, ncomp, f32, *part, out);
    }
    const bool fi = f32 && !inv64;
    // y inverse
    a.nlines = (int64_t)a.px * m[2];
    VPMX_TRY(run_pass<DST_Y>(c, a, ncomp, ax[1], fi));
    // x inverse -> output interior
    a.nlines = (int64_t)m[1] * m[2];
    VPMX_TRY(run_pass<DST_XI>(c, a, ncomp, ax[0], fi));
    if (!bc) VPMX_TRY(zero_faces(c, wn, out, ncomp));
    return VPMX_OK;
}

int zero_faces(Ctx *c, const int n[3], GView v, int ncomp) {
    const int64_t tot = 2 * ((int64_t)n[0] * n[1] + (int64_t)n[0] * n[2] + (int64_t)n[1] * n[2]);
    if ((int64_t)n[0] * n[1] * n[2] <= 0) return VPMX_OK;
    hipLaunchKernelGGL(k_zero_faces, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, v, n[0], n[1], n[2],
                       ncomp);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

}  // namespace vpmx
