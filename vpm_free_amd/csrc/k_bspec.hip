// Spectral evaluation of the infinite-domain boundary sums (Bounds3d -> calc_boundinf_3d,
// pmbound.f90:80-157, 415-515, ibctyp=2).  Same sum as the direct kernels,
//   SOL(c,t) += sum_s a_c(s) / |x_t - x_s|,   a = -sigma * ds / (4 pi),
// over the face-cell sources s of the six block faces and every boundary node t, exact up
// to rounding.
//
// Targets on face B, sources on face A.  Both faces contain an in-plane axis C (for
// perpendicular faces the one axis they share, for parallel faces the higher in-plane
// axis), along which the kernel depends only on the offset e = kt - ks between the target
// node and the source cell:  |x_t - x_s|^2 = K(p, q) + ((e - 1/2) h_C)^2, with p the
// target's other coordinate and q the source's.  So each (B, A) pair is a 1-D convolution
// along C -- real FFTs of length L >= 2 n_C - 2, which leaves the needed outputs free of
// wrap-around -- and a dense sum over q, which for every frequency w is a matrix product
//   O^(p, col; w) = sum_q G^(p, q; w) A^(q, col; w),
// batched over the L/2+1 frequencies.  The kernel is even about e = 1/2 (it depends on
// (e - 1/2)^2), so with the transform X(w) = sum_r x_r exp(-2 pi i w r / L)
//   G^(p, q; w) = exp(-pi i w / L) R(p, q; w),   R = sum_e g(e) cos(2 pi w (e - 1/2) / L)
// real: the table holds R, the product is a real one with the real and imaginary parts of
// A^ as separate columns (half the bytes and flops of a complex product), and the phase
// is applied when the inverse transform reads O^.  Pairs with the same geometry up to a
// reflection of p or q share one table: on a cubic block three (perpendicular, same face,
// opposite face) serve all 36 pairs, whose pairs x components are the columns.  The tables
// depend only on the block geometry and are kept across calls.
//
// For parallel faces p and q run along the same axis and the kernel depends on p - q only,
// so R(p, q; w) = r_w(p - q) and the sum over q is a convolution too: it runs as FFTs of
// length Lq >= M + K - 1 along q (the table holds FFT(r_w) / Lq, nw x Lq values, instead of
// the nw x M x K matrices).  Only the perpendicular family needs a matrix product.
//
// The products run in rocBLAS (FP64, atomics off, so the results do not depend on
// scheduling); the transforms run in LDS with the Stockham FFT of the Poisson passes.  A
// final pass adds, for every boundary node, the six pair contributions of its owner face
// in a fixed order (X faces own their edges, Y faces their z-edges).

#include <climits>
#include <cmath>
#include <vector>

#include "fft_lds.h"
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

constexpr int SPEC_MAXR = 7;  // FFT lengths are chosen 7-smooth

struct SpecFft {  // FFT shape for fft_reg
    int L, npairs, nstages, radix[24];
};

struct SpecTab {  // one Green's-function matrix family
    int type;     // 0: perpendicular faces, 1: parallel faces
    int M, K;     // target coordinates p < M, source coordinates q < K
    int nC, L, nw;
    int poff;     // parallel faces: the table's line p stands for the offset p - poff (0 for matrices)
    int Lq;       // parallel faces: convolution length along q (0: matrix product)
    double hT, hS, hC, d;
};

struct SpecCol {  // one column: a source face line family, one component
    int64_t src_off;  // first source of face A in the source array
    int ustride;      // row length of face A's source array (its lower in-plane axis)
    int s_is_u;       // q runs along A's lower in-plane axis (else C does)
    int flipq;        // q -> K-1-q
    int comp;
    int pair;         // B * 6 + A (the transposed sums read target face B's nodes)
};

struct SpecPair {  // where the contribution of source face A to target face B sits
    int64_t base;  // offset of component 0 in Ores ([col][p][kt])
    int M, nC, tax, cax, flipp, pad;
};

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ int residue_offset(int r, int L, int nC, bool *ok) {
    // circular index r of the kernel -> offset e in [-(nC-2), nC-1]
    *ok = true;
    if (r <= nC - 1) return r;
    if (r >= L - (nC - 2)) return r - L;
    *ok = false;
    return 0;
}

// Two real lines packed as one complex FFT -> their half spectra, k < nw.
__device__ __forceinline__ double2 unpack_half(const double2 *A, int L, int pr, int k, int odd) {
    const double2 z = A[pr * L + k], zc = A[pr * L + (k == 0 ? 0 : L - k)];
    if (!odd) return make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y - zc.y));
    return make_double2(0.5 * (z.y + zc.y), -0.5 * (z.x - zc.x));
}

// exp(-pi i k / L): the phase of the half-sample shift of the kernel's centre.
__device__ __forceinline__ double2 half_shift(int k, int L) {
    double s, c;
    sincospi((double)k / (double)L, &s, &c);
    return make_double2(c, -s);
}

// R[w][q][p] (real, G^ = exp(-pi i w / L) R) for p in the block's lines, one q per blockIdx.y.
__global__ void __launch_bounds__(DST_THREADS) k_bsx_table(SpecFft f, SpecTab t, const double2 *tw, double *G) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    const int L = f.L, np = f.npairs, nl = 2 * np;
    double2 *A = lds, *twl = lds + np * L;
    double *Ad = (double *)A;
    const int tid = threadIdx.x, p0 = blockIdx.x * nl, q = blockIdx.y;
    for (int i = tid; i < L; i += DST_THREADS) twl[i] = tw[i];
    for (int i = tid; i < nl * L; i += DST_THREADS) {
        const int l = i / L, r = i - l * L, p = p0 + l;
        bool ok;
        const int e = residue_offset(r, L, t.nC, &ok);
        double v = 0.0;
        if (ok && p < t.M) {
            const double dc = ((double)e - 0.5) * t.hC;
            double r2;
            if (t.type == 0) {
                const double x = (double)p * t.hT, y = ((double)q + 0.5) * t.hS;
                r2 = (x * x + y * y) + dc * dc;
            } else {
                const double x = ((double)(p - t.poff - q) - 0.5) * t.hT;
                r2 = (x * x + t.d * t.d) + dc * dc;
            }
            v = 1.0 / sqrt(r2);
        }
        Ad[2 * ((l >> 1) * L + r) + (l & 1)] = v;
    }
    __syncthreads();
    fft_reg<SPEC_MAXR>(A, twl, f);
    for (int i = tid; i < t.nw * nl; i += DST_THREADS) {
        const int k = i / nl, l = i - k * nl, p = p0 + l;
        if (p < t.M) {
            const double2 z = unpack_half(A, L, l >> 1, k, l & 1), ph = half_shift(k, L);
            G[((int64_t)k * t.K + q) * t.M + p] = z.x * ph.x + z.y * ph.y;  // Re(z / ph)
        }
    }
}

// Transforms of the forward and inverse kernels.  FFT = -2: the work-group Stockham stages
// (f.npairs sequences, DST_THREADS threads); -1: one wave per sequence with the runtime plan P;
// 0: one wave per sequence with the 576-point compile-time plan.  The wave forms run
// BW_WAVES sequences per work-group and need no work-group barrier between their stages.
constexpr int BW_WAVES = 4, BW_LMAX = 576;
template <int FFT>
constexpr int bsx_threads() { return FFT == -2 ? DST_THREADS : 64 * BW_WAVES; }
template <int FFT>
__device__ __forceinline__ void bsx_fft(double2 *A, const double2 *twl, const SpecFft &f, const WavePlan &P) {
    if constexpr (FFT == -2) {
        fft_reg<SPEC_MAXR>(A, twl, f);  // ends with a barrier
    } else {
        const int lane = threadIdx.x & 63;
        double2 *S = A + (threadIdx.x >> 6) * P.L;
        if constexpr (FFT == 0) FixedFft<0>::template run<double2, double2 *, false>(S, twl, lane);
        else wave_fft<SPEC_MAXR, double2, double2 *, BW_LMAX>(S, twl, P, lane);
        __syncthreads();
    }
}

// A^ of every column as two real columns: Bm[2 col][w][q] = Re, Bm[2 col + 1][w][q] = Im,
// lines q of the block, one column per blockIdx.y.
template <int FFT>
__global__ void __launch_bounds__(bsx_threads<FFT>()) k_bsx_fwd(SpecFft f, WavePlan P, SpecTab t, const double2 *tw,
                                                                 const SpecCol *cols, const double *a, int64_t ns,
                                                                 double *Bm) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    constexpr int NT = bsx_threads<FFT>();
    const int L = f.L, np = FFT == -2 ? f.npairs : BW_WAVES, nl = 2 * np;
    double2 *A = lds, *twl = lds + np * L;
    double *Ad = (double *)A;
    const int tid = threadIdx.x, q0 = blockIdx.x * nl, col = blockIdx.y;
    const SpecCol cd = cols[col];
    const double *src = a + (int64_t)cd.comp * ns + cd.src_off;
    const int ncell = t.nC - 1;  // source cells along C
    for (int i = tid; i < L; i += NT) twl[i] = tw[i];
    if (cd.s_is_u) {
        // q along the face's contiguous axis: neighbouring threads take neighbouring lines
        for (int i = tid; i < nl * L; i += NT) {
            const int r = i / nl, l = i - r * nl, q = q0 + l;
            double v = 0.0;
            if (r < ncell && q < t.K) {
                const int qs = cd.flipq ? t.K - 1 - q : q;
                v = src[(int64_t)r * cd.ustride + qs];
            }
            Ad[2 * ((l >> 1) * L + r) + (l & 1)] = v;
        }
    } else {
        for (int i = tid; i < nl * L; i += NT) {
            const int l = i / L, r = i - l * L, q = q0 + l;
            double v = 0.0;
            if (r < ncell && q < t.K) {
                const int qs = cd.flipq ? t.K - 1 - q : q;
                v = src[(int64_t)qs * cd.ustride + r];
            }
            Ad[2 * ((l >> 1) * L + r) + (l & 1)] = v;
        }
    }
    __syncthreads();
    bsx_fft<FFT>(A, twl, f, P);
    const int64_t cst = (int64_t)t.nw * t.K;
    double *out = Bm + 2 * col * cst;
    for (int i = tid; i < t.nw * nl; i += NT) {
        const int k = i / nl, l = i - k * nl, q = q0 + l;
        if (q < t.K) {
            const double2 z = unpack_half(A, L, l >> 1, k, l & 1);
            out[(int64_t)k * t.K + q] = z.x;
            out[cst + (int64_t)k * t.K + q] = z.y;
        }
    }
}

// Inverse real FFT over w of O^ = exp(-pi i w / L) (Cm[2 col][w][p] + i Cm[2 col + 1][w][p]),
// first nC outputs -> Ores[col][p][kt].
template <int FFT>
__global__ void __launch_bounds__(bsx_threads<FFT>()) k_bsx_inv(SpecFft f, WavePlan P, SpecTab t, const double2 *tw,
                                                                 const double *Cm, double *Ores, int64_t ores_off) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    constexpr int NT = bsx_threads<FFT>();
    const int L = f.L, np = FFT == -2 ? f.npairs : BW_WAVES, nl = 2 * np;
    double2 *A = lds, *twl = lds + np * L;
    const int tid = threadIdx.x, p0 = blockIdx.x * nl, col = blockIdx.y;
    const int64_t cst = (int64_t)t.nw * t.M;
    const double *in = Cm + 2 * col * cst;
    for (int i = tid; i < L; i += NT) twl[i] = tw[i];
    // Z_k = X_k + i Y_k with the Hermitian halves of lines p (X) and p+1 (Y)
    for (int i = tid; i < t.nw * np; i += NT) {
        const int k = i / np, pr = i - k * np, p = p0 + 2 * pr;
        const double2 ph = half_shift(k, L);
        const int64_t o = (int64_t)k * t.M + p;
        const double xr = p < t.M ? in[o] : 0.0, xi = p < t.M ? in[cst + o] : 0.0;
        const double yr = p + 1 < t.M ? in[o + 1] : 0.0, yi = p + 1 < t.M ? in[cst + o + 1] : 0.0;
        const double2 X = make_double2(xr * ph.x - xi * ph.y, xr * ph.y + xi * ph.x);
        const double2 Y = make_double2(yr * ph.x - yi * ph.y, yr * ph.y + yi * ph.x);
        if (k == 0 || 2 * k == L) {
            A[pr * L + k] = make_double2(X.x, Y.x);
        } else {
            A[pr * L + k] = make_double2(X.x - Y.y, X.y + Y.x);
            A[pr * L + L - k] = make_double2(X.x + Y.y, Y.x - X.y);
        }
    }
    __syncthreads();
    bsx_fft<FFT>(A, twl, f, P);
    // z_n = FFT(Z)[(L - n) mod L] / L
    const double invL = 1.0 / (double)L;
    double *out = Ores + ores_off + (int64_t)col * t.M * t.nC;
    for (int i = tid; i < nl * t.nC; i += NT) {
        const int l = i / t.nC, n = i - l * t.nC, p = p0 + l;
        if (p >= t.M) continue;
        const double2 w = A[(l >> 1) * L + (n == 0 ? 0 : L - n)];
        out[(int64_t)p * t.nC + n] = ((l & 1) ? w.y : w.x) * invL;
    }
}

// Parallel-face table: FFT over the offset d = p - q of r_w(d), stored at the circular
// position d mod Lq (d in [-(K-1), M-1], M + K - 1 <= Lq), divided by Lq for the inverse
// transform of the convolution.  Rl[w][d + K - 1] holds r_w(d); one w per sequence.
__global__ void __launch_bounds__(DST_THREADS) k_bsx_toep_fft(SpecFft fq, SpecTab t, const double2 *twq,
                                                              const double *Rl, double2 *T2) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    const int Lq = fq.L, np = fq.npairs, nd = t.M + t.K - 1;
    double2 *A = lds, *twl = lds + np * Lq;
    const int tid = threadIdx.x, w0 = blockIdx.x * np;
    for (int i = tid; i < Lq; i += DST_THREADS) twl[i] = twq[i];
    for (int i = tid; i < np * Lq; i += DST_THREADS) {
        const int s = i / Lq, pos = i - s * Lq, w = w0 + s;
        int d = pos <= t.M - 1 ? pos : (pos >= Lq - (t.K - 1) ? pos - Lq : INT_MIN);
        double v = 0.0;
        if (w < t.nw && d != INT_MIN) v = Rl[(int64_t)w * nd + d + t.K - 1];
        A[i] = make_double2(v, 0.0);
    }
    __syncthreads();
    fft_reg<SPEC_MAXR>(A, twl, fq);
    const double inv = 1.0 / (double)Lq;
    for (int i = tid; i < np * Lq; i += DST_THREADS) {
        const int s = i / Lq, v = i - s * Lq, w = w0 + s;
        if (w < t.nw) T2[(int64_t)w * Lq + v] = make_double2(A[i].x * inv, A[i].y * inv);
    }
}

// Parallel faces: O'(p; w) = sum_q r_w(p - q) A^(q; w) for p < M as a circular convolution
// of length Lq, with the real and imaginary parts of A^ (Bm[2 col], Bm[2 col + 1]) as one
// complex sequence (r_w is real); results to Cm[2 col], Cm[2 col + 1] like the matrix
// product's.  One w per sequence, one column per blockIdx.y.
__global__ void __launch_bounds__(DST_THREADS) k_bsx_qconv(SpecFft fq, SpecTab t, const double2 *twq,
                                                           const double2 *T2, const double *Bm, double *Cm, int cj) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    const int Lq = fq.L, np = fq.npairs;
    double2 *A = lds, *twl = lds + np * Lq;
    const int tid = threadIdx.x, w0 = blockIdx.x * np, col = blockIdx.y;
    const int64_t cb = (int64_t)t.nw * t.K, cc = (int64_t)t.nw * t.M;
    const double *bre = Bm + 2 * col * cb, *bim = bre + cb;
    for (int i = tid; i < Lq; i += DST_THREADS) twl[i] = twq[i];
    for (int i = tid; i < np * Lq; i += DST_THREADS) {
        const int s = i / Lq, q = i - s * Lq, w = w0 + s;
        double2 z = make_double2(0.0, 0.0);
        if (w < t.nw && q < t.K) {
            const int64_t o = (int64_t)w * t.K + q;
            z = make_double2(bre[o], bim[o]);
        }
        A[i] = z;
    }
    __syncthreads();
    fft_reg<SPEC_MAXR>(A, twl, fq);
    for (int i = tid; i < np * Lq; i += DST_THREADS) {
        const int s = i / Lq, v = i - s * Lq, w = w0 + s;
        if (w < t.nw) {
            double2 r = T2[(int64_t)w * Lq + v];
            if (cj) r.y = -r.y;  // correlation: the transposed sums
            A[i] = cmul(A[i], r);
        }
    }
    __syncthreads();
    // inverse transform: FFT again, output n at index (Lq - n) mod Lq (1/Lq is in the table)
    fft_reg<SPEC_MAXR>(A, twl, fq);
    double *ore = Cm + 2 * col * cc, *oim = ore + cc;
    for (int i = tid; i < np * t.M; i += DST_THREADS) {
        const int s = i / t.M, p = i - s * t.M, w = w0 + s;
        if (w >= t.nw) continue;
        const double2 z = A[s * Lq + (p == 0 ? 0 : Lq - p)];
        ore[(int64_t)w * t.M + p] = z.x;
        oim[(int64_t)w * t.M + p] = z.y;
    }
}

// The same convolution with one wave per (column, w) sequence (Lq <= QC_LMAX): no work-group
// barriers after the twiddles are staged, QC_WAVES sequences per work-group.
constexpr int QC_WAVES = 4, QC_LMAX = 576;
template <int FFT>  // 0: the 576-point compile-time plan, -1: the runtime plan P
__global__ void __launch_bounds__(64 * QC_WAVES) k_bsx_qconv_w(WavePlan P, SpecTab t, const double2 *twq,
                                                               const double2 *T2, const double *Bm, double *Cm,
                                                               int ncol, int cj) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    const int Lq = P.L, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double2 *twl = lds + QC_WAVES * Lq;
    for (int i = threadIdx.x; i < Lq; i += 64 * QC_WAVES) twl[i] = twq[i];
    __syncthreads();
    const int seq = blockIdx.x * QC_WAVES + wv;  // col * nw + w
    if (seq >= ncol * t.nw) return;               // no work-group barrier below
    const int col = seq / t.nw, w = seq - col * t.nw;
    const int64_t cb = (int64_t)t.nw * t.K, cc = (int64_t)t.nw * t.M;
    const double *bre = Bm + 2 * col * cb + (int64_t)w * t.K, *bim = bre + cb;
    double2 *A = lds + wv * Lq;
    for (int q = lane; q < Lq; q += 64) A[q] = q < t.K ? make_double2(bre[q], bim[q]) : make_double2(0.0, 0.0);
    __builtin_amdgcn_wave_barrier();
    if constexpr (FFT == 0) FixedFft<0>::template run<double2, double2 *, false>(A, twl, lane);
    else wave_fft<SPEC_MAXR, double2, double2 *, QC_LMAX>(A, twl, P, lane);
    const double2 *tr = T2 + (int64_t)w * Lq;
    for (int v = lane; v < Lq; v += 64) {
        double2 r = tr[v];
        if (cj) r.y = -r.y;  // correlation: the transposed sums
        A[v] = cmul(A[v], r);
    }
    __builtin_amdgcn_wave_barrier();
    if constexpr (FFT == 0) FixedFft<0>::template run<double2, double2 *, false>(A, twl, lane);
    else wave_fft<SPEC_MAXR, double2, double2 *, QC_LMAX>(A, twl, P, lane);
    double *ore = Cm + 2 * col * cc + (int64_t)w * t.M, *oim = ore + cc;
    for (int p = lane; p < t.M; p += 64) {
        const double2 z = A[p == 0 ? 0 : Lq - p];
        ore[p] = z.x;
        oim[p] = z.y;
    }
}

// SOL(t) += sum over the six source faces (fixed order) of the owner face's pair outputs.
// blockIdx.y is the owner face B (X faces whole, Y faces without the x-edges, Z faces without
// the x- and y-edges); each wave takes an 8 x 8 patch of it, so the pair outputs, contiguous
// along one face axis or the other, are read in 64-byte runs either way.
constexpr int GATHER_WAVES = 4;
__host__ __device__ inline void gather_face(const int n[3], int B, int *na, int *nb) {
    const int ax = B >> 1;
    *na = ax == 0 ? n[1] : n[0] - 2;
    *nb = ax == 2 ? n[1] - 2 : n[2];
}
__global__ void __launch_bounds__(64 * GATHER_WAVES) k_bsx_gather(Grid g, const SpecPair *pairs, const double *Ores,
                                                                  GView sol) {
#pragma clang fp contract(off)
    const int B = blockIdx.y, ax = B >> 1, hi = B & 1;
    int na, nb;
    gather_face(g.n, B, &na, &nb);
    const int npa = (na + 7) >> 3, patch = blockIdx.x * GATHER_WAVES + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (patch >= npa * ((nb + 7) >> 3)) return;
    const int a = (patch % npa) * 8 + (l & 7), b = (patch / npa) * 8 + (l >> 3);
    if (a >= na || b >= nb) return;
    int i, j, k;  // 1-based
    if (ax == 0) {
        i = hi ? g.n[0] : 1;
        j = a + 1;
        k = b + 1;
    } else if (ax == 1) {
        j = hi ? g.n[1] : 1;
        i = a + 2;
        k = b + 1;
    } else {
        k = hi ? g.n[2] : 1;
        i = a + 2;
        j = b + 2;
    }
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (int A = 0; A < 6; A++) {
        const SpecPair P = pairs[B * 6 + A];
        int p = P.tax == 0 ? i - 1 : (P.tax == 1 ? j - 1 : k - 1);
        const int kt = P.cax == 0 ? i - 1 : (P.cax == 1 ? j - 1 : k - 1);
        if (P.flipp) p = P.M - 1 - p;
        const int64_t cst = (int64_t)P.M * P.nC;
        const double *o = Ores + P.base + (int64_t)p * P.nC + kt;
        acc0 += o[0];
        acc1 += o[cst];
        acc2 += o[2 * cst];
    }
    const int64_t n = sol.at(i - 1, j - 1, k - 1);
    sol.p[n] += acc0;
    sol.p[sol.cs + n] += acc1;
    sol.p[2 * sol.cs + n] += acc2;
}

// ------------------------------------------------------------------ transposed sums
// The adjoint of the sums: abar(c, s) = sum_t gbar(c, t) / |x_t - x_s| over every boundary
// node t (each on its owner face, as the gather assigns it) for every face-cell source s.
// Per (B, A) pair the forward is a convolution along C and a product (or convolution) over q;
// the transpose is the correlation along C -- the conjugate kernel spectrum, exp(+pi i w / L) R
// with the same real table R -- and the transposed product over p (R^T, or the conjugate of
// the q-convolution's spectrum).  Columns and tables are the forward plan's.

// Node lines p of the column's target face B (owned nodes, else 0), n < nC along C, ->
// their half spectra: Bm[2 col][w][p] = Re, Bm[2 col + 1][w][p] = Im.
template <int FFT>
__global__ void __launch_bounds__(bsx_threads<FFT>()) k_bsxT_fwd(SpecFft f, WavePlan P, SpecTab t, const double2 *tw,
                                                                  const SpecCol *cols, const SpecPair *pairs, Grid g,
                                                                  GView gb, double *Bm) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    constexpr int NT = bsx_threads<FFT>();
    const int L = f.L, np = FFT == -2 ? f.npairs : BW_WAVES, nl = 2 * np;
    double2 *A = lds, *twl = lds + np * L;
    double *Ad = (double *)A;
    const int tid = threadIdx.x, p0 = blockIdx.x * nl, col = blockIdx.y;
    const SpecCol cd = cols[col];
    const SpecPair PP = pairs[cd.pair];
    const int B = cd.pair / 6, ax = B >> 1, hi = B & 1;
    const double *src = gb.p + (int64_t)cd.comp * gb.cs;
    for (int i = tid; i < L; i += NT) twl[i] = tw[i];
    for (int i = tid; i < nl * L; i += NT) {
        const int l = i / L, r = i - l * L, p = p0 + l;
        double v = 0.0;
        if (r < t.nC && p < t.M) {
            int c3[3];  // 0-based node coordinates
            c3[ax] = hi ? g.n[ax] - 1 : 0;
            c3[PP.tax] = PP.flipp ? t.M - 1 - p : p;
            c3[PP.cax] = r;
            // owner faces: X faces whole, Y faces without the x-edges, Z faces without x- and y-edges
            const bool own = ax == 0 || (c3[0] >= 1 && c3[0] <= g.n[0] - 2 && (ax == 1 || (c3[1] >= 1 && c3[1] <= g.n[1] - 2)));
            if (own) v = src[gb.at(c3[0], c3[1], c3[2])];
        }
        Ad[2 * ((l >> 1) * L + r) + (l & 1)] = v;
    }
    __syncthreads();
    bsx_fft<FFT>(A, twl, f, P);
    const int64_t cst = (int64_t)t.nw * t.M;
    double *out = Bm + 2 * col * cst;
    for (int i = tid; i < t.nw * nl; i += NT) {
        const int k = i / nl, l = i - k * nl, p = p0 + l;
        if (p < t.M) {
            const double2 z = unpack_half(A, L, l >> 1, k, l & 1);
            out[(int64_t)k * t.M + p] = z.x;
            out[cst + (int64_t)k * t.M + p] = z.y;
        }
    }
}

// Inverse real FFT over w of exp(+pi i w / L) (Cm[2 col][w][q] + i Cm[2 col + 1][w][q]), the
// first nC - 1 outputs (source cells along C) -> apart[B][comp][s] at the source's index in
// face A's array (every (B, A) column writes its own slice, so no two columns meet).
template <int FFT>
__global__ void __launch_bounds__(bsx_threads<FFT>()) k_bsxT_inv(SpecFft f, WavePlan P, SpecTab t, const double2 *tw,
                                                                  const SpecCol *cols, const double *Cm, int64_t ns,
                                                                  double *apart) {
#pragma clang fp contract(off)
    extern __shared__ double2 lds[];
    constexpr int NT = bsx_threads<FFT>();
    const int L = f.L, np = FFT == -2 ? f.npairs : BW_WAVES, nl = 2 * np;
    double2 *A = lds, *twl = lds + np * L;
    const int tid = threadIdx.x, q0 = blockIdx.x * nl, col = blockIdx.y;
    const SpecCol cd = cols[col];
    const int64_t cst = (int64_t)t.nw * t.K;
    const double *in = Cm + 2 * col * cst;
    for (int i = tid; i < L; i += NT) twl[i] = tw[i];
    for (int i = tid; i < t.nw * np; i += NT) {
        const int k = i / np, pr = i - k * np, q = q0 + 2 * pr;
        double2 ph = half_shift(k, L);
        ph.y = -ph.y;  // exp(+pi i k / L)
        const int64_t o = (int64_t)k * t.K + q;
        const double xr = q < t.K ? in[o] : 0.0, xi = q < t.K ? in[cst + o] : 0.0;
        const double yr = q + 1 < t.K ? in[o + 1] : 0.0, yi = q + 1 < t.K ? in[cst + o + 1] : 0.0;
        const double2 X = make_double2(xr * ph.x - xi * ph.y, xr * ph.y + xi * ph.x);
        const double2 Y = make_double2(yr * ph.x - yi * ph.y, yr * ph.y + yi * ph.x);
        if (k == 0 || 2 * k == L) {
            A[pr * L + k] = make_double2(X.x, Y.x);
        } else {
            A[pr * L + k] = make_double2(X.x - Y.y, X.y + Y.x);
            A[pr * L + L - k] = make_double2(X.x + Y.y, Y.x - X.y);
        }
    }
    __syncthreads();
    bsx_fft<FFT>(A, twl, f, P);
    const double invL = 1.0 / (double)L;
    const int ncell = t.nC - 1;
    double *out = apart + (int64_t)((cd.pair / 6) * 3 + cd.comp) * ns + cd.src_off;
    for (int i = tid; i < nl * ncell; i += NT) {
        const int l = i / ncell, r = i - l * ncell, q = q0 + l;
        if (q >= t.K) continue;
        const double2 w = A[(l >> 1) * L + (r == 0 ? 0 : L - r)];
        const int qs = cd.flipq ? t.K - 1 - q : q;
        const int64_t o = cd.s_is_u ? (int64_t)r * cd.ustride + qs : (int64_t)qs * cd.ustride + r;
        out[o] = ((l & 1) ? w.y : w.x) * invL;
    }
}

// abar(c, s) = sum over the six target faces B (fixed order) of apart[B][c][s].
__global__ void k_bsxT_reduce(int64_t ns, const double *apart, double *abar) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    for (int c = 0; c < 3; c++) {
        double v = 0.0;
        for (int B = 0; B < 6; B++) v += apart[(int64_t)(B * 3 + c) * ns + s];
        abar[(int64_t)c * ns + s] = v;
    }
}

// ------------------------------------------------------------------ host plan
struct SpecTable {
    SpecTab t{};
    SpecFft f{};
    const double2 *tw = nullptr;
    double *G = nullptr;  // perpendicular faces: R[w][q][p]
    SpecFft fq{};         // parallel faces: the convolution along q
    WavePlan wq{};        // the same plan for the one-wave-per-sequence kernel (wq.L = 0: not used)
    WavePlan wf{};        // the transform along C for the one-wave-per-sequence forms
    int form = -2;        // bsx_form of the transforms along C
    const double2 *twq = nullptr;
    double2 *T2 = nullptr;  // parallel faces: FFT(r_w) / Lq, [w][v]
    std::vector<SpecCol> cols;
    SpecCol *cols_d = nullptr;
    int64_t b_off = 0, c_off = 0, o_off = 0;  // column 0 in Bm, Cm (real and imaginary columns) and Ores
};

struct SpecPlan {
    int n[3] = {0, 0, 0};
    double h[3] = {0, 0, 0};
    std::vector<SpecTable> tabs;
    SpecPair *pairs_d = nullptr;
    double *Bm = nullptr, *Cm = nullptr;
    double *Ores = nullptr;
    double flops = 0.0;
    double *BmT = nullptr, *CmT = nullptr, *apart = nullptr;  // transposed sums (allocated on first use)
    int64_t apart_ns = 0;
};

static void free_plan(SpecPlan *P) {
    if (!P) return;
    for (auto &T : P->tabs) {
        if (T.G) (void)hipFree(T.G);
        if (T.T2) (void)hipFree(T.T2);
        if (T.cols_d) (void)hipFree(T.cols_d);
    }
    if (P->pairs_d) (void)hipFree(P->pairs_d);
    if (P->Bm) (void)hipFree(P->Bm);
    if (P->Cm) (void)hipFree(P->Cm);
    if (P->Ores) (void)hipFree(P->Ores);
    if (P->BmT) (void)hipFree(P->BmT);
    if (P->CmT) (void)hipFree(P->CmT);
    if (P->apart) (void)hipFree(P->apart);
    delete P;
}

void free_bspec(Ctx *c) {
    for (int q = 0; q < 2; q++) {
        free_plan((SpecPlan *)c->bsx[q]);
        c->bsx[q] = nullptr;
    }
}

static bool smooth7(int n) {
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Smallest even 7-smooth length >= n (at least 4).
static int smooth_len(int n) {
    int L = n < 4 ? 4 : n;
    while (!(L % 2 == 0 && smooth7(L))) L++;
    return L;
}

static int spec_length(int nC) { return smooth_len(2 * nC - 2); }

static bool same_tab(const SpecTab &a, const SpecTab &b) {
    return a.type == b.type && a.M == b.M && a.K == b.K && a.nC == b.nC && a.L == b.L && a.hT == b.hT &&
           a.hS == b.hS && a.hC == b.hC && a.d == b.d;
}

static int fft_setup(Ctx *c, int L, SpecFft *f, const double2 **tw) {
    FftShape fs;
    VPMX_TRY(fft_shape(c, L, &fs));
    f->L = L;
    f->nstages = fs.nstages;
    for (int s = 0; s < fs.nstages; s++) {
        if (fs.radix[s] > SPEC_MAXR) return fail(c, VPMX_ERR_UNSUP, "spectral boundary sums: FFT radix");
        f->radix[s] = fs.radix[s];
    }
    int np = 1;
    while (np < 8 && 2 * np * L <= DST_EMAX && (2 * np + 1) * L <= DST_LDS_MAX) np *= 2;
    f->npairs = np;
    *tw = fs.tw;
    return VPMX_OK;
}

// The same plan for the one-wave-per-sequence stages.
static WavePlan wave_plan_of(const SpecFft &f) {
    WavePlan P{};
    int ns = 1;
    P.L = f.L;
    P.nstages = f.nstages;
    for (int q = 0; q < f.nstages; q++) {
        P.R[q] = f.radix[q];
        P.Ns[q] = ns;
        P.mg[q] = ns > 1 ? (unsigned)((1ull << 32) / (unsigned long long)ns + 1ull) : 0u;
        ns *= f.radix[q];
    }
    return P;
}

// Transform form of the forward and inverse kernels (bsx_fft): the compile-time plan at 576
// points, one wave per sequence up to BW_LMAX, else the work-group stages.
static int bsx_form(int L) { return L == 576 ? 0 : (L <= BW_LMAX ? -1 : -2); }

static int build_plan(Ctx *c, const Grid &g, SpecPlan **out) {
    SpecPlan *P = new SpecPlan();
    for (int d = 0; d < 3; d++) {
        P->n[d] = g.n[d];
        P->h[d] = g.h[d];
    }
    const int *n = g.n;
    const double *h = g.h;
    // sources of face f start at face_off[f]; face storage is (lower in-plane axis) inner
    int64_t face_off[6], off = 0;
    for (int f = 0; f < 6; f++) {
        const int ax = f / 2, u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
        face_off[f] = off;
        off += (int64_t)(n[u] - 1) * (n[v] - 1);
    }
    struct PairRef {
        int tab, col0, tax, cax, flipp;
    };
    PairRef refs[36];
    for (int B = 0; B < 6; B++)
        for (int A = 0; A < 6; A++) {
            const int At = B / 2, sB = B & 1, As = A / 2, sA = A & 1;
            SpecTab t{};
            int tax, sax, cax, flipp, flipq;
            if (At != As) {
                cax = 3 - At - As;
                tax = As;
                sax = At;
                t.type = 0;
                t.M = n[As];
                t.K = n[At] - 1;
                t.hT = h[As];
                t.hS = h[At];
                t.d = 0.0;
                flipp = sA;
                flipq = sB;
            } else {
                const int u = At == 0 ? 1 : 0, v = At == 2 ? 1 : 2;
                cax = v;
                tax = sax = u;
                t.type = 1;
                t.M = n[u];
                t.K = n[u] - 1;
                t.hT = t.hS = h[u];
                t.d = sA == sB ? 0.0 : (double)(n[At] - 1) * h[At];
                t.Lq = smooth_len(t.M + t.K - 1);
                flipp = flipq = 0;
            }
            t.nC = n[cax];
            t.hC = h[cax];
            t.L = spec_length(t.nC);
            t.nw = t.L / 2 + 1;
            if (t.L > DST_EMAX || 2 * t.L > DST_LDS_MAX || t.Lq > DST_EMAX || 2 * t.Lq > DST_LDS_MAX) {
                free_plan(P);
                return VPMX_ERR_UNSUP;
            }
            int ti = -1;
            for (size_t q = 0; q < P->tabs.size(); q++)
                if (same_tab(P->tabs[q].t, t)) ti = (int)q;
            if (ti < 0) {
                SpecTable T;
                T.t = t;
                P->tabs.push_back(T);
                ti = (int)P->tabs.size() - 1;
            }
            SpecTable &T = P->tabs[ti];
            const int uA = As == 0 ? 1 : 0;
            refs[B * 6 + A] = {ti, (int)T.cols.size(), tax, cax, flipp};
            for (int comp = 0; comp < 3; comp++)
                T.cols.push_back({face_off[A], n[uA] - 1, sax == uA ? 1 : 0, flipq, comp, B * 6 + A});
        }
    // buffers
    int64_t nb = 0, ncm = 0, no = 0;
    for (auto &T : P->tabs) {
        const int64_t N = (int64_t)T.cols.size();
        T.b_off = nb;
        T.c_off = ncm;
        T.o_off = no;
        nb += 2 * N * T.t.nw * T.t.K;
        ncm += 2 * N * T.t.nw * T.t.M;
        no += N * T.t.M * T.t.nC;
    }
    int st = VPMX_OK;
    auto alloc = [&](void **p, size_t bytes) {
        if (st) return;
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) st = fail(c, VPMX_ERR_HIP, std::string("spectral boundary sums: hipMalloc: ") + hipGetErrorString(e));
    };
    alloc((void **)&P->Bm, sizeof(double) * nb);
    alloc((void **)&P->Cm, sizeof(double) * ncm);
    alloc((void **)&P->Ores, sizeof(double) * no);
    alloc((void **)&P->pairs_d, sizeof(SpecPair) * 36);
    std::vector<double *> lines;  // parallel faces: r_w(d) while the tables are built
    for (auto &T : P->tabs) {
        if (T.t.type == 1) {
            double *l = nullptr;
            alloc((void **)&l, sizeof(double) * (size_t)T.t.nw * (T.t.M + T.t.K - 1));
            lines.push_back(l);
            alloc((void **)&T.T2, sizeof(double2) * (size_t)T.t.nw * T.t.Lq);
        } else {
            alloc((void **)&T.G, sizeof(double) * (size_t)T.t.nw * T.t.K * T.t.M);
        }
        alloc((void **)&T.cols_d, sizeof(SpecCol) * T.cols.size());
    }
    auto free_lines = [&]() {
        for (double *l : lines)
            if (l) (void)hipFree(l);
        lines.clear();
    };
    if (st) {
        free_lines();
        free_plan(P);
        return st;
    }
    SpecPair hp[36];
    for (int q = 0; q < 36; q++) {
        const PairRef &R = refs[q];
        const SpecTable &T = P->tabs[R.tab];
        hp[q] = {T.o_off + (int64_t)R.col0 * T.t.M * T.t.nC, T.t.M, T.t.nC, R.tax, R.cax, R.flipp, 0};
    }
    hipError_t e = hipMemcpyAsync(P->pairs_d, hp, sizeof(hp), hipMemcpyHostToDevice, c->stream);
    size_t li = 0;
    for (auto &T : P->tabs) {
        if (e == hipSuccess)
            e = hipMemcpyAsync(T.cols_d, T.cols.data(), sizeof(SpecCol) * T.cols.size(), hipMemcpyHostToDevice,
                               c->stream);
        if (e == hipSuccess) st = fft_setup(c, T.t.L, &T.f, &T.tw);
        if (!st && T.t.type == 1) st = fft_setup(c, T.t.Lq, &T.fq, &T.twq);
        if (st) break;
        if (T.t.type == 1 && T.t.Lq <= QC_LMAX) T.wq = wave_plan_of(T.fq);
        T.wf = wave_plan_of(T.f);
        T.form = bsx_form(T.t.L);
        const size_t lds = sizeof(double2) * (size_t)(T.f.npairs + 1) * T.t.L;
        if (T.t.type == 1) {
            // r_w(d) for d = -(K-1)..M-1: the table kernel with M + K - 1 lines, one q
            SpecTab tl = T.t;
            tl.M = T.t.M + T.t.K - 1;
            tl.K = 1;
            tl.poff = T.t.K - 1;
            double *l = lines[li++];
            dim3 grid((unsigned)((tl.M + 2 * T.f.npairs - 1) / (2 * T.f.npairs)), 1u);
            hipLaunchKernelGGL(k_bsx_table, grid, dim3(DST_THREADS), lds, c->stream, T.f, tl, T.tw, l);
            if (e == hipSuccess) e = hipGetLastError();
            const size_t ldsq = sizeof(double2) * (size_t)(T.fq.npairs + 1) * T.t.Lq;
            hipLaunchKernelGGL(k_bsx_toep_fft, dim3((unsigned)((T.t.nw + T.fq.npairs - 1) / T.fq.npairs)),
                               dim3(DST_THREADS), ldsq, c->stream, T.fq, T.t, T.twq, l, T.T2);
            if (e == hipSuccess) e = hipGetLastError();
            // two FFTs of Lq and the spectral product per column and frequency
            P->flops += (10.0 * T.t.Lq * std::log2((double)T.t.Lq) + 6.0 * T.t.Lq) * (double)T.cols.size() * T.t.nw;
        } else {
            dim3 grid((unsigned)((T.t.M + 2 * T.f.npairs - 1) / (2 * T.f.npairs)), (unsigned)T.t.K);
            hipLaunchKernelGGL(k_bsx_table, grid, dim3(DST_THREADS), lds, c->stream, T.f, T.t, T.tw, T.G);
            if (e == hipSuccess) e = hipGetLastError();
            P->flops += 4.0 * T.t.M * T.t.K * (double)T.cols.size() * T.t.nw;
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // host arrays hp / cols go out of scope
    if (!st && e != hipSuccess) st = fail(c, VPMX_ERR_HIP, std::string("spectral boundary sums: ") + hipGetErrorString(e));
    if (st && e == hipSuccess) (void)hipStreamSynchronize(c->stream);
    free_lines();
    if (st) {
        free_plan(P);
        return st;
    }
    *out = P;
    return VPMX_OK;
}

// The plan of slot for grid g (built, or rebuilt when the grid changed).
static int get_plan(Ctx *c, const Grid &g, int slot, SpecPlan **out) {
    SpecPlan *P = (SpecPlan *)c->bsx[slot];
    if (P && (P->n[0] != g.n[0] || P->n[1] != g.n[1] || P->n[2] != g.n[2] || P->h[0] != g.h[0] ||
              P->h[1] != g.h[1] || P->h[2] != g.h[2])) {
        VPMX_HIP(c, hipStreamSynchronize(c->stream));
        free_plan(P);
        c->bsx[slot] = P = nullptr;
    }
    if (!P) {
        int st = build_plan(c, g, &P);
        if (st) return st;
        c->bsx[slot] = P;
    }
    *out = P;
    return VPMX_OK;
}

int boundary_sums_spectral(Ctx *c, const Grid &g, int64_t ns, const double *a, GView sol, int slot) {
    SpecPlan *P = nullptr;
    VPMX_TRY(get_plan(c, g, slot, &P));
    for (auto &T : P->tabs) {
        const int np = T.form == -2 ? T.f.npairs : BW_WAVES, nl = 2 * np, ncol = (int)T.cols.size();
        const size_t lds = sizeof(double2) * (size_t)(np + 1) * T.t.L;
        const dim3 gf((unsigned)((T.t.K + nl - 1) / nl), (unsigned)ncol), gi((unsigned)((T.t.M + nl - 1) / nl), (unsigned)ncol);
        if (T.form == 0)
            hipLaunchKernelGGL(k_bsx_fwd<0>, gf, dim3(bsx_threads<0>()), lds, c->stream, T.f, T.wf, T.t, T.tw, T.cols_d,
                               a, ns, P->Bm + T.b_off);
        else if (T.form == -1)
            hipLaunchKernelGGL(k_bsx_fwd<-1>, gf, dim3(bsx_threads<-1>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               T.cols_d, a, ns, P->Bm + T.b_off);
        else
            hipLaunchKernelGGL(k_bsx_fwd<-2>, gf, dim3(bsx_threads<-2>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               T.cols_d, a, ns, P->Bm + T.b_off);
        VPMX_HIP(c, hipGetLastError());
        if (T.t.type == 1 && T.wq.L > 0) {
            const size_t ldsq = sizeof(double2) * (size_t)(QC_WAVES + 1) * T.t.Lq;
            const int nseq = ncol * T.t.nw;
            const dim3 gq((unsigned)((nseq + QC_WAVES - 1) / QC_WAVES));
            if (T.t.Lq == 576)
                hipLaunchKernelGGL(k_bsx_qconv_w<0>, gq, dim3(64 * QC_WAVES), ldsq, c->stream, T.wq, T.t, T.twq, T.T2,
                                   P->Bm + T.b_off, P->Cm + T.c_off, ncol, 0);
            else
                hipLaunchKernelGGL(k_bsx_qconv_w<-1>, gq, dim3(64 * QC_WAVES), ldsq, c->stream, T.wq, T.t, T.twq, T.T2,
                                   P->Bm + T.b_off, P->Cm + T.c_off, ncol, 0);
            VPMX_HIP(c, hipGetLastError());
        } else if (T.t.type == 1) {
            const size_t ldsq = sizeof(double2) * (size_t)(T.fq.npairs + 1) * T.t.Lq;
            hipLaunchKernelGGL(k_bsx_qconv, dim3((unsigned)((T.t.nw + T.fq.npairs - 1) / T.fq.npairs), (unsigned)ncol),
                               dim3(DST_THREADS), ldsq, c->stream, T.fq, T.t, T.twq, T.T2, P->Bm + T.b_off,
                               P->Cm + T.c_off, 0);
            VPMX_HIP(c, hipGetLastError());
        } else {
            // per frequency w: C_w (M x 2 ncol) = G_w (M x K) B_w (K x 2 ncol)
            VPMX_TRY(dgemm(c, false, false, T.t.M, 2 * ncol, T.t.K, T.G, T.t.M, (int64_t)T.t.M * T.t.K,
                           P->Bm + T.b_off, T.t.nw * T.t.K, T.t.K, P->Cm + T.c_off, T.t.nw * T.t.M, T.t.M, T.t.nw));
        }
        if (T.form == 0)
            hipLaunchKernelGGL(k_bsx_inv<0>, gi, dim3(bsx_threads<0>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               P->Cm + T.c_off, P->Ores, T.o_off);
        else if (T.form == -1)
            hipLaunchKernelGGL(k_bsx_inv<-1>, gi, dim3(bsx_threads<-1>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               P->Cm + T.c_off, P->Ores, T.o_off);
        else
            hipLaunchKernelGGL(k_bsx_inv<-2>, gi, dim3(bsx_threads<-2>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               P->Cm + T.c_off, P->Ores, T.o_off);
        VPMX_HIP(c, hipGetLastError());
    }
    int maxp = 0;
    for (int B = 0; B < 6; B++) {
        int na, nb;
        gather_face(g.n, B, &na, &nb);
        const int np = ((na + 7) / 8) * ((nb + 7) / 8);
        maxp = np > maxp ? np : maxp;
    }
    hipLaunchKernelGGL(k_bsx_gather, dim3((unsigned)((maxp + GATHER_WAVES - 1) / GATHER_WAVES), 6u),
                       dim3(64 * GATHER_WAVES), 0, c->stream, g, P->pairs_d, P->Ores, sol);
    VPMX_HIP(c, hipGetLastError());
    // FP64 work of the call: the real products and q convolutions, plus ~5 L log2 L flops per
    // complex FFT along C
    double fft = 0.0;
    for (auto &T : P->tabs) {
        const double per = 5.0 * T.t.L * std::log2((double)T.t.L);
        fft += per * (double)T.cols.size() * (T.t.K + T.t.M) / 2.0;
    }
    if (slot == 0) c->k_work[K_BOUNDSUM] = P->flops + fft;
    return VPMX_OK;
}

int boundary_sums_spectral_adj(Ctx *c, const Grid &g, int64_t ns, GView gb, double *abar, int slot) {
    SpecPlan *P = nullptr;
    VPMX_TRY(get_plan(c, g, slot, &P));
    int64_t nbT = 0, ncT = 0;
    for (auto &T : P->tabs) {
        const int64_t N = (int64_t)T.cols.size();
        nbT += 2 * N * T.t.nw * T.t.M;
        ncT += 2 * N * T.t.nw * T.t.K;
    }
    if (!P->BmT) {
        VPMX_HIP(c, hipMalloc((void **)&P->BmT, sizeof(double) * nbT));
        VPMX_HIP(c, hipMalloc((void **)&P->CmT, sizeof(double) * ncT));
    }
    if (P->apart_ns < ns) {
        if (P->apart) {
            VPMX_HIP(c, hipStreamSynchronize(c->stream));
            (void)hipFree(P->apart);
            P->apart = nullptr;
        }
        VPMX_HIP(c, hipMalloc((void **)&P->apart, sizeof(double) * 18 * (size_t)ns));
        P->apart_ns = ns;
    }
    int64_t bo = 0, co = 0;
    for (auto &T : P->tabs) {
        const int np = T.form == -2 ? T.f.npairs : BW_WAVES, nl = 2 * np, ncol = (int)T.cols.size();
        const size_t lds = sizeof(double2) * (size_t)(np + 1) * T.t.L;
        double *Bm = P->BmT + bo, *Cm = P->CmT + co;
        bo += 2 * (int64_t)ncol * T.t.nw * T.t.M;
        co += 2 * (int64_t)ncol * T.t.nw * T.t.K;
        const dim3 gf((unsigned)((T.t.M + nl - 1) / nl), (unsigned)ncol), gi((unsigned)((T.t.K + nl - 1) / nl), (unsigned)ncol);
        if (T.form == 0)
            hipLaunchKernelGGL(k_bsxT_fwd<0>, gf, dim3(bsx_threads<0>()), lds, c->stream, T.f, T.wf, T.t, T.tw, T.cols_d,
                               P->pairs_d, g, gb, Bm);
        else if (T.form == -1)
            hipLaunchKernelGGL(k_bsxT_fwd<-1>, gf, dim3(bsx_threads<-1>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               T.cols_d, P->pairs_d, g, gb, Bm);
        else
            hipLaunchKernelGGL(k_bsxT_fwd<-2>, gf, dim3(bsx_threads<-2>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               T.cols_d, P->pairs_d, g, gb, Bm);
        VPMX_HIP(c, hipGetLastError());
        if (T.t.type == 1) {
            // the q correlation: the forward's convolution with M and K exchanged and the
            // conjugate spectrum
            SpecTab tT = T.t;
            tT.M = T.t.K;
            tT.K = T.t.M;
            if (T.wq.L > 0) {
                const size_t ldsq = sizeof(double2) * (size_t)(QC_WAVES + 1) * T.t.Lq;
                const int nseq = ncol * T.t.nw;
                const dim3 gq((unsigned)((nseq + QC_WAVES - 1) / QC_WAVES));
                if (T.t.Lq == 576)
                    hipLaunchKernelGGL(k_bsx_qconv_w<0>, gq, dim3(64 * QC_WAVES), ldsq, c->stream, T.wq, tT, T.twq, T.T2,
                                       Bm, Cm, ncol, 1);
                else
                    hipLaunchKernelGGL(k_bsx_qconv_w<-1>, gq, dim3(64 * QC_WAVES), ldsq, c->stream, T.wq, tT, T.twq,
                                       T.T2, Bm, Cm, ncol, 1);
            } else {
                const size_t ldsq = sizeof(double2) * (size_t)(T.fq.npairs + 1) * T.t.Lq;
                hipLaunchKernelGGL(k_bsx_qconv, dim3((unsigned)((T.t.nw + T.fq.npairs - 1) / T.fq.npairs), (unsigned)ncol),
                                   dim3(DST_THREADS), ldsq, c->stream, T.fq, tT, T.twq, T.T2, Bm, Cm, 1);
            }
            VPMX_HIP(c, hipGetLastError());
        } else {
            // per frequency w: C_w (K x 2 ncol) = G_w^T (K x M) B_w (M x 2 ncol)
            VPMX_TRY(dgemm(c, true, false, T.t.K, 2 * ncol, T.t.M, T.G, T.t.M, (int64_t)T.t.M * T.t.K, Bm,
                           T.t.nw * T.t.M, T.t.M, Cm, T.t.nw * T.t.K, T.t.K, T.t.nw));
        }
        if (T.form == 0)
            hipLaunchKernelGGL(k_bsxT_inv<0>, gi, dim3(bsx_threads<0>()), lds, c->stream, T.f, T.wf, T.t, T.tw, T.cols_d,
                               Cm, ns, P->apart);
        else if (T.form == -1)
            hipLaunchKernelGGL(k_bsxT_inv<-1>, gi, dim3(bsx_threads<-1>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               T.cols_d, Cm, ns, P->apart);
        else
            hipLaunchKernelGGL(k_bsxT_inv<-2>, gi, dim3(bsx_threads<-2>()), lds, c->stream, T.f, T.wf, T.t, T.tw,
                               T.cols_d, Cm, ns, P->apart);
        VPMX_HIP(c, hipGetLastError());
    }
    hipLaunchKernelGGL(k_bsxT_reduce, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, ns, P->apart, abar);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

}  // namespace vpmx
