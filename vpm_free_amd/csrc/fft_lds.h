// In-LDS complex FFTs shared by the sine-transform Poisson passes and the spectral
// boundary sums: mixed-radix Stockham autosort over `npairs` sequences of length L held
// back to back in LDS, one work-group of DST_THREADS threads, twiddles exp(-2 pi i t / L)
// read from an LDS table.  Args is any struct with fields L, npairs, nstages, radix[].
#pragma once
#include <hip/hip_runtime.h>

namespace vpmx {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Scalar type and constructor of a complex vector type.
template <class C> struct Cx;
template <> struct Cx<double2> {
    using R = double;
    static __device__ __forceinline__ double2 mk(double x, double y) { return make_double2(x, y); }
};
template <> struct Cx<float2> {
    using R = float;
    static __device__ __forceinline__ float2 mk(float x, float y) { return make_float2(x, y); }
};

// ------------------------------------------------------------------ register-staged stages
// Every stage loads each butterfly's R inputs into registers, synchronises, and writes the
// outputs to their Stockham positions.  npairs * L must not exceed DST_EMAX.
constexpr int DST_EMAX = 2304;  // complex line elements held in LDS per work-group (36 KB)
constexpr int DST_LDS_MAX = 4096;  // complex elements of LDS per work-group, lines + twiddles (64 KB)
constexpr int DST_THREADS = 512;

// In-register butterflies of the specialised radices (2, 3, 4, 9).
template <int R, class C>
__device__ __forceinline__ void dft_reg(C (&v)[R]) {
    using T = typename Cx<C>::R;
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T c3 = (T)-0.5, s3 = (T)0.86602540378443864676;
        C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        C base = Cx<C>::mk(v[0].x + c3 * t1.x, v[0].y + c3 * t1.y);
        C rot = Cx<C>::mk(s3 * t2.y, -s3 * t2.x);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(base, rot);
        v[2] = csub(base, rot);
    } else if constexpr (R == 4) {
        C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]), s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        C md = Cx<C>::mk(d13.y, -d13.x);
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, md);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, md);
    } else if constexpr (R == 8) {
        // 8 = 2 x 4: the 4-point DFTs of the even and odd inputs, X_k = E_k + W^k O_k and
        // X_{k+4} = E_k - W^k O_k, W = exp(-2 pi i / 8)
        C e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        dft_reg<4, C>(e);
        dft_reg<4, C>(o);
        const T r2 = (T)0.70710678118654752440;
        const C t1 = Cx<C>::mk(r2 * (o[1].x + o[1].y), r2 * (o[1].y - o[1].x));  // W o1
        const C t2 = Cx<C>::mk(o[2].y, -o[2].x);                                  // W^2 o2 = -i o2
        const C t3 = Cx<C>::mk(r2 * (o[3].y - o[3].x), -r2 * (o[3].x + o[3].y));  // W^3 o3
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], t1);
        v[5] = csub(e[1], t1);
        v[2] = cadd(e[2], t2);
        v[6] = csub(e[2], t2);
        v[3] = cadd(e[3], t3);
        v[7] = csub(e[3], t3);
    } else if constexpr (R == 16) {
        // 16 = 4 x 4 with n = 4a + b, k = k1 + 4 k2: a 4-point DFT over a for each b, the
        // twiddles exp(-2 pi i b k1 / 16), then a 4-point DFT over b for each k1
        C y[4][4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
#pragma unroll
            for (int a = 0; a < 4; a++) y[b][a] = v[4 * a + b];
            dft_reg<4, C>(y[b]);
        }
        const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, r2 = (T)0.70710678118654752440;
        y[1][1] = cmul(y[1][1], Cx<C>::mk(c1, -s1));                                  // W^1
        y[1][2] = Cx<C>::mk(r2 * (y[1][2].x + y[1][2].y), r2 * (y[1][2].y - y[1][2].x));  // W^2
        y[1][3] = cmul(y[1][3], Cx<C>::mk(s1, -c1));                                  // W^3
        y[2][1] = Cx<C>::mk(r2 * (y[2][1].x + y[2][1].y), r2 * (y[2][1].y - y[2][1].x));  // W^2
        y[2][2] = Cx<C>::mk(y[2][2].y, -y[2][2].x);                                   // W^4 = -i
        y[2][3] = Cx<C>::mk(r2 * (y[2][3].y - y[2][3].x), -r2 * (y[2][3].x + y[2][3].y));  // W^6
        y[3][1] = cmul(y[3][1], Cx<C>::mk(s1, -c1));                                  // W^3
        y[3][2] = Cx<C>::mk(r2 * (y[3][2].y - y[3][2].x), -r2 * (y[3][2].x + y[3][2].y));  // W^6
        y[3][3] = cmul(y[3][3], Cx<C>::mk(-c1, s1));                                  // W^9
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            C z[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            dft_reg<4, C>(z);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = z[k2];
        }
    } else if constexpr (R == 9) {
        // 9 = 3 x 3 with n = 3a + b, k = k1 + 3 k2: a 3-point DFT over a for each b, the
        // twiddles exp(-2 pi i b k1 / 9), then a 3-point DFT over b for each k1
        C y0[3] = {v[0], v[3], v[6]}, y1[3] = {v[1], v[4], v[7]}, y2[3] = {v[2], v[5], v[8]};
        dft_reg<3, C>(y0);
        dft_reg<3, C>(y1);
        dft_reg<3, C>(y2);
        const C w1 = Cx<C>::mk((T)0.7660444431189780352024, (T)-0.6427876096865393263226);
        const C w2 = Cx<C>::mk((T)0.1736481776669303488517, (T)-0.9848077530122080593667);
        const C w4 = Cx<C>::mk((T)-0.9396926207859083840541, (T)-0.3420201433256687330441);
        y1[1] = cmul(y1[1], w1);
        y1[2] = cmul(y1[2], w2);
        y2[1] = cmul(y2[1], w2);
        y2[2] = cmul(y2[2], w4);
#pragma unroll
        for (int k1 = 0; k1 < 3; k1++) {
            C z[3] = {y0[k1], y1[k1], y2[k1]};
            dft_reg<3, C>(z);
            v[k1] = z[0];
            v[k1 + 3] = z[1];
            v[k1 + 6] = z[2];
        }
    }
}

// Odd prime R-point DFT of v written to buf[d + k*Ns], k < R, with the conjugate symmetry of
// the roots: for r < R/2, p_r = v_r + v_{R-r} and m_r = v_r - v_{R-r}; with
// exp(-2 pi i r k / R) = c - i s,
//   X_k   = v_0 + sum_r (c p_r.x + s m_r.y,  c p_r.y - s m_r.x)
//   X_R-k = v_0 + sum_r (c p_r.x - s m_r.y,  c p_r.y + s m_r.x)
// so two outputs cost 4 FMA per r.  c and s come from the LDS twiddle table.
template <int R, class C, class Buf>
__device__ __forceinline__ void dft_odd_store(C (&v)[R], Buf buf, int d, int Ns, const C *__restrict__ tw, int L) {
    using T = typename Cx<C>::R;
    constexpr int H = (R - 1) / 2;
    C p[H], mm[H];
    C x0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; r++) {
        p[r - 1] = cadd(v[r], v[R - r]);
        mm[r - 1] = csub(v[r], v[R - r]);
        x0 = cadd(x0, p[r - 1]);
    }
    buf[d] = x0;
    const int Lr = L / R;
    // k is not unrolled: the H twiddles of one k are loaded per iteration, not all H*H at once
#pragma unroll 1
    for (int k = 1; k <= H; k++) {
        T X = v[0].x, Y = v[0].y, U = 0, V = 0;
        int q = 0;  // (r * k) mod R
#pragma unroll
        for (int r = 1; r <= H; r++) {
            q += k;
            if (q >= R) q -= R;
            const C w = tw[q * Lr];  // (c, -s)
            const T c = w.x, s = -w.y;
            X += c * p[r - 1].x;
            Y += c * p[r - 1].y;
            U += s * mm[r - 1].y;
            V += s * mm[r - 1].x;
        }
        buf[d + k * Ns] = Cx<C>::mk(X + U, Y - V);
        buf[d + (R - k) * Ns] = Cx<C>::mk(X - U, Y + V);
    }
}

// cos and sin of 2 pi q / 17, q = 1..8: the roots of the radix-17 stage of the compile-time
// plans as constants (the runtime stages read them from the LDS twiddle table).
__host__ __device__ constexpr double cos17(int q) {
    switch (q) {
        case 1: return 0.932472229404355804573;
        case 2: return 0.739008917220659115925;
        case 3: return 0.445738355776538267396;
        case 4: return 0.0922683594633019952397;
        case 5: return -0.273662990072082863539;
        case 6: return -0.602634636379256389179;
        case 7: return -0.850217135729614152134;
        default: return -0.982973099683901778282;
    }
}
__host__ __device__ constexpr double sin17(int q) {
    switch (q) {
        case 1: return 0.361241666187152948745;
        case 2: return 0.673695643646557211713;
        case 3: return 0.895163291355062322067;
        case 4: return 0.995734176295034521871;
        case 5: return 0.961825643172819070409;
        case 6: return 0.798017227280239503333;
        case 7: return 0.526432162877355800245;
        default: return 0.183749517816570331574;
    }
}

// dft_odd_store for R = 17 with the roots as constants and both loops unrolled: the same sums
// in the same order, without the H * H twiddle reads from LDS.
template <int R, class C, class Buf>
__device__ __forceinline__ void dft_odd_store_c(C (&v)[R], Buf buf, int d, int Ns) {
    static_assert(R == 17, "constant roots are tabulated for radix 17");
    using T = typename Cx<C>::R;
    constexpr int H = (R - 1) / 2;
    C p[H], mm[H];
    C x0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; r++) {
        p[r - 1] = cadd(v[r], v[R - r]);
        mm[r - 1] = csub(v[r], v[R - r]);
        x0 = cadd(x0, p[r - 1]);
    }
    buf[d] = x0;
#pragma unroll
    for (int k = 1; k <= H; k++) {
        T X = v[0].x, Y = v[0].y, U = 0, V = 0;
#pragma unroll
        for (int r = 1; r <= H; r++) {
            const int q = (r * k) % R;  // exp(-2 pi i q / R) = c - i s
            const T c = (T)cos17(q <= H ? q : R - q);
            const T s = q <= H ? (T)sin17(q) : -(T)sin17(R - q);
            X += c * p[r - 1].x;
            Y += c * p[r - 1].y;
            U += s * mm[r - 1].y;
            V += s * mm[r - 1].x;
        }
        buf[d + k * Ns] = Cx<C>::mk(X + U, Y - V);
        buf[d + (R - k) * Ns] = Cx<C>::mk(X - U, Y + V);
    }
}

// One Stockham stage in place: inputs to registers, barrier, outputs to LDS, barrier.
// Radices above 4 form their outputs after the barrier (v holds the twiddled inputs).
template <int R, class Args>
__device__ void stage_reg(double2 *buf, const double2 *__restrict__ tw, const Args &a, int Ns) {
    constexpr int BPT = (DST_EMAX / R + DST_THREADS - 1) / DST_THREADS;
    const int L = a.L, nbf = L / R, total = a.npairs * nbf, twstep = L / (Ns * R);
    double2 v[BPT][R];
    int dst[BPT];
#pragma unroll
    for (int bi = 0; bi < BPT; bi++) {
        const int t = threadIdx.x + bi * DST_THREADS;
        dst[bi] = -1;
        if (t < total) {
            const int pair = t / nbf, j = t - pair * nbf;
            const int k0 = j % Ns;
            const double2 *src = buf + pair * L;
#pragma unroll
            for (int r = 0; r < R; r++) v[bi][r] = src[j + r * nbf];
            if (k0) {
#pragma unroll
                for (int r = 1; r < R; r++) v[bi][r] = cmul(v[bi][r], tw[r * k0 * twstep]);
            }
            if constexpr (R <= 4) dft_reg<R>(v[bi]);
            dst[bi] = pair * L + (j - k0) * R + k0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < BPT; bi++)
        if (dst[bi] >= 0) {
            if constexpr (R <= 4) {
#pragma unroll
                for (int k = 0; k < R; k++) buf[dst[bi] + k * Ns] = v[bi][k];
            } else {
                dft_odd_store<R>(v[bi], buf, dst[bi], Ns, tw, L);
            }
        }
    __syncthreads();
}

// MAXR bounds the radices compiled into a kernel, so the registers of the common 2/3/4
// stages are not sized for a large prime stage.
template <int MAXR, class Args>
__device__ void fft_reg(double2 *buf, const double2 *tw, const Args &a) {
    int Ns = 1;
    for (int s = 0; s < a.nstages; s++) {
        const int R = a.radix[s];
        if (R == 4) stage_reg<4, Args>(buf, tw, a, Ns);
        else if (R == 2) stage_reg<2, Args>(buf, tw, a, Ns);
        else if (R == 3) stage_reg<3, Args>(buf, tw, a, Ns);
        else if constexpr (MAXR >= 7) {
            if (R == 5) stage_reg<5, Args>(buf, tw, a, Ns);
            else if (R == 7) stage_reg<7, Args>(buf, tw, a, Ns);
            else if constexpr (MAXR >= 17) {
                if (R == 11) stage_reg<11, Args>(buf, tw, a, Ns);
                else if (R == 13) stage_reg<13, Args>(buf, tw, a, Ns);
                else if (R == 17) stage_reg<17, Args>(buf, tw, a, Ns);
                else if constexpr (MAXR >= 37) {
                    if (R == 19) stage_reg<19, Args>(buf, tw, a, Ns);
                    else if (R == 23) stage_reg<23, Args>(buf, tw, a, Ns);
                    else if (R == 29) stage_reg<29, Args>(buf, tw, a, Ns);
                    else if (R == 31) stage_reg<31, Args>(buf, tw, a, Ns);
                    else stage_reg<37, Args>(buf, tw, a, Ns);
                }
            }
        }
        Ns *= R;
    }
}

// ------------------------------------------------------------------ one wave per sequence
// FFTs of length L <= WAVE_LMAX where each sequence belongs to one 64-lane wave: the
// stages need no work-group barrier (a wave's LDS operations execute in order; the wave
// barrier only keeps the compiler from moving them).  Per stage: R radix, Ns span, mg a
// multiplier with j / Ns == umulhi(j, mg) for j < 2^16 and Ns > 1 (mg = floor(2^32 / Ns) + 1;
// the first stage, Ns = 1, divides by nothing).
constexpr int WAVE_LMAX = 288;      // double sequences (16 per work-group of 8 waves)
constexpr int WAVE_LMAX_F = 576;    // float sequences (same LDS bytes)

struct WavePlan {
    int L, nstages;
    int R[12], Ns[12];
    unsigned mg[12];
};

// LDS view of one wave's sequence with one unused slot after every 2^S elements (S = 0: no
// slots).  With S = 4 for float2 and 3 for double2 (128 bytes between slots) the first
// Stockham stages, which write at strides of R and R * Ns elements, spread a 64-lane write
// over all LDS banks instead of putting 4 lanes on each.  The slots cost address arithmetic
// and registers on every access, so a kernel takes them only where that pays (measured).
template <int S>
__host__ __device__ constexpr int pad_pos(int e) { return S ? e + (e >> S) : e; }
template <int S>
__host__ __device__ constexpr int pad_len(int n) { return S ? pad_pos<S>(n) + 1 : n; }
template <class C, int S>
struct PadBuf {
    C *p;
    __device__ __forceinline__ C &operator[](int e) const { return p[pad_pos<S>(e)]; }
};

template <int R, int BPL, class C, class Buf>
__device__ __forceinline__ void wave_stage(Buf buf, const C *__restrict__ tw, int L, int Ns, unsigned mg, int lane) {
    const int nbf = L / R, twstep = L / (Ns * R);
    C v[BPL][R];
    int dst[BPL];
#pragma unroll
    for (int i = 0; i < BPL; i++) {
        const int j = lane + 64 * i;
        dst[i] = -1;
        if (j < nbf) {
            const int q = Ns == 1 ? j : (int)__umulhi((unsigned)j, mg), k0 = j - q * Ns;
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = buf[j + r * nbf];
            if (k0) {
                if constexpr (R == 4) {  // one table read, the higher powers in registers
                    const C w1 = tw[k0 * twstep];
                    const C w2 = cmul(w1, w1);
                    v[i][1] = cmul(v[i][1], w1);
                    v[i][2] = cmul(v[i][2], w2);
                    v[i][3] = cmul(v[i][3], cmul(w2, w1));
                } else {
#pragma unroll
                    for (int r = 1; r < R; r++) v[i][r] = cmul(v[i][r], tw[r * k0 * twstep]);
                }
            }
            if constexpr (R <= 4) dft_reg<R>(v[i]);
            dst[i] = q * Ns * R + k0;
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < BPL; i++)
        if (dst[i] >= 0) {
            if constexpr (R <= 4) {
#pragma unroll
                for (int k = 0; k < R; k++) buf[dst[i] + k * Ns] = v[i][k];
            } else {
                dft_odd_store<R>(v[i], buf, dst[i], Ns, tw, L);
            }
        }
    __builtin_amdgcn_wave_barrier();
}

template <int R, int LMAX>
constexpr int wave_bpl() { return (LMAX / R + 63) / 64; }

template <int MAXR, class C, class Buf, int LMAX = (sizeof(C) == 16 ? WAVE_LMAX : WAVE_LMAX_F)>
__device__ void wave_fft(Buf buf, const C *tw, const WavePlan &P, int lane) {
    for (int s = 0; s < P.nstages; s++) {
        const int R = P.R[s], Ns = P.Ns[s];
        const unsigned mg = P.mg[s];
        if (R == 4) wave_stage<4, wave_bpl<4, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
        else if (R == 2) wave_stage<2, wave_bpl<2, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
        else if (R == 3) wave_stage<3, wave_bpl<3, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
        else if constexpr (MAXR >= 7) {
            if (R == 5) wave_stage<5, wave_bpl<5, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
            else if (R == 7) wave_stage<7, wave_bpl<7, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
            else if constexpr (MAXR >= 17) {
                if (R == 11) wave_stage<11, wave_bpl<11, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                else if (R == 13) wave_stage<13, wave_bpl<13, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                else if (R == 17) wave_stage<17, wave_bpl<17, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                else if constexpr (MAXR >= 37) {
                    if (R == 19) wave_stage<19, wave_bpl<19, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                    else if (R == 23) wave_stage<23, wave_bpl<23, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                    else if (R == 29) wave_stage<29, wave_bpl<29, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                    else if (R == 31) wave_stage<31, wave_bpl<31, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                    else wave_stage<37, wave_bpl<37, LMAX>(), C, Buf>(buf, tw, P.L, Ns, mg, lane);
                }
            }
        }
    }
}


// ------------------------------------------------------------------ compile-time plans
// The same Stockham stages with the length and radices known at compile time: spans,
// strides, twiddle steps and divisions by Ns become constants, which removes most of the
// integer work of the runtime-plan stages.  A plan may also merge two radix-3 stages into
// one radix-9 stage (half the LDS round trips of that pair; a 576-point sequence has
// exactly 64 radix-9 butterflies, one per lane), so its rounding differs from the runtime
// plan's.
// CR: a radix-17 stage takes its roots as constants (dft_odd_store_c) instead of from the LDS
// table; it saves the table reads but needs more registers, which costs some kernels their
// occupancy, so each kernel chooses.
template <int R, int L, int Ns, class C, class Buf, bool CR = false>
__device__ __forceinline__ void wave_stage_s(Buf buf, const C *__restrict__ tw, int lane) {
    constexpr int nbf = L / R, twstep = L / (Ns * R), BPL = (nbf + 63) / 64;
    C v[BPL][R];
    int dst[BPL];
#pragma unroll
    for (int i = 0; i < BPL; i++) {
        const int j = lane + 64 * i;
        dst[i] = -1;
        if ((i + 1) * 64 <= nbf || j < nbf) {
            const int q = j / Ns, k0 = j - q * Ns;
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = buf[j + r * nbf];
            if constexpr (Ns > 1) {
                // no branch on k0 == 0: tw[0] = (1, 0) leaves the value unchanged
                if constexpr (R == 4) {
                    // one table read per butterfly, the higher powers by multiplication in
                    // registers: the passes are bound by LDS instructions, not by arithmetic
                    const C w1 = tw[k0 * twstep];
                    const C w2 = cmul(w1, w1);
                    v[i][1] = cmul(v[i][1], w1);
                    v[i][2] = cmul(v[i][2], w2);
                    v[i][3] = cmul(v[i][3], cmul(w2, w1));
                } else if constexpr (R == 8) {
                    // two table reads (w, w^4), the other powers at most two products deep
                    const C w1 = tw[k0 * twstep], w4 = tw[4 * k0 * twstep];
                    const C w2 = cmul(w1, w1), w3 = cmul(w1, w2);
                    v[i][1] = cmul(v[i][1], w1);
                    v[i][2] = cmul(v[i][2], w2);
                    v[i][3] = cmul(v[i][3], w3);
                    v[i][4] = cmul(v[i][4], w4);
                    v[i][5] = cmul(v[i][5], cmul(w1, w4));
                    v[i][6] = cmul(v[i][6], cmul(w2, w4));
                    v[i][7] = cmul(v[i][7], cmul(w3, w4));
                } else if constexpr (R == 9) {
                    // two table reads (w, w^3), the other powers at most two products deep
                    const C w1 = tw[k0 * twstep], w3 = tw[3 * k0 * twstep];
                    const C w2 = cmul(w1, w1), w4 = cmul(w1, w3), w6 = cmul(w3, w3);
                    v[i][1] = cmul(v[i][1], w1);
                    v[i][2] = cmul(v[i][2], w2);
                    v[i][3] = cmul(v[i][3], w3);
                    v[i][4] = cmul(v[i][4], w4);
                    v[i][5] = cmul(v[i][5], cmul(w2, w3));
                    v[i][6] = cmul(v[i][6], w6);
                    v[i][7] = cmul(v[i][7], cmul(w3, w4));
                    v[i][8] = cmul(v[i][8], cmul(w4, w4));
                } else {
#pragma unroll
                    for (int r = 1; r < R; r++) v[i][r] = cmul(v[i][r], tw[r * k0 * twstep]);
                }
            }
            if constexpr (R <= 4 || R == 8 || R == 9 || R == 16) dft_reg<R>(v[i]);
            dst[i] = q * Ns * R + k0;
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < BPL; i++)
        if (dst[i] >= 0) {
            if constexpr (R <= 4 || R == 8 || R == 9 || R == 16) {
#pragma unroll
                for (int k = 0; k < R; k++) buf[dst[i] + k * Ns] = v[i][k];
            } else if constexpr (R == 17 && CR) {
                dft_odd_store_c<R>(v[i], buf, dst[i], Ns);
            } else {
                dft_odd_store<R>(v[i], buf, dst[i], Ns, tw, L);
            }
        }
    __builtin_amdgcn_wave_barrier();
}

// A radix-17 stage with at most 32 butterflies (the 272- and 544-point plans) leaves half or
// three quarters of a wave's lanes idle.  Here SPL = 64 / nbf lanes share a butterfly: each
// reads the 17 inputs (the same LDS addresses: broadcast reads), forms the conjugate-pair sums
// p_r, m_r, and computes H / SPL of the H output pairs (lane 0 of the group also X_0), with the
// same sums in the same order as dft_odd_store and its roots from the LDS table -- so the
// stage's arithmetic and root reads are divided by SPL and its stores spread over the group.
template <int L, int Ns, class C, class Buf>
__device__ __forceinline__ void wave_stage_s17x(Buf buf, const C *__restrict__ tw, int lane) {
    constexpr int R = 17, H = 8, nbf = L / R, twstep = L / (Ns * R);
    static_assert(nbf * R == L && nbf <= 32, "split radix-17 stage: at most 32 butterflies");
    constexpr int SPL = nbf <= 16 ? 4 : 2, KP = H / SPL;
    using T = typename Cx<C>::R;
    const int j = lane / SPL, sub = lane - j * SPL;
    const bool act = j < nbf;
    const int jj = act ? j : 0;  // idle lanes read butterfly 0 and store nothing
    const int q = jj / Ns, k0 = jj - q * Ns;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = buf[jj + r * nbf];
    if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = cmul(v[r], tw[r * k0 * twstep]);
    }
    C p[H], mm[H];
    C x0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; r++) {
        p[r - 1] = cadd(v[r], v[R - r]);
        mm[r - 1] = csub(v[r], v[R - r]);
        x0 = cadd(x0, p[r - 1]);
    }
    const int d = q * Ns * R + k0;
    __builtin_amdgcn_wave_barrier();
    if (act) {
        if (sub == 0) buf[d] = x0;
#pragma unroll
        for (int kk = 0; kk < KP; kk++) {
            const int k = sub * KP + kk + 1;
            T X = v[0].x, Y = v[0].y, U = 0, V = 0;
            int qq = 0;  // (r * k) mod R
#pragma unroll
            for (int r = 1; r <= H; r++) {
                qq += k;
                if (qq >= R) qq -= R;
                const C w = tw[qq * nbf];  // (c, -s)
                const T c = w.x, sn = -w.y;
                X += c * p[r - 1].x;
                Y += c * p[r - 1].y;
                U += sn * mm[r - 1].y;
                V += sn * mm[r - 1].x;
            }
            buf[d + k * Ns] = Cx<C>::mk(X + U, Y - V);
            buf[d + (R - k) * Ns] = Cx<C>::mk(X - U, Y + V);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// SPLIT: radix-17 stages as wave_stage_s17x.
template <class C, class Buf, bool CR, bool SPLIT, int L, int Ns, int R, int... Rest>
__device__ __forceinline__ void wave_fft_sx(Buf buf, const C *tw, int lane) {
    if constexpr (R == 17 && SPLIT) wave_stage_s17x<L, Ns, C, Buf>(buf, tw, lane);
    else wave_stage_s<R, L, Ns, C, Buf, CR>(buf, tw, lane);
    if constexpr (sizeof...(Rest) > 0) wave_fft_sx<C, Buf, CR, SPLIT, L, Ns * R, Rest...>(buf, tw, lane);
}
template <class C, class Buf, bool CR, int L, int Ns, int R, int... Rest>
__device__ __forceinline__ void wave_fft_s(Buf buf, const C *tw, int lane) {
    wave_fft_sx<C, Buf, CR, false, L, Ns, R, Rest...>(buf, tw, lane);
}

// Lengths with a compile-time plan (radices in the planner's order: 4s, 2s, odd primes).
// FIX = -1 selects the runtime plan.
template <int FIX> struct FixedFft;
template <> struct FixedFft<0> {  // 576 = 4.4.4.9 (FP32, 289-node block)
    static constexpr int L = 576;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 576, 1, 4, 4, 4, 9>(b, tw, lane);
    }
};
template <> struct FixedFft<6> {  // 576 = 8.8.9 (FP32, 289-node block): three stages instead of four
    static constexpr int L = 576;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 576, 1, 8, 8, 9>(b, tw, lane);
    }
};
template <> struct FixedFft<13> {  // 576 = 16.4.9 (FP32, 289-node block): a radix-16 first stage (36 butterflies)
    static constexpr int L = 576;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 576, 1, 16, 4, 9>(b, tw, lane);
    }
};
template <> struct FixedFft<7> {  // 288 = 8.4.9 (FP64, 289-node block): three stages instead of four
    static constexpr int L = 288;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 288, 1, 8, 4, 9>(b, tw, lane);
    }
};
template <> struct FixedFft<8> {  // 544 = 17.8.4 (FP32, 273-node inner box): three stages instead of four
    static constexpr int L = 544;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 544, 1, 17, 8, 4>(b, tw, lane);
    }
};
template <> struct FixedFft<9> {  // 272 = 4.4.17 with the radix-17 outputs split over 4 lanes
    static constexpr int L = 272;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_sx<C, Buf, false, true, 272, 1, 4, 4, 17>(b, tw, lane);
    }
};
template <> struct FixedFft<10> {  // 544 = 17.8.4 with the radix-17 outputs split over 2 lanes
    static constexpr int L = 544;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_sx<C, Buf, false, true, 544, 1, 17, 8, 4>(b, tw, lane);
    }
};
template <> struct FixedFft<11> {  // 160 = 8.4.5 (FP64, 161-node block: the 128^3 class)
    static constexpr int L = 160;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 160, 1, 8, 4, 5>(b, tw, lane);
    }
};
template <> struct FixedFft<12> {  // 144 = 4.4.9 (FP64, 145-node inner box of the 128^3 class)
    static constexpr int L = 144;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 144, 1, 4, 4, 9>(b, tw, lane);
    }
};
template <> struct FixedFft<14> {  // 272 = 16.17, the radix-17 outputs split over 4 lanes: two stages
    static constexpr int L = 272;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_sx<C, Buf, false, true, 272, 1, 16, 17>(b, tw, lane);
    }
};
template <> struct FixedFft<2> {  // 288 = 4.4.2.9 (FP64, 289-node block)
    static constexpr int L = 288;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 288, 1, 4, 4, 2, 9>(b, tw, lane);
    }
};
template <> struct FixedFft<5> {  // 544 = 17.4.4.2 (FP32, 273-node inner box; the prime stage first needs
                                  // the fewest registers of the orders tried)
    static constexpr int L = 544;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 544, 1, 17, 4, 4, 2>(b, tw, lane);
    }
};
template <> struct FixedFft<3> {  // 272 = 4.4.17 (FP64, 273-node inner box)
    static constexpr int L = 272;
    template <class C, class Buf, bool CR = false>
    static __device__ __forceinline__ void run(Buf b, const C *tw, int lane) {
        wave_fft_s<C, Buf, CR, 272, 1, 4, 4, 17>(b, tw, lane);
    }
};

}  // namespace vpmx
