/*
 * vpm_oracle.c -- CPU restatement of the reference VPM path (papis84/vpm_free),
 * FP64 throughout, single yaps block (BLOCKS = NBI*NBJ*NBK = 1).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the HIP library.
 * The product (vpm_free_amd/) never links, loads or calls this file.
 *
 * Every routine follows the reference routine cited beside it, operation for
 * operation where the floating-point order matters.  The MKL Poisson Library
 * (d_*_Helmholtz_3D, q=0, 'DDDDDD'; Intel MKL, version not pinned by the reference,
 * 2021.4.0 in this image) is restated by its published algorithm: sine transforms
 * along x and y, then a tridiagonal (Thomas) solve along z, of the standard
 * 7-point system -(u_{i-1}-2u_i+u_{i+1})/hx^2 - ... = f with Dirichlet values.
 *
 * Pinning (tests/test_oracle_*.py):
 *   orc_projection_fun   closed forms, partition of unity, moments; reference projlib (oracle/_ref)
 *   orc_project          reference projlib project_particles_3D / project_vol3d (oracle/_ref)
 *   orc_definepm / orc_define_sizes   hand-derived sizes for the shipped pm.inp and ring boxes,
 *                        divisibility/containment invariants
 *   orc_dst1, orc_solve_window   scipy.fft.dst type 1, sine eigenmodes, residual of the 7-point system
 *   orc_normalderiv      exact on quartic fields (5-point one-sided stencil), face/centroid layout
 *   orc_boundsum         closed-form single-source values; free-space potential of a Gaussian blob
 *   orc_pmesh_inf / orc_yaps3d   free-space potential of a Gaussian blob (discretisation-level);
 *                        exact invariance when the coarse correction vanishes
 *   orc_velocity         exact curl/stretching of polynomial fields
 *   orc_m2p              exact interpolation of quadratics (M4'), brute force
 *   orc_remesh           brute-force threshold scan, circulation/impulse conservation
 *   orc_vpm2             Hill's spherical vortex closed form (test_applic.f90:2-77)
 *   orc_diffuse          exact on quadratic fields (the 7-point Laplacian of a quadratic is its
 *                        constant Laplacian)
 *   orc_vpm45            composition of pinned parts (projection, stretching, diffusion, M2P);
 *                        mode-4/5 semantics (UP zeroed in 5) read off vpm.f90/vpm_time.f90 --
 *                        parity unpinned beyond that composition
 */
#include "vpm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IX(c, i, j, k, nc, nx, ny) \
    ((size_t)(c) + (size_t)(nc) * ((size_t)(i) + (size_t)(nx) * ((size_t)(j) + (size_t)(ny) * (size_t)(k))))

/* Stage timers (seconds) and the target stride of the boundary sums, for timing a bounded
 * sample of a large case: 0 projection, 1 Poisson solves, 2 boundary sums, 3 whole vpm2/remesh. */
static double g_timer[4];
static int g_bs_stride = 1;
static double now(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / CLOCKS_PER_SEC;
#endif
}
void orc_timers(double *out, int reset) {
    for (int i = 0; i < 4; i++) {
        if (out) out[i] = g_timer[i];
        if (reset) g_timer[i] = 0.0;
    }
}
void orc_set_boundsum_stride(int s) { g_bs_stride = s > 0 ? s : 1; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ pm.inp */
/* vpm.f90:132-152: positional free-format read, one record per line. */
int orc_read_pm_inp(const char *path, orc_params *p) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[512];
    memset(p, 0, sizeof(*p));
#define NEXT() do { if (!fgets(line, sizeof line, f)) { fclose(f); return -2; } } while (0)
    NEXT(); if (sscanf(line, "%lf %lf %lf", &p->dpm[0], &p->dpm[1], &p->dpm[2]) != 3) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->interf_iproj) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->ibctyp) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->idvpm) != 1) goto bad;
    NEXT(); if (sscanf(line, "%lf", &p->epsvol) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->ncoarse) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d %d %d", &p->nbi, &p->nbj, &p->nbk) != 3) goto bad;
    NEXT(); if (sscanf(line, "%d %d", &p->nremesh, &p->ncell_rem) != 2) goto bad;
    NEXT(); if (sscanf(line, "%d %d", &p->iyntree, &p->ilevmax) != 2) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->ompthreads) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->idefine) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->iynslice) != 1) goto bad;
    NEXT(); if (sscanf(line, "%d", &p->ipmwrite) != 1) goto bad;
    if (p->ipmwrite > 10) { fclose(f); return -3; } /* vpm.f90:146 */
    for (int i = 0; i < p->ipmwrite; i++) {
        NEXT(); if (sscanf(line, "%d %d", &p->ipmwstart[i], &p->ipmwsteps[i]) != 2) goto bad;
    }
#undef NEXT
    fclose(f);
    return 0;
bad:
    fclose(f);
    return -2;
}

/* ------------------------------------------------------------ projection_fun */
/* pmproject.f90:391-431 */
double orc_projection_fun(int itype, double x) {
    double xabs = fabs(x), r = 0.0;
    if (itype == 2) {
        r = (xabs > 1) ? 0.0 : 1.0 - xabs;
    } else if (itype == 3) {
        if (xabs > 3.0 / 2.0) r = 0.0;
        if (xabs >= 0.5 && xabs <= 3.0 / 2.0) r = 0.5 * ((-xabs + 3.0 / 2.0) * (-xabs + 3.0 / 2.0));
        if (xabs >= 0.0 && xabs < 0.5)
            r = 0.5 * ((xabs + 3.0 / 2.0) * (xabs + 3.0 / 2.0)) - 3.0 / 2.0 * ((xabs + 0.5) * (xabs + 0.5));
    } else if (itype == 4) {
        if (xabs > 2.0) r = 0.0;
        if (xabs >= 1.0 && xabs <= 2.0) r = 0.5 * ((2.0 - xabs) * (2.0 - xabs)) * (1.0 - xabs);
        if (xabs < 1.0) r = 1.0 - 2.5 * (xabs * xabs) + 1.5 * (xabs * xabs * xabs);
    } else if (itype == 5) {
        if (xabs <= 0.5) r = 1.0 - xabs * xabs;
        if (xabs > 0.5 && xabs <= 1.5) r = 0.5 * (1.0 - xabs) * (2.0 - xabs);
        if (xabs > 1.5) r = 0.0;
    } else if (itype == 6) {
        if (xabs <= 1.0) r = 1.0 - xabs * xabs;
        if (xabs > 1.0 && xabs <= 2.0) r = 0.5 * (1.0 - xabs) * (2.0 - xabs);
        if (xabs > 2.0) r = 0.0;
    } else {
        r = NAN; /* 'No such projection function' STOP, pmproject.f90:426-427 */
    }
    return r;
}

/* -------------------------------------------------------------------- definepm */
/* pmlib.f90:216-432 (ND=3). */
void orc_definepm(int itype, double xb[6], double dpm[3], int ndum, const int32_t nsize[3], int32_t nn[3],
                  int32_t nn_bl[6]) {
    int ndum_new[3] = {ndum, ndum, ndum};
    if (itype == 1) {
        for (int d = 0; d < 3; d++) {
            xb[d] = xb[d] - ((double)ndum_new[d] * dpm[d]);
            xb[d + 3] = xb[d + 3] + ((double)ndum_new[d] * dpm[d]);
        }
        for (int d = 0; d < 3; d++) nn[d] = (int32_t)lround(fabs(xb[d + 3] - xb[d]) / dpm[d]) + 1;
    } else if (itype == 3) {
        for (int d = 0; d < 3; d++) {
            xb[d] = xb[d] - ((double)ndum_new[d] * dpm[d]);
            xb[d + 3] = xb[d + 3] + ((double)ndum_new[d] * dpm[d]);
        }
        for (int d = 0; d < 3; d++) nn[d] = (int32_t)lround(fabs(xb[d + 3] - xb[d]) / dpm[d]) + 1;
        for (int d = 0; d < 3; d++) ndum_new[d] = nsize[d] - ((nn[d] - 1) % nsize[d]);
        for (int d = 0; d < 3; d++) {
            if (ndum_new[d] % 2 == 0) {
                ndum_new[d] = ndum_new[d] / 2;
                xb[d] = xb[d] - ((double)ndum_new[d] * dpm[d]);
                xb[d + 3] = xb[d + 3] + ((double)ndum_new[d] * dpm[d]);
                ndum_new[d] = ndum_new[d] + ndum;
            } else {
                int nn1 = ndum_new[d] % 2;
                int nn2 = ndum_new[d] / 2;
                xb[d] = xb[d] - ((double)(nn1 + nn2) * dpm[d]);
                xb[d + 3] = xb[d + 3] + ((double)nn2 * dpm[d]);
            }
            nn[d] = (int32_t)lround(fabs(xb[d + 3] - xb[d]) / dpm[d]) + 1;
            nn_bl[d] = ndum + 1;
            nn_bl[d + 3] = nn[d] - ndum;
        }
        return;
    } else {
        return; /* itype 2/4 are not reached by define_sizes in 3-D */
    }
    for (int d = 0; d < 3; d++) {
        nn_bl[d] = ndum_new[d] + 1;
        nn_bl[d + 3] = nn[d] - ndum_new[d];
    }
}

/* ---------------------------------------------------------------- define_sizes */
/* vpm.f90:477-660 with BLOCKS = NBI*NBJ*NBK (the reference sets BLOCKS = MPI ranks). */
/* Uneven yaps blocks (an extension of the build; the reference's split is even): block
 * extents along each axis in proportion to these weights, in whole coarse cells,
 * cut_i = ncoarse * floor(units * (w_1 + .. + w_i) / W), units = cells / ncoarse.  No weights
 * (the default) is the reference's even split. */
static int g_bw_on = 0, g_bw[3][ORC_MAXB];
void orc_set_block_weights(const int *wx, const int *wy, const int *wz, int nbi, int nbj, int nbk) {
    g_bw_on = wx != NULL;
    if (!g_bw_on) return;
    for (int i = 0; i < nbi; i++) g_bw[0][i] = wx[i];
    for (int i = 0; i < nbj; i++) g_bw[1][i] = wy[i];
    for (int i = 0; i < nbk; i++) g_bw[2][i] = wz[i];
}
static void block_cuts(int cells, int nb, int ncoarse, const int *w, int *cut) {
    long long units = cells / ncoarse, W = 0, acc = 0;
    for (int i = 0; i < nb; i++) W += w ? w[i] : 1;
    for (int i = 0; i <= nb; i++) {
        cut[i] = (int)(ncoarse * ((units * acc) / W));
        if (i < nb) acc += w ? w[i] : 1;
    }
}

int orc_define_sizes(const orc_params *p, orc_geom *g, const double *xp, int64_t nvr, int ntime) {
    int nb = p->nbi * p->nbj * p->nbk;
    if (nb < 1 || nb > ORC_MAXB) return -1;
    int redefine = (ntime == 0 && p->idefine == 1) || p->idefine == 0 || !g->defined;
    double xb[6], dpm[3];
    if (redefine) {
        if (nvr <= 0) return -2;
        double mn[3] = {xp[0], xp[1], xp[2]}, mx[3] = {xp[0], xp[1], xp[2]};
        for (int64_t n = 1; n < nvr; n++)
            for (int d = 0; d < 3; d++) {
                double v = xp[3 * n + d];
                if (v < mn[d]) mn[d] = v;
                if (v > mx[d]) mx[d] = v;
            }
        for (int d = 0; d < 3; d++) {
            xb[d] = mn[d] - (double)p->interf_iproj * p->dpm[d];
            xb[d + 3] = mx[d] + (double)p->interf_iproj * p->dpm[d];
            dpm[d] = p->dpm[d];
        }
        int32_t nsiz[3] = {p->nbi * p->ncoarse, p->nbj * p->ncoarse, p->nbk * p->ncoarse};
        orc_definepm(3, xb, dpm, 0, nsiz, g->nn, g->nn_bl);
        memcpy(g->xbound, xb, sizeof xb);
        memcpy(g->dpm, dpm, sizeof dpm);
        g->defined = 1;
    }
    memcpy(xb, g->xbound, sizeof xb);
    memcpy(dpm, g->dpm, sizeof dpm);
    g->nblocks = nb;
    g->nbi = p->nbi; g->nbj = p->nbj; g->nbk = p->nbk;
    int nxb = (int)lround((xb[3] - xb[0]) / dpm[0]);
    int nyb = (int)lround((xb[4] - xb[1]) / dpm[1]);
    int nzb = (int)lround((xb[5] - xb[2]) / dpm[2]);
    int cx[ORC_MAXB + 1], cy[ORC_MAXB + 1], cz[ORC_MAXB + 1];
    block_cuts(nxb, p->nbi, p->ncoarse, g_bw_on ? g_bw[0] : NULL, cx);
    block_cuts(nyb, p->nbj, p->ncoarse, g_bw_on ? g_bw[1] : NULL, cy);
    block_cuts(nzb, p->nbk, p->ncoarse, g_bw_on ? g_bw[2] : NULL, cz);
    int32_t nsiz_bl[3] = {p->ncoarse, p->ncoarse, p->ncoarse};
    for (int k = 1; k <= p->nbk; k++)
        for (int j = 1; j <= p->nbj; j++)
            for (int i = 1; i <= p->nbi; i++) {
                int b = (k - 1) * p->nbj * p->nbi + (j - 1) * p->nbi + i - 1;
                double t[6];
                t[0] = xb[0] + (double)cx[i - 1] * dpm[0];
                t[3] = xb[0] + (double)cx[i] * dpm[0];
                t[1] = xb[1] + (double)cy[j - 1] * dpm[1];
                t[4] = xb[1] + (double)cy[j] * dpm[1];
                t[2] = xb[2] + (double)cz[k - 1] * dpm[2];
                t[5] = xb[2] + (double)cz[k] * dpm[2];
                double d2[3] = {dpm[0], dpm[1], dpm[2]};
                orc_definepm(1, t, d2, p->ncoarse, nsiz_bl, &g->nnbl[3 * b], &g->nnbl_bl[6 * b]);
                memcpy(&g->xbound_bl[6 * b], t, sizeof t);
            }
    memcpy(g->xbound_coarse, xb, sizeof xb);
    for (int d = 0; d < 3; d++) g->dpm_coarse[d] = (double)p->ncoarse * dpm[d];
    int32_t nsc[3];
    for (int d = 0; d < 3; d++) nsc[d] = 1 << p->ilevmax;
    orc_definepm(1, g->xbound_coarse, g->dpm_coarse, 4, nsc, g->nn_coarse, g->nn_bl_coarse);
    g->ilevmax = p->ilevmax;
    /* vpm.f90:628-645 tests block 1; with uneven blocks the smallest inner box decides */
    int ex[3] = {1 << 30, 1 << 30, 1 << 30};
    for (int b = 0; b < nb; b++)
        for (int d = 0; d < 3; d++)
            if (g->nnbl_bl[6 * b + d + 3] - g->nnbl_bl[6 * b + d] < ex[d]) ex[d] = g->nnbl_bl[6 * b + d + 3] - g->nnbl_bl[6 * b + d];
    for (int lev = 0; lev <= p->ilevmax; lev++) {
        int istep = 1 << lev;
        if (ex[0] / istep == 0 || ex[1] / istep == 0 || ex[2] / istep == 0) {
            g->ilevmax = lev - 1;
            break;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------- FFT / DST */
typedef struct {
    int n, nf, fac[40];
    double complex *tw;
} fftplan;

static void plan_init(fftplan *P, int n) {
    P->n = n;
    P->nf = 0;
    int r = n;
    while (r % 4 == 0) { P->fac[P->nf++] = 4; r /= 4; }
    while (r % 2 == 0) { P->fac[P->nf++] = 2; r /= 2; }
    for (int q = 3; r > 1;) {
        if (r % q == 0) { P->fac[P->nf++] = q; r /= q; }
        else q += 2;
    }
    if (P->nf == 0) P->fac[P->nf++] = 1;
    P->tw = (double complex *)malloc(sizeof(double complex) * (size_t)n);
    for (int t = 0; t < n; t++) {
        double a = 2.0 * M_PI * (double)t / (double)n;
        P->tw[t] = cos(a) - I * sin(a);
    }
}
static void plan_free(fftplan *P) { free(P->tw); }

/* Mixed-radix decimation-in-time: out[k] = sum_j in[j*istride] w_n^{jk}, w_n = exp(-2 pi i/n). */
static void fft_rec(const fftplan *P, double complex *out, const double complex *in, size_t istride, int n,
                    const int *fac, int twmul, double complex *scratch) {
    int p = fac[0], m = n / p, N = P->n;
    if (m == 1) {
        for (int q = 0; q < p; q++) out[q] = in[(size_t)q * istride];
    } else {
        for (int q = 0; q < p; q++)
            fft_rec(P, out + (size_t)q * m, in + (size_t)q * istride, istride * (size_t)p, m, fac + 1, twmul * p,
                    scratch);
    }
    for (int s = 0; s < m; s++) {
        for (int q = 0; q < p; q++) scratch[q] = out[(size_t)q * m + s] * P->tw[((long)q * s * twmul) % N];
        for (int k = 0; k < p; k++) {
            double complex acc = 0;
            for (int q = 0; q < p; q++) acc += scratch[q] * P->tw[((long)q * k * (N / p)) % N];
            out[(size_t)k * m + s] = acc;
        }
    }
}

/* DST-I of two real lines at once (inputs strided, outputs contiguous) through one complex FFT of the odd extension
 * (length L = 2(m+1)): S_k = sum_{j=1..m} x_j sin(pi j k/(m+1)). */
static void dst_pair(const fftplan *P, int m, const double *a, size_t sa, const double *b, size_t sb, double *oa,
                     double *ob, double complex *zin, double complex *zout, double complex *scr) {
    int L = 2 * (m + 1);
    zin[0] = 0;
    zin[m + 1] = 0;
    for (int j = 1; j <= m; j++) {
        double complex v = a[(size_t)(j - 1) * sa] + I * (b ? b[(size_t)(j - 1) * sb] : 0.0);
        zin[j] = v;
        zin[L - j] = -v;
    }
    fft_rec(P, zout, zin, 1, L, P->fac, 1, scr);
    for (int k = 1; k <= m; k++) {
        oa[k - 1] = -cimag(zout[k]) * 0.5;
        if (b) ob[k - 1] = creal(zout[k]) * 0.5;
    }
}

void orc_dst1(int m, const double *in, double *out) {
    fftplan P;
    plan_init(&P, 2 * (m + 1));
    int L = 2 * (m + 1);
    double complex *zi = malloc(sizeof(double complex) * L), *zo = malloc(sizeof(double complex) * L),
                   *sc = malloc(sizeof(double complex) * L);
    double *tmp = malloc(sizeof(double) * m);
    memcpy(tmp, in, sizeof(double) * m);
    dst_pair(&P, m, tmp, 1, NULL, 0, out, NULL, zi, zo, sc);
    free(zi); free(zo); free(sc); free(tmp);
    plan_free(&P);
}

/* In-place DST-I along one axis of w (mx,my,mz x-fastest); scale applied to results. */
static void dst_axis(double *w, int mx, int my, int mz, int axis, double scale) {
    int m = axis == 0 ? mx : (axis == 1 ? my : mz);
    size_t stride = axis == 0 ? 1 : (axis == 1 ? (size_t)mx : (size_t)mx * my);
    int nl1 = axis == 0 ? my : mx, nl2 = axis == 2 ? my : mz;
    size_t s1 = axis == 0 ? (size_t)mx : 1, s2 = axis == 2 ? (size_t)mx : (size_t)mx * my;
    long nlines = (long)nl1 * nl2;
    fftplan P;
    plan_init(&P, 2 * (m + 1));
    int L = 2 * (m + 1);
#pragma omp parallel
    {
        double complex *zi = malloc(sizeof(double complex) * L), *zo = malloc(sizeof(double complex) * L),
                       *sc = malloc(sizeof(double complex) * L);
        double *la = malloc(sizeof(double) * m), *lb = malloc(sizeof(double) * m);
#pragma omp for schedule(static)
        for (long pr = 0; pr < (nlines + 1) / 2; pr++) {
            long l0 = 2 * pr, l1 = 2 * pr + 1;
            double *a = w + (size_t)(l0 % nl1) * s1 + (size_t)(l0 / nl1) * s2;
            double *b = l1 < nlines ? w + (size_t)(l1 % nl1) * s1 + (size_t)(l1 / nl1) * s2 : NULL;
            dst_pair(&P, m, a, stride, b, stride, la, lb, zi, zo, sc);
            for (int k = 0; k < m; k++) a[(size_t)k * stride] = la[k] * scale;
            if (b)
                for (int k = 0; k < m; k++) b[(size_t)k * stride] = lb[k] * scale;
        }
        free(zi); free(zo); free(sc); free(la); free(lb);
    }
    plan_free(&P);
}

/* ------------------------------------------------------ Dirichlet Poisson solve */
/* pmsolve.f90:284-432 with MKL's d_Helmholtz_3D(q=0,'DDDDDD') restated. */
void orc_solve_window(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int comp,
                      const int32_t win[6], const double *rhs, double *sol, int zero_bc, double *out) {
    int s[3] = {win[0], win[1], win[2]}, f[3] = {win[3], win[4], win[5]};
    int n[3], m[3];
    double h[3];
    for (int d = 0; d < 3; d++) {
        double mn = xb[d] + (double)(s[d] - 1) * dpm[d];
        double mx = xb[d] + (double)(f[d] - 1) * dpm[d];
        n[d] = f[d] - s[d] + 1;
        m[d] = n[d] - 2;
        h[d] = (mx - mn) / (double)(n[d] - 1); /* MKL: h = (b-a)/n_intervals */
    }
    double t0 = now();
    const double *bc = zero_bc ? NULL : sol;
    double *dst = zero_bc ? out : sol;
    int NX = nn[0], NY = nn[1];
#define G(i, j, k) IX(comp, (s[0] - 1) + (i), (s[1] - 1) + (j), (s[2] - 1) + (k), nc, NX, NY)
    size_t mt = (size_t)m[0] * m[1] * m[2];
    double *w = malloc(sizeof(double) * (mt ? mt : 1));
    double ih2[3] = {1.0 / (h[0] * h[0]), 1.0 / (h[1] * h[1]), 1.0 / (h[2] * h[2])};
    /* f = -RHS; boundary values moved to the right-hand side (interior node equations) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= m[2]; k++)
        for (int j = 1; j <= m[1]; j++)
            for (int i = 1; i <= m[0]; i++) {
                double v = rhs[G(i, j, k)];
                if (bc) {
                    if (i == 1) v -= bc[G(0, j, k)] * ih2[0];
                    if (i == m[0]) v -= bc[G(m[0] + 1, j, k)] * ih2[0];
                    if (j == 1) v -= bc[G(i, 0, k)] * ih2[1];
                    if (j == m[1]) v -= bc[G(i, m[1] + 1, k)] * ih2[1];
                    if (k == 1) v -= bc[G(i, j, 0)] * ih2[2];
                    if (k == m[2]) v -= bc[G(i, j, m[2] + 1)] * ih2[2];
                }
                w[(size_t)(i - 1) + (size_t)m[0] * ((size_t)(j - 1) + (size_t)m[1] * (k - 1))] = v;
            }
    dst_axis(w, m[0], m[1], m[2], 0, 1.0);
    dst_axis(w, m[0], m[1], m[2], 1, 1.0);
    /* tridiagonal along z for each (kx, ky): (u_{k-1}-2u_k+u_{k+1})/hz^2 - (lx+ly) u_k = r_k */
#pragma omp parallel
    {
        double *cp = malloc(sizeof(double) * (m[2] + 1)), *dp = malloc(sizeof(double) * (m[2] + 1));
#pragma omp for schedule(static)
        for (int ky = 1; ky <= m[1]; ky++)
            for (int kx = 1; kx <= m[0]; kx++) {
                double sx = sin(M_PI * kx / (2.0 * (m[0] + 1))), sy = sin(M_PI * ky / (2.0 * (m[1] + 1)));
                double mu = 4.0 * sx * sx * ih2[0] + 4.0 * sy * sy * ih2[1];
                double a = ih2[2], c = ih2[2], b = -2.0 * ih2[2] - mu;
                double *col = w + (size_t)(kx - 1) + (size_t)m[0] * (ky - 1);
                size_t st = (size_t)m[0] * m[1];
                cp[0] = c / b;
                dp[0] = col[0] / b;
                for (int k = 1; k < m[2]; k++) {
                    double den = b - a * cp[k - 1];
                    cp[k] = c / den;
                    dp[k] = (col[k * st] - a * dp[k - 1]) / den;
                }
                col[(size_t)(m[2] - 1) * st] = dp[m[2] - 1];
                for (int k = m[2] - 2; k >= 0; k--) col[k * st] = dp[k] - cp[k] * col[(k + 1) * st];
            }
        free(cp); free(dp);
    }
    dst_axis(w, m[0], m[1], m[2], 1, 2.0 / (m[1] + 1));
    dst_axis(w, m[0], m[1], m[2], 0, 2.0 / (m[0] + 1));
    /* write the whole window: interior from the solve, boundary = Dirichlet values */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; k++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                int bnd = (i == 0 || j == 0 || k == 0 || i == n[0] - 1 || j == n[1] - 1 || k == n[2] - 1);
                double v;
                if (bnd) v = bc ? bc[G(i, j, k)] : 0.0;
                else v = w[(size_t)(i - 1) + (size_t)m[0] * ((size_t)(j - 1) + (size_t)m[1] * (k - 1))];
                dst[G(i, j, k)] = v;
            }
#undef G
    free(w);
    g_timer[1] += now() - t0;
}

/* ------------------------------------------------------- normal derivatives */
/* calc_normalderiv_3d, pinfdomain.f90:294-547; sources at face-cell centroids. */
int64_t orc_normalderiv(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, const double *s0,
                        double *src, double *cxyz, double *ds) {
    const double a1 = 25.0 / 12.0, a2 = -4.0, a3 = 3.0, a4 = -4.0 / 3.0, a5 = 0.25;
    int NX = nn[0], NY = nn[1], NZ = nn[2];
    int64_t nb = 0;
    double DX = dpm[0], DY = dpm[1], DZ = dpm[2];
#define S(c, i, j, k) s0[IX(c, (i)-1, (j)-1, (k)-1, nc, NX, NY)]
#define XN(i) (xb[0] + (double)((i)-1) * DX)
#define YN(j) (xb[1] + (double)((j)-1) * DY)
#define ZN(k) (xb[2] + (double)((k)-1) * DZ)
#define C4(a, b, c, d) (0.25 * ((((a) + (b)) + (c)) + (d)))
    /* stencil along x from i in direction sgn */
#define DX5(c, i, j, k, sg) \
    ((1.0 / DX) * ((((a1 * S(c, i, j, k) + a2 * S(c, (i) + (sg), j, k)) + a3 * S(c, (i) + 2 * (sg), j, k)) + \
                    a4 * S(c, (i) + 3 * (sg), j, k)) + a5 * S(c, (i) + 4 * (sg), j, k)))
#define DY5(c, i, j, k, sg) \
    ((1.0 / DY) * ((((a1 * S(c, i, j, k) + a2 * S(c, i, (j) + (sg), k)) + a3 * S(c, i, (j) + 2 * (sg), k)) + \
                    a4 * S(c, i, (j) + 3 * (sg), k)) + a5 * S(c, i, (j) + 4 * (sg), k)))
#define DZ5(c, i, j, k, sg) \
    ((1.0 / DZ) * ((((a1 * S(c, i, j, k) + a2 * S(c, i, j, (k) + (sg))) + a3 * S(c, i, j, (k) + 2 * (sg))) + \
                    a4 * S(c, i, j, (k) + 3 * (sg))) + a5 * S(c, i, j, (k) + 4 * (sg))))
    for (int face = 0; face < 2; face++) { /* XMIN, XMAX */
        int i = face == 0 ? 1 : NX, sg = face == 0 ? 1 : -1;
        for (int k = 1; k <= NZ - 1; k++)
            for (int j = 1; j <= NY - 1; j++) {
                for (int c = 0; c < nc; c++) {
                    double p1 = DX5(c, i, j, k, sg), p2 = DX5(c, i, j + 1, k, sg);
                    double p3 = DX5(c, i, j + 1, k + 1, sg), p4 = DX5(c, i, j, k + 1, sg);
                    src[(size_t)nc * nb + c] = 0.25 * (((p1 + p2) + p3) + p4);
                }
                double X = XN(i);
                cxyz[3 * nb + 0] = C4(X, X, X, X);
                if (face == 0) {
                    cxyz[3 * nb + 1] = C4(YN(j), YN(j), YN(j + 1), YN(j + 1));
                    cxyz[3 * nb + 2] = C4(ZN(k), ZN(k + 1), ZN(k + 1), ZN(k));
                } else {
                    cxyz[3 * nb + 1] = C4(YN(j + 1), YN(j + 1), YN(j), YN(j));
                    cxyz[3 * nb + 2] = C4(ZN(k), ZN(k + 1), ZN(k + 1), ZN(k));
                }
                ds[nb] = DY * DZ;
                nb++;
            }
    }
    for (int face = 0; face < 2; face++) { /* YMIN, YMAX */
        int j = face == 0 ? 1 : NY, sg = face == 0 ? 1 : -1;
        for (int k = 1; k <= NZ - 1; k++)
            for (int i = 1; i <= NX - 1; i++) {
                for (int c = 0; c < nc; c++) {
                    double p1 = DY5(c, i, j, k, sg), p2 = DY5(c, i + 1, j, k, sg);
                    double p3 = DY5(c, i + 1, j, k + 1, sg), p4 = DY5(c, i, j, k + 1, sg);
                    src[(size_t)nc * nb + c] = 0.25 * (((p1 + p2) + p3) + p4);
                }
                double Y = YN(j);
                cxyz[3 * nb + 1] = C4(Y, Y, Y, Y);
                if (face == 0) {
                    cxyz[3 * nb + 0] = C4(XN(i), XN(i), XN(i + 1), XN(i + 1));
                    cxyz[3 * nb + 2] = C4(ZN(k), ZN(k + 1), ZN(k + 1), ZN(k));
                } else {
                    cxyz[3 * nb + 0] = C4(XN(i + 1), XN(i + 1), XN(i), XN(i));
                    cxyz[3 * nb + 2] = C4(ZN(k), ZN(k + 1), ZN(k + 1), ZN(k));
                }
                ds[nb] = DX * DZ;
                nb++;
            }
    }
    for (int face = 0; face < 2; face++) { /* ZMIN, ZMAX */
        int k = face == 0 ? 1 : NZ, sg = face == 0 ? 1 : -1;
        for (int j = 1; j <= NY - 1; j++)
            for (int i = 1; i <= NX - 1; i++) {
                for (int c = 0; c < nc; c++) {
                    double p1 = DZ5(c, i, j, k, sg), p2 = DZ5(c, i + 1, j, k, sg);
                    double p3 = DZ5(c, i + 1, j + 1, k, sg), p4 = DZ5(c, i, j + 1, k, sg);
                    src[(size_t)nc * nb + c] = 0.25 * (((p1 + p2) + p3) + p4);
                }
                double Z = ZN(k);
                cxyz[3 * nb + 2] = C4(Z, Z, Z, Z);
                if (face == 0) {
                    cxyz[3 * nb + 0] = C4(XN(i), XN(i), XN(i + 1), XN(i + 1));
                    cxyz[3 * nb + 1] = C4(YN(j), YN(j + 1), YN(j + 1), YN(j));
                } else {
                    cxyz[3 * nb + 0] = C4(XN(i + 1), XN(i + 1), XN(i), XN(i));
                    cxyz[3 * nb + 1] = C4(YN(j), YN(j + 1), YN(j + 1), YN(j));
                }
                ds[nb] = DX * DY;
                nb++;
            }
    }
#undef S
#undef DX5
#undef DY5
#undef DZ5
    return nb;
}

/* -------------------------------------------------------------- boundary sums */
/* Bounds3d (pmbound.f90:80-157) -> calc_boundinf_3d (pmbound.f90:415-515), ibctyp=2.
 * Targets: X faces full, Y faces skip x-edges, Z faces skip x- and y-edges; per target the
 * sources are accumulated in source order, as the reference's nv loop does. */
void orc_boundsum(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                  const double *src, const double *cxyz, const double *ds, double *sol) {
    const double PI = 4.0 * atan(1.0), PI4 = 4.0 * PI;
    const double lg = log((sqrt(2.0) - 1) / (sqrt(2.0) + 1));
    int NX = nn[0], NY = nn[1], NZ = nn[2];
    double DX = dpm[0], DY = dpm[1], DZ = dpm[2];
    /* enumerate targets */
    int64_t nt = 2LL * NY * NZ + 2LL * (NX - 2) * NZ + 2LL * (NX - 2) * (NY - 2);
    int32_t *tg = malloc(sizeof(int32_t) * 4 * nt);
    int64_t t = 0;
    for (int f = 0; f < 2; f++)
        for (int k = 1; k <= NZ; k++)
            for (int j = 1; j <= NY; j++) { tg[4*t] = f ? NX : 1; tg[4*t+1] = j; tg[4*t+2] = k; tg[4*t+3] = 2; t++; }
    for (int f = 0; f < 2; f++)
        for (int k = 1; k <= NZ; k++)
            for (int i = 2; i <= NX - 1; i++) { tg[4*t] = i; tg[4*t+1] = f ? NY : 1; tg[4*t+2] = k; tg[4*t+3] = 1; t++; }
    for (int f = 0; f < 2; f++)
        for (int j = 2; j <= NY - 1; j++)
            for (int i = 2; i <= NX - 1; i++) { tg[4*t] = i; tg[4*t+1] = j; tg[4*t+2] = f ? NZ : 1; tg[4*t+3] = 3; t++; }
    double t0 = now();
    int64_t stride = g_bs_stride;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nt; q += stride) {
        int i = tg[4 * q], j = tg[4 * q + 1], k = tg[4 * q + 2], pl = tg[4 * q + 3];
        double X = xb[0] + (double)(i - 1) * DX, Y = xb[1] + (double)(j - 1) * DY, Z = xb[2] + (double)(k - 1) * DZ;
        double hs = pl == 2 ? (DY + DZ) : (pl == 1 ? (DX + DZ) : (DX + DY));
        double acc[8];
        for (int c = 0; c < nc; c++) acc[c] = sol[IX(c, i - 1, j - 1, k - 1, nc, NX, NY)];
        for (int64_t nv = 0; nv < nbound; nv++) {
            double XR = cxyz[3 * nv] - X, YR = cxyz[3 * nv + 1] - Y, ZR = cxyz[3 * nv + 2] - Z;
            double r = sqrt(XR * XR + YR * YR + ZR * ZR);
            double DS = ds[nv];
            if (r > 1e-05) {
                for (int c = 0; c < nc; c++) acc[c] = acc[c] - src[(size_t)nc * nv + c] * DS / (PI4 * r);
            } else {
                for (int c = 0; c < nc; c++) acc[c] = acc[c] + src[(size_t)nc * nv + c] * hs * lg / PI4;
            }
        }
        for (int c = 0; c < nc; c++) sol[IX(c, i - 1, j - 1, k - 1, nc, NX, NY)] = acc[c];
    }
    free(tg);
    g_timer[2] += now() - t0;
}

/* ------------------------------------------------------------- tree boundary */
/* build_level_nbound_3d (pinfdomain.f90:875-1530): per face a hierarchy of cell groups,
 * level lev grouping 2^lev x 2^lev level-0 cells (centroid = mean of the level-0 centroids,
 * area = sum, source = area-weighted mean); groups that do not fit the face are "dummy"
 * cells that only point at their two children.  Arrays are 1-based like the reference:
 * entry [lev][i], i = 1..nbound_lev[lev]. */
typedef struct {
    int levmax;
    int nlev[16];
    double *xs[16], *ys[16], *zs[16], *dsl[16], *src[16];  /* src: nc per cell */
    int32_t *t[16];                                           /* ilev_t: 6 per cell (1..6) */
    int nc;
} orc_tree;

#define TL(lv, i, m) tr->t[lv][6 * (size_t)(i) + (m) - 1]

/* Refinement factor of the descent (10 in the reference); tests raise it to reach every leaf. */
static double g_tree_factor = 10.0;
void orc_set_tree_factor(double f) { g_tree_factor = f; }

static void tree_face(orc_tree *tr, int lev_unused, const double *xs_tmp, const double *ds_tmp, const double *s_tmp,
                      int S1, int F1, int S2, int F2, int NNrow, int d1, int nn_lev[3][16], int *nbound_lev) {
    (void)lev_unused;
    int nc = tr->nc;
    int ncount[16];
    for (int lev = 0; lev <= tr->levmax; lev++) {
        int istep = 1 << lev;
        int n2s = S2, n2f = F2 - istep, n1s = S1, n1f = F1 - istep;
        int icount = nbound_lev[lev];
        int lf = nbound_lev[0];
        for (int k2 = n2s; k2 <= n2f; k2 += istep) {
            for (int k1 = n1s; k1 <= n1f; k1 += istep) {
                double xc = 0, yc = 0, zc = 0, sc = 0, sourcec[8] = {0};
                for (int mm = 1; mm <= istep; mm++)
                    for (int nn = 1; nn <= istep; nn++) {
                        int leaf = lf + (k2 - S2) * NNrow + (k1 - S1) + (mm - 1) * NNrow + nn; /* 1-based */
                        double x = xs_tmp[3 * (size_t)(leaf - 1)], y = xs_tmp[3 * (size_t)(leaf - 1) + 1],
                               z = xs_tmp[3 * (size_t)(leaf - 1) + 2], sv = ds_tmp[leaf - 1];
                        xc = x + xc; yc = y + yc; zc = z + zc; sc = sv + sc;
                        for (int c = 0; c < nc; c++) sourcec[c] = s_tmp[(size_t)nc * (leaf - 1) + c] * sv + sourcec[c];
                    }
                icount++;
                double fl = (double)(float)(istep * istep);
                tr->xs[lev][icount] = xc / fl;
                tr->ys[lev][icount] = yc / fl;
                tr->zs[lev][icount] = zc / fl;
                tr->dsl[lev][icount] = sc;
                for (int c = 0; c < nc; c++) tr->src[lev][(size_t)nc * icount + c] = sourcec[c] / sc;
                int npre = nn_lev[d1][lev];
                if (lev > 0) {
                    int nnl = nn_lev[d1][lev - 1], nj;
                    if (k1 == n1s && icount - 1 == nbound_lev[lev]) nj = nbound_lev[lev - 1] + 1;
                    else if (k1 == n1s && icount - 1 > nbound_lev[lev]) {
                        nj = TL(lev, icount - 1, 3) + 1;
                        if (TL(lev, icount - 1, 6) < 0) nj = TL(lev, icount - 1, 2) + 1;
                    } else nj = TL(lev, icount - 1, 2) + 1;
                    TL(lev, icount, 1) = nj;
                    TL(lev, icount, 2) = nj + 1;
                    TL(lev, icount, 3) = nj + nnl + 1;
                    TL(lev, icount, 4) = nj + nnl;
                    TL(lev, icount, 5) = 4;
                }
                TL(lev, icount, 6) = npre;
            }
            if ((F1 - S1) % istep != 0 || F1 - S1 <= istep) {
                if (lev == 0) { fprintf(stderr, "tree: level 0 dummy\n"); abort(); }
                icount++;
                int nj = TL(lev, icount - 1, 2) + 1, nnl = nn_lev[d1][lev - 1];
                TL(lev, icount, 1) = nj;
                TL(lev, icount, 2) = nj + nnl;
                TL(lev, icount, 5) = 2;
                TL(lev, icount, 6) = -1;
            }
        }
        if ((F2 - S2) % istep != 0 || F2 - S2 <= istep) {
            if (lev == 0) { fprintf(stderr, "tree: level 0 dummy\n"); abort(); }
            for (int k1 = n1s; k1 <= n1f; k1 += istep) {
                if (k1 == n1f && (F1 - S1) % istep != 0) continue;
                icount++;
                int nj;
                if (k1 == n1s && TL(lev, icount - 1, 6) > 0) nj = TL(lev, icount - 1, 3) + 1;
                else nj = TL(lev, icount - 1, 2) + 1;
                TL(lev, icount, 1) = nj;
                TL(lev, icount, 2) = nj + 1;
                TL(lev, icount, 5) = 2;
                TL(lev, icount, 6) = -2;
            }
        }
        ncount[lev] = icount;
    }
    for (int lev = 0; lev <= tr->levmax; lev++) nbound_lev[lev] = ncount[lev];
}

/* Sources: nb face cells in the reference face order (src nc per cell, cxyz centroids, ds).
 * Node extents of the grid nn (NXs = 1, NXf = NX, ...). */
static void tree_build(orc_tree *tr, int levmax, const int32_t nn[3], int nc, int64_t nb, const double *src,
                       const double *cxyz, const double *ds) {
    memset(tr, 0, sizeof *tr);
    tr->levmax = levmax;
    tr->nc = nc;
    int NX = nn[0], NY = nn[1], NZ = nn[2];
    int NNX = NX - 1, NNY = NY - 1, NNZ = NZ - 1;
    int nn_lev[3][16];
    for (int lev = 0; lev <= levmax; lev++) {
        int istep = 1 << lev;
        nn_lev[0][lev] = NNX / istep + (NNX % istep != 0);
        nn_lev[1][lev] = NNY / istep + (NNY % istep != 0);
        nn_lev[2][lev] = NNZ / istep + (NNZ % istep != 0);
        tr->xs[lev] = calloc((size_t)nb + 2, sizeof(double));
        tr->ys[lev] = calloc((size_t)nb + 2, sizeof(double));
        tr->zs[lev] = calloc((size_t)nb + 2, sizeof(double));
        tr->dsl[lev] = calloc((size_t)nb + 2, sizeof(double));
        tr->src[lev] = calloc(((size_t)nb + 2) * nc, sizeof(double));
        tr->t[lev] = calloc(((size_t)nb + 2) * 6, sizeof(int32_t));
    }
    int nbl[16] = {0};
    /* XMIN, XMAX: outer k, inner j, rows of NNY cells; YMIN, YMAX and ZMIN, ZMAX: inner i */
    tree_face(tr, 0, cxyz, ds, src, 1, NY, 1, NZ, NNY, 1, nn_lev, nbl);
    tree_face(tr, 0, cxyz, ds, src, 1, NY, 1, NZ, NNY, 1, nn_lev, nbl);
    tree_face(tr, 0, cxyz, ds, src, 1, NX, 1, NZ, NNX, 0, nn_lev, nbl);
    tree_face(tr, 0, cxyz, ds, src, 1, NX, 1, NZ, NNX, 0, nn_lev, nbl);
    tree_face(tr, 0, cxyz, ds, src, 1, NX, 1, NY, NNX, 0, nn_lev, nbl);
    tree_face(tr, 0, cxyz, ds, src, 1, NX, 1, NY, NNX, 0, nn_lev, nbl);
    if (nbl[0] != nb) { fprintf(stderr, "tree: nbound_lev problem %d %lld\n", nbl[0], (long long)nb); abort(); }
    for (int lev = 0; lev <= levmax; lev++) tr->nlev[lev] = nbl[lev];
}

static void tree_free(orc_tree *tr) {
    for (int lev = 0; lev <= tr->levmax; lev++) {
        free(tr->xs[lev]); free(tr->ys[lev]); free(tr->zs[lev]); free(tr->dsl[lev]); free(tr->src[lev]);
        free(tr->t[lev]);
    }
}

/* tree_calc_3d (pmbound.f90:1139-1237): SOURCE -= s DS / r over the cells reached by the
 * descent (a cell is refined when r < 10 sqrt(DS) above level 0, and dummies always). */
static void tree_calc(const orc_tree *tr, int nv, int nlev, double X, double Y, double Z, double *S) {
    int list[4], nleaf_n, use_list;
    if (nlev == tr->levmax) { nleaf_n = tr->nlev[nlev]; use_list = 0; }
    else if (TL(nlev + 1, nv, 6) < 0) { list[0] = TL(nlev + 1, nv, 1); list[1] = TL(nlev + 1, nv, 2); nleaf_n = 2; use_list = 1; }
    else { for (int q = 0; q < 4; q++) list[q] = TL(nlev + 1, nv, q + 1); nleaf_n = 4; use_list = 1; }
    const double lg = log((sqrt(2.0) - 1) / (sqrt(2.0) + 1));
    for (int q = 0; q < nleaf_n; q++) {
        int nleaf = use_list ? list[q] : q + 1;
        if (TL(nlev, nleaf, 6) == 0) { fprintf(stderr, "Something is wrong with the tree %d %d\n", nleaf, nlev); abort(); }
        double XR = tr->xs[nlev][nleaf] - X, YR = tr->ys[nlev][nleaf] - Y, ZR = tr->zs[nlev][nleaf] - Z;
        double r = sqrt(XR * XR + YR * YR + ZR * ZR);
        double DS = tr->dsl[nlev][nleaf];
        if ((r < g_tree_factor * sqrt(DS) && nlev > 0) || TL(nlev, nleaf, 6) < 0) {
            tree_calc(tr, nleaf, nlev - 1, X, Y, Z, S);
        } else {
            const double *sb = tr->src[nlev] + (size_t)tr->nc * nleaf;
            if (r > 1e-05) {
                for (int c = 0; c < tr->nc; c++) S[c] = S[c] - sb[c] * DS / r;
            } else {
                for (int c = 0; c < tr->nc; c++) S[c] = S[c] + sb[c] * 2 * sqrt(DS) * lg;
            }
        }
    }
}

/* ------------------------------------------------- panel sources (ibctyp = 3) */
/* Frame of the TARGET plane in calc_boundinf_3d_s / calc_boundinf_3d_lev_s
 * (pmbound.f90:537-620, 1271-1337): every source panel is taken to lie parallel to the
 * plane being evaluated.  iplane = +-1 Y planes, +-2 X planes, +-3 Z planes, negative on
 * the max faces (Bounds3d, pmbound.f90:80-157). */
typedef struct {
    double E1[3], E2[3], E3[3], S[4], T[4], D[4], SINB[4], COSB[4], DIAG, AREA;
} panel_frame;

static void panel_frame_init(panel_frame *f, int iplane, const double dpm[3]) {
    memset(f, 0, sizeof *f);
    const double si = iplane > 0 ? 1.0 : -1.0;
    const double DX = dpm[0], DY = dpm[1], DZ = dpm[2];
    double a = 0, b = 0;
    f->COSB[1] = 1.0; f->COSB[3] = -1.0;
    f->SINB[0] = 1.0; f->SINB[2] = -1.0;
    switch (iplane < 0 ? -iplane : iplane) {
    case 1: f->DIAG = sqrt(DX * DX + DZ * DZ); f->E1[0] = 1.0 * si; f->E2[2] = 1.0; f->E3[1] = -1.0 * si;
            f->AREA = DX * DZ; a = DX; b = DZ; break;
    case 2: f->DIAG = sqrt(DY * DY + DZ * DZ); f->E1[2] = 1.0; f->E2[1] = -1.0 * si; f->E3[0] = -1.0 * si;
            f->AREA = DY * DZ; a = DY; b = DZ; break;
    default: f->DIAG = sqrt(DX * DX + DY * DY); f->E1[0] = 1.0; f->E2[1] = -1.0 * si; f->E3[2] = -1.0 * si;
            f->AREA = DX * DY; a = DX; b = DY; break;
    }
    f->S[0] = -0.5 * a; f->S[1] = -0.5 * a; f->S[2] = 0.5 * a; f->S[3] = 0.5 * a;
    f->T[0] = -0.5 * b; f->T[1] = 0.5 * b; f->T[2] = 0.5 * b; f->T[3] = -0.5 * b;
    f->D[0] = b; f->D[1] = a; f->D[2] = b; f->D[3] = a;
}

/* FSOUR_A4 (pmbound.f90:1541-1632) with NSIDE = 0, ISING = 0, EPSS = 1e-14: potential of a
 * flat constant-source quadrilateral (corners S,T about RG in the frame E1,E2,E3) at XO;
 * far field (RO/DIAG > 10) -AREA/(4 pi RO). */
static double fsour_a4(const double XO[3], const double RG[3], const panel_frame *f, double AREA, double PI4) {
    const double EPSS = 1e-14, TINY = 1e-10;
    const double dx = XO[0] - RG[0], dy = XO[1] - RG[1], dz = XO[2] - RG[2];
    double X = dx * f->E1[0] + dy * f->E1[1] + dz * f->E1[2];
    double Y = dx * f->E2[0] + dy * f->E2[1] + dz * f->E2[2];
    double Z = dx * f->E3[0] + dy * f->E3[1] + dz * f->E3[2];
    const double RO = sqrt(X * X + Y * Y + Z * Z);
    const double RATIO = RO / f->DIAG;
    double FIS = 0;
    if (RATIO > 10) {
        const double AREA1 = AREA / PI4;
        return -AREA1 / RO;
    }
    double UL0 = 0.0, UL1 = 0.0, UL2, FIL = 0.0;
    double ZP = 1.0;
    if (Z < 0.0) ZP = -1.0;
    double AZ = fabs(Z);
    AZ = AZ > EPSS ? AZ : EPSS;
    Z = AZ * ZP;
    const double ZK = Z * Z;
    double R[4], E[4], H[4];
    for (int K = 0; K < 4; K++) {
        const double XK = (X - f->S[K]) * (X - f->S[K]);
        const double YK = (Y - f->T[K]) * (Y - f->T[K]);
        H[K] = (X - f->S[K]) * (Y - f->T[K]);
        R[K] = sqrt(XK + YK + ZK);
        E[K] = ZK + XK;
    }
    for (int K = 0; K < 4; K++) {
        const int K1 = K, K2 = K == 3 ? 0 : K + 1;
        const double A2 = R[K1] + R[K2] + f->D[K];
        const double A1 = R[K1] + R[K2] - f->D[K];
        if (fabs(A1 * A2) < TINY) continue;
        const double AK = log(A1 / A2);
        const double TOO = -(X - f->S[K1]) * f->SINB[K] + (Y - f->T[K1]) * f->COSB[K];
        UL0 = UL0 + (f->T[K2] - f->T[K1]) * AK / f->D[K];
        UL1 = UL1 - (f->S[K2] - f->S[K1]) * AK / f->D[K];
        FIL = FIL - TOO * AK;
    }
    (void)UL0; (void)UL1;
    AZ = AZ > TINY ? AZ : TINY;
    Z = AZ * ZP;
    UL2 = 0.0;
    for (int K = 0; K < 4; K++) {
        if (fabs(f->COSB[K]) >= TINY) {
            const int K1 = K, K2 = K == 3 ? 0 : K + 1;
            UL2 = UL2 + atan((f->SINB[K] * E[K1] - f->COSB[K] * H[K1]) / (Z * f->COSB[K] * R[K1])) -
                  atan((f->SINB[K] * E[K2] - f->COSB[K] * H[K2]) / (Z * f->COSB[K] * R[K2]));
        }
    }
    FIS = FIS + (FIL + Z * UL2) / PI4;
    return FIS;
}

double orc_fsour_a4(const double XO[3], const double RG[3], int iplane, const double dpm[3], double area) {
    panel_frame f;
    panel_frame_init(&f, iplane, dpm);
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    return fsour_a4(XO, RG, &f, area > 0 ? area : f.AREA, PI4);
}

/* tree_calc_3d_s (pmbound.f90:1346-1481): refine while r/sqrt(DS) < 10 above level 0, and
 * always at dummies of kind -1 (dummies of kind -2 are taken as cells of zero area); a cell
 * taken below the top level adds -s DS / (4 pi r); a top-level cell adds FSOUR_A4 (with its
 * own area) times the source of cell nv — the first top-level cell, as the reference passes
 * nv = 1 at the top. */
static void tree_calc_s(const orc_tree *tr, int nv, int nlev, const double XO[3], const panel_frame *f, double PI4,
                        double *S) {
    int list[4], nleaf_n, use_list;
    if (nlev == tr->levmax) { nleaf_n = tr->nlev[nlev]; use_list = 0; }
    else if (TL(nlev + 1, nv, 6) < 0) { list[0] = TL(nlev + 1, nv, 1); list[1] = TL(nlev + 1, nv, 2); nleaf_n = 2; use_list = 1; }
    else { for (int q = 0; q < 4; q++) list[q] = TL(nlev + 1, nv, q + 1); nleaf_n = 4; use_list = 1; }
    for (int q = 0; q < nleaf_n; q++) {
        int nleaf = use_list ? list[q] : q + 1;
        if (TL(nlev, nleaf, 6) == 0) { fprintf(stderr, "Something is wrong with the tree %d %d\n", nleaf, nlev); abort(); }
        double RG[3] = {tr->xs[nlev][nleaf], tr->ys[nlev][nleaf], tr->zs[nlev][nleaf]};
        double a = XO[0] - RG[0], b = XO[1] - RG[1], c3 = XO[2] - RG[2];
        double r = sqrt(a * a + b * b + c3 * c3);
        double DS = tr->dsl[nlev][nleaf];
        double RATIO = r / sqrt(DS);
        if ((RATIO < g_tree_factor && nlev > 0) || TL(nlev, nleaf, 6) == -1) {
            tree_calc_s(tr, nleaf, nlev - 1, XO, f, PI4, S);
        } else if (nlev != tr->levmax) {
            const double *sb = tr->src[nlev] + (size_t)tr->nc * nleaf;
            for (int c = 0; c < tr->nc; c++) S[c] = S[c] - sb[c] * DS / (PI4 * r);
        } else {
            double FIS = fsour_a4(XO, RG, f, DS, PI4);
            const double *sv = tr->src[nlev] + (size_t)tr->nc * nv;
            for (int c = 0; c < tr->nc; c++) S[c] = S[c] + sv[c] * FIS;
        }
    }
}
#undef TL

/* Bounds3d_lev -> calc_boundinf_3d_lev (pmbound.f90:980-1137), ibctyp=2: targets in the
 * order of the direct sums; SOL += SOURCE / (4 pi). */
void orc_boundsum_tree(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                       const double *src, const double *cxyz, const double *ds, double *sol, int levmax) {
    orc_tree tr;
    tree_build(&tr, levmax, nn, nc, nbound, src, cxyz, ds);
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    int NX = nn[0], NY = nn[1], NZ = nn[2];
    int64_t nt = 2LL * NY * NZ + 2LL * (NX - 2) * NZ + 2LL * (NX - 2) * (NY - 2);
    int32_t *tg = malloc(sizeof(int32_t) * 3 * nt);
    int64_t t = 0;
    for (int f = 0; f < 2; f++)
        for (int k = 1; k <= NZ; k++)
            for (int j = 1; j <= NY; j++) { tg[3*t] = f ? NX : 1; tg[3*t+1] = j; tg[3*t+2] = k; t++; }
    for (int f = 0; f < 2; f++)
        for (int k = 1; k <= NZ; k++)
            for (int i = 2; i <= NX - 1; i++) { tg[3*t] = i; tg[3*t+1] = f ? NY : 1; tg[3*t+2] = k; t++; }
    for (int f = 0; f < 2; f++)
        for (int j = 2; j <= NY - 1; j++)
            for (int i = 2; i <= NX - 1; i++) { tg[3*t] = i; tg[3*t+1] = j; tg[3*t+2] = f ? NZ : 1; t++; }
    double t0 = now();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nt; q++) {
        int i = tg[3 * q], j = tg[3 * q + 1], k = tg[3 * q + 2];
        double X = xb[0] + (double)(i - 1) * dpm[0], Y = xb[1] + (double)(j - 1) * dpm[1],
               Z = xb[2] + (double)(k - 1) * dpm[2];
        double S[8] = {0};
        tree_calc(&tr, 1, levmax, X, Y, Z, S);
        for (int c = 0; c < nc; c++) {
            size_t o = IX(c, i - 1, j - 1, k - 1, nc, NX, NY);
            sol[o] = sol[o] + S[c] / PI4;
        }
    }
    g_timer[2] += now() - t0;
    free(tg);
    tree_free(&tr);
}


/* Target nodes of Bounds3d in call order with their plane code (pmbound.f90:80-157):
 * XMIN (2), XMAX (-2), YMIN (1), YMAX (-1), ZMIN (3), ZMAX (-3). */
static int64_t bound_targets(const int32_t nn[3], int32_t **out) {
    int NX = nn[0], NY = nn[1], NZ = nn[2];
    int64_t nt = 2LL * NY * NZ + 2LL * (NX - 2) * NZ + 2LL * (NX - 2) * (NY - 2);
    int32_t *tg = malloc(sizeof(int32_t) * 4 * nt);
    int64_t t = 0;
    for (int f = 0; f < 2; f++)
        for (int k = 1; k <= NZ; k++)
            for (int j = 1; j <= NY; j++) { tg[4*t] = f ? NX : 1; tg[4*t+1] = j; tg[4*t+2] = k; tg[4*t+3] = f ? -2 : 2; t++; }
    for (int f = 0; f < 2; f++)
        for (int k = 1; k <= NZ; k++)
            for (int i = 2; i <= NX - 1; i++) { tg[4*t] = i; tg[4*t+1] = f ? NY : 1; tg[4*t+2] = k; tg[4*t+3] = f ? -1 : 1; t++; }
    for (int f = 0; f < 2; f++)
        for (int j = 2; j <= NY - 1; j++)
            for (int i = 2; i <= NX - 1; i++) { tg[4*t] = i; tg[4*t+1] = j; tg[4*t+2] = f ? NZ : 1; tg[4*t+3] = f ? -3 : 3; t++; }
    *out = tg;
    return nt;
}

/* Bounds3d -> calc_boundinf_3d_s (pmbound.f90:520-648), ibctyp = 3: SOL += FIS * sigma over
 * every source, FIS = FSOUR_A4 in the target plane's frame, sources in order. */
void orc_boundsum_s(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                    const double *src, const double *cxyz, double *sol) {
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    int NX = nn[0], NY = nn[1];
    int32_t *tg;
    int64_t nt = bound_targets(nn, &tg);
    double t0 = now();
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < nt; q++) {
        int i = tg[4 * q], j = tg[4 * q + 1], k = tg[4 * q + 2], pl = tg[4 * q + 3];
        panel_frame f;
        panel_frame_init(&f, pl, dpm);
        double XO[3] = {xb[0] + (double)(i - 1) * dpm[0], xb[1] + (double)(j - 1) * dpm[1], xb[2] + (double)(k - 1) * dpm[2]};
        double acc[8];
        for (int c = 0; c < nc; c++) acc[c] = sol[IX(c, i - 1, j - 1, k - 1, nc, NX, NY)];
        for (int64_t nv = 0; nv < nbound; nv++) {
            double FIS = fsour_a4(XO, &cxyz[3 * nv], &f, f.AREA, PI4);
            for (int c = 0; c < nc; c++) acc[c] = acc[c] + FIS * src[(size_t)nc * nv + c];
        }
        for (int c = 0; c < nc; c++) sol[IX(c, i - 1, j - 1, k - 1, nc, NX, NY)] = acc[c];
    }
    g_timer[2] += now() - t0;
    free(tg);
}

/* Bounds3d_lev -> calc_boundinf_3d_lev_s (pmbound.f90:1255-1345): SOL += SOURCE. */
void orc_boundsum_tree_s(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, int64_t nbound,
                         const double *src, const double *cxyz, const double *ds, double *sol, int levmax) {
    orc_tree tr;
    tree_build(&tr, levmax, nn, nc, nbound, src, cxyz, ds);
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    int NX = nn[0], NY = nn[1];
    int32_t *tg;
    int64_t nt = bound_targets(nn, &tg);
    double t0 = now();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t q = 0; q < nt; q++) {
        int i = tg[4 * q], j = tg[4 * q + 1], k = tg[4 * q + 2], pl = tg[4 * q + 3];
        panel_frame f;
        panel_frame_init(&f, pl, dpm);
        double XO[3] = {xb[0] + (double)(i - 1) * dpm[0], xb[1] + (double)(j - 1) * dpm[1], xb[2] + (double)(k - 1) * dpm[2]};
        double S[8] = {0};
        tree_calc_s(&tr, 1, levmax, XO, &f, PI4, S);
        for (int c = 0; c < nc; c++) {
            size_t o = IX(c, i - 1, j - 1, k - 1, nc, NX, NY);
            sol[o] = sol[o] + S[c];
        }
    }
    g_timer[2] += now() - t0;
    free(tg);
    tree_free(&tr);
}

/* ------------------------------------------------------------------ pmesh */
/* pmlib.f90:69-192 with iynbc=1, ibctyp=2, itree=0: infdomain_3d (pinfdomain.f90:86-158)
 * then solve_eq_3d on the whole grid.  sol accumulates the boundary sums. */
static int g_tree_levmax = -1; /* >= 0: boundary sums by the tree with this levmax */
void orc_set_tree(int levmax) { g_tree_levmax = levmax; }
static int g_ibctyp = 2;       /* 2: point sources (calc_boundinf_3d*), 3: panels (*_s) */
void orc_set_ibctyp(int ibctyp) { g_ibctyp = ibctyp; }

void orc_pmesh_inf(const double xb[6], const double dpm[3], const int32_t nn[3], int nc, const double *rhs,
                   double *sol) {
    size_t G = (size_t)nn[0] * nn[1] * nn[2];
    double *s0 = calloc(G * nc, sizeof(double));
    int32_t win[6] = {1, 1, 1, nn[0], nn[1], nn[2]};
    for (int c = 0; c < nc; c++) orc_solve_window(xb, dpm, nn, nc, c, win, rhs, NULL, 1, s0);
    int64_t maxb = 2LL * nn[0] * nn[1] + 2LL * nn[0] * nn[2] + 2LL * nn[1] * nn[2];
    double *src = malloc(sizeof(double) * nc * maxb), *cx = malloc(sizeof(double) * 3 * maxb),
           *ds = malloc(sizeof(double) * maxb);
    int64_t nb = orc_normalderiv(xb, dpm, nn, nc, s0, src, cx, ds);
    free(s0);
    if (g_ibctyp == 3) {
        if (g_tree_levmax >= 0) orc_boundsum_tree_s(xb, dpm, nn, nc, nb, src, cx, ds, sol, g_tree_levmax);
        else orc_boundsum_s(xb, dpm, nn, nc, nb, src, cx, sol);
    } else if (g_tree_levmax >= 0) orc_boundsum_tree(xb, dpm, nn, nc, nb, src, cx, ds, sol, g_tree_levmax);
    else orc_boundsum(xb, dpm, nn, nc, nb, src, cx, ds, sol);
    free(src); free(cx); free(ds);
    for (int c = 0; c < nc; c++) orc_solve_window(xb, dpm, nn, nc, c, win, rhs, sol, 0, NULL);
}

/* --------------------------------------------------------------------- yaps3d */
/* yaps3d.f90:1-453 for BLOCKS = NBI*NBJ*NBK blocks, the work of every rank in turn.  Block
 * b (0-based, b = (k-1)*NBI*NBJ + (j-1)*NBI + i-1) has nnbl[3b..] nodes; its arrays sit one
 * after another in rhs_bl / sol_bl (nc components, component fastest).
 *   1. each block's infinite-domain solve (yaps3d.f90:36-56);
 *   2. samples of every block on the coarse grid and its map_nodes / NN_coarse_map
 *      (:104-138);
 *   3. coarse RHS = sum over blocks, in block order, of the 7-point Laplacian of the
 *      block's samples on its footprint interior (:179-196, 1990-2016);
 *   4. the coarse infinite-domain solve (:283-299);
 *   5. per block, interp_stencil on its inner-box faces (:302-395, 1914-1983): TSC of the
 *      coarse solution minus the samples of the blocks whose footprint covers the stencil,
 *      plus those neighbours' fine solutions at the node (BBound of mapnodes_bl,
 *      :457-1912: a neighbour sends every node of the receiver's inner faces that lies in
 *      its own block grid; nodes it does not reach stay 0);
 *   6. per block, the Dirichlet solve of its inner box (:397-418). */
/* Planted mistakes for the tests that pin this composition (0: none): 1 drops the TSC
 * correction term, 2 flips its sign, 3 drops the neighbours' fine values (BBound), 4 takes
 * every block as covering the stencil (nnb = 0), 5 samples the coarse Laplacian one coarse
 * node too far (footprint boundary included). */
static int g_yaps_plant = 0;
void orc_set_yaps_plant(int mode) { g_yaps_plant = mode; }

int orc_yaps3d(const orc_params *p, const orc_geom *g, int nc, const double *rhs_bl, double *sol_bl) {
    const int NBL = g->nblocks;
    if (NBL < 1 || NBL > ORC_MAXB || g->nbi * g->nbj * g->nbk != NBL) return -1;
    size_t boff[ORC_MAXB + 1];
    boff[0] = 0;
    for (int b = 0; b < NBL; b++)
        boff[b + 1] = boff[b] + (size_t)g->nnbl[3 * b] * g->nnbl[3 * b + 1] * g->nnbl[3 * b + 2] * nc;
    memset(sol_bl, 0, sizeof(double) * boff[NBL]);
    const int saved_tree = g_tree_levmax, saved_bc = g_ibctyp;
    g_ibctyp = p->ibctyp == 3 ? 3 : 2;
    /* 1. fine infinite-domain solves (tree with lmax = ilevmax when iyntree = 1, :47-48) */
    if (p->iyntree == 1) g_tree_levmax = g->ilevmax;
    for (int b = 0; b < NBL; b++)
        orc_pmesh_inf(&g->xbound_bl[6 * b], g->dpm, &g->nnbl[3 * b], nc, rhs_bl + boff[b], sol_bl + boff[b]);
    /* 2. samples */
    int ncx = g->nn_coarse[0], ncy = g->nn_coarse[1], ncz = g->nn_coarse[2];
    size_t GC = (size_t)ncx * ncy * ncz;
    double *samp = calloc(GC * nc * NBL, sizeof(double));
    unsigned char *map = calloc(GC * NBL, 1);
    int cmap[6 * ORC_MAXB];
    for (int q = 0; q < 6 * NBL; q++) cmap[q] = -1;
    int ncs = p->ncoarse;
    for (int b = 0; b < NBL; b++) {
        const double *xbl = &g->xbound_bl[6 * b];
        const int32_t *nnbl = &g->nnbl[3 * b];
        for (int k = 1; k <= nnbl[2]; k += ncs)
            for (int j = 1; j <= nnbl[1]; j += ncs)
                for (int i = 1; i <= nnbl[0]; i += ncs) {
                    double X0 = xbl[0] + (double)(i - 1) * g->dpm[0];
                    double X1 = xbl[1] + (double)(j - 1) * g->dpm[1];
                    double X2 = xbl[2] + (double)(k - 1) * g->dpm[2];
                    int in = (int)lround((X0 - g->xbound_coarse[0]) / g->dpm_coarse[0]) + 1;
                    int jn = (int)lround((X1 - g->xbound_coarse[1]) / g->dpm_coarse[1]) + 1;
                    int kn = (int)lround((X2 - g->xbound_coarse[2]) / g->dpm_coarse[2]) + 1;
                    if (i == 1) cmap[6 * b + 0] = in;
                    if (i == nnbl[0]) cmap[6 * b + 3] = in;
                    if (j == 1) cmap[6 * b + 1] = jn;
                    if (j == nnbl[1]) cmap[6 * b + 4] = jn;
                    if (k == 1) cmap[6 * b + 2] = kn;
                    if (k == nnbl[2]) cmap[6 * b + 5] = kn;
                    if (in < 1 || in > ncx || jn < 1 || jn > ncy || kn < 1 || kn > ncz) continue;
                    for (int c = 0; c < nc; c++)
                        samp[GC * nc * b + IX(c, in - 1, jn - 1, kn - 1, nc, ncx, ncy)] =
                            sol_bl[boff[b] + IX(c, i - 1, j - 1, k - 1, nc, nnbl[0], nnbl[1])];
                    map[GC * b + (size_t)(in - 1) + (size_t)ncx * ((jn - 1) + (size_t)ncy * (kn - 1))] = 1;
                }
    }
    /* 3. coarse RHS */
    double *rhs_c = calloc(GC * nc, sizeof(double)), *sol_c = calloc(GC * nc, sizeof(double));
    const double *dc = g->dpm_coarse;
    for (int b = 0; b < NBL; b++) {
        const double *sb = samp + GC * nc * b;
        const int *cm = &cmap[6 * b];
#define SC(c, i, j, k) sb[IX(c, (i)-1, (j)-1, (k)-1, nc, ncx, ncy)]
        const int e = g_yaps_plant == 5 ? 0 : 1;
        for (int k = cm[2] + e; k <= cm[5] - e; k++)
            for (int j = cm[1] + e; j <= cm[4] - e; j++)
                for (int i = cm[0] + e; i <= cm[3] - e; i++)
                    for (int c = 0; c < nc; c++) {
                        double v = ((SC(c, i + 1, j, k) - 2 * SC(c, i, j, k)) + SC(c, i - 1, j, k)) / (dc[0] * dc[0]) +
                                   ((SC(c, i, j + 1, k) - 2 * SC(c, i, j, k)) + SC(c, i, j - 1, k)) / (dc[1] * dc[1]) +
                                   ((SC(c, i, j, k + 1) - 2 * SC(c, i, j, k)) + SC(c, i, j, k - 1)) / (dc[2] * dc[2]);
                        size_t o = IX(c, i - 1, j - 1, k - 1, nc, ncx, ncy);
                        rhs_c[o] = rhs_c[o] + v;
                    }
#undef SC
    }
    /* 4. coarse infinite-domain solve */
    double xbc[6];
    memcpy(xbc, g->xbound_coarse, sizeof xbc);
    if (p->iyntree == 1) g_tree_levmax = g->ilevmax - 1;  /* yaps3d.f90:288-294 */
    orc_pmesh_inf(xbc, g->dpm_coarse, g->nn_coarse, nc, rhs_c, sol_c);
    g_tree_levmax = saved_tree;
    g_ibctyp = saved_bc;
    /* 5. interp_stencil; neighbours' fine solutions as they are after step 1 */
    double *fine = malloc(sizeof(double) * boff[NBL]);
    memcpy(fine, sol_bl, sizeof(double) * boff[NBL]);
    const int NBI = g->nbi, NBJ = g->nbj, NBK = g->nbk;
    for (int b = 0; b < NBL; b++) {
        const double *xbl = &g->xbound_bl[6 * b];
        const int32_t *nnbl = &g->nnbl[3 * b], *nnbl_bl = &g->nnbl_bl[6 * b];
        const int nb_i = b % NBI + 1, nb_j = (b / NBI) % NBJ + 1, nb_k = b / (NBI * NBJ) + 1;
        double *sb = sol_bl + boff[b];
        int NXs = nnbl_bl[0], NYs = nnbl_bl[1], NZs = nnbl_bl[2], NXf = nnbl_bl[3], NYf = nnbl_bl[4], NZf = nnbl_bl[5];
        int64_t nface = 2LL * (NYf - NYs + 1) * (NZf - NZs + 1) + 2LL * (NXf - NXs - 1) * (NZf - NZs + 1) +
                        2LL * (NXf - NXs - 1) * (NYf - NYs - 1);
        int32_t *fl = malloc(sizeof(int32_t) * 3 * nface);
        int64_t q = 0;
        for (int f = 0; f < 2; f++)
            for (int k = NZs; k <= NZf; k++)
                for (int j = NYs; j <= NYf; j++) { fl[3*q] = f ? NXf : NXs; fl[3*q+1] = j; fl[3*q+2] = k; q++; }
        for (int f = 0; f < 2; f++)
            for (int k = NZs; k <= NZf; k++)
                for (int i = NXs + 1; i <= NXf - 1; i++) { fl[3*q] = i; fl[3*q+1] = f ? NYf : NYs; fl[3*q+2] = k; q++; }
        for (int f = 0; f < 2; f++)
            for (int j = NYs + 1; j <= NYf - 1; j++)
                for (int i = NXs + 1; i <= NXf - 1; i++) { fl[3*q] = i; fl[3*q+1] = j; fl[3*q+2] = f ? NZf : NZs; q++; }
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < nface; t++) {
            int i = fl[3 * t], j = fl[3 * t + 1], k = fl[3 * t + 2];
            double X0 = xbl[0] + (double)(i - 1) * g->dpm[0];
            double X1 = xbl[1] + (double)(j - 1) * g->dpm[1];
            double X2 = xbl[2] + (double)(k - 1) * g->dpm[2];
            int in = (int)lround((X0 - g->xbound_coarse[0]) / dc[0]) + 1;
            int jn = (int)lround((X1 - g->xbound_coarse[1]) / dc[1]) + 1;
            int kn = (int)lround((X2 - g->xbound_coarse[2]) / dc[2]) + 1;
            int nnb[ORC_MAXB];
            for (int bc = 0; bc < NBL; bc++) {
                const unsigned char *mb = map + GC * bc;
                nnb[bc] = (!mb[(size_t)(in - 2) + (size_t)ncx * ((jn - 2) + (size_t)ncy * (kn - 2))] ||
                           !mb[(size_t)(in) + (size_t)ncx * ((jn) + (size_t)ncy * (kn))]) ? 1 : 0;
                if (g_yaps_plant == 4) nnb[bc] = 0;
            }
            size_t o = IX(0, i - 1, j - 1, k - 1, nc, nnbl[0], nnbl[1]);
            for (int kc = kn - 1; kc <= kn + 1; kc++)
                for (int jc = jn - 1; jc <= jn + 1; jc++)
                    for (int ic = in - 1; ic <= in + 1; ic++) {
                        size_t oc = IX(0, ic - 1, jc - 1, kc - 1, nc, ncx, ncy);
                        double add[8] = {0}, add_sample[8] = {0};
                        for (int bc = 0; bc < NBL; bc++)
                            for (int c = 0; c < nc; c++) {
                                double smp = samp[GC * nc * bc + oc + c];
                                add_sample[c] = smp + add_sample[c];
                                if (nnb[bc]) add[c] = smp + add[c];
                            }
                        double xc = (X0 - g->xbound_coarse[0] - (double)(ic - 1) * dc[0]) / dc[0];
                        double yc = (X1 - g->xbound_coarse[1] - (double)(jc - 1) * dc[1]) / dc[1];
                        double zc = (X2 - g->xbound_coarse[2] - (double)(kc - 1) * dc[2]) / dc[2];
                        double fw = orc_projection_fun(3, xc) * orc_projection_fun(3, yc) * orc_projection_fun(3, zc);
                        const double sgn = g_yaps_plant == 2 ? -1.0 : 1.0;
                        if (g_yaps_plant != 1)
                            for (int c = 0; c < nc; c++)
                                sb[o + c] = sb[o + c] + sgn * (fw * ((sol_c[oc + c] - add_sample[c]) + add[c]));
                    }
            for (int k_nb = nb_k - 1; k_nb <= nb_k + 1; k_nb++)
                for (int j_nb = nb_j - 1; j_nb <= nb_j + 1; j_nb++)
                    for (int i_nb = nb_i - 1; i_nb <= nb_i + 1; i_nb++) {
                        if (i_nb < 1 || i_nb > NBI || j_nb < 1 || j_nb > NBJ || k_nb < 1 || k_nb > NBK) continue;
                        int bc = (k_nb - 1) * NBI * NBJ + (j_nb - 1) * NBI + i_nb - 1;
                        if (nnb[bc] || bc == b || g_yaps_plant == 3) continue;
                        /* the node in block bc's grid (all blocks share the lattice) */
                        const double *xc = &g->xbound_bl[6 * bc];
                        const int32_t *nc3 = &g->nnbl[3 * bc];
                        int ii = i + (int)lround((xbl[0] - xc[0]) / g->dpm[0]);
                        int jj = j + (int)lround((xbl[1] - xc[1]) / g->dpm[1]);
                        int kk = k + (int)lround((xbl[2] - xc[2]) / g->dpm[2]);
                        if (ii < 1 || ii > nc3[0] || jj < 1 || jj > nc3[1] || kk < 1 || kk > nc3[2]) continue;
                        const double *fb = fine + boff[bc];
                        size_t on = IX(0, ii - 1, jj - 1, kk - 1, nc, nc3[0], nc3[1]);
                        for (int c = 0; c < nc; c++) sb[o + c] = sb[o + c] + fb[on + c];
                    }
        }
        free(fl);
    }
    free(fine); free(samp); free(map); free(rhs_c); free(sol_c);
    /* 6. final Dirichlet solves on the inner boxes */
    for (int b = 0; b < NBL; b++)
        for (int c = 0; c < nc; c++)
            orc_solve_window(&g->xbound_bl[6 * b], g->dpm, &g->nnbl[3 * b], nc, c, &g->nnbl_bl[6 * b], rhs_bl + boff[b],
                             sol_bl + boff[b], 0, NULL);
    return 0;
}

/* ------------------------------------------------------- velocity, stretching */
/* calc_velocity_serial_3d (vpm_gcalc.f90:1-136). vel grids are (NX,NY,NZ), 1-component. */
void orc_velocity(const orc_geom *g, int neqpm, const double *rhs, double *sol, double *vx, double *vy,
                  double *vz, int idcalc) {
    int NX = g->nn[0], NY = g->nn[1], NZ = g->nn[2];
    int xs = g->nn_bl[0], ys = g->nn_bl[1], zs = g->nn_bl[2], xf = g->nn_bl[3], yf = g->nn_bl[4], zf = g->nn_bl[5];
    double DX2 = 2 * g->dpm[0], DY2 = 2 * g->dpm[1], DZ2 = 2 * g->dpm[2];
    int nr = neqpm + 1;
#define SL(c, i, j, k) sol[IX(c, (i)-1, (j)-1, (k)-1, neqpm, NX, NY)]
#define RH(c, i, j, k) rhs[IX(c, (i)-1, (j)-1, (k)-1, nr, NX, NY)]
#define V(a, i, j, k) a[IX(0, (i)-1, (j)-1, (k)-1, 1, NX, NY)]
    if (idcalc >= 0) {
#pragma omp parallel for schedule(static)
        for (int k = zs + 1; k <= zf - 1; k++)
            for (int j = ys + 1; j <= yf - 1; j++)
                for (int i = xs + 1; i <= xf - 1; i++) {
                    double dx[3], dy[3], dz[3];
                    for (int c = 0; c < 3; c++) {
                        dx[c] = (SL(c, i + 1, j, k) - SL(c, i - 1, j, k)) / DX2;
                        dy[c] = (SL(c, i, j + 1, k) - SL(c, i, j - 1, k)) / DY2;
                        dz[c] = (SL(c, i, j, k + 1) - SL(c, i, j, k - 1)) / DZ2;
                    }
                    V(vx, i, j, k) = V(vx, i, j, k) + dy[2] - dz[1];
                    V(vy, i, j, k) = V(vy, i, j, k) - (dx[2] - dz[0]);
                    V(vz, i, j, k) = V(vz, i, j, k) + dx[1] - dy[0];
                }
        if (neqpm == 4) {
#pragma omp parallel for schedule(static)
            for (int k = zs + 1; k <= zf - 1; k++)
                for (int j = ys + 1; j <= yf - 1; j++)
                    for (int i = xs + 1; i <= xf - 1; i++) {
                        V(vx, i, j, k) = V(vx, i, j, k) + (SL(3, i + 1, j, k) - SL(3, i - 1, j, k)) / DX2;
                        V(vy, i, j, k) = V(vy, i, j, k) + (SL(3, i, j + 1, k) - SL(3, i, j - 1, k)) / DY2;
                        V(vz, i, j, k) = V(vz, i, j, k) + (SL(3, i, j, k + 1) - SL(3, i, j, k - 1)) / DZ2;
                    }
        }
    }
    if (idcalc == 0) return;
    memset(sol, 0, sizeof(double) * (size_t)neqpm * NX * NY * NZ);
#pragma omp parallel for schedule(static)
    for (int k = zs + 2; k <= zf - 2; k++)
        for (int j = ys + 2; j <= yf - 2; j++)
            for (int i = xs + 2; i <= xf - 2; i++) {
                const double *comp[3] = {vx, vy, vz};
                for (int m = 0; m < 3; m++) {
                    const double *u = comp[m];
                    double pi = V(u, i + 1, j, k), pj = V(u, i, j + 1, k), pk = V(u, i, j, k + 1);
                    double mi = V(u, i - 1, j, k), mj = V(u, i, j - 1, k), mk = V(u, i, j, k - 1);
                    double a = -(RH(0, i + 1, j, k) * pi - (RH(0, i - 1, j, k)) * mi) / DX2;
                    double b = -(RH(1, i, j + 1, k) * pj - (RH(1, i, j - 1, k)) * mj) / DY2;
                    double c = -(RH(2, i, j, k + 1) * pk - (RH(2, i, j, k - 1)) * mk) / DZ2;
                    SL(m, i, j, k) = a + b + c;
                }
            }
#undef SL
#undef RH
#undef V
}

/* ------------------------------------------------------------- interpolation */
/* back_to_particles_3D itype=1 (vpm_time.f90:30-107); UP overwritten (particles_gath copies
 * the scattered UP, which starts from zero). Nodes outside the grid are skipped. */
void orc_m2p(const orc_geom *g, int neqpm, int iproj, const double *sol, const double *vx, const double *vy,
             const double *vz, const double *xp, const double *qp, double *up, double *gp, int64_t nvr) {
    int ips = iproj == 2 ? 0 : 1, ipf = iproj == 2 ? 1 : 2;
    int NX = g->nn[0], NY = g->nn[1], NZ = g->nn[2];
    const double *Xb = g->xbound, *D = g->dpm;
#pragma omp parallel for schedule(static)
    for (int64_t nv = 0; nv < nvr; nv++) {
        const double *x = xp + 3 * nv;
        int in = (int)((x[0] - Xb[0]) / D[0]) + 1, jn = (int)((x[1] - Xb[1]) / D[1]) + 1,
            kn = (int)((x[2] - Xb[2]) / D[2]) + 1;
        double u[3] = {0, 0, 0}, gl[3] = {0, 0, 0};
        for (int k = kn - ips; k <= kn + ipf; k++)
            for (int j = jn - ips; j <= jn + ipf; j++)
                for (int i = in - ips; i <= in + ipf; i++) {
                    double fx = orc_projection_fun(iproj, (x[0] - Xb[0] - (double)(i - 1) * D[0]) / D[0]);
                    double fy = orc_projection_fun(iproj, (x[1] - Xb[1] - (double)(j - 1) * D[1]) / D[1]);
                    double fz = orc_projection_fun(iproj, (x[2] - Xb[2] - (double)(k - 1) * D[2]) / D[2]);
                    double f = fx * fy * fz;
                    if (i < 1 || i > NX || j < 1 || j > NY || k < 1 || k > NZ) continue;
                    size_t o = IX(0, i - 1, j - 1, k - 1, 1, NX, NY);
                    u[0] = u[0] + f * vx[o];
                    u[1] = u[1] + f * vy[o];
                    u[2] = u[2] + f * vz[o];
                    for (int c = 0; c < 3; c++) gl[c] = gl[c] + f * sol[(size_t)neqpm * o + c];
                }
        for (int c = 0; c < 3; c++) {
            up[3 * nv + c] = u[c];
            gp[3 * nv + c] = gl[c] * qp[(size_t)(neqpm + 1) * nv + neqpm];
        }
    }
}

/* ------------------------------------------------------------------ projection */
/* project_particles_3D (pmproject.f90:43-98) + project_vol3d (pmproject.f90:110-153).
 * rhs has ncomp = neq+1 components; the last one carries the volume. */
int orc_project(const double xb[6], const double dpm[3], const int32_t nn[3], const int32_t nn_bl[6],
                const double *xp, const double *qp, int64_t nvr, int ncomp, int itype, int idvpm, double *rhs) {
    int NX = nn[0], NY = nn[1], NZ = nn[2];
    int ips = itype == 2 ? 0 : 1, ipf = itype == 2 ? 1 : 2;
    size_t G = (size_t)NX * NY * NZ;
    memset(rhs, 0, sizeof(double) * G * ncomp);
    int64_t outside = 0;
    double t0 = now();
    for (int64_t nv = 0; nv < nvr; nv++) {
        const double *x = xp + 3 * nv;
        const double *q = qp + (size_t)ncomp * nv;
        int in = (int)((x[0] - xb[0]) / dpm[0]) + 1, jn = (int)((x[1] - xb[1]) / dpm[1]) + 1,
            kn = (int)((x[2] - xb[2]) / dpm[2]) + 1;
        for (int k = kn - ips; k <= kn + ipf; k++)
            for (int j = jn - ips; j <= jn + ipf; j++)
                for (int i = in - ips; i <= in + ipf; i++) {
                    double fx = orc_projection_fun(itype, (x[0] - xb[0] - (double)(i - 1) * dpm[0]) / dpm[0]);
                    double fy = orc_projection_fun(itype, (x[1] - xb[1] - (double)(j - 1) * dpm[1]) / dpm[1]);
                    double fz = orc_projection_fun(itype, (x[2] - xb[2] - (double)(k - 1) * dpm[2]) / dpm[2]);
                    double f = fx * fy * fz;
                    if (i < 1 || i > NX || j < 1 || j > NY || k < 1 || k > NZ) { outside++; continue; }
                    size_t o = IX(0, i - 1, j - 1, k - 1, ncomp, NX, NY);
                    for (int c = 0; c < ncomp; c++) rhs[o + c] = rhs[o + c] + f * q[c];
                }
    }
    g_timer[0] += now() - t0;
    double DV = dpm[0] * dpm[1] * dpm[2];
    if (idvpm == 1) {
        for (size_t n = 0; n < G; n++) {
            for (int c = 0; c < ncomp - 1; c++) rhs[n * ncomp + c] = rhs[n * ncomp + c] / DV;
            rhs[n * ncomp + ncomp - 1] = DV;
        }
    } else if (idvpm == 0) {
        int NXs = nn_bl[0], NYs = nn_bl[1], NZs = nn_bl[2], NXf = nn_bl[3], NYf = nn_bl[4], NZf = nn_bl[5];
        for (int k = 1; k <= NZ; k++)
            for (int j = 1; j <= NY; j++)
                for (int i = 1; i <= NX; i++) {
                    size_t o = IX(0, i - 1, j - 1, k - 1, ncomp, NX, NY);
                    /* the reference tests j.ge.NYf here (pmproject.f90:125) */
                    if (i < NXs || i > NXf || j < NYs || j >= NYf || k < NZs || k > NZf) {
                        for (int c = 0; c < ncomp - 1; c++) rhs[o + c] = 0.0;
                        rhs[o + ncomp - 1] = DV;
                        continue;
                    }
                    for (int c = 0; c < ncomp - 1; c++) rhs[o + c] = rhs[o + c] / rhs[o + ncomp - 1];
                }
    } else {
        return -1; /* 'WRONG IDVPM' */
    }
    return (int)(outside > 0);
}

/* ------------------------------------------------------------------- vpm mode 2 */
/* rhsscat_3d (vpm_mpi.f90:199-238) for every block: the block's inner box takes the global
 * window ixs = (nb_i-1)(NXf-NXs) + NN_bl(1) ..., the rest of the block is 0, and a block that
 * is not first along an axis zeroes its min face on that axis. */
/* First global node of block b's inner box: vpm_mpi.f90:199-238 has (nb_i-1)*(NXf-NXs) +
 * NN_bl(1) for its even blocks; with uneven blocks, the inner extents of the blocks before
 * b along each axis summed. */
static void inner_origin(const orc_geom *g, int b, int *ixs, int *jxs, int *kxs) {
    int nb_i = b % g->nbi + 1, nb_j = (b / g->nbi) % g->nbj + 1, nb_k = b / (g->nbi * g->nbj) + 1;
    *ixs = g->nn_bl[0]; *jxs = g->nn_bl[1]; *kxs = g->nn_bl[2];
    for (int i = 1; i < nb_i; i++) {
        const int32_t *o = &g->nnbl_bl[6 * ((nb_k - 1) * g->nbi * g->nbj + (nb_j - 1) * g->nbi + i - 1)];
        *ixs += o[3] - o[0];
    }
    for (int j = 1; j < nb_j; j++) {
        const int32_t *o = &g->nnbl_bl[6 * ((nb_k - 1) * g->nbi * g->nbj + (j - 1) * g->nbi + nb_i - 1)];
        *jxs += o[4] - o[1];
    }
    for (int k = 1; k < nb_k; k++) {
        const int32_t *o = &g->nnbl_bl[6 * ((k - 1) * g->nbi * g->nbj + (nb_j - 1) * g->nbi + nb_i - 1)];
        *kxs += o[5] - o[2];
    }
}

static void rhsscat(const orc_geom *g, int neqpm, const double *rhs, int nr, double *rhs_bl, size_t *boff) {
    int NX = g->nn[0], NY = g->nn[1];
    boff[0] = 0;
    for (int b = 0; b < g->nblocks; b++)
        boff[b + 1] = boff[b] + (size_t)g->nnbl[3 * b] * g->nnbl[3 * b + 1] * g->nnbl[3 * b + 2] * neqpm;
    memset(rhs_bl, 0, sizeof(double) * boff[g->nblocks]);
    for (int b = 0; b < g->nblocks; b++) {
        const int32_t *nnbl = &g->nnbl[3 * b], *bl = &g->nnbl_bl[6 * b];
        int nb_i = b % g->nbi + 1, nb_j = (b / g->nbi) % g->nbj + 1, nb_k = b / (g->nbi * g->nbj) + 1;
        int ixs, jxs, kxs;
        inner_origin(g, b, &ixs, &jxs, &kxs);
        double *rb = rhs_bl + boff[b];
        for (int k = bl[2]; k <= bl[5]; k++)
            for (int j = bl[1]; j <= bl[4]; j++)
                for (int i = bl[0]; i <= bl[3]; i++) {
                    int zero = (nb_i > 1 && i == bl[0]) || (nb_j > 1 && j == bl[1]) || (nb_k > 1 && k == bl[2]);
                    for (int c = 0; c < neqpm; c++)
                        rb[IX(c, i - 1, j - 1, k - 1, neqpm, nnbl[0], nnbl[1])] =
                            zero ? 0.0 : rhs[IX(c, ixs + (i - bl[0]) - 1, jxs + (j - bl[1]) - 1, kxs + (k - bl[2]) - 1, nr, NX, NY)];
                }
    }
}

/* solget_3d (vpm_mpi.f90:240-328): inner boxes into the global grid in block order (a later
 * block overwrites the faces it shares with an earlier one). */
static void solget(const orc_geom *g, int neqpm, const double *sol_bl, const size_t *boff, double *sol) {
    int NX = g->nn[0], NY = g->nn[1];
    for (int b = 0; b < g->nblocks; b++) {
        const int32_t *nnbl = &g->nnbl[3 * b], *bl = &g->nnbl_bl[6 * b];
        int ixs, jxs, kxs;
        inner_origin(g, b, &ixs, &jxs, &kxs);
        const double *sb = sol_bl + boff[b];
        for (int k = bl[2]; k <= bl[5]; k++)
            for (int j = bl[1]; j <= bl[4]; j++)
                for (int i = bl[0]; i <= bl[3]; i++)
                    for (int c = 0; c < neqpm; c++)
                        sol[IX(c, ixs + (i - bl[0]) - 1, jxs + (j - bl[1]) - 1, kxs + (k - bl[2]) - 1, neqpm, NX, NY)] =
                            sb[IX(c, i - 1, j - 1, k - 1, neqpm, nnbl[0], nnbl[1])];
    }
}

/* vpm.f90:171-313 (WhatToDo=2), every yaps block in turn. */
int orc_vpm2(const orc_params *p, const orc_geom *g, int neqpm, const double *xp, const double *qp, double *up,
             double *gp, int64_t nvr, double *rhs_out, double *psi_out, double *sol_out, double *vel_out) {
    if (nvr == 0) return 0;
    if (g->nblocks < 1 || neqpm != 3) return -1;
    double t0 = now();
    int NX = g->nn[0], NY = g->nn[1], NZ = g->nn[2];
    size_t G = (size_t)NX * NY * NZ;
    int nr = neqpm + 1;
    double *rhs = calloc(G * nr, sizeof(double));
    double *sol = calloc(G * neqpm, sizeof(double));
    orc_project(g->xbound, g->dpm, g->nn, g->nn_bl, xp, qp, nvr, nr, p->interf_iproj, p->idvpm, rhs);
    size_t boff[ORC_MAXB + 1], tot = 0;
    for (int b = 0; b < g->nblocks; b++) tot += (size_t)g->nnbl[3 * b] * g->nnbl[3 * b + 1] * g->nnbl[3 * b + 2] * neqpm;
    double *rhs_bl = calloc(tot, sizeof(double)), *sol_bl = calloc(tot, sizeof(double));
    rhsscat(g, neqpm, rhs, nr, rhs_bl, boff);
    if (orc_yaps3d(p, g, neqpm, rhs_bl, sol_bl) != 0) { free(rhs_bl); free(sol_bl); free(rhs); free(sol); return -1; }
    solget(g, neqpm, sol_bl, boff, sol);
    free(rhs_bl); free(sol_bl);
    if (psi_out) memcpy(psi_out, sol, sizeof(double) * G * neqpm);
    double *vel = calloc(3 * G, sizeof(double));
    orc_velocity(g, neqpm, rhs, sol, vel, vel + G, vel + 2 * G, 1);
    orc_m2p(g, neqpm, p->interf_iproj, sol, vel, vel + G, vel + 2 * G, xp, qp, up, gp, nvr);
    if (rhs_out) memcpy(rhs_out, rhs, sizeof(double) * G * nr);
    if (sol_out) memcpy(sol_out, sol, sizeof(double) * G * neqpm);
    if (vel_out) memcpy(vel_out, vel, sizeof(double) * 3 * G);
    free(vel); free(rhs); free(sol);
    g_timer[3] += now() - t0;
    return 0;
}

/* ------------------------------------------------------------- modes 4 and 5 */
/* diffuse_vort_3d (vpm_gcalc.f90:138-210), neqpm = 3: SOL(1:3) = -NI lap(RHS(1:3)) on the
 * nodes NXs_bl+1..NXf_bl-1 (and y, z), 0 elsewhere. */
void orc_diffuse(const orc_geom *g, int neqpm, double ni, const double *rhs, double *sol) {
    int NX = g->nn[0], NY = g->nn[1], NZ = g->nn[2];
    int xs = g->nn_bl[0], ys = g->nn_bl[1], zs = g->nn_bl[2], xf = g->nn_bl[3], yf = g->nn_bl[4], zf = g->nn_bl[5];
    double DX2 = g->dpm[0] * g->dpm[0], DY2 = g->dpm[1] * g->dpm[1], DZ2 = g->dpm[2] * g->dpm[2];
    int nr = neqpm + 1;
    memset(sol, 0, sizeof(double) * (size_t)neqpm * NX * NY * NZ);
#define SL(c, i, j, k) sol[IX(c, (i)-1, (j)-1, (k)-1, neqpm, NX, NY)]
#define RH(c, i, j, k) rhs[IX(c, (i)-1, (j)-1, (k)-1, nr, NX, NY)]
#pragma omp parallel for schedule(static)
    for (int k = zs + 1; k <= zf - 1; k++)
        for (int j = ys + 1; j <= yf - 1; j++)
            for (int i = xs + 1; i <= xf - 1; i++) {
                double VIS = ni;  /* neqpm = 3 */
                for (int c = 0; c < 3; c++) {
                    double dwxdx = (RH(c, i + 1, j, k) - 2 * RH(c, i, j, k) + RH(c, i - 1, j, k)) / DX2;
                    double dwydy = (RH(c, i, j + 1, k) - 2 * RH(c, i, j, k) + RH(c, i, j - 1, k)) / DY2;
                    double dwzdz = (RH(c, i, j, k + 1) - 2 * RH(c, i, j, k) + RH(c, i, j, k - 1)) / DZ2;
                    SL(c, i, j, k) = -VIS * (dwxdx + dwydy + dwzdz);
                }
            }
#undef SL
#undef RH
}

/* vpm WhatToDo 4 (vpm.f90:316-343) and 5 (vpm.f90:345-373): projection, then for 4 the
 * stretching term from the velocity grids vel (3 x G, from an earlier call) and
 * back_to_particles itype 1; for 5 the diffusion term and back_to_particles itype 2
 * (vpm_time.f90:109-151: UP = 0, GP = sum W SOL * QP(4)).  sol_out: the SOL grid. */
int orc_vpm45(const orc_params *p, const orc_geom *g, int neqpm, int mode, double ni, const double *xp,
              const double *qp, double *up, double *gp, int64_t nvr, const double *vel, double *sol_out) {
    if (nvr == 0) return 0;
    if (neqpm != 3 || (mode != 4 && mode != 5)) return -1;
    int NX = g->nn[0], NY = g->nn[1], NZ = g->nn[2];
    size_t G = (size_t)NX * NY * NZ;
    int nr = neqpm + 1;
    double *rhs = calloc(G * nr, sizeof(double));
    double *sol = calloc(G * neqpm, sizeof(double));
    orc_project(g->xbound, g->dpm, g->nn, g->nn_bl, xp, qp, nvr, nr, p->interf_iproj, p->idvpm, rhs);
    if (mode == 4) {
        double *v = calloc(3 * G, sizeof(double));
        if (vel) memcpy(v, vel, sizeof(double) * 3 * G);
        orc_velocity(g, neqpm, rhs, sol, v, v + G, v + 2 * G, -1);
        orc_m2p(g, neqpm, p->interf_iproj, sol, v, v + G, v + 2 * G, xp, qp, up, gp, nvr);
        free(v);
    } else {
        double *z = calloc(3 * G, sizeof(double));
        orc_diffuse(g, neqpm, ni, rhs, sol);
        orc_m2p(g, neqpm, p->interf_iproj, sol, z, z + G, z + 2 * G, xp, qp, up, gp, nvr);
        for (int64_t nv = 0; nv < nvr; nv++)
            for (int c = 0; c < 3; c++) up[3 * nv + c] = 0.0;
        free(z);
    }
    if (sol_out) memcpy(sol_out, sol, sizeof(double) * G * neqpm);
    free(rhs); free(sol);
    return 0;
}

/* ------------------------------------------------------------------- remesh */
/* remesh_particles_3d(iflag=1), ncell_rem=1 (vpm_remesh.f90:5-225). */
int64_t orc_remesh(const orc_params *p, const orc_geom *g, int neqpm, int mrem, const double *xp, const double *qp,
                   int64_t nvr, double *xp_out, double *qp_out) {
    if (p->ncell_rem != 1) return -1;
    double t0 = now();
    int32_t NN[3], NNb[6];
    double D[3];
    for (int d = 0; d < 3; d++) NN[d] = g->nn[d] * mrem;
    for (int d = 0; d < 6; d++) NNb[d] = g->nn_bl[d] * mrem;
    for (int d = 0; d < 3; d++) D[d] = (g->xbound[d + 3] - g->xbound[d]) / (double)(NN[d] - 1);
    double DV = D[0] * D[1] * D[2];
    int nr = neqpm + 1;
    size_t G = (size_t)NN[0] * NN[1] * NN[2];
    double *rhs = malloc(sizeof(double) * G * nr);
    orc_project(g->xbound, D, NN, NNb, xp, qp, nvr, nr, p->interf_iproj, 2 /* no vol yet */, rhs);
    for (size_t n = 0; n < G; n++) rhs[n * nr + nr - 1] = DV; /* RHS_pm(neqpm+1,:,:,:)=DVpm */
    /* project_vol3d(RHS_pm, ..., IDVPM) */
    {
        int NX = NN[0], NY = NN[1], NZ = NN[2];
        if (p->idvpm == 1) {
            for (size_t n = 0; n < G; n++) {
                for (int c = 0; c < nr - 1; c++) rhs[n * nr + c] = rhs[n * nr + c] / DV;
                rhs[n * nr + nr - 1] = DV;
            }
        } else {
            for (int k = 1; k <= NZ; k++)
                for (int j = 1; j <= NY; j++)
                    for (int i = 1; i <= NX; i++) {
                        size_t o = IX(0, i - 1, j - 1, k - 1, nr, NX, NY);
                        if (i < NNb[0] || i > NNb[3] || j < NNb[1] || j >= NNb[4] || k < NNb[2] || k > NNb[5]) {
                            for (int c = 0; c < nr - 1; c++) rhs[o + c] = 0.0;
                            rhs[o + nr - 1] = DV;
                            continue;
                        }
                        for (int c = 0; c < nr - 1; c++) rhs[o + c] = rhs[o + c] / rhs[o + nr - 1];
                    }
        }
    }
    int h = p->interf_iproj / 2;
    int xs = NNb[0] + h, xf = NNb[3] - h, ys = NNb[1] + h, yf = NNb[4] - h, zs = NNb[2] + h, zf = NNb[5] - h;
    const double thr = (double)1e-09f; /* default-REAL literal, vpm_remesh.f90:184 */
    int64_t np = 0;
    for (int k = zs; k <= zf; k++)
        for (int j = ys; j <= yf; j++)
            for (int i = xs; i <= xf; i++) {
                size_t o = IX(0, i - 1, j - 1, k - 1, nr, NN[0], NN[1]);
                double wmag = sqrt(rhs[o] * rhs[o] + rhs[o + 1] * rhs[o + 1] + rhs[o + 2] * rhs[o + 2]);
                if (wmag < thr) continue;
                if (xp_out) {
                    xp_out[3 * np + 0] = g->xbound[0] + D[0] * (double)(i - 1);
                    xp_out[3 * np + 1] = g->xbound[1] + D[1] * (double)(j - 1);
                    xp_out[3 * np + 2] = g->xbound[2] + D[2] * (double)(k - 1);
                    for (int c = 0; c < neqpm; c++) qp_out[(size_t)nr * np + c] = rhs[o + c] * DV;
                    qp_out[(size_t)nr * np + neqpm] = DV;
                }
                np++;
            }
    free(rhs);
    g_timer[3] += now() - t0;
    return np;
}

void orc_diagnostics(int neqpm, const double *xp, const double *qp, int64_t nvr, double circ[3], double imp[3]) {
    int nr = neqpm + 1;
    double c0 = 0, c1 = 0, c2 = 0, i0 = 0, i1 = 0, i2 = 0;
    for (int64_t n = 0; n < nvr; n++) {
        const double *x = xp + 3 * n, *q = qp + (size_t)nr * n;
        double w0 = -q[0], w1 = -q[1], w2 = -q[2];
        c0 += w0; c1 += w1; c2 += w2;
        i0 += x[1] * w2 - x[2] * w1;
        i1 += x[2] * w0 - x[0] * w2;
        i2 += x[0] * w1 - x[1] * w0;
    }
    circ[0] = c0; circ[1] = c1; circ[2] = c2;
    imp[0] = 0.5 * i0; imp[1] = 0.5 * i1; imp[2] = 0.5 * i2;
}
