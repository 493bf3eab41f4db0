// Panel-source boundary sums (pm.inp ibctyp = 3), direct form:
// Bounds3d -> calc_boundinf_3d_s (pmbound.f90:80-157, 520-648).
//
//   SOL(t) += sum_s FIS(t, s) sigma(s),  FIS = FSOUR_A4 in the frame of t's plane.
//
// FSOUR_A4 is the point-source term -AREA_t / (4 pi RO) beyond RO = 10 DIAG_t and the exact
// panel integral inside.  So the sum is evaluated as
//   AREA_t * sum_s (-sigma(s) / 4 pi) / r(t, s)        (every pair: the point-source forms)
//   + sum_{s: RO <= 10 DIAG_t} (panel(t, s) + AREA_t / (4 pi RO)) sigma(s)
// The first term runs through the same spectral / lattice / direct evaluation as ibctyp = 2
// (sources -sigma / 4 pi) into a scratch grid; k_panel_near adds AREA_t times it and the
// near-field corrections, enumerating for each target only the face cells whose centres
// can lie within 10 DIAG_t.  The near test RO / DIAG > 10 is the reference's, bit for bit.
#include "bnodes.h"
#include "panel.h"
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

namespace {

// Source face f: cells along axes u (A per row) and v (B rows), plane normal n.
struct PanelFace {
    int n, u, v, A, B;
    double plane;
    int64_t off;
};
struct PanelFaces {
    PanelFace f[6];
};

__global__ void k_panel_strengths(int64_t nb, const double *__restrict__ sig, double *__restrict__ a) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    for (int c = 0; c < 3; c++) a[c * nb + s] = -sig[c * nb + s] / PI4;
}

__global__ void k_zero_targets(Grid g, int64_t nt, GView v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i, j, k, pl;
    target_node(g, t, &i, &j, &k, &pl);
    const int64_t o = v.at(i - 1, j - 1, k - 1);
    for (int c = 0; c < 3; c++) v.p[c * v.cs + o] = 0.0;
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

__global__ void __launch_bounds__(256) k_panel_near(Grid g, int64_t nt, PanelFaces pf, int64_t nb,
                                                    const double *__restrict__ sig, const double *__restrict__ cxyz,
                                                    GView farv, GView sol) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    int i, j, k, pl;
    target_node(g, t, &i, &j, &k, &pl);
    const int iplane = pl == 2 ? (i == g.n[0] ? -2 : 2) : (pl == 1 ? (j == g.n[1] ? -1 : 1) : (k == g.n[2] ? -3 : 3));
    const PanelFrame fr = panel_frame(iplane, g.h);
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    const double XO[3] = {g.x0[0] + (double)(i - 1) * g.h[0], g.x0[1] + (double)(j - 1) * g.h[1],
                          g.x0[2] + (double)(k - 1) * g.h[2]};
    const int64_t o = sol.at(i - 1, j - 1, k - 1), of = farv.at(i - 1, j - 1, k - 1);
    double acc0 = sol.p[o] + fr.AREA * farv.p[of];
    double acc1 = sol.p[sol.cs + o] + fr.AREA * farv.p[farv.cs + of];
    double acc2 = sol.p[2 * sol.cs + o] + fr.AREA * farv.p[2 * farv.cs + of];
    const double R = 10.0 * fr.DIAG * (1.0 + 1e-9);  // search radius (the exact test follows)
    const double AP = fr.AREA / PI4;
    for (int q = 0; q < 6; q++) {
        const PanelFace F = pf.f[q];
        const double dn = fabs(XO[F.n] - F.plane);
        if (dn > R) continue;
        const double rho = sqrt(fmax(R * R - dn * dn, 0.0));
        const double hu = g.h[F.u], hv = g.h[F.v];
        // cell u spans [x0 + u h, x0 + (u+1) h]; its centre lies within rho of XO only if
        // u is in [(XO - rho - x0)/h - 1/2, (XO + rho - x0)/h - 1/2]; one cell of margin
        const int ulo = clampi((int)floor((XO[F.u] - rho - g.x0[F.u]) / hu - 0.5) - 1, 0, F.A - 1);
        const int uhi = clampi((int)ceil((XO[F.u] + rho - g.x0[F.u]) / hu - 0.5) + 1, 0, F.A - 1);
        const int vlo = clampi((int)floor((XO[F.v] - rho - g.x0[F.v]) / hv - 0.5) - 1, 0, F.B - 1);
        const int vhi = clampi((int)ceil((XO[F.v] + rho - g.x0[F.v]) / hv - 0.5) + 1, 0, F.B - 1);
        for (int vv = vlo; vv <= vhi; vv++) {
            for (int uu = ulo; uu <= uhi; uu++) {
                const int64_t s = F.off + (int64_t)vv * F.A + uu;
                const double d[3] = {XO[0] - cxyz[s], XO[1] - cxyz[nb + s], XO[2] - cxyz[2 * nb + s]};
                double X, Y, Z, RO;
                panel_local(fr, d, &X, &Y, &Z, &RO);
                const double RATIO = RO / fr.DIAG;
                if (RATIO > 10) continue;  // taken by the point-source sum
                const double w = fsour_near(fr, X, Y, Z, PI4) - (-AP / RO);
                acc0 = acc0 + w * sig[s];
                acc1 = acc1 + w * sig[nb + s];
                acc2 = acc2 + w * sig[2 * nb + s];
            }
        }
    }
    sol.p[o] = acc0;
    sol.p[sol.cs + o] = acc1;
    sol.p[2 * sol.cs + o] = acc2;
}

}  // namespace

int boundary_sums_panels(Ctx *c, const Grid &g, int64_t nb, const double *sig, const double *cxyz, const double *ds,
                         GView sol, bool fine, int levmax, GView scratch) {
    if (levmax >= 0) return boundary_sums_tree(c, g, nb, sig, cxyz, ds, sol, levmax, fine, true);
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t nt = 2 * ((int64_t)NY * NZ + (int64_t)(NX - 2) * NZ + (int64_t)(NX - 2) * (NY - 2));
    int st = VPMX_OK;
    double *a = (double *)ensure(c, c->panel_a, sizeof(double) * 3 * (size_t)nb, &st);
    if (st) return st;
    hipLaunchKernelGGL(k_panel_strengths, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, c->stream, nb, sig, a);
    hipLaunchKernelGGL(k_zero_targets, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, g, nt, scratch);
    VPMX_HIP(c, hipGetLastError());
    VPMX_TRY(point_sums(c, g, nb, cxyz, a, scratch, fine));
    PanelFaces pf;
    const int NNX = NX - 1, NNY = NY - 1, NNZ = NZ - 1;
    const int A[6] = {NNY, NNY, NNX, NNX, NNX, NNX}, B[6] = {NNZ, NNZ, NNZ, NNZ, NNY, NNY};
    const int n[6] = {0, 0, 1, 1, 2, 2}, u[6] = {1, 1, 0, 0, 0, 0}, v[6] = {2, 2, 2, 2, 1, 1};
    int64_t off = 0;
    for (int q = 0; q < 6; q++) {
        PanelFace &F = pf.f[q];
        F.n = n[q];
        F.u = u[q];
        F.v = v[q];
        F.A = A[q];
        F.B = B[q];
        F.plane = g.x0[F.n] + (double)((q & 1) ? g.n[F.n] - 1 : 0) * g.h[F.n];
        F.off = off;
        off += (int64_t)A[q] * B[q];
    }
    if (off != nb) return fail(c, VPMX_ERR_STATE, "panel sums: face cells differ from the sources");
    hipLaunchKernelGGL(k_panel_near, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, g, nt, pf, nb, sig,
                       cxyz, scratch, sol);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

}  // namespace vpmx
