// Host-side grid geometry and pm.inp parsing.
//
//   vpmx_read_pm_inp : vpm.f90:132-152 (positional, one record per line)
//   definepm         : pmlib.f90:216-432 (itype 1 and 3, 3-D)
//   define_sizes     : vpm.f90:477-660 (BLOCKS = NBI*NBJ*NBK instead of the MPI rank count)
//
// These are FP64 operation sequences whose results (node counts, box corners) must be
// bit-identical to the reference; they run on the host in the same order.  Fortran NINT
// is lround (half away from zero), INT is truncation.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/vpmx.h"

namespace {

void definepm(int itype, double xb[6], double dpm[3], int ndum, const int32_t nsize[3], int32_t nn[3],
              int32_t nn_bl[6]) {
    int nd[3] = {ndum, ndum, ndum};
    for (int d = 0; d < 3; d++) {
        xb[d] = xb[d] - ((double)nd[d] * dpm[d]);
        xb[d + 3] = xb[d + 3] + ((double)nd[d] * dpm[d]);
    }
    for (int d = 0; d < 3; d++) nn[d] = (int32_t)std::lround(std::fabs(xb[d + 3] - xb[d]) / dpm[d]) + 1;
    if (itype == 1) {
        for (int d = 0; d < 3; d++) {
            nn_bl[d] = nd[d] + 1;
            nn_bl[d + 3] = nn[d] - nd[d];
        }
        return;
    }
    // itype 3: pad so that (NN-1) is a multiple of nsize; a full extra nsize is added when
    // it already is (nsize - mod(NN-1, nsize)).
    for (int d = 0; d < 3; d++) {
        int add = nsize[d] - ((nn[d] - 1) % nsize[d]);
        if (add % 2 == 0) {
            add /= 2;
            xb[d] = xb[d] - ((double)add * dpm[d]);
            xb[d + 3] = xb[d + 3] + ((double)add * dpm[d]);
        } else {
            int n1 = add % 2, n2 = add / 2;
            xb[d] = xb[d] - ((double)(n1 + n2) * dpm[d]);
            xb[d + 3] = xb[d + 3] + ((double)n2 * dpm[d]);
        }
        nn[d] = (int32_t)std::lround(std::fabs(xb[d + 3] - xb[d]) / dpm[d]) + 1;
        nn_bl[d] = ndum + 1;
        nn_bl[d + 3] = nn[d] - ndum;
    }
}

}  // namespace

extern "C" int32_t vpmx_read_pm_inp(const char *path, vpmx_params *p) {
    if (!path || !p) return VPMX_ERR_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return VPMX_ERR_IO;
    std::memset(p, 0, sizeof(*p));
    char line[512];
    int ok = 1;
    auto next = [&]() -> bool { return std::fgets(line, sizeof line, f) != nullptr; };
    ok = ok && next() && std::sscanf(line, "%lf %lf %lf", &p->dpm[0], &p->dpm[1], &p->dpm[2]) == 3;
    ok = ok && next() && std::sscanf(line, "%d", &p->interf_iproj) == 1;
    ok = ok && next() && std::sscanf(line, "%d", &p->ibctyp) == 1;
    ok = ok && next() && std::sscanf(line, "%d", &p->idvpm) == 1;
    ok = ok && next() && std::sscanf(line, "%lf", &p->epsvol) == 1;
    ok = ok && next() && std::sscanf(line, "%d", &p->ncoarse) == 1;
    ok = ok && next() && std::sscanf(line, "%d %d %d", &p->nbi, &p->nbj, &p->nbk) == 3;
    ok = ok && next() && std::sscanf(line, "%d %d", &p->nremesh, &p->ncell_rem) == 2;
    ok = ok && next() && std::sscanf(line, "%d %d", &p->iyntree, &p->ilevmax) == 2;
    ok = ok && next() && std::sscanf(line, "%d", &p->ompthreads) == 1;
    ok = ok && next() && std::sscanf(line, "%d", &p->idefine) == 1;
    ok = ok && next() && std::sscanf(line, "%d", &p->iynslice) == 1;
    ok = ok && next() && std::sscanf(line, "%d", &p->ipmwrite) == 1;
    if (ok && p->ipmwrite > 10) {  // vpm.f90:146 STOPs
        std::fclose(f);
        return VPMX_ERR_ARG;
    }
    for (int i = 0; ok && i < p->ipmwrite; i++)
        ok = next() && std::sscanf(line, "%d %d", &p->ipmwstart[i], &p->ipmwsteps[i]) == 2;
    std::fclose(f);
    return ok ? VPMX_OK : VPMX_ERR_IO;
}

// Cell offsets of the block boundaries along one axis: ncells cells (a multiple of
// nb * ncoarse, as definepm pads the box) in nb blocks; with weights the block extents are in
// proportion to them, in whole coarse cells (cut_i = ncoarse * floor(units * W_<i / W)),
// which for equal weights is the reference's even split (ncells / nb cells each,
// vpm.f90:556-596).  False if a block would get no coarse cell.
static bool block_cuts(int ncells, int nb, int ncoarse, const int32_t *w, int cut[]) {
    const int64_t units = ncells / ncoarse;
    int64_t W = 0;
    for (int i = 0; i < nb; i++) {
        const int64_t wi = w ? w[i] : 1;
        if (wi < 1) return false;
        W += wi;
    }
    int64_t acc = 0;
    for (int i = 0; i <= nb; i++) {
        cut[i] = (int)(ncoarse * ((units * acc) / W));
        if (i < nb) acc += w ? w[i] : 1;
    }
    for (int i = 0; i < nb; i++)
        if (cut[i + 1] - cut[i] < ncoarse) return false;
    return true;
}

extern "C" int32_t vpmx_define_sizes(const vpmx_params *p, vpmx_geom *g, const double *xp, int64_t nvr,
                                     int32_t ntime) {
    return vpmx_define_sizes_weighted(p, nullptr, nullptr, nullptr, g, xp, nvr, ntime);
}

extern "C" int32_t vpmx_define_sizes_weighted(const vpmx_params *p, const int32_t *wx, const int32_t *wy,
                                              const int32_t *wz, vpmx_geom *g, const double *xp, int64_t nvr,
                                              int32_t ntime) {
    if (!p || !g) return VPMX_ERR_ARG;
    int nb = p->nbi * p->nbj * p->nbk;
    if (p->nbi < 1 || p->nbj < 1 || p->nbk < 1 || nb > VPMX_MAXB || p->ncoarse < 1) return VPMX_ERR_ARG;
    for (int d = 0; d < 3; d++)
        if (!(p->dpm[d] > 0.0)) return VPMX_ERR_ARG;
    bool redefine = (ntime == 0 && p->idefine == 1) || p->idefine == 0 || !g->defined;
    double xb[6], dpm[3];
    if (redefine) {
        if (nvr <= 0 || !xp) return VPMX_ERR_ARG;
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
        definepm(3, xb, dpm, 0, nsiz, g->nn, g->nn_bl);
        std::memcpy(g->xbound, xb, sizeof xb);
        std::memcpy(g->dpm, dpm, sizeof dpm);
        g->defined = 1;
    }
    std::memcpy(xb, g->xbound, sizeof xb);
    std::memcpy(dpm, g->dpm, sizeof dpm);
    g->nblocks = nb;
    g->nbi = p->nbi;
    g->nbj = p->nbj;
    g->nbk = p->nbk;
    int nxb = (int)std::lround((xb[3] - xb[0]) / dpm[0]);
    int nyb = (int)std::lround((xb[4] - xb[1]) / dpm[1]);
    int nzb = (int)std::lround((xb[5] - xb[2]) / dpm[2]);
    // block boundaries in cells: the reference's even split (nxb / NBI cells per block), or
    // extents in proportion to the given weights
    int cx[VPMX_MAXB + 1], cy[VPMX_MAXB + 1], cz[VPMX_MAXB + 1];
    if (!block_cuts(nxb, p->nbi, p->ncoarse, wx, cx) || !block_cuts(nyb, p->nbj, p->ncoarse, wy, cy) ||
        !block_cuts(nzb, p->nbk, p->ncoarse, wz, cz))
        return VPMX_ERR_ARG;
    int32_t nsiz_bl[3] = {p->ncoarse, p->ncoarse, p->ncoarse};
    for (int k = 1; k <= p->nbk; k++)
        for (int j = 1; j <= p->nbj; j++)
            for (int i = 1; i <= p->nbi; i++) {
                int b = (k - 1) * p->nbj * p->nbi + (j - 1) * p->nbi + (i - 1);
                double t[6];
                t[0] = xb[0] + (double)cx[i - 1] * dpm[0];
                t[3] = xb[0] + (double)cx[i] * dpm[0];
                t[1] = xb[1] + (double)cy[j - 1] * dpm[1];
                t[4] = xb[1] + (double)cy[j] * dpm[1];
                t[2] = xb[2] + (double)cz[k - 1] * dpm[2];
                t[5] = xb[2] + (double)cz[k] * dpm[2];
                double d2[3] = {dpm[0], dpm[1], dpm[2]};
                definepm(1, t, d2, p->ncoarse, nsiz_bl, &g->nnbl[3 * b], &g->nnbl_bl[6 * b]);
                std::memcpy(&g->xbound_bl[6 * b], t, sizeof t);
            }
    std::memcpy(g->xbound_coarse, xb, sizeof xb);
    for (int d = 0; d < 3; d++) g->dpm_coarse[d] = (double)p->ncoarse * dpm[d];
    int32_t nsc[3] = {1 << p->ilevmax, 1 << p->ilevmax, 1 << p->ilevmax};
    definepm(1, g->xbound_coarse, g->dpm_coarse, 4, nsc, g->nn_coarse, g->nn_bl_coarse);
    g->ilevmax = p->ilevmax;
    // (every block's inner box: with uneven blocks the smallest decides; the even blocks of
    // the reference are all the first one's size)
    int ext[3] = {1 << 30, 1 << 30, 1 << 30};
    for (int b = 0; b < nb; b++)
        for (int d = 0; d < 3; d++) ext[d] = std::min(ext[d], g->nnbl_bl[6 * b + d + 3] - g->nnbl_bl[6 * b + d]);
    for (int lev = 0; lev <= p->ilevmax; lev++) {
        int istep = 1 << lev;
        if (ext[0] / istep == 0 || ext[1] / istep == 0 || ext[2] / istep == 0) {
            g->ilevmax = lev - 1;
            break;
        }
    }
    return VPMX_OK;
}
