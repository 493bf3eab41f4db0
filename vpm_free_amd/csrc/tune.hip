// Autotuner of launch choices (see tune.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "tune.h"

namespace vpmx {

int tune(Ctx *c, const std::string &key, int nf, const void *const *fns, const char *const *names, size_t lds,
         int threads, const std::function<int(int, int)> &trial, Ctx::Tuned *out, bool persistent) {
    std::vector<size_t> l((size_t)nf, lds);
    std::vector<int> t((size_t)nf, threads);
    return tune_v(c, key, nf, fns, names, l.data(), t.data(), trial, out, persistent);
}

int tune_v(Ctx *c, const std::string &key, int nf, const void *const *fns, const char *const *names,
           const size_t *lds, const int *threads, const std::function<int(int, int)> &trial, Ctx::Tuned *out,
           bool persistent) {
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) {
        *out = it->second;
        return VPMX_OK;
    }
    std::vector<int> occ((size_t)nf, 1);
    for (int fi = 0; fi < nf; fi++) {
        int o = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, fns[fi], threads[fi], lds[fi]) != hipSuccess || o < 1) o = 1;
        occ[fi] = o;
    }
    Ctx::Tuned best;
    best.fi = 0;
    best.per_cu = occ[0];
    best.ms = -1.0;
    char buf[160];
    if (c->tune_variant >= 0 && !c->capturing) {
        best.fi = c->tune_variant < nf ? c->tune_variant : nf - 1;
        best.per_cu = occ[best.fi];
        if (persistent)
            std::snprintf(buf, sizeof buf, "%s, %d work-groups per CU (forced)", names[best.fi], best.per_cu);
        else
            std::snprintf(buf, sizeof buf, "%s (forced)", names[best.fi]);
        best.what = buf;
        c->tuned[key] = best;
        *out = best;
        return VPMX_OK;
    }
    if (!c->tuning || c->capturing) {
        if (persistent)
            std::snprintf(buf, sizeof buf, "%s, %d work-groups per CU (occupancy; not timed)", names[0], occ[0]);
        else
            std::snprintf(buf, sizeof buf, "%s (not timed)", names[0]);
        best.what = buf;
        if (!c->capturing) c->tuned[key] = best;
        *out = best;
        return VPMX_OK;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VPMX_HIP(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(c, VPMX_ERR_HIP, "tuner: hipEventCreate");
    }
    int st = VPMX_OK;
    for (int fi = 0; fi < nf && st == VPMX_OK; fi++) {
        const int pc_end = persistent ? occ[fi] - 4 : occ[fi] - 1;
        for (int pc = occ[fi]; pc >= 1 && pc > pc_end && st == VPMX_OK; pc--) {
            float tmin = 1e30f;
            for (int rep = 0; rep < 3 && st == VPMX_OK; rep++) {
                if (hipEventRecord(e0, c->stream) != hipSuccess) { st = fail(c, VPMX_ERR_HIP, "tuner: event"); break; }
                st = trial(fi, pc);
                if (st) break;
                float ms = 0.f;
                if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    st = fail(c, VPMX_ERR_HIP, "tuner: timing");
                    break;
                }
                tmin = std::min(tmin, ms);
            }
            if (st == VPMX_OK && (best.ms < 0.0 || tmin < best.ms)) {
                best.fi = fi;
                best.per_cu = pc;
                best.ms = tmin;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    if (persistent)
        std::snprintf(buf, sizeof buf, "%s, %d work-groups per CU (%.4f ms)", names[best.fi], best.per_cu, best.ms);
    else
        std::snprintf(buf, sizeof buf, "%s (%.4f ms)", names[best.fi], best.ms);
    best.what = buf;
    c->tuned[key] = best;
    *out = best;
    return VPMX_OK;
}

}  // namespace vpmx
