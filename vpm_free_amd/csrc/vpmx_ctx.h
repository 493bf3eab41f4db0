// The context behind the opaque vpmx_ctx handle.
#pragma once
#include <map>
#include <string>

#include "vpmx_internal.h"

struct vpmx_ctx {};  // opaque at the ABI; the object is a vpmx::Ctx

namespace vpmx {

enum Stage {
    ST_PROJ = 0,       // 'Proj'
    ST_FINE = 1,       // 'Poisson fine'  (block pmesh, infinite-domain BCs)
    ST_BOUNDS = 2,     // 'Bounds'        (fine boundary sums, part of ST_FINE)
    ST_COARSE = 3,     // 'pmesh_coarse'  (sampling + coarse pmesh)
    ST_BC = 4,         // 'pm_bc2'        (face interpolation)
    ST_FINAL = 5,      // 'pmesh_final'
    ST_FD = 6,         // 'fd'            (curl + stretching)
    ST_BACK = 7,       // 'back'          (interpolation to particles)
    ST_REMESH = 8,     // 'remeshing'
    ST_COUNT = 9
};

// Kernel classes timed with events when profiling is on.
enum KClass { K_BOUNDSUM = 0, K_P2M = 1, K_M2P = 2, K_DST = 3, K_COUNT = 4 };

struct Ctx : vpmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    vpmx_params prm{};
    bool have_prm = false;
    vpmx_geom geom{};
    int precision = 0;

    // resident particles
    DBuf xp, qp, up, gp, x0, q0;
    int64_t np = 0;
    int neqpm = 3;

    // grids
    DBuf rhs_bl, sol_bl, sol0, vel, str;
    DBuf src_sig, src_cxyz, src_w, bs_part;
    DBuf samp, rhs_c, sol_c, sol0_c, cmap;
    DBuf dst_work, lam[3];
    DBuf scan_tmp, red_tmp, rm_flag, rm_pos, rm_x, rm_q, rm_grid;
    DBuf stage_a, stage_b, stage_c;
    SortState sort;
    std::map<int, void *> fft_plans;
    bool have_grids = false;  // RHS/SOL/vel of the last mode 1-3 call are valid

    // remesh result waiting for vpmx_remesh_fetch
    int64_t rm_count = -1;
    double *rm_xp = nullptr, *rm_qp = nullptr;

    // timing
    hipEvent_t ev[2 * ST_COUNT] = {};
    bool ev_ok = false;
    double stage_ms[ST_COUNT] = {};
    double last_pairs = 0.0;
    bool profiling = false;
    std::vector<hipEvent_t> kev[K_COUNT];
    size_t kev_used[K_COUNT] = {};
};

void free_fft_plans(Ctx *c);
// Stage timing: the stream is drained before a start marker so that the marker's
// timestamp is taken when the stage's first kernel is issued, not while earlier kernels
// of the stream are still running (a marker does not wait for them to finish).
inline void mark(Ctx *c, int stage, int end) {
    if (!c->ev_ok) return;
    if (!end) (void)hipStreamSynchronize(c->stream);
    (void)hipEventRecord(c->ev[2 * stage + end], c->stream);
}
int exclusive_scan(Ctx *c, const int *in, int *out, int64_t n);

}  // namespace vpmx
