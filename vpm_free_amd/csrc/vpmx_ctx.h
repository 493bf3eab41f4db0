// The context behind the opaque vpmx_ctx handle.
#pragma once
#include <map>
#include <string>
#include <array>
#include <vector>

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
// K_DST: the y passes of the fine block's sine transforms; K_DSTZ: its z passes (forward,
// diagonal, inverse).
enum KClass { K_BOUNDSUM = 0, K_P2M = 1, K_M2P = 2, K_DST = 3, K_DSTZ = 4, K_COUNT = 5 };

struct Ctx : vpmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    vpmx_params prm{};
    bool have_prm = false;
    std::vector<int32_t> block_w[3];  // uneven yaps blocks: extents in proportion (empty: even split)
    vpmx_geom geom{};
    int precision = 0;

    // resident particles
    DBuf xp, qp, up, gp, x0, q0;
    int64_t np = 0;
    // device-count mode (steps captured in a graph): the particle count lives at dn on the
    // device (remesh writes it there), every particle launch is sized for np = the capacity
    // and reads the count; null in the ordinary mode, where np is the count
    int *dn = nullptr;
    int64_t cap = 0;  // particle capacity of the resident arrays in that mode
    bool nodrain = false;  // queued as part of a batch (or captured): no stream drains, no timing events
    bool capturing = false;  // a graph is being captured: nothing may wait on the stream

    // autotuner: per launch key (pass, precision, lengths, sizes) the kernel variant and the
    // resident work-groups per CU chosen by timing the candidates once
    bool tuning = true;
    // >= 0: every launch key takes variant min(tune_variant, last) at its occupancy limit,
    // untimed (for checking that variants meant to give the same bits do)
    int tune_variant = -1;
    struct Tuned {
        int fi = 0, per_cu = 1;
        double ms = 0.0;
        std::string what;
    };
    std::map<std::string, Tuned> tuned;
    DBuf tune_scratch;
    DBuf adj_a, adj_b, adj_c, adj_d, adj_e, adj_f, adj_g, adj_h, adj_i, adj_p, adj_q;  // adjoint work arrays
    DBuf adj_u, adj_s, adj_r, adj_xb, adj_qb, adj_xb2, adj_qb2, adj_xm, adj_qm;
    // roll-out adjoint: one recorded RK2 step (particles at the start, midpoint and end, the two
    // evaluations' grids, the remeshed positions) and the checkpoints (XP, QP)
    struct AdjStep {
        DBuf b[13];
        int64_t n0 = 0, n1 = 0;
    };
    std::vector<AdjStep> adj_steps;
    int adj_bs = 2;  // transposed boundary sums: 2 spectral, 1 direct
    bool adj_m2p_scatter = false;  // M2P^T by FP64 atomics instead of the sorted gather
    DBuf adj_pu, adj_ps;
    std::vector<std::array<DBuf, 2>> adj_ck;
    int neqpm = 3;
    double ni = 0.0;  // kinematic viscosity NI of the diffusion mode (resident calls)

    // grids
    DBuf rhs_bl, sol_bl, sol0, vel, str;
    DBuf src_sig, src_cxyz, src_w, bs_part, lat_pairs, lat_pairs_c, panel_a;
    DBuf nd_face;  // normal derivatives at the face nodes (before the per-cell averages)
    DBuf fd_spec;  // the face sums of the spectral normal derivatives and their partial transforms
    DBuf fp_spec;  // the second fine solve's face and sample values (sums, gathered planes, products)
    int fine_partial = -1;  // second fine solve of a single block: -1 by grid size, 0 every node, 1 faces and samples
    int face_derivs = -1;  // normal derivatives of pmesh: -1 by grid size, 0 from the solved field, 1 spectral
    int bs_mode = 2;  // boundary sums (option boundsum): 2 spectral (default), 0 lattice tables, 1 direct LDS source tiles
    void *bsx[2] = {nullptr, nullptr};  // spectral boundary-sum plans: fine block, coarse grid
    void *tree = nullptr;  // tree boundary-sum plans per (grid, levels)
    int num_cu = 256;       // compute units of the device
    int64_t max_lds = 65536;  // LDS bytes a work-group may hold (device attribute)
    bool dst_reuse = true;  // second fine solve from the first one's forward transforms (option dst_reuse)
    int dst_kernel = 4;   // DST passes: 4 half-length, one wave per line pair (default), 3 half-length
                          // work-group FFTs, 2 odd extension, 1 two-buffer
    std::map<int, void *> sin_cache;  // sin(j pi / N) per half-length FFT size N
    std::map<std::pair<int, double>, void *> eig_cache;  // DST-I eigenvalues per (m, h)
    DBuf samp, rhs_c, sol_c, sol0_c, cmap;
    std::vector<int> cmap_host;   // contents of cmap as last uploaded (kept alive for the async copy)
    const void *cmap_dev = nullptr;  // device address they were uploaded to
    DBuf rhs_g, sol_g, mb_blocks, mb_tmp;  // several yaps blocks: global RHS / SOL grids, block table
    std::vector<char> mb_host;             // the block table as last uploaded
    const void *mb_dev = nullptr;          // where it was uploaded to
    DBuf dst_work, lam[3];
    DBuf dst_work64;  // FP64 work array of the mixed form's accurate solves (after the diagonal)
    DBuf dst_save, fold_tab;  // forward-pass reuse: the kept work array, the face transforms
    int save_key[6] = {0, 0, 0, 0, 0, -1};  // m[3], px, ncomp, f32 of the kept array (f32 -1: none)
    std::map<int, void *> sinm_cache;  // sin(pi (t+1)(k+1) / (m+1)), t, k < m, per m
    std::map<int, void *> sinx_cache;  // sin(pi r (k+1) / (m+1)), r <= m+1, k < m, per m
    std::map<std::pair<int, double>, void *> fdw_cache;  // normal-derivative weights per (m, h)
    std::map<std::pair<int, int>, void *> sinr_cache;  // sin(pi p k / (m+1)), k = 1..m, per (m, p)
    std::map<std::pair<int, int>, void *> sinl_cache;  // rows p = a s, a = 1..m/s, of the DST-I matrix, per (m, s)
    DBuf scan_tmp, red_tmp, rm_flag, rm_pos, rm_x, rm_q, rm_grid;
    DBuf tile_flag, tile_pos, tile_list;  // interpolation: cell tiles holding particles
    // projection without zeros for empty tiles: tile_flag maps the node tiles of a view of
    // tile_map_vn nodes (ntx x nty tiles per plane) that hold values; valid until the next use
    bool tile_map = false;
    int tile_map_vn[3] = {0, 0, 0}, tile_map_ntx = 0, tile_map_nty = 0;
    DBuf stage_a, stage_b, stage_c;
    SortState sort;
    int64_t sort_n = 0;     // particles of the last sort
    int64_t alloc_gen = 0;  // counts (re)allocations by ensure()
    // projection into the block RHS with the zeros kept: the empty tiles of the last such
    // projection into zt_ptr (view zt_vn, offsets zt_wlo) are still zero, so the next one
    // zeroes only the tiles that were occupied then (zt_prev) and are empty now
    bool zt_valid = false;
    const void *zt_ptr = nullptr;
    int zt_vn[3] = {0, 0, 0}, zt_wlo[3] = {0, 0, 0};
    int64_t zt_gen = -1;
    DBuf zt_prev;

    std::map<int, void *> fft_plans;
    double *scratch = nullptr;  // the device's fixed array of the allocation-free entry points
    std::string scratch_err;    // why it could not be set up
    bool have_grids = false;  // RHS/SOL/vel of the last mode 1-3 call are valid

    // remesh result waiting for vpmx_remesh_fetch
    int64_t rm_count = -1;
    double *rm_xp = nullptr, *rm_qp = nullptr;

    // timing: stages by events recorded on the stream at their boundaries, kernels of class
    // K_* by the dispatch's own start/stop events (hipExtLaunchKernel); both read back when
    // the outermost API call ends (TimingScope), so no stage drains the stream
    bool timing = true;
    int timing_depth = 0;
    std::vector<hipEvent_t> sev0[ST_COUNT], sev1[ST_COUNT];
    int sev_used[ST_COUNT] = {};
    double stage_ms[ST_COUNT] = {};
    double last_pairs = 0.0;
    std::vector<hipEvent_t> kev0[K_COUNT], kev1[K_COUNT];
    int kev_pending[K_COUNT] = {};
    bool kev_ok = false;
    double k_ms[K_COUNT] = {};
    int64_t k_launches[K_COUNT] = {};
    double k_work[K_COUNT] = {};  // algorithmic work of the last timed launch (flops or bytes)
    int dst_timed_m = -1;         // interior length of the DST passes timed as K_DST (fine block)
};

void free_fft_plans(Ctx *c);
double wall_ms();
// Stage timing: drain the stream at both ends so the interval covers exactly the stage's
// kernels (enqueueing is asynchronous).
// Add the durations of the recorded dispatches of every class (the stream is drained).
inline void resolve_kernel_times(Ctx *c) {
    for (int k = 0; k < K_COUNT; k++) {
        for (int i = 0; i < c->kev_pending[k]; i++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, c->kev0[k][i], c->kev1[k][i]) == hipSuccess) {
                c->k_ms[k] += ms;
                c->k_launches[k] += 1;
            }
        }
        c->kev_pending[k] = 0;
    }
}
// Stage timing: an event on the stream at the stage's start and end (a stage may run several
// times in one call; each run gets its own pair).
inline void mark(Ctx *c, int stage, int end) {
    if (!c->timing || !c->kev_ok || c->timing_depth == 0 || c->nodrain) return;
    if (!end) {
        const int i = c->sev_used[stage];
        if (i == (int)c->sev0[stage].size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess) return;
            if (hipEventCreate(&b) != hipSuccess) {
                (void)hipEventDestroy(a);
                return;
            }
            c->sev0[stage].push_back(a);
            c->sev1[stage].push_back(b);
        }
        if (hipEventRecord(c->sev0[stage][i], c->stream) == hipSuccess) c->sev_used[stage] = i + 1;
    } else if (c->sev_used[stage] > 0) {
        (void)hipEventRecord(c->sev1[stage][c->sev_used[stage] - 1], c->stream);
    }
}
// The outermost API call that runs stages: zero the stage times at its start, drain the
// stream at its end and add up the recorded intervals (and the timed kernels).
struct TimingScope {
    Ctx *c;
    explicit TimingScope(Ctx *cc) : c(cc) {
        if (c->timing_depth++ == 0) {
            for (int s = 0; s < ST_COUNT; s++) {
                c->stage_ms[s] = 0.0;
                c->sev_used[s] = 0;
            }
        }
    }
    ~TimingScope() {
        if (--c->timing_depth != 0) return;
        if (c->nodrain) return;  // (nothing was recorded; a drain is not allowed under capture)
        (void)hipStreamSynchronize(c->stream);
        for (int s = 0; s < ST_COUNT; s++) {
            for (int i = 0; i < c->sev_used[s]; i++) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, c->sev0[s][i], c->sev1[s][i]) == hipSuccess) c->stage_ms[s] += ms;
            }
            c->sev_used[s] = 0;
        }
        resolve_kernel_times(c);
    }
};
// Start/stop events for the next timed dispatch of class k (false: time nothing).
inline bool kernel_events(Ctx *c, int k, hipEvent_t *e0, hipEvent_t *e1) {
    if (!c->kev_ok || c->nodrain) return false;
    const int i = c->kev_pending[k];
    if (i >= 64) return false;  // more than 64 undrained dispatches: leave the rest untimed
    if (i == (int)c->kev0[k].size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return false;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return false;
        }
        c->kev0[k].push_back(a);
        c->kev1[k].push_back(b);
    }
    *e0 = c->kev0[k][i];
    *e1 = c->kev1[k][i];
    c->kev_pending[k] = i + 1;
    return true;
}
int exclusive_scan(Ctx *c, const int *in, int *out, int64_t n);

}  // namespace vpmx
