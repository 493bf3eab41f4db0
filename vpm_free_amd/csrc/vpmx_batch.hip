// Batched clouds: many independent particle clouds stepped together, each with a context of
// its own (parameters, geometry, grids, plans and HIP stream).  The reference runs such a
// set as one vpm() call after another (vpm.f90:54-475, per cloud: the application's RK2
// step with remesh_particles_3d, test.f90:83-150); here one step of every cloud is one fork
// from a master stream onto the clouds' streams and one join, so the clouds' kernels overlap
// on the device, and that whole step can be captured once as a hipGraph and replayed.
//
// Capture needs a step whose launches and buffers do not change from step to step, so a
// prepared batch runs its clouds in the device-count mode: every particle array is sized for a
// capacity (the larger of the cloud's count and the nodes of its remesh window, the most a
// remesh can keep), particle launches are sized for it and read the live count from the
// device, and the remesh writes the particles and their count in place.  Preparing runs one
// step of every cloud eagerly (every buffer and plan reaches its size) and restores the
// particles, so the first real step -- eager or captured -- starts from the uploaded state.
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "vpmx_ctx.h"
#include "vpmx_internal.h"

struct vpmx_batch {};

namespace vpmx {

struct Batch : vpmx_batch {
    int device = 0;
    std::vector<Ctx *> clouds;
    hipStream_t master = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> done;
    int *counts_d = nullptr;      // the clouds' live particle counts (device-count mode)
    std::vector<int> counts_h;
    bool prepared = false;
    int prep_mrem = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double key_dt = 0.0, key_u[3] = {0, 0, 0};
    int key_mrem = 0, key_rm = -1;
    std::string err;
    double step_ms = 0.0, enqueue_ms = 0.0;
    int last_graph = 0;
};

static int bfail(Batch *b, int code, const std::string &m) {
    b->err = m;
    return code;
}
#define BATCH_HIP(b, expr)                                                                                 \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return bfail((b), VPMX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
// a cloud's failure, with its message
static int cfail(Batch *b, int i, int code) {
    b->err = "cloud " + std::to_string(i) + ": " + b->clouds[i]->err;
    return code;
}

static void drop_graph(Batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Grow a device buffer to `bytes`, keeping its first `keep` bytes.
static int grow_keep(Ctx *c, DBuf &buf, size_t bytes, size_t keep) {
    if (buf.bytes >= bytes) return VPMX_OK;
    void *p = nullptr;
    VPMX_HIP(c, hipMalloc(&p, bytes));
    if (buf.p && keep) VPMX_HIP(c, hipMemcpyAsync(p, buf.p, keep, hipMemcpyDeviceToDevice, c->stream));
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    if (buf.p) (void)hipFree(buf.p);
    buf.p = p;
    buf.bytes = bytes;
    c->alloc_gen++;
    return VPMX_OK;
}

// The step of every cloud, forked from the master stream and joined back to it (no cloud
// drains its stream: the host queues every cloud's launches before any waits).
static int enqueue_all(Batch *b, double dt, const double uinf[3], int mrem, int do_remesh) {
    struct NoDrain {
        Batch *b;
        std::vector<bool> was;
        explicit NoDrain(Batch *bb) : b(bb) {
            for (Ctx *c : b->clouds) {
                was.push_back(c->nodrain);
                c->nodrain = true;
            }
        }
        ~NoDrain() {
            for (size_t i = 0; i < b->clouds.size(); i++) b->clouds[i]->nodrain = was[i];
        }
    } nd(b);
    BATCH_HIP(b, hipEventRecord(b->fork, b->master));
    for (size_t i = 0; i < b->clouds.size(); i++) {
        Ctx *c = b->clouds[i];
        BATCH_HIP(b, hipStreamWaitEvent(c->stream, b->fork, 0));
        // ntime > 0: the box stays fixed (idefine = 1, checked when prepared)
        const int r = step_rk2_enqueue(c, dt, uinf, 1, mrem, do_remesh);
        if (r != VPMX_OK) return cfail(b, (int)i, r);
        BATCH_HIP(b, hipEventRecord(b->done[i], c->stream));
        BATCH_HIP(b, hipStreamWaitEvent(b->master, b->done[i], 0));
    }
    return VPMX_OK;
}

static int read_counts(Batch *b) {
    const size_t K = b->clouds.size();
    BATCH_HIP(b, hipMemcpyAsync(b->counts_h.data(), b->counts_d, sizeof(int) * K, hipMemcpyDeviceToHost, b->master));
    BATCH_HIP(b, hipStreamSynchronize(b->master));
    for (size_t i = 0; i < K; i++) b->clouds[i]->np = b->counts_h[i];
    return VPMX_OK;
}

// Device-count mode for every cloud, then one eager step of all (buffers, plans and the
// projection's tile records reach their steady state) and the particles restored.
static int prepare(Batch *b, int mrem) {
    const size_t K = b->clouds.size();
    drop_graph(b);
    struct Saved {  // copies of the uploaded particles, freed on every way out
        std::vector<DBuf> x, q;
        explicit Saved(size_t k) : x(k), q(k) {}
        ~Saved() {
            for (DBuf &d : x)
                if (d.p) (void)hipFree(d.p);
            for (DBuf &d : q)
                if (d.p) (void)hipFree(d.p);
        }
    } saved(K);
    std::vector<DBuf> &save_x = saved.x, &save_q = saved.q;
    for (size_t i = 0; i < K; i++) {
        Ctx *c = b->clouds[i];
        if (!c->have_prm || !c->geom.defined)
            return bfail(b, VPMX_ERR_STATE, "cloud " + std::to_string(i) + ": parameters and geometry (vpm mode 0) first");
        if (c->prm.idefine != 1)
            return bfail(b, VPMX_ERR_UNSUP, "cloud " + std::to_string(i) + ": a batch needs a fixed box (idefine = 1)");
        if (c->np <= 0) return bfail(b, VPMX_ERR_STATE, "cloud " + std::to_string(i) + ": no particles");
        const int nr = c->neqpm + 1;
        const int64_t live = c->np;
        int64_t cap = remesh_window_nodes(c, mrem);
        if (cap < live) cap = live;
        int r = VPMX_OK;
        if ((r = grow_keep(c, c->xp, sizeof(double) * 3 * (size_t)cap, sizeof(double) * 3 * (size_t)live)) ||
            (r = grow_keep(c, c->qp, sizeof(double) * nr * (size_t)cap, sizeof(double) * nr * (size_t)live)) ||
            (r = grow_keep(c, c->up, sizeof(double) * 3 * (size_t)cap, 0)) ||
            (r = grow_keep(c, c->gp, sizeof(double) * 3 * (size_t)cap, 0)))
            return cfail(b, (int)i, r);
        c->cap = cap;
        c->dn = b->counts_d + i;
        b->counts_h[i] = (int)live;
        // the uploaded particles, restored after the warm-up step
        void *sx = nullptr, *sq = nullptr;
        BATCH_HIP(b, hipMalloc(&sx, sizeof(double) * 3 * (size_t)live));
        save_x[i].p = sx;
        BATCH_HIP(b, hipMalloc(&sq, sizeof(double) * nr * (size_t)live));
        save_q[i].p = sq;
        BATCH_HIP(b, hipMemcpy(sx, c->xp.p, sizeof(double) * 3 * (size_t)live, hipMemcpyDeviceToDevice));
        BATCH_HIP(b, hipMemcpy(sq, c->qp.p, sizeof(double) * nr * (size_t)live, hipMemcpyDeviceToDevice));
    }
    BATCH_HIP(b, hipMemcpy(b->counts_d, b->counts_h.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    const double u0[3] = {0.0, 0.0, 0.0};
    int r = enqueue_all(b, 0.0, u0, mrem, 1);
    BATCH_HIP(b, hipStreamSynchronize(b->master));
    if (r != VPMX_OK) return r;
    for (size_t i = 0; i < K; i++) {
        Ctx *c = b->clouds[i];
        const int nr = c->neqpm + 1;
        const size_t live = (size_t)b->counts_h[i];
        BATCH_HIP(b, hipMemcpy(c->xp.p, save_x[i].p, sizeof(double) * 3 * live, hipMemcpyDeviceToDevice));
        BATCH_HIP(b, hipMemcpy(c->qp.p, save_q[i].p, sizeof(double) * nr * live, hipMemcpyDeviceToDevice));
        c->np = (int64_t)live;
    }
    BATCH_HIP(b, hipMemcpy(b->counts_d, b->counts_h.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    b->prepared = true;
    b->prep_mrem = mrem;
    return VPMX_OK;
}

static int capture(Batch *b, double dt, const double uinf[3], int mrem, int do_remesh) {
    drop_graph(b);
    for (Ctx *c : b->clouds) c->nodrain = c->capturing = true;
    BATCH_HIP(b, hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal));
    int r = enqueue_all(b, dt, uinf, mrem, do_remesh);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->master, &g);
    for (Ctx *c : b->clouds) c->nodrain = c->capturing = false;
    if (r != VPMX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess || !g) return bfail(b, VPMX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = g;
    BATCH_HIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->key_dt = dt;
    for (int d = 0; d < 3; d++) b->key_u[d] = uinf[d];
    b->key_mrem = mrem;
    b->key_rm = do_remesh;
    return VPMX_OK;
}

}  // namespace vpmx

using namespace vpmx;

static Batch *B(vpmx_batch *p) { return static_cast<Batch *>(p); }

extern "C" {

int32_t vpmx_batch_create(vpmx_batch **out, int32_t device, int32_t nclouds) {
    if (!out) return VPMX_ERR_ARG;
    *out = nullptr;
    if (nclouds < 1 || nclouds > 4096) return VPMX_ERR_ARG;
    Batch *b = new Batch();
    b->device = device;
    int r = VPMX_OK;
    for (int i = 0; i < nclouds && r == VPMX_OK; i++) {
        vpmx_ctx *c = nullptr;
        r = vpmx_create(&c, device);
        if (r == VPMX_OK) b->clouds.push_back(static_cast<Ctx *>(c));
    }
    if (r == VPMX_OK && hipSetDevice(device) != hipSuccess) r = VPMX_ERR_HIP;
    if (r == VPMX_OK && hipStreamCreateWithFlags(&b->master, hipStreamNonBlocking) != hipSuccess) r = VPMX_ERR_HIP;
    if (r == VPMX_OK && (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
                         hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess))
        r = VPMX_ERR_HIP;
    b->done.resize((size_t)nclouds, nullptr);
    for (int i = 0; i < nclouds && r == VPMX_OK; i++)
        if (hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming) != hipSuccess) r = VPMX_ERR_HIP;
    if (r == VPMX_OK && hipMalloc(&b->counts_d, sizeof(int) * (size_t)nclouds) != hipSuccess) r = VPMX_ERR_HIP;
    b->counts_h.assign((size_t)nclouds, 0);
    if (r != VPMX_OK) {
        vpmx_batch_destroy(b);
        return r;
    }
    *out = b;
    return VPMX_OK;
}

void vpmx_batch_destroy(vpmx_batch *p) {
    Batch *b = B(p);
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->master) (void)hipStreamSynchronize(b->master);
    drop_graph(b);
    for (Ctx *c : b->clouds) vpmx_destroy(c);
    for (hipEvent_t e : b->done)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->counts_d) (void)hipFree(b->counts_d);
    if (b->master) (void)hipStreamDestroy(b->master);
    delete b;
}

const char *vpmx_batch_last_error(const vpmx_batch *p) {
    return p ? static_cast<const Batch *>(p)->err.c_str() : "null batch";
}

int32_t vpmx_batch_cloud(vpmx_batch *p, int32_t i, vpmx_ctx **ctx) {
    Batch *b = B(p);
    if (!b || !ctx) return VPMX_ERR_ARG;
    if (i < 0 || i >= (int)b->clouds.size()) return bfail(b, VPMX_ERR_ARG, "cloud index out of range");
    *ctx = b->clouds[i];
    return VPMX_OK;
}

int32_t vpmx_batch_step_rk2(vpmx_batch *p, double dt, const double uinf[3], int32_t mrem, int32_t do_remesh,
                            int32_t use_graph) {
    Batch *b = B(p);
    if (!b || !uinf) return VPMX_ERR_ARG;
    if (mrem < 1) return bfail(b, VPMX_ERR_ARG, "mrem must be >= 1");
    BATCH_HIP(b, hipSetDevice(b->device));
    if (!b->prepared || b->prep_mrem != mrem) VPMX_TRY(prepare(b, mrem));
    const auto w0 = std::chrono::steady_clock::now();
    if (use_graph) {
        const bool same = b->exec && b->key_dt == dt && b->key_u[0] == uinf[0] && b->key_u[1] == uinf[1] &&
                          b->key_u[2] == uinf[2] && b->key_mrem == mrem && b->key_rm == do_remesh;
        if (!same) VPMX_TRY(capture(b, dt, uinf, mrem, do_remesh));
        BATCH_HIP(b, hipEventRecord(b->t0, b->master));
        BATCH_HIP(b, hipGraphLaunch(b->exec, b->master));
    } else {
        BATCH_HIP(b, hipEventRecord(b->t0, b->master));
        VPMX_TRY(enqueue_all(b, dt, uinf, mrem, do_remesh));
    }
    BATCH_HIP(b, hipEventRecord(b->t1, b->master));
    b->enqueue_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    VPMX_TRY(read_counts(b));  // (drains the master stream)
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->t0, b->t1) == hipSuccess) b->step_ms = ms;
    b->last_graph = use_graph ? 1 : 0;
    return VPMX_OK;
}

int32_t vpmx_batch_times(const vpmx_batch *p, double *step_ms, double *enqueue_ms) {
    if (!p) return VPMX_ERR_ARG;
    const Batch *b = static_cast<const Batch *>(p);
    if (step_ms) *step_ms = b->step_ms;
    if (enqueue_ms) *enqueue_ms = b->enqueue_ms;
    return VPMX_OK;
}

}  // extern "C"
