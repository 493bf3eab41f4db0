// Tree form of the infinite-domain boundary sums (pm.inp "tree = 1"):
//
//   links       : build_level_nbound_3d (pinfdomain.f90:875-1530) — per face, level lev
//                 groups 2^lev x 2^lev face cells; groups that do not fit are "dummy" cells
//                 that only link two finer cells.  The links follow the reference's rules
//                 exactly (including the finer cells those rules leave unreached when both
//                 face extents are not multiples of 2^lev), because the sums must be the
//                 reference's sums.
//   tree sums   : Bounds3d_lev -> calc_boundinf_3d_lev -> tree_calc_3d
//                 (pmbound.f90:980-1250), ibctyp=2; calc_boundinf_3d_lev_s -> tree_calc_3d_s
//                 (pmbound.f90:1255-1481), ibctyp=3 (panel sources, k_tree_eval_s).
//
// Device form.  The forest is flattened on the host into one preorder list per geometry:
// entry e is a (level, cell) node, its children follow it, and skip[e] is the first entry
// after its subtree.  The reference's recursion visits the children of a refined cell in
// link order, depth first — exactly this preorder — so walking the list with "refine:
// e+1, accept: skip[e]" takes the reference's terms in the reference's order.  A wave owns
// an 8x8 patch of target nodes and walks the list with a wave-uniform cursor (every entry
// load is a broadcast): a lane that accepts a node waits until the cursor has passed that
// node's subtree while lanes that refine descend.
//
// The refinement test r < 10 sqrt(DS) decides which terms are taken, so r, the centroids
// and DS are formed with the reference's operation order and no contraction: for
// lattice-aligned centroids r can equal 10 sqrt(DS) to the last bit.  The accepted term
// is s * (DS / r) (the reference forms (s * DS) / r; the two differ in the last bit).
#include <hip/hip_ext.h>

#include <array>
#include <map>

#include "bnodes.h"
#include "panel.h"
#include "vpmx_ctx.h"
#include "vpmx_internal.h"

namespace vpmx {

namespace {

struct TreeLink {
    int32_t base;   // first level-0 cell of the group (0-based source index); -1 for a dummy
    int32_t pitch;  // level-0 cells per face row
    int32_t kind;   // +1 group, -1 dummy closing a row, -2 dummy in the closing row of a face
    int32_t ch[4];  // children (indices into the next finer level)
    int32_t nch;
};

// Target patches: wave w of the evaluation covers tile w; the tiles of face q are
// [off[q], off[q+1]), nu[q] tiles along the face's inner target axis.
struct TreeTiles {
    int off[7];
    int nu[6];
};

struct TreePlan {
    int64_t nb = 0;
    int64_t nent = 0;
    TreeTiles tiles{};
    int4 *desc = nullptr;     // per entry: base, pitch, level, kind
    int *skip = nullptr;      // per entry
    double4 *pos = nullptr;   // per entry: centroid, refinement parameter (k_tree_entries)
    double4 *val = nullptr;   // per entry: DS, source per component
};

using PlanKey = std::array<int, 4>;
using PlanMap = std::map<PlanKey, TreePlan *>;

struct FaceDims {
    int64_t off;  // first level-0 cell of the face in the source order
    int A, B;     // cells per row, rows
};

// Level-0 face cells in the order of calc_normalderiv_3d: X faces rows of NNY cells (k outer),
// Y faces rows of NNX cells (k outer), Z faces rows of NNX cells (j outer).
void face_dims(const int n[3], FaceDims f[6]) {
    const int NNX = n[0] - 1, NNY = n[1] - 1, NNZ = n[2] - 1;
    const int A[6] = {NNY, NNY, NNX, NNX, NNX, NNX};
    const int B[6] = {NNZ, NNZ, NNZ, NNZ, NNY, NNY};
    int64_t off = 0;
    for (int q = 0; q < 6; q++) {
        f[q].off = off;
        f[q].A = A[q];
        f[q].B = B[q];
        off += (int64_t)A[q] * B[q];
    }
}

int ceil_div(int a, int b) { return a / b + (a % b != 0); }

// Links of every level, built face by face and, within a face, level by level (the order
// in which the reference numbers the cells of each level).
int build_links(Ctx *c, const int n[3], int levmax, std::vector<std::vector<TreeLink>> &lv) {
    FaceDims fd[6];
    face_dims(n, fd);
    lv.assign(levmax + 1, {});
    for (int q = 0; q < 6; q++) {
        const int A = fd[q].A, B = fd[q].B;
        std::vector<int32_t> start(levmax + 1, 0);  // first cell of this face per level
        for (int lev = 0; lev <= levmax; lev++) {
            const int s = 1 << lev;
            std::vector<TreeLink> &L = lv[lev];
            start[lev] = (int32_t)L.size();
            const int qa = A / s, qb = B / s;                  // whole groups per row, rows of groups
            const int nnl = lev > 0 ? ceil_div(A, s / 2) : 0;  // cells per row one level finer
            for (int b = 0; b < qb; b++) {
                for (int a = 0; a < qa; a++) {
                    TreeLink t{};
                    t.base = (int32_t)(fd[q].off + (int64_t)b * s * A + (int64_t)a * s);
                    t.pitch = A;
                    t.kind = 1;
                    if (lev > 0) {
                        int32_t nj;
                        if (a == 0 && (int32_t)L.size() == start[lev]) nj = start[lev - 1];  // face start
                        else if (a == 0) nj = L.back().kind < 0 ? L.back().ch[1] + 1 : L.back().ch[2] + 1;
                        else nj = L.back().ch[1] + 1;
                        t.ch[0] = nj;
                        t.ch[1] = nj + 1;
                        t.ch[2] = nj + nnl + 1;
                        t.ch[3] = nj + nnl;
                        t.nch = 4;
                    }
                    L.push_back(t);
                }
                if (A % s != 0 || A <= s) {
                    if (lev == 0 || L.empty()) return fail(c, VPMX_ERR_UNSUP, "tree: face narrower than 2^levmax cells");
                    TreeLink t{};
                    const int32_t nj = L.back().ch[1] + 1;
                    t.base = -1;
                    t.kind = -1;
                    t.ch[0] = nj;
                    t.ch[1] = nj + nnl;
                    t.nch = 2;
                    L.push_back(t);
                }
            }
            if (B % s != 0 || B <= s) {
                if (lev == 0 || L.empty()) return fail(c, VPMX_ERR_UNSUP, "tree: face narrower than 2^levmax cells");
                for (int a = 0; a < qa; a++) {
                    // the reference leaves out the corner when a*s == A-s with A % s != 0,
                    // which never holds; the test is kept so the cell count is its own
                    if (a * s == A - s && A % s != 0) continue;
                    const TreeLink &p = L.back();
                    const int32_t nj = (a == 0 && p.kind > 0) ? p.ch[2] + 1 : p.ch[1] + 1;
                    TreeLink t{};
                    t.base = -1;
                    t.kind = -2;
                    t.ch[0] = nj;
                    t.ch[1] = nj + 1;
                    t.nch = 2;
                    L.push_back(t);
                }
            }
        }
    }
    return VPMX_OK;
}

// Preorder of the forest rooted at every cell of level levmax.
int flatten(Ctx *c, const std::vector<std::vector<TreeLink>> &lv, int levmax, std::vector<int4> &desc,
            std::vector<int> &skip) {
    struct Frame {
        int lev, idx, next;
        size_t pos;
    };
    std::vector<Frame> st;
    for (int r = 0; r < (int)lv[levmax].size(); r++) {
        st.push_back({levmax, r, 0, desc.size()});
        const TreeLink &t0 = lv[levmax][r];
        desc.push_back(make_int4(t0.base, t0.pitch, levmax, t0.kind));
        skip.push_back(0);
        while (!st.empty()) {
            Frame &f = st.back();
            const TreeLink &t = lv[f.lev][f.idx];
            if (f.lev > 0 && f.next < t.nch) {
                const int ch = t.ch[f.next++];
                const int fl = f.lev - 1;
                if (ch < 0 || ch >= (int)lv[fl].size())
                    return fail(c, VPMX_ERR_UNSUP, "tree: a cell links outside its level (faces too narrow for the levels)");
                const TreeLink &u = lv[fl][ch];
                st.push_back({fl, ch, 0, desc.size()});
                desc.push_back(make_int4(u.base, u.pitch, fl, u.kind));
                skip.push_back(0);
                if (desc.size() > (size_t)INT32_MAX / 2) return fail(c, VPMX_ERR_UNSUP, "tree: too many nodes");
                continue;
            }
            skip[f.pos] = (int)desc.size();
            st.pop_back();
        }
    }
    return VPMX_OK;
}

// Entry aggregates from the level-0 cells (build_level_nbound_3d's sums, same order:
// rows of the group outer, cells of a row inner; area-weighted source).  pos.w encodes the
// refinement test of the traversal:
//   point sources (tree_calc_3d):   refine iff r < pos.w  — 10 sqrt(DS) above level 0, -1 at
//                                   level 0 (never), +inf at dummies (always);
//   panels (tree_calc_3d_s):        refine iff r / pos.w < 10 — sqrt(DS) above level 0, NaN at
//                                   level 0 (never), -1 at dummies of kind -1 (always), 0 at
//                                   dummies of kind -2 (never: the reference takes them as
//                                   cells of zero area, r / 0 = inf).
// Dummies sit at the origin with zero area and source, as the reference's zeroed level
// arrays leave them.
__global__ void k_tree_entries(int64_t nent, const int4 *__restrict__ desc, int64_t nb, const double *__restrict__ cxyz,
                               const double *__restrict__ sig, const double *__restrict__ ds, double4 *__restrict__ pos,
                               double4 *__restrict__ val, int panels) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    const int4 d = desc[e];
    if (d.w < 0) {
        const double w = panels ? (d.w == -1 ? -1.0 : 0.0) : __builtin_inf();
        pos[e] = make_double4(0.0, 0.0, 0.0, w);
        val[e] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const int s = 1 << d.z;
    double xc = 0, yc = 0, zc = 0, sc = 0, q0 = 0, q1 = 0, q2 = 0;
    for (int mm = 0; mm < s; mm++) {
        for (int nn = 0; nn < s; nn++) {
            const int64_t l = (int64_t)d.x + (int64_t)mm * d.y + nn;
            const double sv = ds[l];
            xc = cxyz[l] + xc;
            yc = cxyz[nb + l] + yc;
            zc = cxyz[2 * nb + l] + zc;
            sc = sv + sc;
            q0 = sig[l] * sv + q0;
            q1 = sig[nb + l] * sv + q1;
            q2 = sig[2 * nb + l] * sv + q2;
        }
    }
    const double fl = (double)(s * s);
    double w;
    if (panels) w = d.z > 0 ? sqrt(sc) : __builtin_nan("");
    else w = d.z > 0 ? 10.0 * sqrt(sc) : -1.0;
    pos[e] = make_double4(xc / fl, yc / fl, zc / fl, w);
    val[e] = make_double4(sc, q0 / sc, q1 / sc, q2 / sc);
}

constexpr int TR_WAVES = 4;

// One wave per 8x8 patch of target nodes.  SOL(target) += SOURCE / (4 pi).
__global__ void __launch_bounds__(64 * TR_WAVES)
    k_tree_eval(Grid g, TreeTiles tt, int nent, const double4 *__restrict__ pos, const double4 *__restrict__ val,
                const int *__restrict__ skip, GView sol) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * TR_WAVES + (threadIdx.x >> 6);
    if (w >= tt.off[6]) return;  // whole waves
    int q = 0;
    while (w >= tt.off[q + 1]) q++;
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int tl = w - tt.off[q];
    const int u = (tl % tt.nu[q]) * 8 + (lane & 7);
    const int v = (tl / tt.nu[q]) * 8 + (lane >> 3);
    const int hi = q & 1;
    int i, j, k;
    bool ok;
    if (q < 2) {  // X faces: all (j, k)
        i = hi ? NX : 1; j = u + 1; k = v + 1;
        ok = u < NY && v < NZ;
    } else if (q < 4) {  // Y faces: i = 2..NX-1, all k
        j = hi ? NY : 1; i = u + 2; k = v + 1;
        ok = u < NX - 2 && v < NZ;
    } else {  // Z faces: i = 2..NX-1, j = 2..NY-1
        k = hi ? NZ : 1; i = u + 2; j = v + 2;
        ok = u < NX - 2 && v < NY - 2;
    }
    const double X = g.x0[0] + (double)(i - 1) * g.h[0];
    const double Y = g.x0[1] + (double)(j - 1) * g.h[1];
    const double Z = g.x0[2] + (double)(k - 1) * g.h[2];
    const double lg2 = 2.0 * log((sqrt(2.0) - 1) / (sqrt(2.0) + 1));
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    int until = ok ? 0 : 0x7fffffff;  // the lane takes no term before entry `until`
    int e = 0;
    while (e < nent) {
        const double4 p = pos[e];
        const double4 a = val[e];
        const int sk = skip[e];
        bool refine = false;
        if (e >= until) {
            const double XR = p.x - X, YR = p.y - Y, ZR = p.z - Z;
            const double r = sqrt(XR * XR + YR * YR + ZR * ZR);
            if (r < p.w) {
                refine = true;
            } else {
                if (r > 1e-05) {
                    const double wdr = a.x / r;
                    S0 = S0 - a.y * wdr;
                    S1 = S1 - a.z * wdr;
                    S2 = S2 - a.w * wdr;
                } else {  // self term of a cell whose centroid is the node
                    const double f = sqrt(a.x) * lg2;
                    S0 = S0 + a.y * f;
                    S1 = S1 + a.z * f;
                    S2 = S2 + a.w * f;
                }
                until = sk;
            }
        }
        e = __builtin_amdgcn_readfirstlane(__any(refine) ? e + 1 : sk);
    }
    if (!ok) return;
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    const int64_t o = sol.at(i - 1, j - 1, k - 1);
    sol.p[o] = sol.p[o] + S0 / PI4;
    sol.p[sol.cs + o] = sol.p[sol.cs + o] + S1 / PI4;
    sol.p[2 * sol.cs + o] = sol.p[2 * sol.cs + o] + S2 / PI4;
}

// tree_calc_3d_s (pmbound.f90:1346-1481) under calc_boundinf_3d_lev_s (:1255-1345):
// refine iff r / sqrt(DS) < 10 above level 0 or at a dummy of kind -1; a cell taken below the
// top level adds -s DS / (4 pi r); a top-level cell adds FSOUR_A4 (the target plane's frame,
// the cell's area) times the source of the FIRST top-level cell, as the reference passes
// nv = 1 at the top.  SOL += SOURCE.
__global__ void __launch_bounds__(64 * TR_WAVES)
    k_tree_eval_s(Grid g, TreeTiles tt, int nent, int levmax, const int4 *__restrict__ desc,
                  const double4 *__restrict__ pos, const double4 *__restrict__ val, const int *__restrict__ skip,
                  GView sol) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * TR_WAVES + (threadIdx.x >> 6);
    if (w >= tt.off[6]) return;
    int q = 0;
    while (w >= tt.off[q + 1]) q++;
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int tl = w - tt.off[q];
    const int u = (tl % tt.nu[q]) * 8 + (lane & 7);
    const int v = (tl / tt.nu[q]) * 8 + (lane >> 3);
    const int hi = q & 1;
    int i, j, k, iplane;
    bool ok;
    if (q < 2) {
        i = hi ? NX : 1; j = u + 1; k = v + 1; iplane = hi ? -2 : 2;
        ok = u < NY && v < NZ;
    } else if (q < 4) {
        j = hi ? NY : 1; i = u + 2; k = v + 1; iplane = hi ? -1 : 1;
        ok = u < NX - 2 && v < NZ;
    } else {
        k = hi ? NZ : 1; i = u + 2; j = v + 2; iplane = hi ? -3 : 3;
        ok = u < NX - 2 && v < NY - 2;
    }
    const PanelFrame fr = panel_frame(iplane, g.h);
    const double XO[3] = {g.x0[0] + (double)(i - 1) * g.h[0], g.x0[1] + (double)(j - 1) * g.h[1],
                          g.x0[2] + (double)(k - 1) * g.h[2]};
    const double PI4 = 4.0 * (4.0 * atan(1.0));
    const double4 first = val[0];  // the first top-level cell
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    int until = ok ? 0 : 0x7fffffff;
    int e = 0;
    while (e < nent) {
        const double4 p = pos[e];
        const double4 a = val[e];
        const int lev = desc[e].z;
        const int sk = skip[e];
        bool refine = false;
        if (e >= until) {
            const double d[3] = {XO[0] - p.x, XO[1] - p.y, XO[2] - p.z};
            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r / p.w < 10) {
                refine = true;
            } else {
                if (lev != levmax) {
                    const double PR = PI4 * r;
                    S0 = S0 - a.y * a.x / PR;
                    S1 = S1 - a.z * a.x / PR;
                    S2 = S2 - a.w * a.x / PR;
                } else {
                    const double FIS = fsour_a4(fr, d, a.x, PI4);
                    S0 = S0 + first.y * FIS;
                    S1 = S1 + first.z * FIS;
                    S2 = S2 + first.w * FIS;
                }
                until = sk;
            }
        }
        e = __builtin_amdgcn_readfirstlane(__any(refine) ? e + 1 : sk);
    }
    if (!ok) return;
    const int64_t o = sol.at(i - 1, j - 1, k - 1);
    sol.p[o] = sol.p[o] + S0;
    sol.p[sol.cs + o] = sol.p[sol.cs + o] + S1;
    sol.p[2 * sol.cs + o] = sol.p[2 * sol.cs + o] + S2;
}

void free_plan(TreePlan *p) {
    if (!p) return;
    (void)hipFree(p->desc);
    (void)hipFree(p->skip);
    (void)hipFree(p->pos);
    (void)hipFree(p->val);
    delete p;
}

int get_plan(Ctx *c, const Grid &g, int levmax, int64_t nb, TreePlan **out) {
    if (!c->tree) c->tree = new PlanMap();
    PlanMap &m = *(PlanMap *)c->tree;
    const PlanKey key{g.n[0], g.n[1], g.n[2], levmax};
    auto it = m.find(key);
    if (it != m.end()) {
        *out = it->second;
        return VPMX_OK;
    }
    if (levmax < 0 || levmax > 12) return fail(c, VPMX_ERR_UNSUP, "tree: levels must be 0..12");
    std::vector<std::vector<TreeLink>> lv;
    VPMX_TRY(build_links(c, g.n, levmax, lv));
    if ((int64_t)lv[0].size() != nb) return fail(c, VPMX_ERR_STATE, "tree: level-0 cells differ from the face sources");
    std::vector<int4> desc;
    std::vector<int> skip;
    desc.reserve((size_t)(nb * 4 / 3 + 64));
    skip.reserve(desc.capacity());
    VPMX_TRY(flatten(c, lv, levmax, desc, skip));
    TreePlan *p = new TreePlan();
    p->nb = nb;
    p->nent = (int64_t)desc.size();
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int U[6] = {NY, NY, NX - 2, NX - 2, NX - 2, NX - 2};
    const int V[6] = {NZ, NZ, NZ, NZ, NY - 2, NY - 2};
    p->tiles.off[0] = 0;
    for (int q = 0; q < 6; q++) {
        p->tiles.nu[q] = (U[q] + 7) / 8;
        p->tiles.off[q + 1] = p->tiles.off[q] + p->tiles.nu[q] * ((V[q] + 7) / 8);
    }
    const size_t ne = (size_t)p->nent;
    hipError_t e = hipMalloc((void **)&p->desc, sizeof(int4) * ne);
    if (e == hipSuccess) e = hipMalloc((void **)&p->skip, sizeof(int) * ne);
    if (e == hipSuccess) e = hipMalloc((void **)&p->pos, sizeof(double4) * ne);
    if (e == hipSuccess) e = hipMalloc((void **)&p->val, sizeof(double4) * ne);
    if (e == hipSuccess) e = hipMemcpy(p->desc, desc.data(), sizeof(int4) * ne, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->skip, skip.data(), sizeof(int) * ne, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_plan(p);
        return fail(c, VPMX_ERR_HIP, std::string("tree plan: ") + hipGetErrorString(e));
    }
    m[key] = p;
    *out = p;
    return VPMX_OK;
}

}  // namespace

int boundary_sums_tree(Ctx *c, const Grid &g, int64_t nb, const double *sig, const double *cxyz, const double *ds,
                       GView sol, int levmax, bool fine, bool panels) {
    if (g.n[0] < 3 || g.n[1] < 3 || g.n[2] < 3) return fail(c, VPMX_ERR_ARG, "tree: grid too small");
    TreePlan *p = nullptr;
    VPMX_TRY(get_plan(c, g, levmax, nb, &p));
    hipLaunchKernelGGL(k_tree_entries, dim3((unsigned)((p->nent + 255) / 256)), dim3(256), 0, c->stream, p->nent,
                       p->desc, nb, cxyz, sig, ds, p->pos, p->val, panels ? 1 : 0);
    VPMX_HIP(c, hipGetLastError());
    const unsigned blocks = (unsigned)((p->tiles.off[6] + TR_WAVES - 1) / TR_WAVES);
    if (panels) {
        hipLaunchKernelGGL(k_tree_eval_s, dim3(blocks), dim3(64 * TR_WAVES), 0, c->stream, g, p->tiles, (int)p->nent,
                           levmax, p->desc, p->pos, p->val, p->skip, sol);
        VPMX_HIP(c, hipGetLastError());
        return VPMX_OK;
    }
    hipEvent_t e0, e1;
    if (fine && kernel_events(c, K_BOUNDSUM, &e0, &e1))
        hipExtLaunchKernelGGL(k_tree_eval, dim3(blocks), dim3(64 * TR_WAVES), 0, c->stream, e0, e1, 0, g, p->tiles,
                              (int)p->nent, p->pos, p->val, p->skip, sol);
    else
        hipLaunchKernelGGL(k_tree_eval, dim3(blocks), dim3(64 * TR_WAVES), 0, c->stream, g, p->tiles, (int)p->nent,
                           p->pos, p->val, p->skip, sol);
    VPMX_HIP(c, hipGetLastError());
    return VPMX_OK;
}

int64_t tree_entries(Ctx *c, const int n[3], int levmax) {
    if (!c->tree) return -1;
    PlanMap &m = *(PlanMap *)c->tree;
    auto it = m.find(PlanKey{n[0], n[1], n[2], levmax});
    return it == m.end() ? -1 : it->second->nent;
}

void free_tree(Ctx *c) {
    if (!c->tree) return;
    PlanMap *m = (PlanMap *)c->tree;
    for (auto &kv : *m) free_plan(kv.second);
    delete m;
    c->tree = nullptr;
}

}  // namespace vpmx
