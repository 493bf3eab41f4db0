// Enumeration of the boundary nodes that receive the infinite-domain boundary sums, in the
// reference's target order (calc_boundinf_3d, pmbound.f90:415-515): X faces full, Y faces
// without the x-edges, Z faces without the x- and y-edges.
#pragma once
#include "vpmx_internal.h"

namespace vpmx {

// Boundary node t (reference face order) -> 1-based (i,j,k) and plane code.
__device__ __forceinline__ void target_node(const Grid &g, int64_t t, int *i, int *j, int *k, int *pl) {
    const int NX = g.n[0], NY = g.n[1], NZ = g.n[2];
    const int64_t fX = (int64_t)NY * NZ, fY = (int64_t)(NX - 2) * NZ, fZ = (int64_t)(NX - 2) * (NY - 2);
    if (t < 2 * fX) {
        int f = (int)(t / fX); int64_t r = t % fX;
        *i = f ? NX : 1; *j = (int)(r % NY) + 1; *k = (int)(r / NY) + 1; *pl = 2;
    } else if (t < 2 * (fX + fY)) {
        int64_t u = t - 2 * fX; int f = (int)(u / fY); int64_t r = u % fY;
        *j = f ? NY : 1; *i = (int)(r % (NX - 2)) + 2; *k = (int)(r / (NX - 2)) + 1; *pl = 1;
    } else {
        int64_t u = t - 2 * (fX + fY); int f = (int)(u / fZ); int64_t r = u % fZ;
        *k = f ? NZ : 1; *i = (int)(r % (NX - 2)) + 2; *j = (int)(r / (NX - 2)) + 2; *pl = 3;
    }
}

}  // namespace vpmx
