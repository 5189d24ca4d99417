// Ghost-position enumeration and source lookup of the closed-form
// updatePadding kernel. Host- and device-callable so the index math can be
// checked exhaustively on the CPU (tests/cpp_host/index_check.cpp).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OPF_HD __host__ __device__ __forceinline__
#else
#define OPF_HD inline
#endif

namespace opf {

// Ghost-filling rule of one (dim, side, layer): value = f(mirror value).
// type: 0 none (no bc), 1 dirichlet-intp, 2 neumann, 3 symm, 4 asymm, 5 periodic
struct GhostRule {
    int type;
    double c1, c2, c3;  // dirichlet: (c1*v - c2)/c3 ; neumann: v + c1
};

// Per-field ghost table consumed by the device ghost kernels.
struct GhostInfo {
    int dim;
    int ls[3], le[3];         // localRange
    int gs[3], ge[3];         // logicalRange (ghost box)
    int off[3];               // storage origin
    int ext[3];               // storage extents
    int per[3];               // periodic in dim
    int corner[3];            // corner-located in dim (mirror about the boundary node)
    int nobc[3][2];           // side has no fill (undefined bc) in dim
    int ngl[3][2];            // ghost layers per side
    GhostRule rule[3][2][4];  // up to 4 layers per side
};

enum { GHOST_SKIP = 0, GHOST_COPY = 1, GHOST_RULE = 2 };

OPF_HD long long ghost_lin(const GhostInfo& gi, const int p[3]) {
    return (p[0] - gi.off[0]) + (long long)gi.ext[0] * ((p[1] - gi.off[1]) + (long long)gi.ext[1] * (p[2] - gi.off[2]));
}

// Thread t -> ghost position p (slab k: coords j<k interior, coord k ghost,
// coords j>k anywhere in the ghost box), its periodic-wrapped source `src`,
// the interior mirror `mir` and per-dim (layer, side) of the rules to apply.
// Returns GHOST_SKIP (nothing to do), GHOST_COPY (copy src -> p: a never-filled
// position reached by a periodic wrap) or GHOST_RULE.
// I: the thread-index type (32-bit unsigned where the launch allows: its division is
// several times cheaper than a 64-bit one)
template <typename I>
OPF_HD int ghost_locate(const GhostInfo& gi, I t, I n0, I n1, I n2, int p[3],
                        int src[3], int mir[3], int lay[3], int side[3]) {
    int k;
    if (t < n0) k = 0;
    else if (t < n0 + n1) { k = 1; t -= n0; }
    else if (t < n0 + n1 + n2) { k = 2; t -= n0 + n1; }
    else return GHOST_SKIP;
    for (int j = 0; j < 3; ++j) {
        int len;
        if (j < k) len = gi.le[j] - gi.ls[j];
        else if (j == k) len = (gi.ls[j] - gi.gs[j]) + (gi.ge[j] - gi.le[j]);
        else len = gi.ge[j] - gi.gs[j];
        int r = (int)(t % len);
        t /= len;
        if (j < k) p[j] = gi.ls[j] + r;
        else if (j == k) {
            int nlo = gi.ls[j] - gi.gs[j];
            p[j] = r < nlo ? gi.gs[j] + r : gi.le[j] + (r - nlo);
        } else p[j] = gi.gs[j] + r;
    }
    bool wrapped = false;
    for (int j = 0; j < 3; ++j) {
        src[j] = p[j];
        if (!gi.per[j]) continue;
        int L = gi.le[j] - gi.ls[j];
        if (src[j] < gi.ls[j]) { src[j] += L; wrapped = true; }
        else if (src[j] >= gi.le[j]) { src[j] -= L; wrapped = true; }
    }
    for (int j = 0; j < 3; ++j) {
        mir[j] = src[j];
        lay[j] = -1;
        side[j] = 0;
        if (src[j] < gi.ls[j]) { side[j] = 0; lay[j] = gi.ls[j] - 1 - src[j]; }
        else if (src[j] >= gi.le[j]) { side[j] = 1; lay[j] = src[j] - gi.le[j]; }
        else continue;
        if (gi.nobc[j][side[j]]) return wrapped ? GHOST_COPY : GHOST_SKIP;
        if (side[j] == 0) mir[j] = gi.corner[j] ? 2 * gi.ls[j] - src[j] : 2 * gi.ls[j] - 1 - src[j];
        else mir[j] = gi.corner[j] ? 2 * gi.le[j] - 2 - src[j] : 2 * gi.le[j] - 1 - src[j];
    }
    return GHOST_RULE;
}

}  // namespace opf
