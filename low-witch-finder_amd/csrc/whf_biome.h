// whf_biome.h — device functions of the biome stage: quantization, the depth gradient, the FP32 terrain spline
// and the biome tree walk. Kept apart from the kernels so that they also compile for the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "whf_game_data.h"

namespace whf {

constexpr int kSplineMaxDepth = 6;     // nesting the device spline evaluator unrolls
constexpr int kTreeMaxDepth = 8;       // levels of the device tree walk

// Device copy of the overworld terrain spline.
struct SplineTab {
    whf_gd_spline_node node[128];
    float f[640];
    int32_t child[384];
    int root[3];            // offset, factor, peaks
};

__device__ __forceinline__ int64_t quantize(float f) { return (int64_t)(f * 10000.0f); }   // biome.c:71-73

__device__ __forceinline__ double depth_gradient(double y) {
    // biome.c:77-80: clamp((y + 64) / 384, 0, 1) mapped from 1.5 to -1.5
    double t = (y - -64.0) / (320.0 - -64.0);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return 1.5 + t * (-1.5 - 1.5);
}

__device__ __forceinline__ float normalized_weirdness(float w) {
    // spline.c:63-70, in float
    const float a = w < 0.0f ? -w : w;
    const float b = a - 0.6666667f;
    const float c = b < 0.0f ? -b : b;
    const float d = c - 0.33333334f;
    return -d * 3.0f;
}

// The nested spline at a point, in FP32 (spline.c:28-62). One instance per nesting level, so the recursion of
// the data is unrolled at compile time; a constant node ends it. Level 0 is never reached for data whose depth
// the host has checked against kSplineMaxDepth.
template <int D>
__device__ __noinline__ float spline_at(const SplineTab *__restrict__ S, int node, float c, float e, float nw,
                                        float w) {
    const whf_gd_spline_node n = S->node[node];
    if constexpr (D == 0) {
        return n.constant;
    } else {
    if (n.kind == 1) return n.constant;
    const float f = n.coord == 0 ? c : (n.coord == 1 ? e : (n.coord == 2 ? nw : w));
    const float *loc = S->f + n.loc;
    const float *der = S->f + n.der;
    const int32_t *kid = S->child + n.child;
    int lo = 0, hi = n.n;   // first location above f (n when none, also for NaN)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (f < loc[mid]) hi = mid;
        else lo = mid + 1;
    }
    const int i = lo - 1, last = n.n - 1;
    if (i < 0 || i == last) {
        const int k = i < 0 ? 0 : last;
        return spline_at<D - 1>(S, kid[k], c, e, nw, w) + der[k] * (f - loc[k]);
    }
    const float x0 = loc[i], x1 = loc[i + 1];
    const float t = (f - x0) / (x1 - x0);
    const float y0 = spline_at<D - 1>(S, kid[i], c, e, nw, w);
    const float y1 = spline_at<D - 1>(S, kid[i + 1], c, e, nw, w);
    const float a = der[i] * (x1 - x0) - (y1 - y0);
    const float b = -der[i + 1] * (x1 - x0) + (y1 - y0);
    const float base = y0 + t * (y1 - y0);
    const float bend = a + t * (b - a);
    return base + t * (1.0f - t) * bend;
    }
}

// An R-tree node as the device walks it: the game data's node with its children made contiguous (the children of
// an internal node are nodes first .. first + kids - 1, in the game's order), so that a node's children load
// together instead of one after another through the pre-order end links. The host builds it from whf_gd_node.
struct TreeNode {
    int16_t range[12];
    uint32_t first;
    uint8_t kids;
    uint8_t biome;
    uint16_t pad;
};
static_assert(sizeof(TreeNode) == 32, "tree node layout");

// The six quantized climate targets, by value.
struct Target {
    int64_t t[6];
};

__device__ __forceinline__ int64_t node_dist(const TreeNode &n, const Target &q) {
    // biome.c:23-31: squared distance from the target to the node's box, summed over the six parameters
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int64_t above = q.t[k] - (int64_t)n.range[2 * k + 1];
        const int64_t below = (int64_t)n.range[2 * k] - q.t[k];
        const int64_t d = above > 0 ? above : (below > 0 ? below : 0);
        s += d * d;
    }
    return s;
}

// Climate.RTree.SubTree.search as lysh walks it (biome.c:34-47), for an internal node: its children in order; a
// child whose box is nearer than the best leaf so far is entered (a leaf child is its own answer), and its answer
// replaces the best only when strictly nearer. `best` is always the distance of `res` (or the maximum while there
// is none), which is what the reference's callee recomputes from the leaf it is handed. Children are read 8 at a
// time; D bounds the depth (one function per level).
constexpr int kTreeGroup = 8;
template <int D>
__device__ __noinline__ void tree_walk(const TreeNode *__restrict__ g, int node, Target q, int64_t &best, int &res) {
    const TreeNode nd = g[node];
    const int first = (int)nd.first, kids = nd.kids;
    for (int k0 = 0; k0 < kids; k0 += kTreeGroup) {
        int64_t d[kTreeGroup];
        int ck[kTreeGroup];
#pragma unroll
        for (int k = 0; k < kTreeGroup; k++) {
            d[k] = 0;
            ck[k] = 0;
            if (k0 + k < kids) {
                const TreeNode c = g[first + k0 + k];
                d[k] = node_dist(c, q);
                ck[k] = c.kids;
            }
        }
#pragma unroll
        for (int k = 0; k < kTreeGroup; k++) {
            if (k0 + k < kids && best > d[k]) {
                const int c = first + k0 + k;
                if (ck[k] == 0) {
                    best = d[k];
                    res = c;
                } else if constexpr (D > 1) {
                    int64_t b2 = best;
                    int r2 = res;
                    tree_walk<D - 1>(g, c, q, b2, r2);
                    if (best > b2) {
                        best = b2;
                        res = r2;
                    }
                }
            }
        }
    }
}

// The leaf the reference's search returns for the target (its node index in the device layout).
__device__ __forceinline__ int tree_leaf(const TreeNode *__restrict__ g, const Target &q) {
    if (g[0].kids == 0) return 0;
    int64_t best = INT64_MAX;
    int res = -1;
    tree_walk<kTreeMaxDepth>(g, 0, q, best, res);
    return res;
}

}  // namespace whf
