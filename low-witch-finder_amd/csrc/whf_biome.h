// whf_biome.h — device functions of the biome stage: quantization, the depth gradient, the FP32 terrain spline
// and the biome tree walk. Kept apart from the kernels so that they also compile for the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "whf_game_data.h"

namespace whf {

constexpr int kSplineMaxDepth = 6;     // nesting the device spline evaluator unrolls
constexpr int kTreeMaxDepth = 12;      // frames of the device tree walk

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

__device__ __forceinline__ int64_t node_dist(const whf_gd_node &n, const int64_t t[6]) {
    // biome.c:23-31: squared distance from the target to the node's box, summed over the six parameters
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int64_t above = t[k] - (int64_t)n.range[2 * k + 1];
        const int64_t below = (int64_t)n.range[2 * k] - t[k];
        const int64_t d = above > 0 ? above : (below > 0 ? below : 0);
        s += d * d;
    }
    return s;
}

// Climate.RTree.SubTree.search as lysh walks it (biome.c:34-47), without recursion: a stack of frames, one per
// internal node on the path. A frame scans its children in order; a child whose box is nearer than the best leaf
// so far is entered (a leaf child is its own answer), and its answer replaces the best only when strictly nearer.
// A frame's best distance is always the distance of its best leaf (or "none"), so a child frame starts from it.
__device__ int tree_leaf(const whf_gd_node *__restrict__ g, const int64_t t[6]) {
    if (g[0].kids == 0) return 0;
    int fnode[kTreeMaxDepth], fc[kTreeMaxDepth], fk[kTreeMaxDepth], fres[kTreeMaxDepth];
    int64_t fbest[kTreeMaxDepth], fd[kTreeMaxDepth];
    int sp = 0;
    fnode[0] = 0; fc[0] = 1; fk[0] = 0; fres[0] = -1; fbest[0] = INT64_MAX; fd[0] = 0;
    int ret = -1;
    bool back = false;
    for (;;) {
        if (back) {
            // the child at fc[sp] (distance fd[sp]) returned leaf `ret`
            const int64_t dg = ret == fc[sp] ? fd[sp] : node_dist(g[ret], t);
            if (fbest[sp] > dg) { fbest[sp] = dg; fres[sp] = ret; }
            fc[sp] = g[fc[sp]].end;
            fk[sp]++;
            back = false;
        }
        const int kids = g[fnode[sp]].kids;
        bool down = false;
        while (fk[sp] < kids) {
            const int c = fc[sp];
            const whf_gd_node nc = g[c];
            const int64_t d = node_dist(nc, t);
            if (fbest[sp] > d) {
                if (nc.kids == 0) {
                    fbest[sp] = d;
                    fres[sp] = c;
                } else {
                    fd[sp] = d;
                    const int s1 = sp + 1;
                    fnode[s1] = c; fc[s1] = c + 1; fk[s1] = 0; fres[s1] = fres[sp]; fbest[s1] = fbest[sp];
                    sp = s1;
                    down = true;
                    break;
                }
            }
            fc[sp] = nc.end;
            fk[sp]++;
        }
        if (down) continue;
        ret = fres[sp];
        if (sp == 0) return ret;
        sp--;
        back = true;
    }
}

}  // namespace whf
