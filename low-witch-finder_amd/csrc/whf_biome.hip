// whf_biome.hip — the biome stage on gfx950: the swamp-hut biome check of the multi-noise biome source for a
// batch of (world seed, hut chunk, occupied height) queries, and the check's verdict at every quart height.
//
//   k_build      the 46 Perlin samplers of the six climate noises per distinct seed (the Seed Check's builder)
//   k_bio_point  one lane per query: shifted coordinates from the offset noise (two 3D samples), continentalness,
//                erosion, weirdness, temperature and humidity at y = 0 (terrain.c:22-44, biome.c:121-122), the
//                overworld offset through the FP32 terrain spline (column_top.c:57-77, spline.c:28-62), and the five
//                height-independent quantized parameters
//   k_bio_tree   one 128-lane block per query: lane i < 97 walks the version's biome tree for quart height -17 + i,
//                lane 97 for the query's own height; the swamp verdicts are gathered by wave ballot
//   k_bio_tree1  with WHF_BIO_OCC_ONLY: one lane per query, its own height only
//
// The trees (2.25 MB for all eight versions, about 290 KB each) do not fit the 160 KB LDS; a walk touches a few
// hundred 32-byte nodes near the root, which stay in L2 across the batch. The host renumbers each tree breadth
// first at init so that a node's children are consecutive and load together (relay_trees). All of this is rare
// work next to the phase-1 gates: it runs for survivors, not for every region.
#include <cstring>
#include <unordered_map>
#include <vector>

#include "whf_biome.h"
#include "whf_internal.h"

using namespace whf;

namespace {

constexpr uint32_t kBioBatch = 8192;   // queries per device round
constexpr int kBioSamplers = 46;       // offset 6, continentalness 18, erosion 8, ridge 6, temperature 4, vegetation 4
constexpr int kQyLanes = WHF_BIOME_QY_N;

enum BioNoise { B_OFFSET = 0, B_CONT, B_EROSION, B_RIDGE, B_TEMPERATURE, B_VEGETATION, B_COUNT };

struct BioTables {
    NoiseDesc noise[B_COUNT];
};

struct BioPoint {
    int64_t t[6];       // t[4] (depth) is filled per height
    double offset;      // overworld offset at the quart point
    int32_t qx, qz;
};

// Device buffers of the stage, carved from one allocation.
struct BioBufs {
    TreeNode *nodes;          // all trees, children contiguous
    SplineTab *spline;
    BioTables *tabs;          // [2]: NORMAL, LARGE_BIOMES
    BuildDesc *bd;            // [2][kBioSamplers]
    uint8_t *tab;             // [kBioBatch][kBioSamplers][256]
    SPar *par;                // [kBioBatch][kBioSamplers]
    uint64_t *seeds;          // [kBioBatch] distinct seeds of the round
    whf_biome_query *q;       // [kBioBatch]
    uint32_t *slot;           // [kBioBatch] seed slot of each query
    BioPoint *pt;             // [kBioBatch]
    whf_biome_result *res;    // [kBioBatch]
    uint8_t *win;             // [kBioBatch][97]
};

__global__ __launch_bounds__(64) void k_bio_point(const BioTables *__restrict__ bt, const SplineTab *__restrict__ S,
                                                  const uint8_t *__restrict__ tab, const SPar *__restrict__ par,
                                                  const whf_biome_query *__restrict__ q,
                                                  const uint32_t *__restrict__ slot_of, uint32_t nq,
                                                  BioPoint *__restrict__ out) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= nq) return;
    const SetView sv{tab, par, kBioSamplers};
    const uint32_t slot = slot_of[i];
    const whf_biome_query qi = q[i];
    // chunk middle -> quart position (biome.c:108-119)
    const int qx = ((qi.hut_x >> 4) * 16 + 8) >> 2;
    const int qz = ((qi.hut_z >> 4) * 16 + 8) >> 2;
    // shifted coordinates: offset noise x 4 at (qx, 0, qz) and at (qz, qx, 0) (terrain.c:18-31)
    const double sx = noise_3d(bt->noise[B_OFFSET], sv, slot, (double)qx, 0.0, (double)qz) * 4.0;
    const double sz = noise_3d(bt->noise[B_OFFSET], sv, slot, (double)qz, (double)qx, 0.0) * 4.0;
    const double x = (double)qx + sx, z = (double)qz + sz;
    const double cont = noise_y0(bt->noise[B_CONT], sv, slot, x, z);
    const double weird = noise_y0(bt->noise[B_RIDGE], sv, slot, x, z);
    const double ero = noise_y0(bt->noise[B_EROSION], sv, slot, x, z);
    const double temp = noise_y0(bt->noise[B_TEMPERATURE], sv, slot, x, z);
    const double hum = noise_y0(bt->noise[B_VEGETATION], sv, slot, x, z);
    // overworld offset at block (4 qx, 4 qz): its terrain point is this quart point's (column_top.c:80-83, 124-130)
    const float cf = (float)cont, ef = (float)ero, wf = (float)weird;
    const double raw = (double)spline_at<kSplineMaxDepth>(S, S->root[0], cf, ef, normalized_weirdness(wf), wf);
    const double alpha = 1.0, blend = 0.0;   // blend alpha and offset of an unblended chunk (column_top.c:71-76)
    BioPoint p;
    p.offset = blend * (1.0 + -1.0 * alpha) + (-0.5037500262260437 + raw) * alpha;
    p.t[0] = quantize((float)temp);
    p.t[1] = quantize((float)hum);
    p.t[2] = quantize(cf);
    p.t[3] = quantize(ef);
    p.t[4] = 0;
    p.t[5] = quantize(wf);
    p.qx = qx;
    p.qz = qz;
    out[i] = p;
}

__global__ __launch_bounds__(128) void k_bio_tree(const TreeNode *__restrict__ g, const BioPoint *__restrict__ pts,
                                                  const whf_biome_query *__restrict__ q,
                                                  whf_biome_result *__restrict__ res, uint8_t *__restrict__ win,
                                                  unsigned long long *__restrict__ checks) {
    const uint32_t qi = blockIdx.x;
    const int lane = threadIdx.x;
    const BioPoint p = pts[qi];
    const int occ_y = q[qi].occ_y;
    const bool active = lane <= kQyLanes;   // lanes 0..96: quart heights; lane 97: the query's own height
    int winner = -1;
    Target t{{p.t[0], p.t[1], p.t[2], p.t[3], 0, p.t[5]}};
    const int qy = lane < kQyLanes ? WHF_BIOME_QY0 + lane : (occ_y >> 2);
    if (active) {
        // depth = gradient at block y 4 qy + overworld offset, in double, quantized from float (biome.c:131-141)
        t.t[4] = quantize((float)(depth_gradient((double)(qy * 4)) + p.offset));
        winner = g[tree_leaf(g, t)].biome;
        if (lane < kQyLanes) win[(size_t)qi * kQyLanes + lane] = (uint8_t)winner;
    }
    const unsigned long long m = __ballot(active && lane < kQyLanes && winner == WHF_BIOME_SWAMP);
    if ((lane & 63) == 0) res[qi].swamp_qy[lane >> 6] = m;
    if (lane == kQyLanes) {
        whf_biome_result &r = res[qi];
        for (int k = 0; k < 6; k++) r.target[k] = t.t[k];
        r.qx = p.qx;
        r.qy = qy;
        r.qz = p.qz;
        r.occ_y = occ_y;
        r.winner = winner;
        r.ok = winner == WHF_BIOME_SWAMP;
        atomicAdd(checks, (unsigned long long)(kQyLanes + 1));
    }
}

// WHF_BIO_OCC_ONLY: only the answer at each query's own height, one lane per query (a map needs no per-height
// verdicts, which take the other 97 lanes of k_bio_tree).
__global__ __launch_bounds__(64) void k_bio_tree1(const TreeNode *__restrict__ g, const BioPoint *__restrict__ pts,
                                                  const whf_biome_query *__restrict__ q,
                                                  whf_biome_result *__restrict__ res, uint32_t n,
                                                  unsigned long long *__restrict__ checks) {
    const uint32_t qi = blockIdx.x * 64 + threadIdx.x;
    const bool live = qi < n;
    if (live) {
        const BioPoint p = pts[qi];
        const int occ_y = q[qi].occ_y;
        const int qy = occ_y >> 2;
        Target t{{p.t[0], p.t[1], p.t[2], p.t[3], 0, p.t[5]}};
        t.t[4] = quantize((float)(depth_gradient((double)(qy * 4)) + p.offset));
        const int winner = g[tree_leaf(g, t)].biome;
        whf_biome_result &r = res[qi];
        for (int k = 0; k < 6; k++) r.target[k] = t.t[k];
        r.qx = p.qx;
        r.qy = qy;
        r.qz = p.qz;
        r.occ_y = occ_y;
        r.winner = winner;
        r.ok = winner == WHF_BIOME_SWAMP;
        r.swamp_qy[0] = r.swamp_qy[1] = 0;
    }
    const unsigned long long m = __ballot(live);
    if (threadIdx.x == 0 && m) atomicAdd(checks, (unsigned long long)__popcll(m));
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Carves the stage's buffers from one allocation; with base == nullptr returns the size.
size_t bio_carve(uint8_t *base, BioBufs &b) {
    size_t o = 0;
    auto take = [&](size_t bytes) { uint8_t *p = base ? base + o : nullptr; o += align256(bytes); return p; };
    const uint32_t nn = whf_gd_tree_base[WHF_GD_TREES];
    b.nodes = (TreeNode *)take(nn * sizeof(TreeNode));
    b.spline = (SplineTab *)take(sizeof(SplineTab));
    b.tabs = (BioTables *)take(2 * sizeof(BioTables));
    b.bd = (BuildDesc *)take(2 * kBioSamplers * sizeof(BuildDesc));
    b.tab = take((size_t)kBioBatch * kBioSamplers * 256);
    b.par = (SPar *)take((size_t)kBioBatch * kBioSamplers * sizeof(SPar));
    b.seeds = (uint64_t *)take(kBioBatch * sizeof(uint64_t));
    b.q = (whf_biome_query *)take(kBioBatch * sizeof(whf_biome_query));
    b.slot = (uint32_t *)take(kBioBatch * sizeof(uint32_t));
    b.pt = (BioPoint *)take(kBioBatch * sizeof(BioPoint));
    b.res = (whf_biome_result *)take(kBioBatch * sizeof(whf_biome_result));
    b.win = take((size_t)kBioBatch * kQyLanes);
    return o;
}

// The six noises of the stage for a rule-flag set, their samplers in one set, and the build descriptors.
void bio_layout(int large, BioTables &bt, std::vector<BuildDesc> &bd) {
    const int pid[B_COUNT] = {N_OFFSET,
                              large ? N_CONTINENTALNESS_LARGE : N_CONTINENTALNESS,
                              large ? N_EROSION_LARGE : N_EROSION,
                              N_RIDGE,
                              large ? N_TEMPERATURE_LARGE : N_TEMPERATURE,
                              large ? N_VEGETATION_LARGE : N_VEGETATION};
    std::vector<SamplerRef> set;
    for (int n = 0; n < B_COUNT; n++)
        for (int s = 0; s < 2; s++)
            for (int i = 0; i < kNoise[pid[n]].n; i++)
                if (kNoise[pid[n]].amps[i] != 0.0) set.push_back({pid[n], s, i});
    std::memset(&bt, 0, sizeof bt);
    for (int n = 0; n < B_COUNT; n++) fill_noise_desc(bt.noise[n], pid[n], set, SET_COUNT);
    bd.clear();
    for (const SamplerRef &r : set) bd.push_back(make_build_desc(r));
}

// The trees in the device layout: each tree keeps its node range, with the nodes renumbered breadth first so that
// the children of a node are consecutive, in the game's order (the pre-order data chains them through end links).
std::vector<TreeNode> relay_trees() {
    const uint32_t nn = whf_gd_tree_base[WHF_GD_TREES];
    std::vector<TreeNode> out(nn);
    std::vector<uint32_t> order;
    for (int t = 0; t < WHF_GD_TREES; t++) {
        const uint32_t base = whf_gd_tree_base[t];
        const whf_gd_node *g = whf_gd_nodes + base;
        order.assign(1, 0);   // tree-local pre-order indices in the new order
        for (size_t i = 0; i < order.size(); i++) {
            const whf_gd_node &src = g[order[i]];
            TreeNode &dst = out[base + i];
            std::memcpy(dst.range, src.range, sizeof dst.range);
            dst.kids = src.kids;
            dst.biome = src.biome;
            dst.pad = 0;
            dst.first = (uint32_t)order.size();
            uint32_t c = order[i] + 1;
            for (int k = 0; k < src.kids; k++, c = g[c].end) order.push_back(c);
        }
    }
    return out;
}

int bio_init(whf_ctx *c, BioBufs &b) {
    if (c->d_bio) {
        bio_carve((uint8_t *)c->d_bio, b);
        return WHF_OK;
    }
    if (whf_gd_spline_depth > kSplineMaxDepth || whf_gd_tree_depth > kTreeMaxDepth ||
        whf_gd_n_spline_nodes > 128 || whf_gd_n_spline_floats > 640 || whf_gd_n_spline_child > 384) {
        std::snprintf(g_last_error, sizeof g_last_error, "biome game data exceeds the device layout");
        return WHF_ENOTSUP;
    }
    const size_t bytes = bio_carve(nullptr, b);
    void *base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return WHF_ENOMEM;
    }
    c->d_bio = base;
    c->bio_bytes = bytes;
    bio_carve((uint8_t *)base, b);
    const hipStream_t st = c->stream;
    const std::vector<TreeNode> dev_nodes = relay_trees();
    HIP_TRY(hipMemcpyAsync(b.nodes, dev_nodes.data(), dev_nodes.size() * sizeof(TreeNode), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    SplineTab sp;
    std::memset(&sp, 0, sizeof sp);
    std::memcpy(sp.node, whf_gd_spline_nodes, whf_gd_n_spline_nodes * sizeof(whf_gd_spline_node));
    std::memcpy(sp.f, whf_gd_spline_floats, whf_gd_n_spline_floats * sizeof(float));
    std::memcpy(sp.child, whf_gd_spline_child, whf_gd_n_spline_child * sizeof(int32_t));
    for (int r = 0; r < 3; r++) sp.root[r] = whf_gd_spline_root[r];
    HIP_TRY(hipMemcpyAsync(b.spline, &sp, sizeof sp, hipMemcpyHostToDevice, st));
    BioTables bt[2];
    std::vector<BuildDesc> bd[2];
    for (int l = 0; l < 2; l++) {
        bio_layout(l, bt[l], bd[l]);
        if ((int)bd[l].size() != kBioSamplers) {
            std::snprintf(g_last_error, sizeof g_last_error, "biome sampler set has %zu samplers", bd[l].size());
            return WHF_EDEVICE;
        }
        HIP_TRY(hipMemcpyAsync(b.bd + l * kBioSamplers, bd[l].data(), kBioSamplers * sizeof(BuildDesc),
                               hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(b.tabs, bt, sizeof bt, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WHF_OK;
}

}  // namespace

extern "C" {

int whf_biome_tree(const char *version, int *tree) {
    if (!version || !tree) return WHF_EINVAL;
    for (int i = 0; i < whf_gd_n_versions; i++)
        if (!std::strcmp(whf_gd_versions[i].name, version)) { *tree = whf_gd_versions[i].tree; return WHF_OK; }
    int vf = 0;
    if (whf_version_flags(version, &vf) != WHF_OK) return WHF_EINVAL;
    const char *dflt = (vf & WHF_V_118) ? "1.18.2" : ((vf & WHF_V_262) ? "26.2" : "1.21.5");
    for (int i = 0; i < whf_gd_n_versions; i++)
        if (!std::strcmp(whf_gd_versions[i].name, dflt)) { *tree = whf_gd_versions[i].tree; return WHF_OK; }
    return WHF_EINVAL;
}

int whf_biome(whf_ctx *c, const whf_biome_query *queries, uint64_t n, int tree, int vflags, whf_biome_result *out,
              uint8_t *winners) {
    if (!c || (n && (!queries || !out))) return WHF_EINVAL;
    if (tree < 0 || tree >= WHF_GD_TREES) return WHF_EINVAL;
    const bool occ_only = (vflags & WHF_BIO_OCC_ONLY) != 0;
    if (occ_only && winners) return WHF_EINVAL;
    vflags &= ~WHF_BIO_OCC_ONLY;
    g_last_error[0] = 0;
    if (n == 0) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    BioBufs b;
    int rc = bio_init(c, b);
    if (rc != WHF_OK) return rc;
    const hipStream_t st = c->stream;
    const int large = (vflags & WHF_V_LARGE) ? 1 : 0;
    const TreeNode *g = b.nodes + whf_gd_tree_base[tree];
    std::vector<uint64_t> seeds;
    std::vector<uint32_t> slot;
    std::unordered_map<uint64_t, uint32_t> seen;
    std::vector<whf_biome_query> qb;
    std::vector<whf_column> cols;
    std::vector<int32_t> heights;
    std::vector<uint32_t> need;
    for (uint64_t q0 = 0; q0 < n; q0 += kBioBatch) {
        const uint32_t nq = (uint32_t)std::min<uint64_t>(kBioBatch, n - q0);
        // queries that ask for the surface: the pre-carver height at the chunk middle (biome.c:108-114 with the
        // height callback of eval.c:62-67), occupied height = surface - 1
        qb.assign(queries + q0, queries + q0 + nq);
        cols.clear();
        need.clear();
        for (uint32_t i = 0; i < nq; i++)
            if (qb[i].occ_y == WHF_OCC_SURFACE) {
                cols.push_back(whf_column{qb[i].seed, (qb[i].hut_x >> 4) * 16 + 8, (qb[i].hut_z >> 4) * 16 + 8});
                need.push_back(i);
            }
        if (!cols.empty()) {
            heights.resize(cols.size());
            rc = terrain_heights(c, cols.data(), cols.size(), vflags, heights.data(), nullptr);
            if (rc != WHF_OK) return rc;
            for (size_t k = 0; k < need.size(); k++) qb[need[k]].occ_y = heights[k] - 1;
        }
        seeds.clear();
        slot.resize(nq);
        seen.clear();
        for (uint32_t i = 0; i < nq; i++) {
            const uint64_t s = (uint64_t)qb[i].seed;
            auto it = seen.find(s);
            if (it == seen.end()) {
                it = seen.emplace(s, (uint32_t)seeds.size()).first;
                seeds.push_back(s);
            }
            slot[i] = it->second;
        }
        const uint32_t ns = (uint32_t)seeds.size();
        HIP_TRY(hipMemcpyAsync(b.seeds, seeds.data(), ns * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b.q, qb.data(), nq * sizeof(whf_biome_query), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b.slot, slot.data(), nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        launch_build(c, b.seeds, b.bd + large * kBioSamplers, kBioSamplers, ns, b.tab, b.par, c->d_cnt + CNT_BIO_BUILT);
        hipLaunchKernelGGL(k_bio_point, dim3((nq + 63) / 64), dim3(64), 0, st, b.tabs + large, b.spline, b.tab, b.par,
                           b.q, b.slot, nq, b.pt);
        if (occ_only)
            hipLaunchKernelGGL(k_bio_tree1, dim3((nq + 63) / 64), dim3(64), 0, st, g, b.pt, b.q, b.res, nq,
                               c->d_cnt + CNT_BIO_CHECKS);
        else
            hipLaunchKernelGGL(k_bio_tree, dim3(nq), dim3(128), 0, st, g, b.pt, b.q, b.res, b.win,
                               c->d_cnt + CNT_BIO_CHECKS);
        rc = check_launch();
        if (rc != WHF_OK) return rc;
        HIP_TRY(hipMemcpyAsync(out + q0, b.res, nq * sizeof(whf_biome_result), hipMemcpyDeviceToHost, st));
        if (winners)
            HIP_TRY(hipMemcpyAsync(winners + q0 * kQyLanes, b.win, (size_t)nq * kQyLanes, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    return WHF_OK;
}

}  // extern "C"
