// whf_terrain.hip — the terrain stage on gfx950: the pre-carver surface height of a batch of (seed, column) queries
// from the overworld's interpolated final density and aquifer.
//
//   k_build          the 144 named-noise samplers of a seed (climate, density, noodle and aquifer noises)
//   k_build_blended  one lane per seed: the blended noise's 40 samplers, built in sequence from one random as the
//                    reference builds them, and the aquifer's positional random factory
//   k_surface        one 256-lane block per column:
//                      A  the 121 preliminary-surface samples of the column chunk's aquifer grid (their maximum gives
//                         skipSamplingAboveY) and the four corner terrain points;
//                      B  the 4 x 49 corner samples of the noise column and the 4 x 3 x 49 noodle children;
//                      C  the final density at every y (trilinear over the corners, squeezed, min with the noodle);
//                         the top solid block y_d is its highest positive density;
//                      D  above y_d, the aquifer decides each block up to skipSamplingAboveY, the sea above it; the
//                         height is one above the highest block that is not air.
// This is work per phase-1 survivor (one in 1e5..1e7 regions), so it is laid out for clarity of the reference's
// order of operations, not for throughput: every lane calls the same pointwise functions the reference calls.
#include <cstring>
#include <unordered_map>
#include <vector>

#include "whf_internal.h"
#include "whf_terrain.h"

using namespace whf;

namespace {

constexpr uint32_t kTerBatch = 1024;   // columns per device round (distinct seeds per round at most this)
constexpr int kTerSamplers = 144;      // named-noise samplers per seed
constexpr int kCornerJobs = 4 * 49;    // main corner samples
constexpr int kNoodleJobs = 4 * 3 * 49;

struct TerBufs {
    TerrainTables *tt;        // [2]: NORMAL, LARGE_BIOMES
    BuildDesc *bd;            // [2][kTerSamplers]
    SplineTab *spline;
    uint8_t *tab;             // [kTerBatch][kTerSamplers][256]
    SPar *par;
    uint8_t *ltab;            // [kTerBatch][40][256]
    SPar *lpar;
    ulonglong2 *aq;           // [kTerBatch]
    uint64_t *seeds;          // [kTerBatch]
    whf_column *q;            // [kTerBatch]
    uint32_t *slot;           // [kTerBatch]
    int2 *out;                // [kTerBatch]: height, top state
};

__global__ __launch_bounds__(64) void k_build_blended(const uint64_t *__restrict__ seeds, uint32_t ns, uint64_t t_hi,
                                                      uint64_t t_lo, uint64_t a_hi, uint64_t a_lo,
                                                      uint8_t *__restrict__ ltab, SPar *__restrict__ lpar,
                                                      ulonglong2 *__restrict__ aq) {
    __shared__ uint32_t words[64 * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t s = blockIdx.x * 64 + lane;
    if (s >= ns) return;   // each lane shuffles in its own LDS column only
    const uint64_t seed = seeds[s];
    // deriver of the world random, createRandom("minecraft:terrain"); then lower, upper and main stacks, each
    // from its coarsest octave down, all from this one random (interp_noise.c:19-48, noise.c:167-209)
    Xr r = xr_seed(seed);
    Xr der = xr_fork(r);
    Xr pr = xr_make(t_hi ^ der.lo, t_lo ^ der.hi);
    for (int k = 0; k < kLegacyPerlins; k++)
        build_sampler(pr, words, lane, ltab + ((size_t)s * kLegacyPerlins + k) * 256, lpar[(size_t)s * kLegacyPerlins + k]);
    // the aquifer's random: forkPositional of the world random, fromHashOf("minecraft:aquifer"), forkPositional
    // (aquifer.c:123-135)
    Xr w = xr_seed(seed);
    const uint64_t b_lo = xr_next(w);
    const uint64_t b_hi = xr_next(w);
    Xr a = xr_make(a_hi ^ b_lo, a_lo ^ b_hi);
    ulonglong2 f;
    f.x = xr_next(a);
    f.y = xr_next(a);
    aq[s] = f;
}

__global__ __launch_bounds__(256) void k_surface(const TerrainTables *__restrict__ tt, const SplineTab *__restrict__ S,
                                                 const uint8_t *__restrict__ tab, const SPar *__restrict__ par,
                                                 const uint8_t *__restrict__ ltab, const SPar *__restrict__ lpar,
                                                 const ulonglong2 *__restrict__ aq, const whf_column *__restrict__ q,
                                                 const uint32_t *__restrict__ slot_of, int2 *__restrict__ out) {
    __shared__ double s_main[4][49];
    __shared__ double s_nood[4][3][49];
    __shared__ double s_dens[384];
    __shared__ float s_tp[4][3];
    __shared__ int s_psl, s_top_solid, s_best, s_state;
    const uint32_t qi = blockIdx.x;
    const int t = threadIdx.x;
    const whf_column col = q[qi];
    const uint32_t slot = slot_of[qi];
    const ulonglong2 f = aq[slot];
    const Terr T{tt, SetView{tab, par, kTerSamplers}, slot, ltab + (size_t)slot * kLegacyPerlins * 256,
                 lpar + (size_t)slot * kLegacyPerlins, S, f.x, f.y};
    const int x = col.x, z = col.z;
    if (t == 0) {
        s_psl = -0x7FFFFFFF - 1;
        s_top_solid = -65;
        s_best = -0x7FFFFFFF - 1;
        s_state = BS_AIR;
    }
    __syncthreads();

    // A: skipSamplingAboveY of the column's chunk (aquifer.c:160-196): preliminary surfaces on the 11 x 11 grid of
    // step 4 from the chunk's aquifer grid corner; corner terrain points of the column's cell (phase2.c:71-83)
    const int min_gx = ((x >> 4) * 16 - 5) >> 4, min_gz = ((z >> 4) * 16 - 5) >> 4;
    if (t < 121) {
        atomicMax(&s_psl, T.prelim_surface(min_gx * 16 + 4 * (t % 11), min_gz * 16 + 4 * (t / 11)));
    } else if (t >= 128 && t < 132) {
        const int c = t - 128;
        T.point261((x >> 2) + (c & 1), (z >> 2) + (c >> 1), s_tp[c]);
    }
    __syncthreads();
    const int skip_above = (floor_div((s_psl + 8) + 12, 12) + 1) * 12 + 11 - 1;

    // B: corner samples at y = -64 + 8 i of the corners (4 cx, 4 cz) .. (4 cx + 4, 4 cz + 4)
    for (int j = t; j < kCornerJobs + kNoodleJobs; j += 256) {
        if (j < kCornerJobs) {
            const int c = j / 49, i = j - c * 49;
            const int bx = ((x >> 2) + (c & 1)) * 4, bz = ((z >> 2) + (c >> 1)) * 4, y = -64 + 8 * i;
            s_main[c][i] = T.column(bx, y, bz, s_tp[c], T.blended(bx, y, bz));
        } else {
            const int jj = j - kCornerJobs;
            const int c = jj / 147, r = jj - c * 147, w = r / 49, i = r - w * 49;
            const int bx = ((x >> 2) + (c & 1)) * 4, bz = ((z >> 2) + (c >> 1)) * 4;
            s_nood[c][w][i] = T.noodle(w + 1, bx, -64 + 8 * i, bz);
        }
    }
    __syncthreads();

    // C: the final density of every block of the column (phase2.c:161-194), and the highest solid one
    const double dx = (double)(x & 3) / 4.0, dz = (double)(z & 3) / 4.0;
    for (int yi = t; yi < 384; yi += 256) {
        const int y = -64 + yi;
        const int cy = yi / 8;
        const double dy = (double)(yi - cy * 8) / 8.0;
        const double v = cell_lerp(dx, dy, dz, s_main[0], s_main[1], s_main[2], s_main[3], cy);
        const double a = jclamp(v * 0.64, -1.0, 1.0);
        const double squeezed = a / 2.0 - a * a * a / 24.0;
        double ch[3];
        for (int w = 0; w < 3; w++) ch[w] = cell_lerp(dx, dy, dz, s_nood[0][w], s_nood[1][w], s_nood[2][w], s_nood[3][w], cy);
        const double n = T.noodle(0, x, y, z);
        const double noodle = (n >= -1000000.0 && n < 0.0) ? 64.0 : ch[0] + 1.5 * jmax(fabs(ch[1]), fabs(ch[2]));
        const double d = jmin(squeezed, noodle);
        s_dens[yi] = d;
        if (d > 0.0) atomicMax(&s_top_solid, y);
    }
    __syncthreads();

    // D: blocks above the top solid one are air or fluid: at or below skipSamplingAboveY the aquifer decides
    // (a barrier can also make one solid), above it the sea (water below y = 63). Top-down rounds of 256 blocks.
    const int y_d = s_top_solid;
    const int y_hi = min(skip_above, 319);
    for (int y_top = y_hi; y_top > y_d; y_top -= 256) {
        const int y = y_top - t;
        int st = BS_AIR;
        if (y > y_d) {
            st = T.substance(x, y, z, s_dens[y + 64], skip_above);
            if (st != BS_AIR) atomicMax(&s_best, y);
        }
        __syncthreads();
        const int best = s_best;
        __syncthreads();   // every lane has read s_best before a next round could raise it
        if (best != -0x7FFFFFFF - 1) {
            if (y == best) s_state = st;
            break;
        }
    }
    __syncthreads();
    if (t == 0) {
        int h = y_d, st = BS_SOLID;
        if (s_best != -0x7FFFFFFF - 1 && s_best > h) { h = s_best; st = s_state; }
        if (skip_above < 62 && 62 > h) { h = 62; st = BS_WATER; }   // the sea above the aquifer's reach
        int2 o;
        o.x = h >= -64 ? h + 1 : -64;
        o.y = h >= -64 ? st : BS_AIR;
        out[qi] = o;
    }
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

size_t ter_carve(uint8_t *base, TerBufs &b) {
    size_t o = 0;
    auto take = [&](size_t bytes) { uint8_t *p = base ? base + o : nullptr; o += align256(bytes); return p; };
    b.tt = (TerrainTables *)take(2 * sizeof(TerrainTables));
    b.bd = (BuildDesc *)take(2 * kTerSamplers * sizeof(BuildDesc));
    b.spline = (SplineTab *)take(sizeof(SplineTab));
    b.tab = take((size_t)kTerBatch * kTerSamplers * 256);
    b.par = (SPar *)take((size_t)kTerBatch * kTerSamplers * sizeof(SPar));
    b.ltab = take((size_t)kTerBatch * kLegacyPerlins * 256);
    b.lpar = (SPar *)take((size_t)kTerBatch * kLegacyPerlins * sizeof(SPar));
    b.aq = (ulonglong2 *)take(kTerBatch * sizeof(ulonglong2));
    b.seeds = (uint64_t *)take(kTerBatch * sizeof(uint64_t));
    b.q = (whf_column *)take(kTerBatch * sizeof(whf_column));
    b.slot = (uint32_t *)take(kTerBatch * sizeof(uint32_t));
    b.out = (int2 *)take(kTerBatch * sizeof(int2));
    return o;
}

// The 31 noises of the stage for a rule-flag set, their samplers in one set (the biome stage's six first), and
// their build descriptors.
void ter_layout(int large, TerrainTables &tt, std::vector<BuildDesc> &bd) {
    const int pid[T_COUNT] = {
        N_OFFSET, large ? N_CONTINENTALNESS_LARGE : N_CONTINENTALNESS, large ? N_EROSION_LARGE : N_EROSION, N_RIDGE,
        large ? N_TEMPERATURE_LARGE : N_TEMPERATURE, large ? N_VEGETATION_LARGE : N_VEGETATION,
        N_JAGGED, N_CAVE_ENTRANCE, N_SPAG_ROUGH, N_SPAG_ROUGH_MOD, N_SPAG3D_1, N_SPAG3D_2, N_SPAG3D_RARITY,
        N_SPAG3D_THICKNESS, N_SPAG2D, N_SPAG2D_ELEVATION, N_SPAG2D_MODULATOR, N_SPAG2D_THICKNESS, N_PILLAR,
        N_PILLAR_RARENESS, N_PILLAR_THICKNESS, N_CAVE_LAYER, N_CAVE_CHEESE, N_NOODLE, N_NOODLE_THICKNESS,
        N_NOODLE_RIDGE_A, N_NOODLE_RIDGE_B, N_AQUIFER_BARRIER, N_FLOODEDNESS, N_AQUIFER_SPREAD, N_AQUIFER_LAVA};
    std::vector<SamplerRef> set;
    for (int n = 0; n < T_COUNT; n++)
        for (int s = 0; s < 2; s++)
            for (int i = 0; i < kNoise[pid[n]].n; i++)
                if (kNoise[pid[n]].amps[i] != 0.0) set.push_back({pid[n], s, i});
    std::memset(&tt, 0, sizeof tt);
    for (int n = 0; n < T_COUNT; n++) fill_noise_desc(tt.noise[n], pid[n], set, SET_COUNT);
    bd.clear();
    for (const SamplerRef &r : set) bd.push_back(make_build_desc(r));
}

int ter_init(whf_ctx *c, TerBufs &b) {
    if (c->d_ter) {
        ter_carve((uint8_t *)c->d_ter, b);
        return WHF_OK;
    }
    if (whf_gd_spline_depth > kSplineMaxDepth || whf_gd_n_spline_nodes > 128 || whf_gd_n_spline_floats > 640 ||
        whf_gd_n_spline_child > 384) {
        std::snprintf(g_last_error, sizeof g_last_error, "terrain spline exceeds the device layout");
        return WHF_ENOTSUP;
    }
    const size_t bytes = ter_carve(nullptr, b);
    void *base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return WHF_ENOMEM;
    }
    c->d_ter = base;
    c->ter_bytes = bytes;
    ter_carve((uint8_t *)base, b);
    const hipStream_t st = c->stream;
    TerrainTables tt[2];
    std::vector<BuildDesc> bd[2];
    for (int l = 0; l < 2; l++) {
        ter_layout(l, tt[l], bd[l]);
        if ((int)bd[l].size() != kTerSamplers) {
            std::snprintf(g_last_error, sizeof g_last_error, "terrain sampler set has %zu samplers", bd[l].size());
            return WHF_EDEVICE;
        }
        HIP_TRY(hipMemcpyAsync(b.bd + l * kTerSamplers, bd[l].data(), kTerSamplers * sizeof(BuildDesc),
                               hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(b.tt, tt, sizeof tt, hipMemcpyHostToDevice, st));
    SplineTab sp;
    std::memset(&sp, 0, sizeof sp);
    std::memcpy(sp.node, whf_gd_spline_nodes, whf_gd_n_spline_nodes * sizeof(whf_gd_spline_node));
    std::memcpy(sp.f, whf_gd_spline_floats, whf_gd_n_spline_floats * sizeof(float));
    std::memcpy(sp.child, whf_gd_spline_child, whf_gd_n_spline_child * sizeof(int32_t));
    for (int r = 0; r < 3; r++) sp.root[r] = whf_gd_spline_root[r];
    HIP_TRY(hipMemcpyAsync(b.spline, &sp, sizeof sp, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WHF_OK;
}

}  // namespace

namespace whf {

int terrain_heights(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, int32_t *height, int32_t *state) {
    if (n == 0) return WHF_OK;
    TerBufs b;
    int rc = ter_init(c, b);
    if (rc != WHF_OK) return rc;
    const hipStream_t st = c->stream;
    const int large = (vflags & WHF_V_LARGE) ? 1 : 0;
    uint64_t t_hi, t_lo, a_hi, a_lo;
    md5_longs("minecraft:terrain", t_hi, t_lo);
    md5_longs("minecraft:aquifer", a_hi, a_lo);
    std::vector<uint64_t> seeds;
    std::vector<uint32_t> slot;
    std::vector<int2> res(kTerBatch);
    std::unordered_map<uint64_t, uint32_t> seen;
    for (uint64_t q0 = 0; q0 < n; q0 += kTerBatch) {
        const uint32_t nq = (uint32_t)std::min<uint64_t>(kTerBatch, n - q0);
        seeds.clear();
        seen.clear();
        slot.resize(nq);
        for (uint32_t i = 0; i < nq; i++) {
            const uint64_t s = (uint64_t)cols[q0 + i].seed;
            auto it = seen.find(s);
            if (it == seen.end()) {
                it = seen.emplace(s, (uint32_t)seeds.size()).first;
                seeds.push_back(s);
            }
            slot[i] = it->second;
        }
        const uint32_t ns = (uint32_t)seeds.size();
        HIP_TRY(hipMemcpyAsync(b.seeds, seeds.data(), ns * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b.q, cols + q0, nq * sizeof(whf_column), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(b.slot, slot.data(), nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        launch_build(c, b.seeds, b.bd + large * kTerSamplers, kTerSamplers, ns, b.tab, b.par, c->d_cnt + CNT_BIO_BUILT);
        hipLaunchKernelGGL(k_build_blended, dim3((ns + 63) / 64), dim3(64), 0, st, b.seeds, ns, t_hi, t_lo, a_hi, a_lo,
                           b.ltab, b.lpar, b.aq);
        hipLaunchKernelGGL(k_surface, dim3(nq), dim3(256), 0, st, b.tt + large, b.spline, b.tab, b.par, b.ltab, b.lpar,
                           b.aq, b.q, b.slot, b.out);
        rc = check_launch();
        if (rc != WHF_OK) return rc;
        HIP_TRY(hipMemcpyAsync(res.data(), b.out, nq * sizeof(int2), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < nq; i++) {
            height[q0 + i] = res[i].x;
            if (state) state[q0 + i] = res[i].y;
        }
    }
    return WHF_OK;
}

}  // namespace whf

extern "C" int whf_surface_heights(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, int32_t *height,
                                   int32_t *state) {
    if (!c || (n && (!cols || !height))) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (n == 0) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    return terrain_heights(c, cols, n, vflags, height, state);
}
