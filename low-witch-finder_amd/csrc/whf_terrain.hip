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
//                         height is one above the highest block that is not air;
//                      E  (hut evaluation) below that surface, the cells the carvers claim turn into the carver's
//                         air or fluid where their block allows it; the post-carver height is one above the highest
//                         block that is then not air; and whether water stands above y = 62 in the column.
//   k_claims         one 256-lane block per target chunk: the cells its 17 x 17 x 3 carvers claim (whf_carver.h)
// This is work per phase-1 survivor (one in 1e5..1e7 regions), so it is laid out for clarity of the reference's
// order of operations, not for throughput: every lane calls the same pointwise functions the reference calls.
#include <cstring>
#include <unordered_map>
#include <vector>

#include "whf_carver.h"
#include "whf_internal.h"
#include "whf_terrain.h"

using namespace whf;

namespace {

constexpr uint32_t kTerBatch = 1024;   // columns per device round (distinct seeds per round at most this)
constexpr int kTerSamplers = 144;      // named-noise samplers per seed
constexpr int kCornerJobs = 4 * 49;    // main corner samples
constexpr int kNoodleJobs = 4 * 3 * 49;
constexpr uint32_t kHutBatch = 16;     // huts per device round (16 x 63 columns fit one round of columns)
constexpr int kClaimWords = 256 * 6;   // a chunk's claimed-cell bits

// One column of a round: its seed slot, position, and the target chunk whose carver claims apply (-1: none).
struct ColJob {
    uint32_t slot;
    int32_t x, z;
    int32_t chunk;
};

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
    ColJob *jobs;             // [kTerBatch]
    int4 *out;                // [kTerBatch]: pre-carver height, its top state, post-carver height, water above 62
    float *sin_tab;           // Mth.SIN
    int4 *chunks;             // [kHutBatch]: seed slot, chunk x, chunk z
    unsigned long long *claims;   // [kHutBatch][kClaimWords]
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

template <bool kPost>
__global__ __launch_bounds__(256) void k_surface(const TerrainTables *__restrict__ tt, const SplineTab *__restrict__ S,
                                                 const uint8_t *__restrict__ tab, const SPar *__restrict__ par,
                                                 const uint8_t *__restrict__ ltab, const SPar *__restrict__ lpar,
                                                 const ulonglong2 *__restrict__ aq, const ColJob *__restrict__ jobs,
                                                 const unsigned long long *__restrict__ claims, int4 *__restrict__ out) {
    __shared__ double s_main[4][49];
    __shared__ double s_nood[4][3][49];
    __shared__ double s_dens[384];
    __shared__ float s_tp[4][3];
    __shared__ int s_psl, s_top_solid, s_best, s_state, s_wet;
    const uint32_t qi = blockIdx.x;
    const int t = threadIdx.x;
    const ColJob col = jobs[qi];
    const uint32_t slot = col.slot;
    const ulonglong2 f = aq[slot];
    const Terr T{tt, SetView{tab, par, kTerSamplers}, slot, ltab + (size_t)slot * kLegacyPerlins * 256,
                 lpar + (size_t)slot * kLegacyPerlins, S, f.x, f.y};
    const int x = col.x, z = col.z;
    if (t == 0) {
        s_psl = -0x7FFFFFFF - 1;
        s_top_solid = -65;
        s_best = -0x7FFFFFFF - 1;
        s_state = BS_AIR;
        s_wet = 0;
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
    } else if (kPost && t == 192) {
        // the hut's flood test for this column: water above y = 62 (eval.c:151-168, computeFluid aquifer.c:286-331)
        int lv, ty;
        T.fluid(x, 62, z, lv, ty);
        s_wet = (ty == BS_WATER && 62 < lv) ? 1 : 0;
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
    int h = y_d, st = BS_SOLID;   // uniform: every lane reads the same shared values
    if (s_best != -0x7FFFFFFF - 1 && s_best > h) { h = s_best; st = s_state; }
    if (skip_above < 62 && 62 > h) { h = 62; st = BS_WATER; }   // the sea above the aquifer's reach
    const int pre = h >= -64 ? h + 1 : -64;
    int post = pre;

    if constexpr (kPost) {
        // E: post-carver surface (phase2.c:337-345, 438-456): below the pre-carver surface, a claimed cell whose
        // block is solid or water becomes lava at y <= -56 and elsewhere the aquifer's substance at density 0,
        // unless that is solid (carver.c:145-175); other cells keep their block state. Top-down rounds of 256.
        const unsigned long long *cw =
            col.chunk >= 0 ? claims + (size_t)col.chunk * kClaimWords + (size_t)(((z & 15) * 16 + (x & 15)) * 6) : nullptr;
        __syncthreads();
        if (t == 0) s_best = -0x7FFFFFFF - 1;
        __syncthreads();
        int found = -0x7FFFFFFF - 1;
        for (int y_top = pre - 1; y_top >= -64; y_top -= 256) {
            const int y = y_top - t;
            if (y >= -64) {
                const double d = s_dens[y + 64];
                int bs = d > 0.0 ? BS_SOLID : T.substance(x, y, z, d, skip_above);
                const int b = y + 64;
                if (cw && ((cw[b >> 6] >> (b & 63)) & 1ull) && (bs == BS_SOLID || bs == BS_WATER)) {
                    const int cs = y <= -56 ? BS_LAVA : T.substance(x, y, z, 0.0, skip_above);
                    if (cs != BS_SOLID) bs = cs;
                }
                if (bs != BS_AIR) atomicMax(&s_best, y);
            }
            __syncthreads();
            found = s_best;
            __syncthreads();
            if (found != -0x7FFFFFFF - 1) break;
        }
        post = found != -0x7FFFFFFF - 1 ? found + 1 : -64;
    }
    if (t == 0) {
        int4 o;
        o.x = pre;
        o.y = h >= -64 ? st : BS_AIR;
        o.z = post;
        o.w = s_wet;
        out[qi] = o;
    }
}

__global__ __launch_bounds__(256) void k_claims(const float *__restrict__ sin_tab, const uint64_t *__restrict__ seeds,
                                                const int4 *__restrict__ chunks,
                                                unsigned long long *__restrict__ claims) {
    // the cells of one target chunk that its 17 x 17 origin chunks' three carvers claim (carver.c:498-523)
    __shared__ unsigned long long bits[kClaimWords];
    const int4 ch = chunks[blockIdx.x];
    for (int i = threadIdx.x; i < kClaimWords; i += 256) bits[i] = 0ull;
    __syncthreads();
    const CarveClaims cc{sin_tab, bits, ch.y * 16, ch.z * 16, (double)(ch.y * 16 + 8), (double)(ch.z * 16 + 8)};
    const uint64_t world = seeds[ch.x];
    for (int job = threadIdx.x; job < 17 * 17 * 3; job += 256) {
        const int o = job / 3, ci = job - o * 3;
        cc.carver(world, ch.y + o / 17 - 8, ch.z + o % 17 - 8, ci);
    }
    __syncthreads();
    unsigned long long *dst = claims + (size_t)blockIdx.x * kClaimWords;
    for (int i = threadIdx.x; i < kClaimWords; i += 256) dst[i] = bits[i];
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
    b.jobs = (ColJob *)take(kTerBatch * sizeof(ColJob));
    b.out = (int4 *)take(kTerBatch * sizeof(int4));
    b.sin_tab = (float *)take(65536 * sizeof(float));
    b.chunks = (int4 *)take(kHutBatch * sizeof(int4));
    b.claims = (unsigned long long *)take((size_t)kHutBatch * kClaimWords * sizeof(unsigned long long));
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
    HIP_TRY(hipMemcpyAsync(b.sin_tab, whf_gd_mc_sin, 65536 * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WHF_OK;
}

// Sampler tables of a round's distinct seeds (slots 0..ns-1 of the round).
int build_round(whf_ctx *c, TerBufs &b, const std::vector<uint64_t> &seeds, int large) {
    const hipStream_t st = c->stream;
    const uint32_t ns = (uint32_t)seeds.size();
    uint64_t t_hi, t_lo, a_hi, a_lo;
    md5_longs("minecraft:terrain", t_hi, t_lo);
    md5_longs("minecraft:aquifer", a_hi, a_lo);
    HIP_TRY(hipMemcpyAsync(b.seeds, seeds.data(), ns * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    launch_build(c, b.seeds, b.bd + large * kTerSamplers, kTerSamplers, ns, b.tab, b.par, c->d_cnt + CNT_BIO_BUILT);
    hipLaunchKernelGGL(k_build_blended, dim3((ns + 63) / 64), dim3(64), 0, st, b.seeds, ns, t_hi, t_lo, a_hi, a_lo,
                       b.ltab, b.lpar, b.aq);
    return check_launch();
}

// Slot of a seed in the round, adding it when new.
uint32_t slot_for(uint64_t seed, std::vector<uint64_t> &seeds, std::unordered_map<uint64_t, uint32_t> &seen) {
    auto it = seen.find(seed);
    if (it == seen.end()) {
        it = seen.emplace(seed, (uint32_t)seeds.size()).first;
        seeds.push_back(seed);
    }
    return it->second;
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
    std::vector<uint64_t> seeds;
    std::vector<ColJob> jobs;
    std::vector<int4> res(kTerBatch);
    std::unordered_map<uint64_t, uint32_t> seen;
    for (uint64_t q0 = 0; q0 < n; q0 += kTerBatch) {
        const uint32_t nq = (uint32_t)std::min<uint64_t>(kTerBatch, n - q0);
        seeds.clear();
        seen.clear();
        jobs.resize(nq);
        for (uint32_t i = 0; i < nq; i++)
            jobs[i] = ColJob{slot_for((uint64_t)cols[q0 + i].seed, seeds, seen), cols[q0 + i].x, cols[q0 + i].z, -1};
        if ((rc = build_round(c, b, seeds, large)) != WHF_OK) return rc;
        HIP_TRY(hipMemcpyAsync(b.jobs, jobs.data(), nq * sizeof(ColJob), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_surface<false>, dim3(nq), dim3(256), 0, st, b.tt + large, b.spline, b.tab, b.par, b.ltab,
                           b.lpar, b.aq, b.jobs, (const unsigned long long *)nullptr, b.out);
        if ((rc = check_launch()) != WHF_OK) return rc;
        HIP_TRY(hipMemcpyAsync(res.data(), b.out, nq * sizeof(int4), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < nq; i++) {
            height[q0 + i] = res[i].x;
            if (state) state[q0 + i] = res[i].y;
        }
    }
    return WHF_OK;
}

// Pre- and post-carver heights and the water-above-62 flag of columns grouped by target chunk: chunk k of the round
// is (chunks[k].seed, chunk x, chunk z); the claims of each chunk are gathered once.
int carved_columns(whf_ctx *c, TerBufs &b, int large, const std::vector<uint64_t> &seeds,
                   const std::vector<int4> &chunks, const std::vector<ColJob> &jobs, std::vector<int4> &res) {
    const hipStream_t st = c->stream;
    int rc = build_round(c, b, seeds, large);
    if (rc != WHF_OK) return rc;
    HIP_TRY(hipMemcpyAsync(b.chunks, chunks.data(), chunks.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b.jobs, jobs.data(), jobs.size() * sizeof(ColJob), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_claims, dim3((uint32_t)chunks.size()), dim3(256), 0, st, b.sin_tab, b.seeds, b.chunks,
                       b.claims);
    hipLaunchKernelGGL(k_surface<true>, dim3((uint32_t)jobs.size()), dim3(256), 0, st, b.tt + large, b.spline, b.tab,
                       b.par, b.ltab, b.lpar, b.aq, b.jobs, (const unsigned long long *)b.claims, b.out);
    if ((rc = check_launch()) != WHF_OK) return rc;
    res.resize(jobs.size());
    HIP_TRY(hipMemcpyAsync(res.data(), b.out, jobs.size() * sizeof(int4), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
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

namespace {

// Rounds of columns grouped by target chunk: at most kHutBatch chunks and kTerBatch columns per round.
struct CarveRound {
    std::vector<uint64_t> seeds;
    std::unordered_map<uint64_t, uint32_t> seen;
    std::vector<int4> chunks;
    std::vector<ColJob> jobs;
    void clear() { seeds.clear(); seen.clear(); chunks.clear(); jobs.clear(); }
};

}  // namespace

extern "C" int whf_post_carver_heights(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, int32_t *post,
                                       int32_t *pre) {
    if (!c || (n && (!cols || !post))) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (n == 0) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    TerBufs b;
    int rc = ter_init(c, b);
    if (rc != WHF_OK) return rc;
    const int large = (vflags & WHF_V_LARGE) ? 1 : 0;
    CarveRound r;
    std::vector<int4> res;
    std::vector<uint64_t> idx;
    auto flush = [&]() -> int {
        if (r.jobs.empty()) return WHF_OK;
        const int rc2 = carved_columns(c, b, large, r.seeds, r.chunks, r.jobs, res);
        if (rc2 != WHF_OK) return rc2;
        for (size_t k = 0; k < r.jobs.size(); k++) {
            post[idx[k]] = res[k].z;
            if (pre) pre[idx[k]] = res[k].x;
        }
        r.clear();
        idx.clear();
        return WHF_OK;
    };
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t seed = (uint64_t)cols[i].seed;
        const int cx = cols[i].x >> 4, cz = cols[i].z >> 4;
        int chunk = -1;
        for (size_t k = 0; k < r.chunks.size(); k++)
            if (r.seeds[r.chunks[k].x] == seed && r.chunks[k].y == cx && r.chunks[k].z == cz) chunk = (int)k;
        if (chunk < 0 && (r.chunks.size() == kHutBatch || r.jobs.size() == kTerBatch)) {
            if ((rc = flush()) != WHF_OK) return rc;
        } else if (r.jobs.size() == kTerBatch) {
            if ((rc = flush()) != WHF_OK) return rc;
            chunk = -1;
        }
        const uint32_t slot = slot_for(seed, r.seeds, r.seen);
        if (chunk < 0) {
            chunk = (int)r.chunks.size();
            r.chunks.push_back(make_int4((int)slot, cx, cz, 0));
        }
        r.jobs.push_back(ColJob{slot, cols[i].x, cols[i].z, chunk});
        idx.push_back(i);
    }
    return flush();
}

extern "C" int whf_hut_eval(whf_ctx *c, const whf_hut_query *huts, uint64_t n, int tree, int vflags, int max_y,
                            whf_hut_result *out) {
    if (!c || (n && (!huts || !out))) return WHF_EINVAL;
    if (tree < 0 || tree >= WHF_GD_TREES) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (n == 0) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    // the biome check at the chunk middle's pre-carver surface (eval.c:141-143)
    std::vector<whf_biome_query> bq(n);
    std::vector<whf_biome_result> br(n);
    for (uint64_t i = 0; i < n; i++) bq[i] = whf_biome_query{huts[i].seed, huts[i].cx * 16, huts[i].cz * 16, WHF_OCC_SURFACE, 0};
    int rc = whf_biome(c, bq.data(), n, tree, vflags, br.data(), nullptr);
    if (rc != WHF_OK) return rc;
    TerBufs b;
    if ((rc = ter_init(c, b)) != WHF_OK) return rc;
    const int large = (vflags & WHF_V_LARGE) ? 1 : 0;
    CarveRound r;
    std::vector<int4> res;
    for (uint64_t h0 = 0; h0 < n; h0 += kHutBatch) {
        const uint32_t nh = (uint32_t)std::min<uint64_t>(kHutBatch, n - h0);
        r.clear();
        for (uint32_t k = 0; k < nh; k++) {
            const whf_hut_query &q = huts[h0 + k];
            whf_hut_result &o = out[h0 + k];
            std::memset(&o, 0, sizeof o);
            // orientation and footprint (eval.c:41-50, 129-132)
            JRand jr;
            jr.large_feature((uint64_t)q.seed, q.cx, q.cz);
            o.dir = jr.next_int(4);
            const bool axis_z = o.dir == 0 || o.dir == 2;
            o.size_x = axis_z ? 7 : 9;
            o.size_z = axis_z ? 9 : 7;
            const uint32_t slot = slot_for((uint64_t)q.seed, r.seeds, r.seen);
            r.chunks.push_back(make_int4((int)slot, q.cx, q.cz, 0));
            for (int dx = 0; dx < o.size_x; dx++)
                for (int dz = 0; dz < o.size_z; dz++)
                    r.jobs.push_back(ColJob{slot, q.cx * 16 + dx, q.cz * 16 + dz, (int)k});
        }
        if ((rc = carved_columns(c, b, large, r.seeds, r.chunks, r.jobs, res)) != WHF_OK) return rc;
        size_t j = 0;
        for (uint32_t k = 0; k < nh; k++) {
            whf_hut_result &o = out[h0 + k];
            const whf_biome_result &bz = br[h0 + k];
            o.biome_ok = bz.ok;
            o.winner = bz.winner;
            o.occ_y = bz.occ_y;
            long long sum_pre = 0, sum_post = 0;
            bool blocked = false;
            for (int col = 0; col < 63; col++, j++) {
                o.wet += res[j].w;
                sum_pre += res[j].x;
                sum_post += res[j].z;
                if (res[j].x - 1 >= 64) blocked = true;
            }
            o.flooded = o.wet == 63;
            // a flooded hut whose pre-carver top blocks are all at or below 63 keeps the pre-carver sum
            // (eval.c:186-205)
            o.carve_skipped = o.flooded && !blocked;
            o.sum_h = o.carve_skipped ? sum_pre : sum_post;
            o.avg_y = (int32_t)(o.sum_h / 63);
            if (!o.biome_ok) { o.ok = 0; o.reject = WHF_HUT_REJECT_BIOME; }
            else if (o.flooded) { o.ok = 0; o.reject = WHF_HUT_REJECT_FLOOD; }
            else if (o.avg_y > max_y) { o.ok = 0; o.reject = WHF_HUT_REJECT_Y; }
            else { o.ok = 1; o.reject = WHF_HUT_OK; }
        }
    }
    return WHF_OK;
}
