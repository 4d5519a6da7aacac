// whf_terrain.hip — the terrain stage on gfx950: pre- and post-carver surface heights of (seed, column) queries from
// the overworld's interpolated final density, the aquifer and the carvers, and lysh's hut evaluation over them.
//
// Columns are grouped by chunk; what the columns of a chunk share is computed once per chunk:
//   k_build          the 144 named-noise samplers of a seed (climate, density, noodle and aquifer noises)
//   k_build_blended  a wave per seed: the blended noise's 40 samplers, which the reference draws one after another
//                    from one random, a lane each after a jump ahead in that random; and the aquifer's positional
//                    random factory
//   k_quarts         a lane per quart of the chunk's 27 x 19 aquifer quart table (preliminary surface, deep-dark
//                    offset), and a lane per corner column the chunk's columns use (its terrain point)
//   k_cells          a block per chunk: skipSamplingAboveY from the 11 x 11 surface grid, then a lane per aquifer
//                    grid cell in reach (anchor and computeFluid)
//   k_corners        a lane per corner sample: 49 noise-column values and 3 x 49 noodle children per corner column
//   k_claims         a block per chunk: the cells its 17 x 17 x 3 carvers claim (whf_carver.h)
//   k_column         a wave per column: the final density top-down in windows of 64 blocks, the top solid block,
//                    the aquifer's blocks above it (pre-carver height), then below it the carved cells' states
//                    (post-carver height) and the flood test's fluid at y = 62
// Work here is per phase-1 survivor, a hut at a time: a round holds up to 512 chunks so that the device is full.
#include <cstring>
#include <unordered_map>
#include <vector>

#include "whf_carver.h"
#include "whf_internal.h"
#include "whf_terrain.h"

using namespace whf;

namespace {

constexpr uint32_t kChunkBatch = 512;           // chunks (so also distinct seeds) per device round
constexpr uint32_t kColBatch = kChunkBatch * 64; // columns per device round
constexpr int kTerSamplers = 144;                // named-noise samplers per seed
constexpr int kCornerCols = 25;                  // corner columns of a chunk's 4 x 4 cells, (qx - 4 cx) * 5 + qz - 4 cz
constexpr int kClaimWords = 256 * 6;             // a chunk's claimed-cell bits
constexpr int kTunnelQ = 256;                    // cave tunnels queued per chunk (64 bytes each, in LDS)
constexpr int kEllQ = 640;                       // ellipsoids queued per pass (48 bytes each)
constexpr int kCanyonSlots = 4;                  // canyon width tables per chunk in LDS

// One column of a round: its seed slot, position, and its chunk in the round.
struct ColJob {
    uint32_t slot;
    int32_t x, z;
    int32_t chunk;
};

// A seed slot's sampler tables, by value into every kernel.
struct TerArgs {
    const TerrainTables *tt;
    const SplineTab *spline;
    const uint8_t *tab;
    const SPar *par;
    const uint8_t *ltab;
    const SPar *lpar;
    const ulonglong2 *aq;
    __device__ Terr at(uint32_t slot) const {
        const ulonglong2 f = aq[slot];
        return Terr{tt, SetView{tab, par, kTerSamplers}, slot, ltab + (size_t)slot * kLegacyPerlins * 256,
                    lpar + (size_t)slot * kLegacyPerlins, spline, f.x, f.y};
    }
};

// Per-chunk tables of a round; chunk k's rows start at k times the row size.
struct ChunkBufs {
    const int4 *chunks;        // seed slot, chunk x, chunk z, corner-column mask
    int *qsurf;                // [kAqQuarts]
    double *qdd;               // [kAqQuarts]
    int4 *cells;               // [kAqCells]
    int *skip;                 // skipSamplingAboveY
    float4 *tp;                // [kCornerCols] terrain points
    double *cmain;             // [kCornerCols][49]
    double *cnood;             // [kCornerCols][3][49]
    unsigned long long *claims;   // [kClaimWords]
    __device__ ChunkAq aq(uint32_t k) const {
        const int4 ch = chunks[k];
        return ChunkAq{qsurf + (size_t)k * kAqQuarts, qdd + (size_t)k * kAqQuarts, cells + (size_t)k * kAqCells, ch.y,
                       ch.z};
    }
};

struct TerBufs {
    TerrainTables *tt;        // [2]: NORMAL, LARGE_BIOMES
    BuildDesc *bd;            // [2][kTerSamplers]
    SplineTab *spline;
    uint8_t *tab;             // [kChunkBatch][kTerSamplers][256]
    SPar *par;
    uint8_t *ltab;            // [kChunkBatch][40][256]
    SPar *lpar;
    ulonglong2 *aq;           // [kChunkBatch]
    uint64_t *seeds;          // [kChunkBatch]
    float *sin_tab;           // Mth.SIN
    ulonglong2 *jumps;        // [kLegacyPerlins][128]: the engine's 259 k-step jump matrices
    int4 *chunks;             // [kChunkBatch]
    int2 *pairs;              // [kChunkBatch * kCornerCols]: (chunk, corner column) of the round's corner columns
    int *qsurf;
    double *qdd;
    int4 *cells;
    int *skip;
    float4 *tp;
    double *cmain;
    double *cnood;
    unsigned long long *claims;
    ColJob *jobs;             // [kColBatch]
    int4 *out;                // [kColBatch]: pre-carver height, its top state, post-carver height, water above 62
    TerArgs args(int large) const { return TerArgs{tt + large, spline, tab, par, ltab, lpar, aq}; }
    ChunkBufs chunk_bufs() const { return ChunkBufs{chunks, qsurf, qdd, cells, skip, tp, cmain, cnood, claims}; }
};

// The state of an Xoroshiro128 engine after n steps, n fixed: a 128 x 128 matrix over GF(2) (the engine is linear),
// stored as its 128 columns; column j is the image of state bit j (bits 0..63 of lo, then of hi).
__device__ __forceinline__ Xr xr_jump(const ulonglong2 *__restrict__ cols, Xr s) {
    uint64_t lo = 0, hi = 0;
#pragma unroll 4
    for (int j = 0; j < 128; j++) {
        const uint64_t bit = j < 64 ? (s.lo >> j) & 1ull : (s.hi >> (j - 64)) & 1ull;
        const ulonglong2 c = cols[j];
        lo ^= c.x & (0ull - bit);
        hi ^= c.y & (0ull - bit);
    }
    return Xr{lo, hi};
}

__global__ __launch_bounds__(64) void k_build_blended(const uint64_t *__restrict__ seeds, uint64_t t_hi, uint64_t t_lo,
                                                      uint64_t a_hi, uint64_t a_lo, const ulonglong2 *__restrict__ jumps,
                                                      uint8_t *__restrict__ ltab, SPar *__restrict__ lpar,
                                                      ulonglong2 *__restrict__ aq) {
    // a wave per seed. The game builds the blended noise's 40 samplers (lower, upper and main stacks, each from its
    // coarsest octave down) one after another from one random, createRandom("minecraft:terrain") of the world's
    // deriver (interp_noise.c:19-48, noise.c:167-209). A sampler takes 3 + 256 draws of one engine step each unless
    // a bounded draw rejects (odds below 2^-24 per draw), so lane k jumps to step 259 k and builds sampler k. A lane
    // whose engine did not end where the next lane's began saw a rejection (exactly: the engine's period is
    // 2^128 - 1); its end is where the next sampler starts, so the lanes after it jump again from there.
    __shared__ uint32_t words[64 * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t s = blockIdx.x;
    const uint64_t seed = seeds[s];
    Xr r = xr_seed(seed);
    Xr der = xr_fork(r);
    Xr base = xr_make(t_hi ^ der.lo, t_lo ^ der.hi);   // start of sampler `first`
    uint8_t *tab = ltab + (size_t)s * kLegacyPerlins * 256;
    SPar *par = lpar + (size_t)s * kLegacyPerlins;
    for (uint32_t first = 0;;) {
        Xr start = base, pr = base;
        const bool mine = lane >= first && lane < (uint32_t)kLegacyPerlins;
        if (mine) {
            start = xr_jump(jumps + (size_t)(lane - first) * 128, base);
            pr = start;
            build_sampler(pr, words, lane, tab + (size_t)lane * 256, par[lane]);
        }
        const uint64_t next_lo = __shfl_down(start.lo, 1), next_hi = __shfl_down(start.hi, 1);
        const bool off = mine && lane + 1 < (uint32_t)kLegacyPerlins && (pr.lo != next_lo || pr.hi != next_hi);
        const unsigned long long m = __ballot(off);
        if (!m) break;
        const int k = __ffsll((long long)m) - 1;   // the first lane that ended off: its samplers are right
        base.lo = __shfl(pr.lo, k);
        base.hi = __shfl(pr.hi, k);
        first = (uint32_t)k + 1;
    }
    if (lane == 0) {
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
}

__global__ __launch_bounds__(64) void k_quarts(TerArgs ta, ChunkBufs cb, uint32_t nch, const int2 *__restrict__ pairs,
                                               uint32_t npairs) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i < nch * kAqQuarts) {
        // the preliminary surface and deep-dark offset of a quart (column_top.c:85-122, aquifer.c:218-228)
        const uint32_t k = i / kAqQuarts, q = i - k * kAqQuarts;
        const int4 ch = cb.chunks[k];
        const int qi = (int)q / kAqQz, qj = (int)q - qi * kAqQz;
        int surface;
        double dd;
        ta.at((uint32_t)ch.x).quart_surface(4 * ch.y - 16 + qi, 4 * ch.z - 8 + qj, surface, dd);
        cb.qsurf[i] = surface;
        cb.qdd[i] = dd;
    } else if (i < nch * kAqQuarts + npairs) {
        // the terrain point of a corner column (phase2.c:71-83)
        const int2 p = pairs[i - nch * kAqQuarts];
        const int4 ch = cb.chunks[p.x];
        float t[3];
        ta.at((uint32_t)ch.x).point261(4 * ch.y + p.y / 5, 4 * ch.z + p.y % 5, t);
        cb.tp[(size_t)p.x * kCornerCols + p.y] = make_float4(t[0], t[1], t[2], 0.0f);
    }
}

__global__ __launch_bounds__(256) void k_cells(TerArgs ta, ChunkBufs cb) {
    __shared__ int s_psl;
    const uint32_t k = blockIdx.x;
    const int t = threadIdx.x;
    const int4 ch = cb.chunks[k];
    if (t == 0) s_psl = -0x7FFFFFFF - 1;
    __syncthreads();
    // skipSamplingAboveY (aquifer.c:160-196): the preliminary surfaces on the 11 x 11 grid of step 4 from the
    // chunk's aquifer grid corner (cx - 1, cz - 1), quarts 4 cx - 4 + u, 4 cz - 4 + v of the quart table
    const int *qs = cb.qsurf + (size_t)k * kAqQuarts;
    if (t < 121) atomicMax(&s_psl, qs[(12 + t % 11) * kAqQz + 4 + t / 11]);
    __syncthreads();
    const int skip_above = (floor_div((s_psl + 8) + 12, 12) + 1) * 12 + 11 - 1;
    if (t == 0) cb.skip[k] = skip_above;
    // anchors and fluids of the grid cells computeSubstance can read: y <= min(skipSamplingAboveY, 319)
    const int gy_hi = min(kAqGy0 + kAqGyN - 1, floor_div(min(skip_above, 319) + 1, 12) + 1);
    const int n = (gy_hi - kAqGy0 + 1) * 9;
    const Terr T = ta.at((uint32_t)ch.x);
    const ChunkAq A = cb.aq(k);
    for (int c = t; c < n; c += 256) {
        const int gy = kAqGy0 + c / 9, r = c % 9;
        int ax, ay, az, lv, ty;
        T.anchor(ch.y - 1 + r % 3, gy, ch.z - 1 + r / 3, ax, ay, az);
        T.fluid(A, ax, ay, az, lv, ty);
        cb.cells[(size_t)k * kAqCells + c] = make_int4(ax, ay, az, lv * 4 + ty);
    }
}

__global__ __launch_bounds__(64) void k_corners(TerArgs ta, ChunkBufs cb, const int2 *__restrict__ pairs,
                                                uint32_t npairs) {
    // corner samples at y = -64 + 8 i (phase2.c:108-130): the noise column over the blended noise, and the three
    // noodle children
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const uint32_t nm = npairs * 49;
    if (i < nm) {
        const uint32_t p = i / 49, yi = i - p * 49;
        const int2 pr = pairs[p];
        const int4 ch = cb.chunks[pr.x];
        const Terr T = ta.at((uint32_t)ch.x);
        const int bx = (4 * ch.y + pr.y / 5) * 4, bz = (4 * ch.z + pr.y % 5) * 4, y = -64 + 8 * (int)yi;
        const float4 t4 = cb.tp[(size_t)pr.x * kCornerCols + pr.y];
        const float tp[3] = {t4.x, t4.y, t4.z};
        cb.cmain[((size_t)pr.x * kCornerCols + pr.y) * 49 + yi] = T.column(bx, y, bz, tp, T.blended(bx, y, bz));
    } else if (i < nm + npairs * 147) {
        const uint32_t j = i - nm, p = j / 147, r = j - p * 147, w = r / 49, yi = r - w * 49;
        const int2 pr = pairs[p];
        const int4 ch = cb.chunks[pr.x];
        const Terr T = ta.at((uint32_t)ch.x);
        const int bx = (4 * ch.y + pr.y / 5) * 4, bz = (4 * ch.z + pr.y % 5) * 4;
        cb.cnood[((size_t)pr.x * kCornerCols + pr.y) * 147 + w * 49 + yi] = T.noodle((int)w + 1, bx, -64 + 8 * (int)yi, bz);
    }
}

__global__ __launch_bounds__(256) void k_claims(const float *__restrict__ sin_tab, const uint64_t *__restrict__ seeds,
                                                const int4 *__restrict__ chunks,
                                                unsigned long long *__restrict__ claims) {
    // the cells of one target chunk that its 17 x 17 origin chunks' three carvers claim (carver.c:498-523). No lane
    // walks a whole cave system or claims a whole ellipsoid: (1) the carvers' start rolls, cave rooms and canyon walks,
    // queueing each cave tunnel; (2) the queued tunnels, queueing their branch tunnels; (3) the branch tunnels. Every
    // pass queues the ellipsoids within reach, and then each wave claims the cells of every fourth queued ellipsoid,
    // a column per lane. The claimed set is a union, so the order changes nothing; a piece that finds its queue full
    // runs in the lane that made it.
    __shared__ unsigned long long bits[kClaimWords];
    __shared__ TunnelJob tq[kTunnelQ];
    __shared__ Ell eq[kEllQ];
    __shared__ float cw[kCanyonSlots][384];
    __shared__ int s_top, s_all, s_ell, s_can;
    const int t = threadIdx.x;
    const int4 ch = chunks[blockIdx.x];
    for (int i = t; i < kClaimWords; i += 256) bits[i] = 0ull;
    if (t == 0) {
        s_top = 0;
        s_ell = 0;
        s_can = 0;
    }
    __syncthreads();
    const CarveClaims cc{sin_tab, bits, ch.y * 16, ch.z * 16, (double)(ch.y * 16 + 8), (double)(ch.z * 16 + 8)};
    const uint64_t world = seeds[ch.x];
    auto push_ell = [&](const Ell &e, const float *width) {
        if (e.canyon != kOwnWidths) {
            const int at = atomicAdd(&s_ell, 1);
            if (at < kEllQ) {
                eq[at] = e;
                return;
            }
        }
        cc.claim(e, width, 0, 1);
    };
    auto widths = [&](int &id) -> float * {
        const int at = atomicAdd(&s_can, 1);
        if (at >= kCanyonSlots) return nullptr;
        id = at;
        return cw[at];
    };
    auto flush = [&]() {
        __syncthreads();
        const int n = min(s_ell, kEllQ);
        for (int e = t >> 6; e < n; e += 4) {
            const Ell el = eq[e];
            cc.claim(el, el.canyon >= 0 ? cw[el.canyon] : nullptr, t & 63, 64);
        }
        __syncthreads();
        if (t == 0) s_ell = 0;
        __syncthreads();
    };
    auto no_branch = [](const TunnelJob &) {};
    auto run_branch = [&](const TunnelJob &b) { cc.tunnel<false>(b, no_branch, push_ell); };
    auto push_top = [&](const TunnelJob &j) {
        const int at = atomicAdd(&s_top, 1);
        if (at < kTunnelQ) tq[at] = j;
        else cc.tunnel<true>(j, run_branch, push_ell);
    };
    for (int job = t; job < 17 * 17 * 3; job += 256) {
        const int o = job / 3, ci = job - o * 3;
        cc.carver(world, ch.y + o / 17 - 8, ch.z + o % 17 - 8, ci, push_top, push_ell, widths);
    }
    flush();
    const int n_top = min(s_top, kTunnelQ);
    if (t == 0) s_all = n_top;
    __syncthreads();
    auto push_branch = [&](const TunnelJob &b) {
        const int at = atomicAdd(&s_all, 1);
        if (at < kTunnelQ) tq[at] = b;
        else cc.tunnel<false>(b, no_branch, push_ell);
    };
    for (int i = t; i < n_top; i += 256) {
        const TunnelJob j = tq[i];
        cc.tunnel<true>(j, push_branch, push_ell);
    }
    flush();
    const int n_all = min(s_all, kTunnelQ);
    for (int i = n_top + t; i < n_all; i += 256) {
        const TunnelJob j = tq[i];
        cc.tunnel<false>(j, no_branch, push_ell);
    }
    flush();
    unsigned long long *dst = claims + (size_t)blockIdx.x * kClaimWords;
    for (int i = t; i < kClaimWords; i += 256) dst[i] = bits[i];
}

template <bool kPost>
__global__ __launch_bounds__(64) void k_column(TerArgs ta, ChunkBufs cb, const ColJob *__restrict__ jobs,
                                               int4 *__restrict__ out) {
    __shared__ double s_dens[384];   // final density of y = -64 + i, filled top-down
    const ColJob col = jobs[blockIdx.x];
    const int lane = threadIdx.x;
    const uint32_t k = (uint32_t)col.chunk;
    const int4 ch = cb.chunks[k];
    const Terr T = ta.at(col.slot);
    const ChunkAq A = cb.aq(k);
    const int skip_above = cb.skip[k];
    const int x = col.x, z = col.z;
    // the cell's corner columns in x-then-z order (phase2.c:132-150)
    const int c0 = ((x >> 2) - 4 * ch.y) * 5 + ((z >> 2) - 4 * ch.z);
    const double *m00 = cb.cmain + ((size_t)k * kCornerCols + c0) * 49;
    const double *n00 = cb.cnood + ((size_t)k * kCornerCols + c0) * 147;
    const double dx = (double)(x & 3) / 4.0, dz = (double)(z & 3) / 4.0;

    // the final density (phase2.c:161-194) of y in [lo, 319], computed in windows of 64 from the top as needed
    int dens_lo = 320;
    auto fill_to = [&](int y_lo) {
        y_lo = max(y_lo, -64);
        while (dens_lo > y_lo) {
            const int y = dens_lo - 1 - lane;
            if (y >= -64) {
                const int yi = y + 64, cy = yi / 8;
                const double dy = (double)(yi - cy * 8) / 8.0;
                const double v = cell_lerp(dx, dy, dz, m00, m00 + 5 * 49, m00 + 49, m00 + 6 * 49, cy);
                const double a = jclamp(v * 0.64, -1.0, 1.0);
                const double squeezed = a / 2.0 - a * a * a / 24.0;
                const double n = T.noodle(0, x, y, z);
                double noodle = 64.0;
                if (!(n >= -1000000.0 && n < 0.0)) {
                    double chv[3];
                    for (int w = 0; w < 3; w++)
                        chv[w] = cell_lerp(dx, dy, dz, n00 + w * 49, n00 + 5 * 147 + w * 49, n00 + 147 + w * 49,
                                           n00 + 6 * 147 + w * 49, cy);
                    noodle = chv[0] + 1.5 * jmax(fabs(chv[1]), fabs(chv[2]));
                }
                s_dens[yi] = jmin(squeezed, noodle);
            }
            dens_lo = max(dens_lo - 64, -64);
        }
        __syncthreads();
    };

    // the top solid block: the highest positive density
    int y_d = -65;
    for (int y_top = 319; y_top >= -64; y_top -= 64) {
        fill_to(y_top - 63);
        const int y = y_top - lane;
        const unsigned long long m = __ballot(y >= -64 && s_dens[y + 64] > 0.0);
        if (m) {
            y_d = y_top - (__ffsll((long long)m) - 1);
            break;
        }
    }

    // above it the aquifer decides each block up to skipSamplingAboveY (a barrier can make one solid), the sea
    // above that (water below y = 63); the height is one above the highest block that is not air
    int h = y_d, st = BS_SOLID;
    for (int y_top = min(skip_above, 319); y_top > y_d; y_top -= 64) {
        const int y = y_top - lane;
        const int s = y > y_d ? T.substance(A, x, y, z, s_dens[y + 64], skip_above) : BS_AIR;
        const unsigned long long m = __ballot(s != BS_AIR);
        if (m) {
            const int l = __ffsll((long long)m) - 1;
            h = y_top - l;
            st = __shfl(s, l);
            break;
        }
    }
    if (skip_above < 62 && 62 > h) { h = 62; st = BS_WATER; }
    const int pre = h >= -64 ? h + 1 : -64;
    int post = pre, wet = 0;

    if constexpr (kPost) {
        // post-carver surface (phase2.c:337-345, 438-456): below the pre-carver surface, a claimed cell whose block
        // is solid or water becomes lava at y <= -56 and elsewhere the aquifer's substance at density 0, unless that
        // is solid (carver.c:145-175); other cells keep their block state
        const unsigned long long *cw = cb.claims + (size_t)k * kClaimWords + (size_t)(((z & 15) * 16 + (x & 15)) * 6);
        post = -64;
        for (int y_top = pre - 1; y_top >= -64; y_top -= 64) {
            fill_to(y_top - 63);
            const int y = y_top - lane;
            int bs = BS_AIR;
            if (y >= -64) {
                const double d = s_dens[y + 64];
                bs = d > 0.0 ? BS_SOLID : T.substance(A, x, y, z, d, skip_above);
                const int b = y + 64;
                if (((cw[b >> 6] >> (b & 63)) & 1ull) && (bs == BS_SOLID || bs == BS_WATER)) {
                    const int cs = y <= -56 ? BS_LAVA : T.substance(A, x, y, z, 0.0, skip_above);
                    if (cs != BS_SOLID) bs = cs;
                }
            }
            const unsigned long long m = __ballot(bs != BS_AIR);
            if (m) {
                post = y_top - (__ffsll((long long)m) - 1) + 1;
                break;
            }
        }
        // the hut's flood test for this column: water above y = 62 (eval.c:151-168, computeFluid aquifer.c:286-331)
        if (lane == 0) {
            int lv, ty;
            T.fluid(A, x, 62, z, lv, ty);
            wet = (ty == BS_WATER && 62 < lv) ? 1 : 0;
        }
    }
    if (lane == 0) out[blockIdx.x] = make_int4(pre, h >= -64 ? st : BS_AIR, post, wet);
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

size_t ter_carve(uint8_t *base, TerBufs &b) {
    size_t o = 0;
    auto take = [&](size_t bytes) { uint8_t *p = base ? base + o : nullptr; o += align256(bytes); return p; };
    const size_t nc = kChunkBatch;
    b.tt = (TerrainTables *)take(2 * sizeof(TerrainTables));
    b.bd = (BuildDesc *)take(2 * kTerSamplers * sizeof(BuildDesc));
    b.spline = (SplineTab *)take(sizeof(SplineTab));
    b.tab = take(nc * kTerSamplers * 256);
    b.par = (SPar *)take(nc * kTerSamplers * sizeof(SPar));
    b.ltab = take(nc * kLegacyPerlins * 256);
    b.lpar = (SPar *)take(nc * kLegacyPerlins * sizeof(SPar));
    b.aq = (ulonglong2 *)take(nc * sizeof(ulonglong2));
    b.seeds = (uint64_t *)take(nc * sizeof(uint64_t));
    b.sin_tab = (float *)take(65536 * sizeof(float));
    b.jumps = (ulonglong2 *)take((size_t)kLegacyPerlins * 128 * sizeof(ulonglong2));
    b.chunks = (int4 *)take(nc * sizeof(int4));
    b.pairs = (int2 *)take(nc * kCornerCols * sizeof(int2));
    b.qsurf = (int *)take(nc * kAqQuarts * sizeof(int));
    b.qdd = (double *)take(nc * kAqQuarts * sizeof(double));
    b.cells = (int4 *)take(nc * kAqCells * sizeof(int4));
    b.skip = (int *)take(nc * sizeof(int));
    b.tp = (float4 *)take(nc * kCornerCols * sizeof(float4));
    b.cmain = (double *)take(nc * kCornerCols * 49 * sizeof(double));
    b.cnood = (double *)take(nc * kCornerCols * 147 * sizeof(double));
    b.claims = (unsigned long long *)take(nc * kClaimWords * sizeof(unsigned long long));
    b.jobs = (ColJob *)take((size_t)kColBatch * sizeof(ColJob));
    b.out = (int4 *)take((size_t)kColBatch * sizeof(int4));
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

// Host side of the jump matrices: the Xoroshiro128 engine step (whf_device.h xr_next) as a GF(2) matrix, its
// 259th power P, and P^k for k < kLegacyPerlins, each checked against stepping the engine.
struct Gf2Mat {
    uint64_t lo[128], hi[128];   // column j
};
void gf2_apply(const Gf2Mat &m, uint64_t lo, uint64_t hi, uint64_t &olo, uint64_t &ohi) {
    olo = ohi = 0;
    for (int j = 0; j < 128; j++)
        if (j < 64 ? (lo >> j) & 1 : (hi >> (j - 64)) & 1) {
            olo ^= m.lo[j];
            ohi ^= m.hi[j];
        }
}
Gf2Mat gf2_mul(const Gf2Mat &a, const Gf2Mat &b) {   // a after b
    Gf2Mat c;
    for (int j = 0; j < 128; j++) gf2_apply(a, b.lo[j], b.hi[j], c.lo[j], c.hi[j]);
    return c;
}
void engine_step(uint64_t &lo, uint64_t &hi) {
    const uint64_t h = hi ^ lo;
    lo = ((lo << 49) | (lo >> 15)) ^ h ^ (h << 21);
    hi = (h << 28) | (h >> 36);
}
int jump_matrices(std::vector<ulonglong2> &out) {
    Gf2Mat step, id;
    for (int j = 0; j < 128; j++) {
        uint64_t lo = j < 64 ? 1ull << j : 0, hi = j < 64 ? 0 : 1ull << (j - 64);
        id.lo[j] = lo;
        id.hi[j] = hi;
        engine_step(lo, hi);
        step.lo[j] = lo;
        step.hi[j] = hi;
    }
    Gf2Mat p = id, sq = step;
    for (unsigned e = 259; e; e >>= 1) {
        if (e & 1) p = gf2_mul(sq, p);
        sq = gf2_mul(sq, sq);
    }
    out.assign((size_t)kLegacyPerlins * 128, ulonglong2{});
    Gf2Mat m = id;
    uint64_t tlo = 0x0123456789ABCDEFull, thi = 0xFEDCBA9876543210ull;   // a test state stepped alongside
    for (int k = 0; k < kLegacyPerlins; k++) {
        uint64_t jlo, jhi;
        gf2_apply(m, 0x0123456789ABCDEFull, 0xFEDCBA9876543210ull, jlo, jhi);
        if (jlo != tlo || jhi != thi) return WHF_EDEVICE;
        for (int j = 0; j < 128; j++) out[(size_t)k * 128 + j] = ulonglong2{m.lo[j], m.hi[j]};
        m = gf2_mul(p, m);
        for (int i = 0; i < 259; i++) engine_step(tlo, thi);
    }
    return WHF_OK;
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
    std::vector<ulonglong2> jm;
    if (jump_matrices(jm) != WHF_OK) {
        std::snprintf(g_last_error, sizeof g_last_error, "engine jump matrices do not match stepping");
        return WHF_EDEVICE;
    }
    HIP_TRY(hipMemcpyAsync(b.jumps, jm.data(), jm.size() * sizeof(ulonglong2), hipMemcpyHostToDevice, st));
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
    hipLaunchKernelGGL(k_build_blended, dim3(ns), dim3(64), 0, st, (const uint64_t *)b.seeds, t_hi, t_lo, a_hi, a_lo,
                       (const ulonglong2 *)b.jumps, b.ltab, b.lpar, b.aq);
    return check_launch();
}

// A device round: up to kChunkBatch chunks (with their seeds) and kColBatch columns, each column in one of the
// round's chunks.
struct Round {
    struct Key {
        uint64_t seed;
        int cx, cz;
        bool operator==(const Key &o) const { return seed == o.seed && cx == o.cx && cz == o.cz; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            return std::hash<uint64_t>()(k.seed ^ ((uint64_t)(uint32_t)k.cx * 0x9E3779B97F4A7C15ull) ^
                                         ((uint64_t)(uint32_t)k.cz << 32));
        }
    };
    std::vector<uint64_t> seeds;
    std::unordered_map<uint64_t, uint32_t> slot_of;
    std::vector<int4> chunks;   // slot, chunk x, chunk z, mask of the corner columns its columns use
    std::unordered_map<Key, uint32_t, KeyHash> chunk_of;
    std::vector<ColJob> jobs;
    void clear() {
        seeds.clear();
        slot_of.clear();
        chunks.clear();
        chunk_of.clear();
        jobs.clear();
    }
    bool empty() const { return jobs.empty(); }
    // Adds a column; false when the round is full (then nothing is added).
    bool add(uint64_t seed, int x, int z) {
        const Key key{seed, x >> 4, z >> 4};
        auto it = chunk_of.find(key);
        if (jobs.size() == kColBatch || (it == chunk_of.end() && chunks.size() == kChunkBatch)) return false;
        if (it == chunk_of.end()) {
            auto s = slot_of.find(seed);
            if (s == slot_of.end()) {
                s = slot_of.emplace(seed, (uint32_t)seeds.size()).first;
                seeds.push_back(seed);
            }
            it = chunk_of.emplace(key, (uint32_t)chunks.size()).first;
            chunks.push_back(make_int4((int)s->second, key.cx, key.cz, 0));
        }
        int4 &ch = chunks[it->second];
        const int c0 = ((x >> 2) - 4 * ch.y) * 5 + ((z >> 2) - 4 * ch.z);
        ch.w |= (1 << c0) | (1 << (c0 + 1)) | (1 << (c0 + 5)) | (1 << (c0 + 6));
        jobs.push_back(ColJob{(uint32_t)ch.x, x, z, (int32_t)it->second});
        return true;
    }
};

// Runs a round: res[j] = {pre-carver height, its top state, post-carver height (kPost), water above 62 (kPost)}.
int run_round(whf_ctx *c, TerBufs &b, int large, bool post, const Round &r, std::vector<int4> &res) {
    const hipStream_t st = c->stream;
    int rc = build_round(c, b, r.seeds, large);
    if (rc != WHF_OK) return rc;
    std::vector<int2> pairs;
    for (size_t k = 0; k < r.chunks.size(); k++)
        for (int q = 0; q < kCornerCols; q++)
            if ((r.chunks[k].w >> q) & 1) pairs.push_back(make_int2((int)k, q));
    const uint32_t nch = (uint32_t)r.chunks.size(), np = (uint32_t)pairs.size(), nj = (uint32_t)r.jobs.size();
    HIP_TRY(hipMemcpyAsync(b.chunks, r.chunks.data(), nch * sizeof(int4), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b.pairs, pairs.data(), np * sizeof(int2), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b.jobs, r.jobs.data(), nj * sizeof(ColJob), hipMemcpyHostToDevice, st));
    const TerArgs ta = b.args(large);
    const ChunkBufs cb = b.chunk_bufs();
    hipLaunchKernelGGL(k_quarts, dim3((nch * kAqQuarts + np + 63) / 64), dim3(64), 0, st, ta, cb, nch,
                       (const int2 *)b.pairs, np);
    hipLaunchKernelGGL(k_cells, dim3(nch), dim3(256), 0, st, ta, cb);
    hipLaunchKernelGGL(k_corners, dim3((np * 196 + 63) / 64), dim3(64), 0, st, ta, cb, (const int2 *)b.pairs, np);
    if (post) {
        hipLaunchKernelGGL(k_claims, dim3(nch), dim3(256), 0, st, (const float *)b.sin_tab, (const uint64_t *)b.seeds,
                           (const int4 *)b.chunks, b.claims);
        hipLaunchKernelGGL(k_column<true>, dim3(nj), dim3(64), 0, st, ta, cb, (const ColJob *)b.jobs, b.out);
    } else {
        hipLaunchKernelGGL(k_column<false>, dim3(nj), dim3(64), 0, st, ta, cb, (const ColJob *)b.jobs, b.out);
    }
    if ((rc = check_launch()) != WHF_OK) return rc;
    res.resize(nj);
    HIP_TRY(hipMemcpyAsync(res.data(), b.out, nj * sizeof(int4), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return WHF_OK;
}

// Columns in rounds; emit(i, res) receives column i's result.
template <class Emit>
int columns_in_rounds(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, bool post, Emit emit) {
    TerBufs b;
    int rc = ter_init(c, b);
    if (rc != WHF_OK) return rc;
    const int large = (vflags & WHF_V_LARGE) ? 1 : 0;
    Round r;
    std::vector<uint64_t> idx;
    std::vector<int4> res;
    auto flush = [&]() -> int {
        if (r.empty()) return WHF_OK;
        const int rc2 = run_round(c, b, large, post, r, res);
        if (rc2 != WHF_OK) return rc2;
        for (size_t j = 0; j < idx.size(); j++) emit(idx[j], res[j]);
        r.clear();
        idx.clear();
        return WHF_OK;
    };
    for (uint64_t i = 0; i < n; i++) {
        if (!r.add((uint64_t)cols[i].seed, cols[i].x, cols[i].z)) {
            if ((rc = flush()) != WHF_OK) return rc;
            r.add((uint64_t)cols[i].seed, cols[i].x, cols[i].z);
        }
        idx.push_back(i);
    }
    return flush();
}

}  // namespace

namespace whf {

int terrain_heights(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, int32_t *height, int32_t *state) {
    if (n == 0) return WHF_OK;
    return columns_in_rounds(c, cols, n, vflags, false, [&](uint64_t i, const int4 &o) {
        height[i] = o.x;
        if (state) state[i] = o.y;
    });
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

extern "C" int whf_post_carver_heights(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, int32_t *post,
                                       int32_t *pre) {
    if (!c || (n && (!cols || !post))) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (n == 0) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    return columns_in_rounds(c, cols, n, vflags, true, [&](uint64_t i, const int4 &o) {
        post[i] = o.z;
        if (pre) pre[i] = o.x;
    });
}

extern "C" int whf_hut_eval(whf_ctx *c, const whf_hut_query *huts, uint64_t n, int tree, int vflags, int max_y,
                            whf_hut_result *out) {
    if (!c || (n && (!huts || !out))) return WHF_EINVAL;
    if (tree < 0 || tree >= WHF_GD_TREES) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (n == 0) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    TerBufs b;
    int rc = ter_init(c, b);
    if (rc != WHF_OK) return rc;
    const int large = (vflags & WHF_V_LARGE) ? 1 : 0;
    // per round of huts: the 63 footprint columns of each, and its chunk middle, whose pre-carver surface is the
    // biome check's height (eval.c:141-143); all in the hut's chunk
    std::vector<whf_biome_query> bq(n);
    Round r;
    std::vector<int4> res;
    for (uint64_t h0 = 0; h0 < n; h0 += kChunkBatch) {
        const uint32_t nh = (uint32_t)std::min<uint64_t>(kChunkBatch, n - h0);
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
            for (int dx = 0; dx < o.size_x; dx++)
                for (int dz = 0; dz < o.size_z; dz++) r.add((uint64_t)q.seed, q.cx * 16 + dx, q.cz * 16 + dz);
            r.add((uint64_t)q.seed, q.cx * 16 + 8, q.cz * 16 + 8);
        }
        if ((rc = run_round(c, b, large, true, r, res)) != WHF_OK) return rc;
        size_t j = 0;
        for (uint32_t k = 0; k < nh; k++) {
            const whf_hut_query &q = huts[h0 + k];
            whf_hut_result &o = out[h0 + k];
            long long sum_pre = 0, sum_post = 0;
            bool blocked = false;
            for (int col = 0; col < 63; col++, j++) {
                o.wet += res[j].w;
                sum_pre += res[j].x;
                sum_post += res[j].z;
                if (res[j].x - 1 >= 64) blocked = true;
            }
            bq[h0 + k] = whf_biome_query{q.seed, q.cx * 16, q.cz * 16, res[j].x - 1, 0};
            j++;
            o.flooded = o.wet == 63;
            // a flooded hut whose pre-carver top blocks are all at or below 63 keeps the pre-carver sum
            // (eval.c:186-205)
            o.carve_skipped = o.flooded && !blocked;
            o.sum_h = o.carve_skipped ? sum_pre : sum_post;
            o.avg_y = (int32_t)(o.sum_h / 63);
        }
    }
    std::vector<whf_biome_result> br(n);
    if ((rc = whf_biome(c, bq.data(), n, tree, vflags, br.data(), nullptr)) != WHF_OK) return rc;
    for (uint64_t i = 0; i < n; i++) {
        whf_hut_result &o = out[i];
        o.biome_ok = br[i].ok;
        o.winner = br[i].winner;
        o.occ_y = br[i].occ_y;
        // lysh's reject order: biome, flood, height (eval.c:221-224)
        if (!o.biome_ok) { o.ok = 0; o.reject = WHF_HUT_REJECT_BIOME; }
        else if (o.flooded) { o.ok = 0; o.reject = WHF_HUT_REJECT_FLOOD; }
        else if (o.avg_y > max_y) { o.ok = 0; o.reject = WHF_HUT_REJECT_Y; }
        else { o.ok = 1; o.reject = WHF_HUT_OK; }
    }
    return WHF_OK;
}
