// whf_scan.hip — the MI355X (gfx950) witch-hut phase-1 Seed Check: kernels and the C ABI (include/whf.h).
//
// Pipeline per batch of seeds x regions (one HIP stream, no host round trip inside a scan):
//   k_seeds      seed values of the batch (SplitMix64 / counter form on the device, or an uploaded list)
//   k_build      Perlin samplers, one lane per (seed slot, sampler): Xoroshiro derivation + Fisher-Yates in LDS,
//                tables and origins written to HBM. Run for set E (all seeds), set C (seeds with erosion
//                survivors) and set L (seeds with Entrance@50/60 survivors).
//   k_erosion    placement + tiered erosion per (seed, region) with the seed tile's set-E tables in LDS; a
//                per-lane work queue keeps lanes busy through the tier early exits; survivors -> queue Q1.
//   k_compact    seeds flagged by a gate -> dense slot list for the next build.
//   k_gate2a     temperature and ridge per Q1 entry; survivors -> Q2.
//   k_gate2b     Entrance@50, Entrance@60 per Q2 entry; survivors -> Q1 (free again by then).
//   k_gate3      cave@maxHeight, both ladders, continentalness, floodedness per gate-2 survivor; survivors -> hits.
// Every decision follows lysh_phase1_check_ex (engine/tools/lysh/src/phase1.c:130-200) bit for bit, so the
// survivors and the funnel counters equal the reference CPU scanner's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <dlfcn.h>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <type_traits>

#include "whf_internal.h"

using namespace whf;

// ================================================================================================ device

namespace {

// ---------------------------------------------------------------------------------------------- seeds

__global__ void k_seeds(int mode, uint64_t stream, uint64_t offset, uint32_t n, uint64_t *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t idx = offset + i;
    out[i] = mode == WHF_SEEDS_SPLITMIX ? mix64(stream + (idx + 1) * kGolden) : mix64(stream + idx + kGolden);
}

// Zero the per-batch queue counts and gate flags (a fill kernel, so the batch stays capturable). The hit
// count is not reset: hits of all batches of a scan accumulate in one buffer.
__global__ void k_reset(uint32_t *ctl, uint32_t *needC, uint32_t *needL, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < CTL_HITS) ctl[i] = 0;
    if (i < n) { needC[i] = 0; needL[i] = 0; }
}

// ---------------------------------------------------------------------------------------------- build
// One lane per (slot, sampler). The permutation is shuffled in LDS in a lane-interleaved layout (word w of
// lane l at word w * 64 + l), so every byte access of the 64 lanes hits 64 distinct banks whatever the
// random swap index. Job j writes table j of the output (slot-major, set_n tables per slot).

__global__ __launch_bounds__(64) void k_build(const uint64_t *__restrict__ seeds, const BuildDesc *__restrict__ bd,
                                              int set_n, const uint32_t *__restrict__ slot_seed,
                                              const uint32_t *__restrict__ n_slots_dev, uint32_t n_slots_host,
                                              uint8_t *__restrict__ out_tab, SPar *__restrict__ out_par,
                                              unsigned long long *__restrict__ built_cnt,
                                              const uint8_t *__restrict__ lay_of) {
    // lay_of (served batches): the rule layout of each seed slot, whose descriptors follow bd at layout *
    // kBdPerLayout; built_cnt may be null (served batches count their samplers per request elsewhere)
    __shared__ uint32_t words[64 * 64];
    const uint32_t lane = threadIdx.x;
    // with a device count, n_slots_host caps it at the table capacity
    const uint32_t n_slots = n_slots_dev ? min(*n_slots_dev, n_slots_host) : n_slots_host;
    const uint64_t n_jobs = (uint64_t)n_slots * (uint32_t)set_n;
    uint32_t built = 0;
    for (uint64_t j0 = (uint64_t)blockIdx.x * 64; j0 < n_jobs; j0 += (uint64_t)gridDim.x * 64) {
        const uint64_t j = j0 + lane;
        if (j >= n_jobs) break;
        const uint32_t slot = (uint32_t)(j / (uint32_t)set_n);
        const int sid = (int)(j - (uint64_t)slot * (uint32_t)set_n);
        const uint32_t si = slot_seed ? slot_seed[slot] : slot;
        const uint64_t seed = seeds[si];
        const BuildDesc d = bd[(lay_of ? (uint32_t)lay_of[si] * (uint32_t)kBdPerLayout : 0u) + (uint32_t)sid];

        // worldSeed -> deriver -> createRandom(identifier) -> octave deriver (second one for sub 1)
        // -> createRandom("octave_<n>")  (noise.c:80-143, rng.c:157-170)
        Xr r = xr_seed(seed);
        Xr der = xr_fork(r);
        Xr nr = xr_make(d.na ^ der.lo, d.nb ^ der.hi);
        if (d.sub) { xr_next(nr); xr_next(nr); }
        Xr od = xr_fork(nr);
        Xr pr = xr_make(d.oa ^ od.lo, d.ob ^ od.hi);

        // PerlinNoiseSampler(random): origin, then Fisher-Yates over 0..255 (noise.c:61-78)
        build_sampler(pr, words, lane, out_tab + j * 256, out_par[j]);
        built++;
    }
    // one atomic per wave
    unsigned long long tot = built;
    for (int o = 32; o > 0; o >>= 1) tot += __shfl_down(tot, o);
    if (lane == 0 && tot && built_cnt) atomicAdd(built_cnt, tot);
}

// ---------------------------------------------------------------------------------------------- erosion gate
// Tiered erosion (noise.c:432-459) over (seed, region) items, level-synchronously: level L evaluates octave slot
// L of both sub-samplers (evaluations 2L and 2L + 1) for every item still alive, as one interleaved pair, then
// applies the tier checks that fall after those two evaluations in the reference's order. Items that survive are
// compacted in place into one LDS list (each level reads its entries into registers before the first append). The
// second evaluation of a pair is computed even when the check
// after the first one rejects; it is then discarded, so the decisions and the tier histogram are the reference's.

struct LPair {
    const uint16_t *p;
    __device__ __forceinline__ uint32_t operator()(int i) const { return p[i & 255]; }
    template <int SEL>
    __device__ __forceinline__ uint32_t next(uint32_t pair, int a) const { return p[byte_add<SEL>(pair, (uint32_t)a)]; }
};

// Items per pass: a whole tile of 4 seeds x 256 regions, so that levels 2 and 3 still fill whole waves (measured
// against 512, 768, 1536 and 2048: 1024 is fastest; 768 splits such a tile into a full pass and a thin one). The
// workgroup size (256 or 512 lanes, 4 or 2 items per lane) is a tuning choice (whf_tuning.ero_wg).
constexpr int kEroPass = 1024;
constexpr int kCzBias = 1 << 22;       // chunk z is stored biased so that 4 tile bits fit above it

__device__ __forceinline__ void unpack_item(int2 e, int &cx, int &cz, int &sl) {
    cx = e.x;
    sl = e.y >> 23;
    cz = (e.y & ((1 << 23) - 1)) - kCzBias;
}

// Tier decision of one check. Plain FP64: the reference's comparison. FP32 gate: decided only when the estimate
// is farther than the margin from the threshold; otherwise the item is re-evaluated in FP64 (returns 2).
template <bool F32>
__device__ __forceinline__ int tier_decide(double val, double thr, double margin) {
    if (!F32) return val < thr ? 0 : 1;
    if (val < thr - margin) return 0;
    if (val >= thr + margin) return 1;
    return 2;
}

// Queue an undecided item for the FP64 re-evaluation (collective over the wave).
__device__ __forceinline__ void push_recheck(bool pred, uint32_t slot, int cx, int cz, uint4 *qx, uint32_t *ctl,
                                             unsigned long long *cnt) {
    const uint32_t pos = wave_append(pred, &ctl[CTL_QX]);
    if (pred) {
        if (pos < kQxCap) qx[pos] = make_uint4(slot, (uint32_t)cx, (uint32_t)cz, 0u);
        else atomicAdd(&cnt[CNT_QX_OVER], 1ull);
    }
}

// Evaluations k0 and k0 + 1 of IL items (tile seed sl[u], point (x[u], z[u])): v[2u] and v[2u + 1]. The 2 * IL
// evaluations are independent, so their lookup chains interleave.
template <bool F32, int IL, typename Par>
__device__ __forceinline__ void ero_evals(const uint16_t *s_tab, const Par *s_par, const EroDesc &ed, int k0,
                                          const uint32_t *sl, const double *x, const double *z, double *v) {
    const int k1 = k0 + 1;
    LPair Q[2 * IL];
    const Par *P[2 * IL];
    double xs[2 * IL], zs[2 * IL];
#pragma unroll
    for (int u = 0; u < IL; u++) {
        Q[2 * u] = LPair{s_tab + (sl[u] * 8 + k0) * 256};
        Q[2 * u + 1] = LPair{s_tab + (sl[u] * 8 + k1) * 256};
        P[2 * u] = &s_par[sl[u] * 8 + k0];
        P[2 * u + 1] = &s_par[sl[u] * 8 + k1];
        xs[2 * u] = x[u] * ed.sc[k0] * ed.lac[k0];
        xs[2 * u + 1] = x[u] * ed.sc[k1] * ed.lac[k1];
        zs[2 * u] = z[u] * ed.sc[k0] * ed.lac[k0];
        zs[2 * u + 1] = z[u] * ed.sc[k1] * ed.lac[k1];
        if (!(fabs(x[u]) < ed.coord_lim && fabs(z[u]) < ed.coord_lim)) {
            for (int t = 2 * u; t < 2 * u + 2; t++) { xs[t] = keep_precision(xs[t]); zs[t] = keep_precision(zs[t]); }
        }
    }
    if constexpr (F32) perlin_y0_xn_q<2 * IL>(Q, P, xs, zs, v);
    else perlin_y0_x2(Q, P, xs, zs, v);   // the plain kernel takes one item per lane
}

// ST seeds per workgroup. Items of a pass are (seed-in-tile, region) pairs laid out seed-major. F32 selects the
// FP32-gated fixed-point corner blend (perlin_y0_xn_q, values scaled by 2^23, weights w23) with FP64 re-evaluation of
// undecided items. IL items per lane are evaluated at once (the FP32 gate's evaluations are short, and without a
// second item their LDS lookup chains leave the device waiting; IL = 1 for the plain kernel).
//
// SERVE (served batches): the workgroups steal tiles from a device list (tiles[0 .. *n_tiles), the next one at
// *tile_next) instead of striding over the launch's (tile, region chunk) blocks: tiles of different requests differ
// widely in work. A tile is a run of at most ST slots of one request, whose rect, salt and rule layout come from the
// request table (srv.req); every region of the rect is one chunk (at most kSrvMaxRegions), and the tile's tier counts
// go to its request's counters when the tile is done.
template <int ST, bool F32, int WG, int IL, bool SERVE = false>
__global__ __launch_bounds__(WG) void k_erosion(const uint64_t *__restrict__ seeds, uint32_t n_seeds,
                                                    const uint8_t *__restrict__ tabE, const SPar *__restrict__ parE,
                                                    const GateTables *__restrict__ gt_, int rx0_, int rz0_, FastDiv Wd_,
                                                    FastDiv Rd, uint64_t flat0, uint32_t flat_n, uint32_t chunk,
                                                    uint32_t n_chunks, int salt_, int single_,
                                                    uint4 *__restrict__ q1, uint32_t *__restrict__ ctl,
                                                    uint32_t *__restrict__ needC, uint4 *__restrict__ qx,
                                                    unsigned long long *__restrict__ cnt, SrvView srv,
                                                    const STile *__restrict__ tiles,
                                                    const uint32_t *__restrict__ n_tiles,
                                                    uint32_t *__restrict__ tile_next) {
    static_assert(F32 || IL == 1, "the plain kernel evaluates one item per lane");
    __shared__ __align__(16) uint16_t s_tab[ST * 8 * 256];   // byte-pair tables: P(i) | P(i + 1) << 8
    using Par = typename std::conditional<F32, EPar, SPar>::type;
    __shared__ Par s_par[ST * 8];
    __shared__ uint64_t s_seed[ST];
    __shared__ int2 s_item[kEroPass];
    // alive list, compacted in place: item index and running sum (the FP32 gate keeps its sum as a float between
    // levels, which its margins allow for, and so fits five workgroups per CU)
    using Sum = typename std::conditional<F32, float, double>::type;
    __shared__ uint16_t s_li[kEroPass];
    __shared__ Sum s_ls[kEroPass];
    __shared__ uint32_t s_n[4];                  // list sizes after levels 0, 1, 2
    __shared__ uint32_t s_cnt[8];
    __shared__ uint32_t s_blk;                   // SERVE: the tile this workgroup stole

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    if (tid < 8) s_cnt[tid] = 0;
    // persistent workgroups: each takes blocks (tile, region chunk) blockIdx.x, + gridDim.x, ... in turn; the barrier
    // at the end of a block's last pass keeps its tables until every wave is done with them
    const uint32_t n_blocks = SERVE ? *n_tiles : ((n_seeds + ST - 1) / ST) * n_chunks;
    // the next block: stolen from the cursor when serving (both barriers keep s_blk one value per round), else strided
    auto next_block = [&](uint32_t prev) -> uint32_t {
        if constexpr (SERVE) {
            if (tid == 0) s_blk = atomicAdd(tile_next, 1u);
            __syncthreads();
            // one value for the whole workgroup: read as a scalar, so the tile's and its request's fields load
            // into scalar registers, as the launch's block index does for the scan
            const uint32_t b = __builtin_amdgcn_readfirstlane(s_blk);
            __syncthreads();
            return b;
        } else {
            return prev + gridDim.x;
        }
    };
    for (uint32_t blk = SERVE ? next_block(0) : blockIdx.x; blk < n_blocks; blk = next_block(blk)) {
    uint32_t seed0, ns, rbase, rn, req = 0;
    FastDiv Rn, Wd;
    const GateTables *gt;
    int rx0, rz0, salt, single;
    if constexpr (SERVE) {
        // the request table is written by other kernels of the loop, so its fields load per lane; each is one
        // value for the workgroup and is kept as a scalar, and the rule layout's tables are read through gt_ (the
        // server's table array, read-only here) so that they load as scalars, as the scan's do
        auto u = [](uint32_t v) -> uint32_t { return __builtin_amdgcn_readfirstlane(v); };
        const STile t = tiles[blk];
        const SReq &r = srv.req[t.req];
        req = t.req;
        seed0 = t.slot0;
        ns = t.ns;
        rbase = 0;
        rn = u(r.R);
        Rn = FastDiv{u(r.Rdiv.d), u(r.Rdiv.m), u(r.Rdiv.l)};
        Wd = FastDiv{u(r.Wdiv.d), u(r.Wdiv.m), u(r.Wdiv.l)};
        const int vflags = (int)u((uint32_t)r.vflags);
        gt = gt_ + vflags;
        rx0 = (int)u((uint32_t)r.rx0);
        rz0 = (int)u((uint32_t)r.rz0);
        salt = (int)u((uint32_t)r.salt);
        single = (vflags & WHF_V_SINGLE) ? 1 : 0;
    } else {
        const uint32_t tile = blk / n_chunks;
        const uint32_t ck = blk - tile * n_chunks;
        seed0 = tile * ST;
        ns = min((uint32_t)ST, n_seeds - seed0);
        rbase = ck * chunk;
        rn = min(chunk, flat_n - rbase);
        // regions per seed in this workgroup: the full chunk (Rd) unless this is the last, partial chunk
        Rn = rn == Rd.d ? Rd : FastDiv{rn, 0u, 0u};
        Wd = Wd_;
        gt = gt_;
        rx0 = rx0_;
        rz0 = rz0_;
        salt = salt_;
        single = single_;
    }
    const uint32_t W = Wd.d;
    // FP32-gate queue overflows count into the request's counters when serving
    unsigned long long *const cnt_blk = SERVE ? srv.rcnt + (size_t)req * CNT_N : cnt;

    {   // stage the tile's erosion tables as byte pairs, and the origins: a lane takes 16 bytes of a table with one
        // 16-byte load, and the byte after them (the next 16 bytes' first, or the table's first after its last), and
        // writes their 16 pairs
        const uint8_t *src = tabE + (size_t)seed0 * 8 * 256;
        for (uint32_t q = tid; q < ns * 8 * 16; q += WG) {
            const uint32_t t = q >> 4, c = q & 15;
            const uint4 v = reinterpret_cast<const uint4 *>(src)[q];
            const uint32_t w[5] = {v.x, v.y, v.z, v.w, (uint32_t)src[(t << 8) | (((c + 1) & 15) << 4)]};
            uint32_t o[8];
#pragma unroll
            for (int k = 0; k < 4; k++) {   // bytes 4k .. 4k + 3 and the next one
                const uint32_t cur = w[k], nxt = w[k + 1] & 255u;
                o[2 * k] = (cur & 0xFFFFu) | (((cur >> 8) & 0xFFFFu) << 16);
                o[2 * k + 1] = (cur >> 16) | (((cur >> 24) | (nxt << 8)) << 16);
            }
            uint4 *dst = reinterpret_cast<uint4 *>(s_tab + (size_t)q * 16);
            dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
            dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
        }
        for (uint32_t w = tid; w < ns * 8; w += WG) {
            if constexpr (F32) s_par[w] = make_epar(parE[(size_t)seed0 * 8 + w]);
            else s_par[w] = parE[(size_t)seed0 * 8 + w];
        }
        if (tid < ns) s_seed[tid] = seeds[seed0 + tid];
        if (tid < 4) s_n[tid] = 0;
    }
    const EroDesc &ed = gt->ero;
    const double amplitude = ed.amplitude;
    const double *w = F32 ? ed.w23 : ed.w;   // the FP32 gate's values are scaled by 2^23
    // first region of this chunk as (row, column); a region offset ro then needs one 32-bit division
    const uint64_t fbase = SERVE ? 0 : flat0 + rbase;
    const uint32_t row0 = (uint32_t)(fbase / W);
    const uint32_t col0 = (uint32_t)(fbase - (uint64_t)row0 * W);
    __syncthreads();

    const uint64_t total = (uint64_t)ns * rn;
    for (uint64_t pass0 = 0; pass0 < total; pass0 += kEroPass) {
        const uint32_t npass = (uint32_t)min((uint64_t)kEroPass, total - pass0);
        // ---- level 0: placement, evaluations 0 and 1, tier 0; IL items per lane ----
        for (uint32_t i0 = tid & ~63u; i0 < npass; i0 += WG * IL) {
            uint32_t i[IL], sl[IL];
            int cx[IL], cz[IL];
            bool in[IL], alive[IL], undecided[IL];
            double sum[IL], x[IL], z[IL];
#pragma unroll
            for (int u = 0; u < IL; u++) {
                i[u] = i0 + (uint32_t)u * WG + lane;
                in[u] = i[u] < npass;
                alive[u] = undecided[u] = false;
                sum[u] = 0.0;
                cx[u] = cz[u] = 0;
                sl[u] = 0;
                if (in[u]) {
                    const uint32_t it = (uint32_t)(pass0 + i[u]);
                    sl[u] = Rn.m ? Rn.div(it) : it / rn;
                    const uint32_t c = col0 + (it - sl[u] * rn);
                    const uint32_t drow = Wd.div(c);
                    hut_chunk(s_seed[sl[u]], rx0 + (int)(c - drow * W), rz0 + (int)(row0 + drow), salt, cx[u], cz[u]);
                    s_item[i[u]] = make_int2(cx[u], (cz[u] + kCzBias) | (int)(sl[u] << 23));
                }
                x[u] = (double)(cx[u] * 16 + 8) / 4.0;
                z[u] = (double)(cz[u] * 16 + 8) / 4.0;
            }
            if (single) {
#pragma unroll
                for (int u = 0; u < IL; u++) alive[u] = in[u];
            } else {
                double v[2 * IL];
                ero_evals<F32, IL>(s_tab, s_par, ed, 0, sl, x, z, v);
#pragma unroll
                for (int u = 0; u < IL; u++) {
                    if (!in[u]) continue;
                    sum[u] += w[0] * v[2 * u];
                    sum[u] += w[1] * v[2 * u + 1];
                    const int dcs = tier_decide<F32>(sum[u] * amplitude, ed.thr[0], ed.margin[1]);
                    if (dcs == 0) atomicAdd(&s_cnt[0], 1u);
                    else if (dcs == 1) alive[u] = true;
                    else undecided[u] = true;
                }
            }
#pragma unroll
            for (int u = 0; u < IL; u++) {
                if constexpr (F32) push_recheck(undecided[u], seed0 + sl[u], cx[u], cz[u], qx, ctl, cnt_blk);
                if (single) {
                    const uint32_t pos = wave_append(alive[u], &ctl[CTL_Q1]);
                    if (alive[u]) {
                        q1[pos] = make_uint4(seed0 + sl[u], (uint32_t)cx[u], (uint32_t)cz[u], 0u);
                        needC[seed0 + sl[u]] = 1u;
                    }
                } else {
                    const uint32_t pos = wave_append(alive[u], &s_n[0]);
                    if (alive[u]) {
                        s_li[pos] = (uint16_t)i[u];
                        s_ls[pos] = (Sum)sum[u];
                    }
                }
            }
        }
        __syncthreads();
        if (!single) {
            // ---- levels 1..3: evaluations 2L, 2L + 1 and tiers 2L - 1, 2L ----
            // a lane holds list entries tid + k * WG; all are read before the barrier, so the survivors can be
            // appended to the same list
            constexpr int kPer = kEroPass / WG;
            static_assert(kPer % IL == 0, "list entries per lane come in groups of IL");
            for (int L = 1; L <= 3; L++) {
                const uint32_t n = s_n[L - 1];
                const int k0 = 2 * L, k1 = 2 * L + 1;
                uint32_t idx[kPer];
                double run[kPer];
#pragma unroll
                for (int k = 0; k < kPer; k++) {
                    const uint32_t q = tid + (uint32_t)k * WG;
                    idx[k] = q < n ? s_li[q] : 0u;
                    run[k] = q < n ? (double)s_ls[q] : 0.0;
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < kPer; k += IL) {
                    if ((tid & ~63u) + (uint32_t)k * WG >= n) break;   // wave-uniform
                    uint32_t sl[IL];
                    int cx[IL], cz[IL];
                    bool in[IL], alive[IL], undecided[IL];
                    double sum[IL], x[IL], z[IL];
#pragma unroll
                    for (int u = 0; u < IL; u++) {
                        in[u] = tid + (uint32_t)(k + u) * WG < n;
                        alive[u] = undecided[u] = false;
                        sum[u] = run[k + u];
                        int s = 0;
                        cx[u] = cz[u] = 0;
                        if (in[u]) unpack_item(s_item[idx[k + u]], cx[u], cz[u], s);
                        sl[u] = (uint32_t)s;
                        x[u] = (double)(cx[u] * 16 + 8) / 4.0;
                        z[u] = (double)(cz[u] * 16 + 8) / 4.0;
                    }
                    double v[2 * IL];
                    ero_evals<F32, IL>(s_tab, s_par, ed, k0, sl, x, z, v);
#pragma unroll
                    for (int u = 0; u < IL; u++) {
                        if (!in[u]) continue;
                        sum[u] += w[k0] * v[2 * u];
                        int dcs = tier_decide<F32>(sum[u] * amplitude, ed.thr[k0 - 1], ed.margin[k0]);
                        if (dcs == 0) {
                            atomicAdd(&s_cnt[k0 - 1], 1u);
                        } else if (dcs == 1) {
                            sum[u] += w[k1] * v[2 * u + 1];
                            dcs = tier_decide<F32>(sum[u] * amplitude, ed.thr[k1 - 1], ed.margin[k1]);
                            if (dcs == 0) atomicAdd(&s_cnt[k1 - 1], 1u);
                            else if (dcs == 1) alive[u] = true;
                            else undecided[u] = true;
                        } else {
                            undecided[u] = true;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < IL; u++) {
                        if constexpr (F32) push_recheck(undecided[u], seed0 + sl[u], cx[u], cz[u], qx, ctl, cnt_blk);
                        if (L < 3) {
                            const uint32_t pos = wave_append(alive[u], &s_n[L]);
                            if (alive[u]) {
                                s_li[pos] = (uint16_t)idx[k + u];
                                s_ls[pos] = (Sum)sum[u];
                            }
                        } else {
                            // survivors of the erosion gate -> Q1
                            const uint32_t pos = wave_append(alive[u], &ctl[CTL_Q1]);
                            if (alive[u]) {
                                q1[pos] = make_uint4(seed0 + sl[u], (uint32_t)cx[u], (uint32_t)cz[u], 0u);
                                needC[seed0 + sl[u]] = 1u;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        // every thread has read the list sizes of this pass: clear them for the next one
        if (tid < 4) s_n[tid] = 0;
        __syncthreads();
    }
    if constexpr (SERVE) {
        // the tile's tier counts go to its request; the next tile's staging barrier orders the reset before its
        // first count
        if (tid < 7) {
            if (s_cnt[tid]) atomicAdd(&srv.rcnt[(size_t)req * CNT_N + CNT_TIER0 + tid], (unsigned long long)s_cnt[tid]);
            s_cnt[tid] = 0;
        }
    }
    }   // blocks
    if constexpr (!SERVE)
        if (tid < 7 && s_cnt[tid]) atomicAdd(&cnt[CNT_TIER0 + tid], (unsigned long long)s_cnt[tid]);
}

// FP64 re-evaluation of the items the FP32 gate left undecided: the reference's tiered gate from the first
// evaluation on, with the set-E tables read from HBM (noise.c:432-459).
// SERVE: each entry's rule layout and request come from its slot (srv), and its tier and re-evaluation counts go to
// the request's counters.
template <bool SERVE = false>
__global__ __launch_bounds__(256) void k_erosion_exact(const uint8_t *__restrict__ tabE,
                                                       const SPar *__restrict__ parE,
                                                       const GateTables *__restrict__ gt,
                                                       const uint4 *__restrict__ qx, uint4 *__restrict__ q1,
                                                       uint32_t *__restrict__ ctl, uint32_t *__restrict__ needC,
                                                       unsigned long long *__restrict__ cnt, SrvView srv) {
    __shared__ uint32_t s_cnt[8];
    if (threadIdx.x < 8) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = min(ctl[CTL_QX], kQxCap);
    for (uint32_t q0 = blockIdx.x * 256 + (threadIdx.x & ~63u); q0 < n; q0 += gridDim.x * 256) {
        const uint32_t q = q0 + lane_id();
        bool alive = false;
        int tier = -1;
        uint32_t req = 0;
        uint4 e = make_uint4(0, 0, 0, 0);
        if (q < n) {
            e = qx[q];
            const EroDesc &ed = SERVE ? gt[srv.lay_of[e.x]].ero : gt->ero;
            if constexpr (SERVE) req = srv.rq_of[e.x];
            const double x = (double)((int)e.y * 16 + 8) / 4.0, z = (double)((int)e.z * 16 + 8) / 4.0;
            double sum = 0.0;
            alive = true;
            for (int k = 0; k < 8; k++) {
                const size_t t = (size_t)e.x * 8 + k;
                const double v = perlin_y0(GPerm{tabE + t * 256}, parE[t], keep_precision(x * ed.sc[k] * ed.lac[k]),
                                           keep_precision(z * ed.sc[k] * ed.lac[k]));
                sum += ed.w[k] * v;
                if (k >= 1 && sum * ed.amplitude < ed.thr[k - 1]) {
                    tier = k - 1;
                    alive = false;
                    break;
                }
            }
        }
        if constexpr (SERVE) {
            wave_count_req(tier >= 0 ? CNT_TIER0 + tier : -1, req, srv.rcnt);
            wave_count_req(q < n ? (int)CNT_QX : -1, req, srv.rcnt);
        } else if (tier >= 0) {
            atomicAdd(&s_cnt[tier], 1u);
        }
        const uint32_t pos = wave_append(alive, &ctl[CTL_Q1]);
        if (alive) {
            q1[pos] = e;
            needC[e.x] = 1u;
        }
    }
    if constexpr (!SERVE) {
        __syncthreads();
        if (threadIdx.x < 7 && s_cnt[threadIdx.x])
            atomicAdd(&cnt[CNT_TIER0 + threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&cnt[CNT_QX], (unsigned long long)ctl[CTL_QX]);
    }
}

// Measured error of the FP32-gate corner blend against FP64, per set-E sampler (the erosion octaves): for
// n_pts points per seed slot spread over the +-30M world, max |v32 - v64| as the bits of a non-negative double.
__global__ void k_noise_error(const uint8_t *__restrict__ tabE, const SPar *__restrict__ parE,
                              const GateTables *__restrict__ gt, uint32_t n_seeds, uint32_t n_pts,
                              unsigned long long *__restrict__ maxbits) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)n_seeds * n_pts) return;
    const uint32_t slot = (uint32_t)(i / n_pts);
    const uint64_t h = mix64(i * kGolden + 0x5EEDull);
    const int hx = (int)(h % 60000001ull) - 30000000, hz = (int)((h >> 32) % 60000001ull) - 30000000;
    const double x = (double)(hx + 8) / 4.0, z = (double)(hz + 8) / 4.0;
    const EroDesc &ed = gt->ero;
    for (int L = 0; L < 4; L++) {
        const int k0 = 2 * L, k1 = k0 + 1;
        const size_t t0 = (size_t)slot * 8 + k0, t1 = t0 + 1;
        const GPair Q[2] = {GPair{tabE + t0 * 256}, GPair{tabE + t1 * 256}};
        const SPar *P[2] = {&parE[t0], &parE[t1]};
        const double xs[2] = {keep_precision(x * ed.sc[k0] * ed.lac[k0]), keep_precision(x * ed.sc[k1] * ed.lac[k1])};
        const double zs[2] = {keep_precision(z * ed.sc[k0] * ed.lac[k0]), keep_precision(z * ed.sc[k1] * ed.lac[k1])};
        const EPar e0 = make_epar(parE[t0]), e1 = make_epar(parE[t1]);
        const EPar *E[2] = {&e0, &e1};
        double v64[2], v32[2];
        perlin_y0_x2(Q, P, xs, zs, v64);
        perlin_y0_xn_q<2>(Q, E, xs, zs, v32);
        for (int t = 0; t < 2; t++) {
            const double err = fabs(v32[t] * kFixInv - v64[t]);
            atomicMax(&maxbits[k0 + t], (unsigned long long)__double_as_longlong(err));
        }
    }
}

// ---------------------------------------------------------------------------------------------- compaction

__global__ void k_compact(const uint32_t *__restrict__ need, uint32_t n, uint32_t *__restrict__ slot_of,
                          uint32_t *__restrict__ seed_of, uint32_t *__restrict__ count, uint32_t cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool f = i < n && need[i];
    uint32_t pos = wave_append(f, count);
    if (f) {
        slot_of[i] = pos;   // a slot >= cap has no tables: the gates count it as overflow
        if (pos < cap) seed_of[pos] = i;
    }
}

// ---------------------------------------------------------------------------------------------- gates 2 and 3

__device__ __forceinline__ void block_count(uint32_t *s_cnt, int kind) { atomicAdd(&s_cnt[kind], 1u); }

constexpr int kGateWG = 256;

// Gate 2 in two kernels. k_gate2a: temperature and ridge with the version rules on Q1 -> Q2 (count CTL_QM); it
// needs few registers, so many waves hide its gathers. k_gate2b: the cave ladder's Entrance@50 and @60 on those
// survivors -> Q1, which gate 3 then reads (count CTL_Q2).
// SERVE (both gate-2 kernels and gate 3): each entry's rule layout (= its request's vflags) and request come from its
// seed slot; counts go to the request's counters.
template <bool SERVE = false>
__global__ __launch_bounds__(kGateWG) void k_gate2a(const GateTables *__restrict__ gt_, const uint8_t *__restrict__ tabC,
                                                     const SPar *__restrict__ parC, const uint32_t *__restrict__ slotC,
                                                     const uint4 *__restrict__ q1, uint4 *__restrict__ qm,
                                                     uint32_t *__restrict__ ctl, int vflags_,
                                                     unsigned long long *__restrict__ cnt, SrvView srv) {
    __shared__ uint32_t s_cnt[4];
    if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = ctl[CTL_Q1];
    const SetView C{tabC, parC, kSetSize[SET_C]};
    for (uint32_t q0 = blockIdx.x * kGateWG + (threadIdx.x & ~63u); q0 < n; q0 += gridDim.x * kGateWG) {
        const uint32_t q = q0 + lane_id();
        bool survive = false;
        int kind = -1;
        uint32_t req = 0;
        uint4 e = make_uint4(0, 0, 0, 0);
        int lay = vflags_;
        if (q < n) {
            e = q1[q];
            if constexpr (SERVE) {
                lay = (int)srv.lay_of[e.x];
                req = srv.rq_of[e.x];
            }
        }
        // the temperature and ridge checks of this lane's entry under the rules of layout vflags
        auto check = [&](const GateTables *gt, int vflags) {
            const bool single = (vflags & WHF_V_SINGLE) != 0;
            const uint32_t slot = slotC[e.x];
            if (slot >= kMaxSeeds) {
                atomicAdd(&cnt[CNT_SLOT_OVER], 1ull);
                kind = 0;   // no tables: neither counted nor kept
            } else if (!single) {
                const double qx = ((int)e.y * 16 + 8) / 4.0, qz = ((int)e.z * 16 + 8) / 4.0;
                const double t = noise_y0(gt->noise[G_TEMPERATURE], C, slot, qx, qz);
                const bool t_bad = (vflags & WHF_V_118) ? (t < -0.45) : (t > 0.2 || t < -0.45);
                if (t_bad) kind = 1;
                else {
                    const double r = noise_y0(gt->noise[G_RIDGE], C, slot, qx, qz);
                    if ((r > 0.42 && r < 0.91) || (r < -0.42 && r > -0.91)) kind = 2;
                    else if ((vflags & WHF_V_262) && r <= -0.91) kind = 2;
                }
            }
        };
        if constexpr (SERVE) {
            // a wave whose entries share one rule layout (the usual case) reads that layout's noise descriptors with
            // scalar loads; a mixed wave reads each lane's own
            const unsigned long long act = __ballot(q < n);
            const int lay0 = act ? __shfl(lay, __ffsll((long long)act) - 1) : 0;
            if (__ballot(q < n && lay != lay0) == 0) {
                const int L = __builtin_amdgcn_readfirstlane(lay0);
                if (q < n) check(gt_ + L, L);
            } else if (q < n) {
                check(gt_ + lay, lay);
            }
        } else if (q < n) {
            check(gt_, vflags_);
        }
        if (q < n) {
            if (!SERVE && kind > 0) block_count(s_cnt, kind);
            else if (kind < 0) survive = true;
        }
        if constexpr (SERVE) wave_count_req(kind > 0 ? CNT_KIND0 + kind : -1, req, srv.rcnt);
        const uint32_t pos = wave_append(survive, &ctl[CTL_QM]);
        if (survive) qm[pos] = e;
    }
    if constexpr (!SERVE) {
        __syncthreads();
        if (threadIdx.x >= 1 && threadIdx.x < 3 && s_cnt[threadIdx.x])
            atomicAdd(&cnt[CNT_KIND0 + threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
    }
}

template <bool SERVE = false>
__global__ __launch_bounds__(kGateWG) void k_gate2b(const GateTables *__restrict__ gt_, const uint8_t *__restrict__ tabC,
                                                     const SPar *__restrict__ parC, const uint32_t *__restrict__ slotC,
                                                     const uint4 *__restrict__ qm, uint4 *__restrict__ qout,
                                                     uint32_t *__restrict__ ctl, uint32_t *__restrict__ needL,
                                                     unsigned long long *__restrict__ cnt, SrvView srv) {
    __shared__ uint32_t s_cnt[8];
    if (threadIdx.x < 8) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = ctl[CTL_QM];
    const SetView C{tabC, parC, kSetSize[SET_C]};
    for (uint32_t q0 = blockIdx.x * kGateWG + (threadIdx.x & ~63u); q0 < n; q0 += gridDim.x * kGateWG) {
        const uint32_t q = q0 + lane_id();
        bool survive = false;
        int kind = -1;
        uint32_t req = 0;
        uint4 e = make_uint4(0, 0, 0, 0);
        if (q < n) {
            e = qm[q];
            const GateTables *gt = SERVE ? gt_ + srv.lay_of[e.x] : gt_;
            if constexpr (SERVE) req = srv.rq_of[e.x];
            else atomicAdd(&s_cnt[0], 1u);   // reached the cave ladder
            CaveCtx cv{gt, C, SetView{nullptr, nullptr, 0}, slotC[e.x], 0};
            const int x = (int)e.y * 16 + 3, z = (int)e.z * 16 + 3;
            if (cv.entrance_nonneg(x, 50, z)) kind = 3;
            else if (cv.entrance_nonneg(x, 60, z)) kind = 4;
            if (!SERVE && kind > 0) block_count(s_cnt, kind);
            else if (kind < 0) survive = true;
        }
        if constexpr (SERVE) {
            wave_count_req(q < n ? (int)CNT_LADDER : -1, req, srv.rcnt);
            wave_count_req(kind > 0 ? CNT_KIND0 + kind : -1, req, srv.rcnt);
        }
        const uint32_t pos = wave_append(survive, &ctl[CTL_Q2]);
        if (survive) {
            qout[pos] = e;
            needL[e.x] = 1u;
        }
    }
    if constexpr (!SERVE) {
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt[0]) atomicAdd(&cnt[CNT_LADDER], (unsigned long long)s_cnt[0]);
        if ((threadIdx.x == 3 || threadIdx.x == 4) && s_cnt[threadIdx.x])
            atomicAdd(&cnt[CNT_KIND0 + threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
    }
}

// Gate 3 per gate-2 survivor, in two stages per chunk of kG3Entries entries. Stage A: cave@maxHeight, one lane per
// entry (it decides 3 in 4 entries). Stage B: every remaining check of the survivors at once, one lane per (entry,
// check): the negative ladder, the positive ladder, continentalness and the floodedness heights, in lysh's order
// (phase1.c:151-196). Each check is a pure function of the entry, so the entry's reject kind is its first failing
// check, as in the sequential scan, and the survivors are the entries with no failing check. A chunk is a quarter of
// the workgroup's lanes: a batch holds a few thousand entries, and small chunks spread them over more workgroups and
// give each lane fewer of stage B's dependent noise lookups.
constexpr int kG3Checks = 32;
constexpr int kG3Entries = kGateWG / 4;

__device__ __forceinline__ int g3_layout(int max_h, bool single, int &n_neg, int &n_flood) {
    int floor_y = max_h < -40 ? max_h : -40;
    if (floor_y < -50) floor_y = -50;
    n_neg = 0;
    for (int y = 0; y >= floor_y; y -= 10)
        if (max_h < y) n_neg++;
    n_flood = max_h <= 60 ? (60 - max_h) / 10 + 1 : 0;
    return n_neg + 4 + (single ? 0 : 1) + n_flood;
}

// SERVE: every entry carries its own request's maxHeight and rule layout, so stage B gives each entry kG3Checks lanes
// (the most any layout needs) and lanes past the entry's own check count idle; survivors go to the request's queue.
template <bool SERVE = false>
__global__ __launch_bounds__(kGateWG) void k_gate3(const uint64_t *__restrict__ seeds, const GateTables *__restrict__ gt_,
                                                    const uint8_t *__restrict__ tabC, const SPar *__restrict__ parC,
                                                    const uint8_t *__restrict__ tabL, const SPar *__restrict__ parL,
                                                    const uint32_t *__restrict__ slotC,
                                                    const uint32_t *__restrict__ slotL, const uint4 *__restrict__ surv2,
                                                    const uint32_t *__restrict__ ctl, int vflags_, int max_h_,
                                                    DevHit *__restrict__ hits, uint32_t *__restrict__ hit_ctr,
                                                    uint32_t hit_cap, unsigned long long *__restrict__ cnt, SrvView srv) {
    __shared__ uint32_t s_cnt[16];
    __shared__ uint32_t s_list[kG3Entries];
    __shared__ uint32_t s_fail[kG3Entries];
    __shared__ uint32_t s_n;
    if (threadIdx.x < 16) s_cnt[threadIdx.x] = 0;
    const uint32_t n = ctl[CTL_Q2];
    // the rules of an entry: its layout's tables, maxHeight, SINGLE_BIOME flag and check layout
    struct Rules {
        const GateTables *gt;
        int max_h, n_neg, n_flood, n_checks;
        bool single;
    };
    auto rules_of = [&](uint32_t slot) {
        Rules r;
        const int vflags = SERVE ? (int)srv.lay_of[slot] : vflags_;
        r.gt = SERVE ? gt_ + vflags : gt_;
        r.max_h = SERVE ? srv.req[srv.rq_of[slot]].max_height : max_h_;
        r.single = (vflags & WHF_V_SINGLE) != 0;
        r.n_checks = g3_layout(r.max_h, r.single, r.n_neg, r.n_flood);
        return r;
    };
    // the launch's rules; a served batch reads every entry's own (slot 0 of the tables need not be written)
    const Rules launch_rules = SERVE ? Rules{gt_, 0, 0, 0, 0, false} : rules_of(0);
    const uint32_t stride = SERVE ? (uint32_t)kG3Checks : (uint32_t)launch_rules.n_checks;
    const SetView C{tabC, parC, kSetSize[SET_C]}, L{tabL, parL, kSetSize[SET_L]};
    for (uint32_t base = blockIdx.x * kG3Entries; base < n; base += gridDim.x * kG3Entries) {
        if (threadIdx.x == 0) s_n = 0;
        if (threadIdx.x < kG3Entries) s_fail[threadIdx.x] = 0;
        __syncthreads();
        // ---- stage A: cave@maxHeight, the chunk's entries on its first waves ----
        if (threadIdx.x < kG3Entries) {
            const uint32_t q = base + threadIdx.x;
            bool keep = false;
            int idx = -1;
            uint32_t req = 0;
            if (q < n) {
                const uint4 e = surv2[q];
                const Rules r = SERVE ? rules_of(e.x) : launch_rules;
                if constexpr (SERVE) req = srv.rq_of[e.x];
                const CaveCtx cv{r.gt, C, L, slotC[e.x], slotL[e.x]};
                const int x = (int)e.y * 16 + 3, z = (int)e.z * 16 + 3;
                if (cv.slotC >= kMaxSeeds || cv.slotL >= kMaxSeeds) atomicAdd(&cnt[CNT_SLOT_OVER], 1ull);
                else if (cv.spaghetti(x, r.max_h, z) >= 0 && cv.cheese(x, r.max_h, z) >= 0) idx = CNT_KIND0 + 5;
                else keep = true;
            }
            if constexpr (SERVE) wave_count_req(idx, req, srv.rcnt);
            else if (idx >= 0) atomicAdd(&s_cnt[5], 1u);
            const uint32_t pos = wave_append(keep, &s_n);
            if (keep) s_list[pos] = q;
        }
        __syncthreads();
        // ---- stage B: all remaining checks of the survivors, one lane per (entry, check) ----
        const uint32_t ns = s_n;
        for (uint32_t w = threadIdx.x; w < ns * stride; w += kGateWG) {
            const uint32_t li = w / stride;
            const int j = (int)(w - li * stride);
            const uint4 e = surv2[s_list[li]];
            const Rules r = SERVE ? rules_of(e.x) : launch_rules;
            if (SERVE && j >= r.n_checks) continue;
            const CaveCtx cv{r.gt, C, L, slotC[e.x], slotL[e.x]};
            const int hut_x = (int)e.y * 16, hut_z = (int)e.z * 16, x = hut_x + 3, z = hut_z + 3;
            bool fail;
            if (j < r.n_neg) {
                const int y = -10 * j;
                fail = cv.spaghetti(x, y, z) >= 0 && cv.cheese(x, y, z) >= 0;
            } else if (j < r.n_neg + 4) {
                const int y = 10 * (j - r.n_neg + 1);
                fail = cv.entrance_nonneg(x, y, z) && cv.cheese(x, y, z) >= 0;
            } else if (!r.single && j == r.n_neg + 4) {
                const NoiseDesc &d = r.gt->noise[G_CONTINENTALNESS];
                fail = noise_y0(d, cv.L, cv.slotL, (hut_x + 8) / 4.0, (hut_z + 8) / 4.0) < -0.11;
            } else {
                const int y = r.max_h + 10 * (j - r.n_neg - 4 - (r.single ? 0 : 1));
                fail = cv.flooded(hut_x, hut_z, y) > 0.41;
            }
            if (fail) atomicOr(&s_fail[li], 1u << j);
        }
        __syncthreads();
        // ---- verdicts: first failing check, in order ----
        if (threadIdx.x < kG3Entries) {
            const uint32_t li = threadIdx.x;
            bool accept = false;
            int idx = -1;
            uint32_t req = 0;
            uint4 e = make_uint4(0, 0, 0, 0);
            if (li < ns) {
                e = surv2[s_list[li]];
                const Rules r = SERVE ? rules_of(e.x) : launch_rules;
                if constexpr (SERVE) req = srv.rq_of[e.x];
                const uint32_t f = s_fail[li];
                int kind = 10;
                if (f == 0) {
                    accept = true;
                } else {
                    const int j = __ffs((int)f) - 1;
                    kind = j < r.n_neg ? 6 : (j < r.n_neg + 4 ? 7 : ((!r.single && j == r.n_neg + 4) ? 8 : 9));
                }
                idx = CNT_KIND0 + kind;
            }
            if constexpr (SERVE) {
                wave_count_req(idx, req, srv.rcnt);
                if (accept) {
                    // the request's survivor queue: counted past its capacity, written up to it
                    const uint32_t pos = atomicAdd(&srv.state[req].hits, 1u);
                    const SReq &rq = srv.req[req];
                    if (pos < rq.hit_cap) {
                        DevHit h;
                        h.seed = (int64_t)seeds[e.x];
                        h.cx = (int)e.y;
                        h.cz = (int)e.z;
                        h.rx = h.cx >> 5;
                        h.rz = h.cz >> 5;
                        h.floor_y = INT32_MIN;
                        h.flags = 0;
                        srv.hit_pool[(size_t)rq.hit_base + pos] = h;
                    }
                }
            } else {
                if (idx >= 0) atomicAdd(&s_cnt[idx - CNT_KIND0], 1u);
                const uint32_t pos = wave_append(accept, hit_ctr);
                if (accept && pos < hit_cap) {
                    DevHit h;
                    h.seed = (int64_t)seeds[e.x];
                    h.cx = (int)e.y;
                    h.cz = (int)e.z;
                    h.rx = h.cx >> 5;
                    h.rz = h.cz >> 5;
                    h.floor_y = INT32_MIN;
                    h.flags = 0;
                    hits[pos] = h;
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (!SERVE && threadIdx.x < 16 && threadIdx.x > 0 && s_cnt[threadIdx.x])
        atomicAdd(&cnt[CNT_KIND0 + threadIdx.x], (unsigned long long)s_cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------------------------- probe
// The 38 values of lysh_phase1_probe (phase1.c:39-80) per (seed slot, hut position); all three sets are
// built for every slot. Erosion is the exact double-noise sample here, not the tiered gate.

__global__ void k_probe(const GateTables *__restrict__ gt, const uint8_t *tabE, const SPar *parE,
                        const uint8_t *tabC, const SPar *parC, const uint8_t *tabL, const SPar *parL,
                        const int32_t *__restrict__ hx, const int32_t *__restrict__ hz, uint32_t n, int max_h,
                        double *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SetView E{tabE, parE, kSetSize[SET_E]};
    const CaveCtx cv{gt, SetView{tabC, parC, kSetSize[SET_C]}, SetView{tabL, parL, kSetSize[SET_L]}, i, i};
    const int hut_x = hx[i], hut_z = hz[i];
    const double qx = (hut_x + 8) / 4.0, qz = (hut_z + 8) / 4.0;
    const int x = hut_x + 3, z = hut_z + 3;
    double *o = out + (size_t)i * 38;
    int k = 0;
    o[k++] = noise_y0(gt->noise[G_EROSION], E, i, qx, qz);
    o[k++] = noise_y0(gt->noise[G_TEMPERATURE], cv.C, i, qx, qz);
    o[k++] = noise_y0(gt->noise[G_RIDGE], cv.C, i, qx, qz);
    o[k++] = noise_y0(gt->noise[G_CONTINENTALNESS], cv.L, i, qx, qz);
    o[k++] = cv.entrance(x, 50, z);
    o[k++] = cv.entrance(x, 60, z);
    o[k++] = cv.spaghetti(x, max_h, z);
    o[k++] = cv.cheese(x, max_h, z);
    for (int y = 0; y >= -40; y -= 10) { o[k++] = cv.spaghetti(x, y, z); o[k++] = cv.cheese(x, y, z); }
    for (int y = 10; y <= 40; y += 10) { o[k++] = cv.entrance(x, y, z); o[k++] = cv.cheese(x, y, z); }
    for (int y = -50; y <= 60; y += 10) o[k++] = cv.flooded(hut_x, hut_z, y);
}

}  // namespace

// ================================================================================================ host

namespace {

// ---- MD5 (RFC 1321), used once per context to key the sampler derivations ----
struct Md5 {
    static uint32_t rol(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }
    static void block(uint32_t st[4], const uint8_t *p, const uint32_t K[64]) {
        static const int S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
        uint32_t m[16];
        for (int i = 0; i < 16; i++) m[i] = p[4 * i] | p[4 * i + 1] << 8 | p[4 * i + 2] << 16 | (uint32_t)p[4 * i + 3] << 24;
        uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
        for (int i = 0; i < 64; i++) {
            uint32_t f;
            int g, r = i >> 4;
            if (r == 0) { f = (b & c) | (~b & d); g = i; }
            else if (r == 1) { f = (d & b) | (~d & c); g = (5 * i + 1) & 15; }
            else if (r == 2) { f = b ^ c ^ d; g = (3 * i + 5) & 15; }
            else { f = c ^ (b | ~d); g = (7 * i) & 15; }
            uint32_t t = d;
            d = c;
            c = b;
            b = b + rol(a + f + K[i] + m[g], S[r][i & 3]);
            a = t;
        }
        st[0] += a; st[1] += b; st[2] += c; st[3] += d;
    }
    static void digest(const std::string &s, uint64_t &hi, uint64_t &lo) {
        uint32_t K[64];
        for (int i = 0; i < 64; i++) K[i] = (uint32_t)std::floor(std::fabs(std::sin((double)(i + 1))) * 4294967296.0);
        uint32_t st[4] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u};
        std::vector<uint8_t> msg(s.begin(), s.end());
        uint64_t bits = (uint64_t)msg.size() * 8;
        msg.push_back(0x80);
        while (msg.size() % 64 != 56) msg.push_back(0);
        for (int i = 0; i < 8; i++) msg.push_back((uint8_t)(bits >> (8 * i)));
        for (size_t o = 0; o < msg.size(); o += 64) block(st, msg.data() + o, K);
        uint8_t out[16];
        for (int i = 0; i < 16; i++) out[i] = (uint8_t)(st[i / 4] >> (8 * (i % 4)));
        hi = lo = 0;
        for (int i = 0; i < 8; i++) hi = hi << 8 | out[i];
        for (int i = 8; i < 16; i++) lo = lo << 8 | out[i];
    }
};



}  // namespace

namespace whf {

void md5_longs(const char *name, uint64_t &hi, uint64_t &lo) { Md5::digest(name, hi, lo); }

BuildDesc make_build_desc(const SamplerRef &r) {
    BuildDesc b{};
    Md5::digest(kNoise[r.noise].name, b.na, b.nb);
    Md5::digest("octave_" + std::to_string(kNoise[r.noise].first_octave + r.slot), b.oa, b.ob);
    b.sub = r.sub;
    return b;
}

void fill_noise_desc(NoiseDesc &d, int pid, const std::vector<SamplerRef> &set, int set_id) {
    const NoiseParam &np = kNoise[pid];
    d.set = set_id;
    int p = 0, first = -1, last = -1;
    const double per0 = std::ldexp(1.0, np.n - 1) / (std::ldexp(1.0, np.n) - 1.0);
    for (int i = 0; i < np.n; i++) {
        if (np.amps[i] == 0.0) continue;
        if (first < 0) first = i;
        last = i;
        for (int s = 0; s < 2; s++)
            for (size_t t = 0; t < set.size(); t++)
                if (set[t].noise == pid && set[t].sub == s && set[t].slot == i) d.tab[s][p] = (int)t;
        d.lac[p] = std::ldexp(1.0, np.first_octave + i);
        d.amp[p] = np.amps[i];
        d.per[p] = std::ldexp(per0, -i);   // persistence halves per slot (noise.c:277-278), exactly
        p++;
    }
    d.np = p;
    d.amplitude = 0.16666666666666666 / (0.1 * (1.0 + 1.0 / ((double)(last - first) + 1.0)));
}

}  // namespace whf

namespace {

// Which noise stands in each gate role for these rule flags (LARGE_BIOMES swaps three climate noises).
int gate_noise_param(int g, int vflags) {
    const bool large = (vflags & WHF_V_LARGE) != 0;
    switch (g) {
    case G_EROSION: return large ? N_EROSION_LARGE : N_EROSION;
    case G_TEMPERATURE: return large ? N_TEMPERATURE_LARGE : N_TEMPERATURE;
    case G_RIDGE: return N_RIDGE;
    case G_CAVE_ENTRANCE: return N_CAVE_ENTRANCE;
    case G_SPAG_RARITY: return N_SPAG3D_RARITY;
    case G_SPAG_THICKNESS: return N_SPAG3D_THICKNESS;
    case G_SPAG_1: return N_SPAG3D_1;
    case G_SPAG_2: return N_SPAG3D_2;
    case G_ROUGH_MOD: return N_SPAG_ROUGH_MOD;
    case G_ROUGH: return N_SPAG_ROUGH;
    case G_CAVE_LAYER: return N_CAVE_LAYER;
    case G_CAVE_CHEESE: return N_CAVE_CHEESE;
    case G_CONTINENTALNESS: return large ? N_CONTINENTALNESS_LARGE : N_CONTINENTALNESS;
    case G_FLOODEDNESS: return N_FLOODEDNESS;
    }
    return -1;
}

int gate_set(int g) {
    switch (g) {
    case G_EROSION: return SET_E;
    case G_TEMPERATURE: case G_RIDGE: case G_CAVE_ENTRANCE: case G_SPAG_RARITY: case G_SPAG_THICKNESS:
    case G_SPAG_1: case G_SPAG_2: case G_ROUGH_MOD: case G_ROUGH: return SET_C;
    default: return SET_L;
    }
}

// Table layouts of the host-side gate tables and build descriptors for a rule-flag set.
struct Layout {
    std::vector<SamplerRef> set[SET_COUNT];
    GateTables gt;
    std::vector<BuildDesc> bd[SET_COUNT];
};

void build_layout(int vflags, Layout &L) {
    for (int s = 0; s < SET_COUNT; s++) { L.set[s].clear(); L.bd[s].clear(); }
    std::memset(&L.gt, 0, sizeof L.gt);
    // set E: erosion in tier order A0 B0 A1 B1 ... over its non-empty slots
    {
        const NoiseParam &np = kNoise[gate_noise_param(G_EROSION, vflags)];
        for (int i = 0; i < np.n; i++)
            if (np.amps[i] != 0.0) { L.set[SET_E].push_back({gate_noise_param(G_EROSION, vflags), 0, i});
                                     L.set[SET_E].push_back({gate_noise_param(G_EROSION, vflags), 1, i}); }
    }
    // sets C and L: per gate noise, all non-empty slots of sub 0 then of sub 1
    for (int g = 1; g < G_COUNT; g++) {
        const int pid = gate_noise_param(g, vflags);
        const NoiseParam &np = kNoise[pid];
        for (int s = 0; s < 2; s++)
            for (int i = 0; i < np.n; i++)
                if (np.amps[i] != 0.0) L.set[gate_set(g)].push_back({pid, s, i});
    }
    // noise descriptors
    for (int g = 0; g < G_COUNT; g++)
        fill_noise_desc(L.gt.noise[g], gate_noise_param(g, vflags), L.set[gate_set(g)], gate_set(g));
    // tiered erosion
    {
        const NoiseDesc &d = L.gt.noise[G_EROSION];
        EroDesc &e = L.gt.ero;
        for (int k = 0; k < 8; k++) {
            const int p = k >> 1;
            e.sc[k] = (k & 1) ? 1.0181268882175227 : 1.0;
            e.lac[k] = d.lac[p];
            e.w[k] = d.amp[p] * d.per[p];
            e.w23[k] = e.w[k] * kFixInv;
        }
        e.amplitude = d.amplitude;
        const double thr[7] = {0.05, 0.22, 0.42, 0.45, 0.49, 0.51, 0.55};
        for (int t = 0; t < 7; t++) e.thr[t] = thr[t];
        // FP32 gate margin after evaluation k: amplitude x sum of |weights| so far x the per-evaluation bound,
        // plus slack for the FP64 accumulation and the final product's rounding
        double smax = 0.0;
        for (int k = 0; k < 8; k++) smax = std::max(smax, e.sc[k] * e.lac[k]);
        // x * sc * lac stays below 1e7 in magnitude (maintainPrecision's identity range) when |x| < coord_lim;
        // the 0.999 factor leaves room for the rounding of the two products
        e.coord_lim = 1.0e7 / smax * 0.999;
        // and the FP32 gate keeps the running sum as a float between levels: up to three roundings of a sum no larger
        // than sum |w| x 1.1 (|Perlin| <= 1.07), each within 2^-24 of it
        double wsum = 0.0, wall = 0.0;
        for (int k = 0; k < 8; k++) wall += std::fabs(e.w[k]);
        for (int k = 0; k < 8; k++) {
            wsum += std::fabs(e.w[k]);
            e.margin[k] = e.amplitude * (wsum * kFp32PerlinBound + 3.0 * 0x1p-24 * 1.1 * wall) + 1e-12;
        }
    }
    // build descriptors
    for (int s = 0; s < SET_COUNT; s++)
        for (const SamplerRef &r : L.set[s]) L.bd[s].push_back(make_build_desc(r));
}


}  // namespace

namespace whf {
thread_local char g_last_error[512];
}

namespace {

int ctx_alloc(whf_ctx *c) {
    const size_t S = kMaxSeeds, B = kMaxBatchSeeds;
    HIP_TRY(hipMalloc(&c->d_seeds, B * sizeof(uint64_t)));
    for (int s = 0; s < SET_COUNT; s++) {
        HIP_TRY(hipMalloc(&c->d_tab[s], S * kSetSize[s] * 256));
        HIP_TRY(hipMalloc(&c->d_par[s], S * kSetSize[s] * sizeof(SPar)));
    }
    HIP_TRY(hipMalloc(&c->d_q1, (size_t)kMaxItems * sizeof(uint4)));
    HIP_TRY(hipMalloc(&c->d_q2, (size_t)kMaxItems * sizeof(uint4)));
    HIP_TRY(hipMalloc(&c->d_qx, (size_t)kQxCap * sizeof(uint4)));
    HIP_TRY(hipMalloc(&c->d_needC, B * 4));
    HIP_TRY(hipMalloc(&c->d_needL, B * 4));
    HIP_TRY(hipMalloc(&c->d_slotC, B * 4));
    HIP_TRY(hipMalloc(&c->d_slotL, B * 4));
    HIP_TRY(hipMalloc(&c->d_seedC, S * 4));
    HIP_TRY(hipMalloc(&c->d_seedL, S * 4));
    HIP_TRY(hipMalloc(&c->d_ctl, 64));
    static_assert(CTL_N * sizeof(uint32_t) <= 64, "control words");
    c->batch_set[0] = BatchSet{c->d_seeds, c->d_tab[SET_C], c->d_par[SET_C], c->d_q1, c->d_q2, c->d_needC,
                               c->d_slotC, c->d_seedC, c->d_needL, c->d_ctl};
    for (int k = 1; k < 3; k++) {   // per-seed arrays for the largest batch (a Deep Dive expansion's 2^20 seeds)
        BatchSet &b = c->batch_set[k];
        HIP_TRY(hipMalloc(&b.seeds, B * sizeof(uint64_t)));
        HIP_TRY(hipMalloc(&b.tabC, S * kSetSize[SET_C] * 256));
        HIP_TRY(hipMalloc(&b.parC, S * kSetSize[SET_C] * sizeof(SPar)));
        HIP_TRY(hipMalloc(&b.q1, (size_t)kMaxItems * sizeof(uint4)));
        HIP_TRY(hipMalloc(&b.q2, (size_t)kMaxItems * sizeof(uint4)));
        HIP_TRY(hipMalloc(&b.needC, B * 4));
        HIP_TRY(hipMalloc(&b.slotC, B * 4));
        HIP_TRY(hipMalloc(&b.seedC, S * 4));
        HIP_TRY(hipMalloc(&b.needL, B * 4));
        HIP_TRY(hipMalloc(&b.ctl, 64));
    }
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < 3; k++) HIP_TRY(hipEventCreateWithFlags(&c->ev_stage[s][k], hipEventDisableTiming));
    HIP_TRY(hipMalloc(&c->d_cnt, CNT_N * sizeof(unsigned long long)));
    HIP_TRY(hipMalloc(&c->d_hits, (size_t)kHitCap * sizeof(DevHit)));
    HIP_TRY(hipMalloc(&c->d_gt, kLayouts * sizeof(GateTables)));
    HIP_TRY(hipMalloc(&c->d_bd, kLayouts * kBdPerLayout * sizeof(BuildDesc)));
    // layouts for every rule-flag combination, uploaded once
    std::vector<GateTables> gts(kLayouts);
    std::vector<BuildDesc> bds(kLayouts * kBdPerLayout);
    for (int v = 0; v < kLayouts; v++) {
        Layout L;
        build_layout(v, L);
        gts[v] = L.gt;
        size_t o = (size_t)v * kBdPerLayout;
        for (int s = 0; s < SET_COUNT; s++) {
            if ((int)L.bd[s].size() != kSetSize[s]) {
                std::snprintf(g_last_error, sizeof g_last_error, "sampler set %d has %zu entries", s, L.bd[s].size());
                return WHF_EINVAL;
            }
            std::copy(L.bd[s].begin(), L.bd[s].end(), bds.begin() + o);
            o += kSetSize[s];
        }
    }
    HIP_TRY(hipMemcpy(c->d_gt, gts.data(), gts.size() * sizeof(GateTables), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_bd, bds.data(), bds.size() * sizeof(BuildDesc), hipMemcpyHostToDevice));
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
    return WHF_OK;
}

}  // namespace

__attribute__((visibility("hidden"))) uint64_t whf_ctx_device_bytes(const whf_ctx *c) {
    const uint64_t S = kMaxSeeds, B = kMaxBatchSeeds;
    uint64_t b = B * 8 + 4 * B * 4 + 2 * S * 4 + 2ull * kMaxItems * sizeof(uint4) + (uint64_t)kQxCap * sizeof(uint4) +
                 (uint64_t)kHitCap * sizeof(DevHit) + kLayouts * (sizeof(GateTables) + kBdPerLayout * sizeof(BuildDesc));
    for (int s = 0; s < SET_COUNT; s++) b += S * kSetSize[s] * (256 + sizeof(SPar));
    // the second and third buffer sets of the pipelined Seed Check
    b += 2 * (B * 8 + S * kSetSize[SET_C] * (256 + sizeof(SPar)) + 2ull * kMaxItems * sizeof(uint4) + 3 * B * 4 +
              S * 4 + 64);
    return b + (c ? c->dd_bytes + c->bio_bytes + c->ter_bytes : 0);
}

namespace whf {

const BuildDesc *bd_of(whf_ctx *c, int layout, int set) {
    size_t o = (size_t)layout * kBdPerLayout;
    for (int s = 0; s < set; s++) o += kSetSize[s];
    return c->d_bd + o;
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::snprintf(g_last_error, sizeof g_last_error, "kernel launch: %s", hipGetErrorString(e));
        return WHF_EDEVICE;
    }
    return WHF_OK;
}

void launch_build(whf_ctx *c, const uint64_t *seeds, const BuildDesc *bd, int set_n, uint32_t n_slots,
                  uint8_t *tab, SPar *par, unsigned long long *built) {
    const uint64_t jobs = (uint64_t)n_slots * (uint32_t)set_n;
    hipLaunchKernelGGL(k_build, dim3((uint32_t)std::max<uint64_t>(1, (jobs + 63) / 64)), dim3(64), 0, c->stream, seeds,
                       bd, set_n, (const uint32_t *)nullptr, (const uint32_t *)nullptr, n_slots, tab, par, built,
                       (const uint8_t *)nullptr);
}

void launch_reset(whf_ctx *c, uint32_t ns) {
    Prof p(c, F_SEEDS);
    hipLaunchKernelGGL(k_reset, dim3(std::max(1u, (ns + 255) / 256)), dim3(256), 0, c->stream, c->d_ctl, c->d_needC,
                       c->d_needL, ns);
}

// Set C for the seeds with erosion survivors (slots by compaction, then tables) and gate 2.
void stage_c(whf_ctx *c, const BatchSet &b, int layout, uint32_t ns, int vflags, hipStream_t st) {
    const uint32_t blocks_n = std::max(1u, (ns + 255) / 256);
    const uint32_t persist = (uint32_t)c->n_cu * (uint32_t)c->tune[layout].gate_wg_per_cu;
    const uint32_t nb = std::min(ns, kMaxSeeds);   // at most kMaxSeeds seeds get tables; gates count the rest
    {
        Prof p(c, F_COMPACT, st);
        hipLaunchKernelGGL(k_compact, dim3(blocks_n), dim3(256), 0, st, b.needC, ns, b.slotC, b.seedC,
                           b.ctl + CTL_NC, kMaxSeeds);
    }
    {
        Prof p(c, F_BUILD_C, st);
        hipLaunchKernelGGL(k_build, dim3(std::min<uint32_t>(persist * 4, (nb * 28 + 63) / 64 + 1)), dim3(64), 0, st,
                           b.seeds, bd_of(c, layout, SET_C), 28, b.seedC, b.ctl + CTL_NC, kMaxSeeds, b.tabC, b.parC,
                           c->d_cnt + CNT_BUILT + SET_C, (const uint8_t *)nullptr);
    }
    {
        Prof p(c, F_GATE2, st);
        hipLaunchKernelGGL(k_gate2a<false>, dim3(persist), dim3(kGateWG), 0, st, c->d_gt + layout, b.tabC, b.parC,
                           b.slotC, b.q1, b.q2, b.ctl, vflags, c->d_cnt, SrvView{});
        hipLaunchKernelGGL(k_gate2b<false>, dim3(persist), dim3(kGateWG), 0, st, c->d_gt + layout, b.tabC, b.parC,
                           b.slotC, b.q2, b.q1, b.ctl, b.needL, c->d_cnt, SrvView{});
    }
}

// Set L for the seeds with gate-2 survivors, and gate 3 -> hits.
void stage_l(whf_ctx *c, const BatchSet &b, int layout, uint32_t ns, int vflags, int max_height, hipStream_t st) {
    const uint32_t blocks_n = std::max(1u, (ns + 255) / 256);
    const uint32_t persist = (uint32_t)c->n_cu * (uint32_t)c->tune[layout].gate_wg_per_cu;
    const uint32_t nb = std::min(ns, kMaxSeeds);
    {
        Prof p(c, F_COMPACT, st);
        hipLaunchKernelGGL(k_compact, dim3(blocks_n), dim3(256), 0, st, b.needL, ns, c->d_slotL, c->d_seedL,
                           b.ctl + CTL_NL, kMaxSeeds);
    }
    {
        Prof p(c, F_BUILD_L, st);
        hipLaunchKernelGGL(k_build, dim3(std::min<uint32_t>(persist, (nb * 36 + 63) / 64 + 1)), dim3(64), 0, st,
                           b.seeds, bd_of(c, layout, SET_L), 36, c->d_seedL, b.ctl + CTL_NL, kMaxSeeds,
                           c->d_tab[SET_L], c->d_par[SET_L], c->d_cnt + CNT_BUILT + SET_L, (const uint8_t *)nullptr);
    }
    {
        Prof p(c, F_GATE3, st);
        hipLaunchKernelGGL(k_gate3<false>, dim3(persist), dim3(kGateWG), 0, st, b.seeds, c->d_gt + layout, b.tabC,
                           b.parC, c->d_tab[SET_L], c->d_par[SET_L], b.slotC, c->d_slotL, b.q1, b.ctl, vflags, max_height,
                           c->d_hits, c->d_ctl + CTL_HITS, kHitCap, c->d_cnt, SrvView{});
    }
}

int run_tail(whf_ctx *c, int layout, uint32_t ns, int vflags, int max_height) {
    const BatchSet &b = c->batch_set[0];
    stage_c(c, b, layout, ns, vflags, c->stream);
    stage_l(c, b, layout, ns, vflags, max_height, c->stream);
    return check_launch();
}

void reset_set(whf_ctx *c, const BatchSet &b, uint32_t ns) {
    Prof p(c, F_SEEDS);
    hipLaunchKernelGGL(k_reset, dim3(std::max(1u, (ns + 255) / 256)), dim3(256), 0, c->stream, b.ctl, b.needC, b.needL,
                       ns);
}

int Pipeline::begin(uint64_t n_batches) {
    const hipStream_t st = c->stream;
    // one batch has nothing to overlap; then the stage streams, which share the device's few hardware queues with
    // other contexts' streams, stay idle
    piped = n_batches > 1 && !c->one_stream;
    if (piped)   // created on first use, of default priority like the library's stream
        for (hipStream_t &s : c->stage_stream)
            if (!s) HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ss[0] = st;
    ss[1] = piped ? c->stage_stream[0] : st;
    ss[2] = piped ? c->stage_stream[1] : st;
    if (piped) {
        HIP_TRY(hipEventRecord(c->ev_stage[0][0], st));   // the stage streams start after what st has so far
        for (int s = 1; s < 3; s++) HIP_TRY(hipStreamWaitEvent(ss[s], c->ev_stage[0][0], 0));
    }
    return WHF_OK;
}

int Pipeline::acquire(const BatchSet *&b) {
    b = &c->batch_set[set];
    // stage L of the batch three back is the last reader of this set
    if (piped && used[set]) HIP_TRY(hipStreamWaitEvent(ss[0], c->ev_stage[2][set], 0));
    return WHF_OK;
}

int Pipeline::tail(const BatchSet &b, int layout, uint32_t ns, int vflags, int max_height) {
    if (piped) {
        HIP_TRY(hipEventRecord(c->ev_stage[0][set], ss[0]));
        HIP_TRY(hipStreamWaitEvent(ss[1], c->ev_stage[0][set], 0));
    }
    stage_c(c, b, layout, ns, vflags, ss[1]);
    if (piped) {
        HIP_TRY(hipEventRecord(c->ev_stage[1][set], ss[1]));
        HIP_TRY(hipStreamWaitEvent(ss[2], c->ev_stage[1][set], 0));
    }
    stage_l(c, b, layout, ns, vflags, max_height, ss[2]);
    const int rc = check_launch();
    if (rc != WHF_OK) return rc;
    if (piped) HIP_TRY(hipEventRecord(c->ev_stage[2][set], ss[2]));
    used[set] = true;
    set = (set + 1) % 3;
    return WHF_OK;
}

// ---- served batches (packed on the device by whf_serve.hip's claim kernel) ----
// A serving lane runs whole batches on its own stream, with its own sampler tables, re-evaluation queue and set-L
// slots (ln), so that several lanes' batches run side by side.

// The batch's queue counts and the seed flags of every slot, on the lane's stream.
void serve_reset(const BatchSet &b, hipStream_t st) {
    hipLaunchKernelGGL(k_reset, dim3(kMaxSeeds / 256), dim3(256), 0, st, b.ctl, b.needC, b.needL, kMaxSeeds);
}

// Set E for the batch's slots, each with its request's rule layout, then the erosion gate over the tile classes of
// class_mask (bit k: class k, 0/1 FP64 with 4/8 seeds per tile, 2/3 FP32 gate with 4/8) and, with an FP32 class, the
// FP64 re-evaluation of the items the FP32 gate left undecided. Every grid is fixed; the kernels read the batch's slot
// and tile counts on the device.
void serve_stage_e(whf_ctx *c, const BatchSet &b, const SrvLane &ln, SBatch *sb, const SrvView &v, hipStream_t st,
                   int class_mask) {
    const uint32_t persist = (uint32_t)c->n_cu * (uint32_t)c->tune[0].gate_wg_per_cu;
    hipLaunchKernelGGL(k_build, dim3(persist * 4), dim3(64), 0, st, b.seeds, bd_of(c, 0, SET_E), 8,
                       (const uint32_t *)nullptr, &sb->n_slots, kMaxSeeds, ln.tabE, ln.parE,
                       (unsigned long long *)nullptr, v.lay_of);
    const int per_cu = c->tune[0].ero_wg_per_cu > 0 ? c->tune[0].ero_wg_per_cu : 8;
    const dim3 grid((uint32_t)c->n_cu * (uint32_t)per_cu);
    const FastDiv one{1u, 0u, 0u};
#define WHF_SRV_EROSION_ARGS(cls)                                                                                    \
    grid, dim3(256), 0, st, b.seeds, 0u, ln.tabE, ln.parE, c->d_gt, 0, 0, one, one, (uint64_t)0, 0u, 1u, 1u, 0, 0,   \
        b.q1, b.ctl, b.needC, ln.qx, c->d_cnt, v, &sb->tiles[cls][0], &sb->n_tiles[cls], &sb->tile_next[cls]
    if (class_mask & 1) hipLaunchKernelGGL((k_erosion<4, false, 256, 1, true>), WHF_SRV_EROSION_ARGS(0));
    if (class_mask & 2) hipLaunchKernelGGL((k_erosion<8, false, 256, 1, true>), WHF_SRV_EROSION_ARGS(1));
    if (class_mask & 4) hipLaunchKernelGGL((k_erosion<4, true, 256, 2, true>), WHF_SRV_EROSION_ARGS(2));
    if (class_mask & 8) hipLaunchKernelGGL((k_erosion<8, true, 256, 2, true>), WHF_SRV_EROSION_ARGS(3));
#undef WHF_SRV_EROSION_ARGS
    if (class_mask & 12)
        hipLaunchKernelGGL(k_erosion_exact<true>, dim3((uint32_t)c->n_cu * 2), dim3(256), 0, st, ln.tabE, ln.parE,
                           c->d_gt, ln.qx, b.q1, b.ctl, b.needC, c->d_cnt, v);
}

// Set C for the slots with erosion survivors and gate 2, with each slot's rule layout.
void serve_stage_c(whf_ctx *c, const BatchSet &b, const SrvView &v, hipStream_t st) {
    const uint32_t persist = (uint32_t)c->n_cu * (uint32_t)c->tune[0].gate_wg_per_cu;
    hipLaunchKernelGGL(k_compact, dim3(kMaxSeeds / 256), dim3(256), 0, st, b.needC, kMaxSeeds, b.slotC, b.seedC,
                       b.ctl + CTL_NC, kMaxSeeds);
    hipLaunchKernelGGL(k_build, dim3(persist * 4), dim3(64), 0, st, b.seeds, bd_of(c, 0, SET_C), 28, b.seedC,
                       b.ctl + CTL_NC, kMaxSeeds, b.tabC, b.parC, (unsigned long long *)nullptr, v.lay_of);
    hipLaunchKernelGGL(k_gate2a<true>, dim3(persist), dim3(kGateWG), 0, st, c->d_gt, b.tabC, b.parC, b.slotC, b.q1,
                       b.q2, b.ctl, 0, c->d_cnt, v);
    hipLaunchKernelGGL(k_gate2b<true>, dim3(persist), dim3(kGateWG), 0, st, c->d_gt, b.tabC, b.parC, b.slotC, b.q2,
                       b.q1, b.ctl, b.needL, c->d_cnt, v);
}

// Set L for the slots with gate-2 survivors and gate 3, whose survivors go to their requests' queues.
void serve_stage_l(whf_ctx *c, const BatchSet &b, const SrvLane &ln, const SrvView &v, hipStream_t st) {
    const uint32_t persist = (uint32_t)c->n_cu * (uint32_t)c->tune[0].gate_wg_per_cu;
    hipLaunchKernelGGL(k_compact, dim3(kMaxSeeds / 256), dim3(256), 0, st, b.needL, kMaxSeeds, ln.slotL, ln.seedL,
                       b.ctl + CTL_NL, kMaxSeeds);
    hipLaunchKernelGGL(k_build, dim3(persist), dim3(64), 0, st, b.seeds, bd_of(c, 0, SET_L), 36, ln.seedL,
                       b.ctl + CTL_NL, kMaxSeeds, ln.tabL, ln.parL, (unsigned long long *)nullptr, v.lay_of);
    hipLaunchKernelGGL(k_gate3<true>, dim3(persist), dim3(kGateWG), 0, st, b.seeds, c->d_gt, b.tabC, b.parC, ln.tabL,
                       ln.parL, b.slotC, ln.slotL, b.q1, b.ctl, 0, 0, (DevHit *)nullptr, (uint32_t *)nullptr, 0u,
                       c->d_cnt, v);
}

int Pipeline::end() {
    for (int s = 1; s < 3 && piped; s++) {   // everything on the stage streams so far
        HIP_TRY(hipEventRecord(c->ev_stage[s][0], ss[s]));
        HIP_TRY(hipStreamWaitEvent(ss[0], c->ev_stage[s][0], 0));
    }
    return WHF_OK;
}

}  // namespace whf

namespace {

// The first stage of one batch (seeds already in b.seeds): seeds [s0, s0 + ns) of the request, flat region range
// [f0, f0 + fn): reset, set E, the erosion gate.
int stage_e(whf_ctx *c, const whf_request *req, int layout, uint32_t ns, uint64_t f0, uint32_t fn, uint32_t W,
            const BatchSet &b) {
    const hipStream_t st = c->stream;
    reset_set(c, b, ns);
    {
        Prof p(c, F_BUILD_E);
        const uint32_t jobs = ns * 8;
        hipLaunchKernelGGL(k_build, dim3((jobs + 63) / 64), dim3(64), 0, st, b.seeds, bd_of(c, layout, SET_E), 8,
                           (const uint32_t *)nullptr, (const uint32_t *)nullptr, ns, c->d_tab[SET_E], c->d_par[SET_E],
                           c->d_cnt + CNT_BUILT + SET_E, (const uint8_t *)nullptr);
    }
    {
        Prof p(c, F_EROSION);
        const int salt = req->salt ? req->salt : 14357620;
        const int single = (req->vflags & WHF_V_SINGLE) ? 1 : 0;
        const GateTables *gt = c->d_gt + layout;
        // seeds per tile: narrow rects want many seeds per workgroup to fill the lanes
        const bool wide = fn >= 64;
        const int ST = wide ? 4 : 8;
        const uint32_t tiles = (ns + ST - 1) / ST;
        // regions per block: up to 2^16, but a batch of few seeds (few tiles) is cut finer so that it still has 8
        // blocks per CU (one seed over the whole world would otherwise be 512 blocks of 2^16 regions a batch)
        const uint32_t want = (uint32_t)c->n_cu * 8;
        uint32_t chunk = std::min<uint32_t>(fn, 1u << 16);
        if ((uint64_t)tiles * ((fn + chunk - 1) / chunk) < want)
            chunk = std::min(chunk, std::max<uint32_t>((uint32_t)kEroPass, (uint32_t)(((uint64_t)fn * tiles + want - 1) / want)));
        const uint32_t n_chunks = (fn + chunk - 1) / chunk;
        // persistent: ero_wg_per_cu workgroups per compute unit loop over the blocks (-1: one workgroup per block)
        const uint32_t blocks = tiles * n_chunks;
        const int per_cu = c->tune[layout].ero_wg_per_cu;
        const dim3 grid(per_cu > 0 ? std::min<uint32_t>(blocks, (uint32_t)c->n_cu * (uint32_t)per_cu) : blocks);
        const bool f32 = req->precision == WHF_PREC_FP32_GATE;
        const FastDiv Wd = make_fastdiv(W), Rd = make_fastdiv(chunk);
        const int wg = c->tune[layout].ero_wg == 512 ? 512 : 256;
        const int il = c->tune[layout].ero_ilp == 1 ? 1 : 2;   // items per lane of the FP32 gate
#define WHF_EROSION_ARGS(WG)                                                                                      \
    grid, dim3(WG), 0, st, b.seeds, ns, c->d_tab[SET_E], c->d_par[SET_E], gt, req->rx0, req->rz0, Wd, Rd, f0, fn,      \
        chunk, n_chunks, salt, single, b.q1, b.ctl, b.needC, c->d_qx, c->d_cnt, SrvView{}, (const STile *)nullptr,     \
        (const uint32_t *)nullptr, (uint32_t *)nullptr
        if (!f32) {
            if (wide && wg == 512) hipLaunchKernelGGL((k_erosion<4, false, 512, 1>), WHF_EROSION_ARGS(512));
            else if (wide) hipLaunchKernelGGL((k_erosion<4, false, 256, 1>), WHF_EROSION_ARGS(256));
            else if (wg == 512) hipLaunchKernelGGL((k_erosion<8, false, 512, 1>), WHF_EROSION_ARGS(512));
            else hipLaunchKernelGGL((k_erosion<8, false, 256, 1>), WHF_EROSION_ARGS(256));
        } else if (wg == 512) {
            if (wide) hipLaunchKernelGGL((k_erosion<4, true, 512, 1>), WHF_EROSION_ARGS(512));
            else hipLaunchKernelGGL((k_erosion<8, true, 512, 1>), WHF_EROSION_ARGS(512));
        } else if (il == 2) {
            if (wide) hipLaunchKernelGGL((k_erosion<4, true, 256, 2>), WHF_EROSION_ARGS(256));
            else hipLaunchKernelGGL((k_erosion<8, true, 256, 2>), WHF_EROSION_ARGS(256));
        } else {
            if (wide) hipLaunchKernelGGL((k_erosion<4, true, 256, 1>), WHF_EROSION_ARGS(256));
            else hipLaunchKernelGGL((k_erosion<8, true, 256, 1>), WHF_EROSION_ARGS(256));
        }
#undef WHF_EROSION_ARGS
        if (f32 && !single)
            hipLaunchKernelGGL(k_erosion_exact<false>, dim3((uint32_t)c->n_cu * 2), dim3(256), 0, st, c->d_tab[SET_E],
                               c->d_par[SET_E], gt, c->d_qx, b.q1, b.ctl, b.needC, c->d_cnt, SrvView{});
    }
    return check_launch();
}

}  // namespace

namespace whf {

// lysh's phase 2 on phase-1 survivors (lysh_grade_scan, eval.c:381-477): the hut verdict of each, keeping the accepted
// huts (in place, in order) with their floor; counters evaluated, accepted, rejected for the biome, the height
// (avg_y above the threshold, or not equal to it with 'exactly') and the flood.
int phase2_tree(const whf_request *req, int *tree) {
    *tree = req->biome_tree - 1;
    if (req->biome_tree == 0) {
        const char *dflt = (req->vflags & WHF_V_118) ? "1.18.2" : ((req->vflags & WHF_V_262) ? "26.2" : "1.21.5");
        if (whf_biome_tree(dflt, tree) != WHF_OK) return WHF_EINVAL;
    }
    return WHF_OK;
}

int phase2_filter(whf_ctx *c, const whf_request *req, whf_hit *hits, uint64_t n, uint64_t *kept, uint64_t p2[5],
                  double *ms) {
    const auto t0 = std::chrono::steady_clock::now();
    *kept = 0;
    int tree;
    if (phase2_tree(req, &tree) != WHF_OK) return WHF_EINVAL;
    std::vector<whf_hut_query> hq(n);
    std::vector<whf_hut_result> hr(n);
    for (uint64_t i = 0; i < n; i++) hq[i] = whf_hut_query{hits[i].seed, hits[i].cx, hits[i].cz};
    const int rc = whf_hut_eval(c, hq.data(), n, tree, req->vflags & WHF_V_LARGE, req->phase2_max_y, hr.data());
    if (rc != WHF_OK) return rc;
    phase2_apply(req, hits, n, hr.data(), kept, p2);
    *ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return WHF_OK;
}

void phase2_apply(const whf_request *req, whf_hit *hits, uint64_t n, const whf_hut_result *hr, uint64_t *kept,
                  uint64_t p2[5]) {
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; i++) {
        const whf_hut_result &h = hr[i];
        p2[0]++;
        // the verdict's floor test is its last step (biome, flood, height: eval.c:221-224), so it is applied here
        // with the request's own threshold whatever max_y the evaluation ran with
        bool ok = h.biome_ok && !h.flooded;
        int reject = !h.biome_ok ? WHF_HUT_REJECT_BIOME : (h.flooded ? WHF_HUT_REJECT_FLOOD : WHF_HUT_OK);
        if (ok && h.avg_y > req->phase2_max_y) { ok = false; reject = WHF_HUT_REJECT_Y; }
        if (ok && req->phase2 == WHF_P2_EXACTLY && h.avg_y != req->phase2_max_y) { ok = false; reject = WHF_HUT_REJECT_Y; }
        if (ok) {
            p2[1]++;
            whf_hit d = hits[i];
            d.floor_y = h.avg_y;
            d.flags |= WHF_HIT_PHASE2;
            hits[k++] = d;
        } else {
            p2[reject == WHF_HUT_REJECT_BIOME ? 2 : (reject == WHF_HUT_REJECT_Y ? 3 : 4)]++;
        }
    }
    *kept = k;
}

}  // namespace whf

// ================================================================================================ C ABI

namespace {

#ifndef WHF_BUILD_ID
#define WHF_BUILD_ID "unknown"
#endif

void tune_defaults(whf_tuning &t) {
    std::memset(&t, 0, sizeof t);
    t.ero_wg = 256;
    t.ero_ilp = 2;
    t.ero_wg_per_cu = -1;
    // gates 2 and 3 wait on gathers: more resident workgroups than fit at once keep the CUs fed between batches
    // (2^24-seed, 256-region scan: 4 per CU 308, 8 299, 12 297, 16 291, 20 292, 24 294, 32 293 ms)
    t.gate_wg_per_cu = 16;
    t.chunk_checks = 1ull << 24;
}

// Fills the zero fields of t with the defaults; false when a field is out of range.
bool tune_complete(whf_tuning &t) {
    whf_tuning d;
    tune_defaults(d);
    if (!t.ero_wg) t.ero_wg = d.ero_wg;
    if (!t.ero_ilp) t.ero_ilp = d.ero_ilp;
    if (!t.ero_wg_per_cu) t.ero_wg_per_cu = d.ero_wg_per_cu;
    if (!t.gate_wg_per_cu) t.gate_wg_per_cu = d.gate_wg_per_cu;
    if (!t.chunk_checks) t.chunk_checks = d.chunk_checks;
    return (t.ero_wg == 256 || t.ero_wg == 512) && (t.ero_ilp == 1 || t.ero_ilp == 2) && t.ero_wg_per_cu >= -1 &&
           t.ero_wg_per_cu <= 64 && t.gate_wg_per_cu >= 1 && t.gate_wg_per_cu <= 32 &&
           t.chunk_checks >= 1024 && t.chunk_checks <= (1ull << 32);
}

// A small reader for the tuned-configuration file: a JSON object whose keys are device names, each an object of
// numeric launch-shape fields plus "build" (the library build it was measured with) and anything else, skipped.
struct Json {
    const std::string &s;
    size_t i = 0;
    explicit Json(const std::string &t) : s(t) {}
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) i++; }
    bool str(std::string &out) {
        ws();
        if (i >= s.size() || s[i] != '"') return false;
        out.clear();
        for (i++; i < s.size() && s[i] != '"'; i++) {
            if (s[i] == '\\' && i + 1 < s.size()) i++;
            out += s[i];
        }
        if (i >= s.size()) return false;
        i++;
        return true;
    }
    bool skip() {   // any value
        ws();
        if (i >= s.size()) return false;
        if (s[i] == '"') { std::string t; return str(t); }
        if (s[i] == '{' || s[i] == '[') {
            const char open = s[i], close = open == '{' ? '}' : ']';
            int depth = 0;
            for (; i < s.size(); i++) {
                if (s[i] == '"') { std::string t; if (!str(t)) return false; i--; continue; }
                if (s[i] == open) depth++;
                else if (s[i] == close && --depth == 0) { i++; return true; }
            }
            return false;
        }
        while (i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ']') i++;
        return true;
    }
    // Calls f(key) at each member of the object at the cursor, with the cursor on the member's value; f consumes it.
    template <typename F>
    bool object(F f) {
        ws();
        if (i >= s.size() || s[i] != '{') return false;
        i++;
        for (;;) {
            ws();
            if (i < s.size() && s[i] == '}') { i++; return true; }
            std::string k;
            if (!str(k)) return false;
            ws();
            if (i >= s.size() || s[i] != ':') return false;
            i++;
            if (!f(k)) return false;
            ws();
            if (i < s.size() && s[i] == ',') i++;
        }
    }
    bool number(double &v) {
        ws();
        char *end = nullptr;
        v = std::strtod(s.c_str() + i, &end);
        if (end == s.c_str() + i) return false;
        i = (size_t)(end - s.c_str());
        return true;
    }
};

// The path of the tuned-configuration file: $WHF_TUNED, else tuned.json beside this library.
std::string tuned_path() {
    if (const char *e = std::getenv("WHF_TUNED")) return e;
    Dl_info info;
    if (dladdr((void *)&tune_defaults, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t k = p.rfind('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/tuned.json";
    }
    return "tuned.json";
}

// The tuning of every rule layout for `device` from the file at `path`: the device's entry, else the "default"
// entry, when it was measured with this build; its "versions" object overrides fields per game version (for every
// preset of that version's rules). Returns whether an entry applied; out[] holds the defaults otherwise.
bool tune_load(whf_tuning out[kLayouts], const std::string &path, const char *device) {
    for (int l = 0; l < kLayouts; l++) tune_defaults(out[l]);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    std::fclose(f);
    struct Entry {
        whf_tuning base;
        std::vector<std::pair<int, whf_tuning>> per_version;   // (rule flags, fields given)
        bool ok = false;
    } found[2];
    Json j(text);
    auto fields = [&](whf_tuning &e, const std::string &k) -> int {   // 1 parsed, 0 not a field, -1 error
        double v = 0.0;
        int *ip = k == "ero_wg" ? &e.ero_wg : k == "ero_ilp" ? &e.ero_ilp : k == "ero_wg_per_cu" ? &e.ero_wg_per_cu
                  : k == "gate_wg_per_cu" ? &e.gate_wg_per_cu : nullptr;
        if (ip) { if (!j.number(v)) return -1; *ip = (int)v; return 1; }
        if (k == "chunk_checks") { if (!j.number(v)) return -1; e.chunk_checks = (uint64_t)v; return 1; }
        return 0;
    };
    const bool parsed = j.object([&](const std::string &dev) {
        const int which = dev == device ? 0 : (dev == "default" ? 1 : -1);
        if (which < 0) return j.skip();
        Entry e;
        std::memset(&e.base, 0, sizeof e.base);
        std::string build;
        const bool ok = j.object([&](const std::string &k) {
            if (k == "build") return j.str(build);
            if (k == "versions") {
                return j.object([&](const std::string &ver) {
                    int vf = 0;
                    if (whf_version_flags(ver.c_str(), &vf) != WHF_OK) return j.skip();
                    whf_tuning t;
                    std::memset(&t, 0, sizeof t);
                    const bool vok = j.object([&](const std::string &k2) {
                        const int r = fields(t, k2);
                        return r < 0 ? false : (r == 0 ? j.skip() : true);
                    });
                    if (vok) e.per_version.emplace_back(vf, t);
                    return vok;
                });
            }
            const int r = fields(e.base, k);
            return r < 0 ? false : (r == 0 ? j.skip() : true);
        });
        e.ok = ok && build == WHF_BUILD_ID;
        if (e.ok) found[which] = e;
        return ok;
    });
    if (!parsed) return false;
    const Entry *e = found[0].ok ? &found[0] : (found[1].ok ? &found[1] : nullptr);
    if (!e) return false;
    for (int l = 0; l < kLayouts; l++) {
        whf_tuning t = e->base;
        for (const auto &pv : e->per_version) {
            if ((l & (WHF_V_118 | WHF_V_262)) != pv.first) continue;
            const whf_tuning &o = pv.second;
            if (o.ero_wg) t.ero_wg = o.ero_wg;
            if (o.ero_ilp) t.ero_ilp = o.ero_ilp;
            if (o.ero_wg_per_cu) t.ero_wg_per_cu = o.ero_wg_per_cu;
            if (o.gate_wg_per_cu) t.gate_wg_per_cu = o.gate_wg_per_cu;
            if (o.chunk_checks) t.chunk_checks = o.chunk_checks;
        }
        if (!tune_complete(t)) return false;
        out[l] = t;
    }
    return true;
}

}  // namespace

extern "C" {

const char *whf_strerror(int code) {
    switch (code) {
    case WHF_OK: return "ok";
    case WHF_EINVAL: return "invalid argument";
    case WHF_ENOMEM: return "out of memory or over the HBM budget";
    case WHF_EDEVICE: return g_last_error[0] ? g_last_error : "device error";
    case WHF_EOVERFLOW: return "more survivors than the sink holds";
    case WHF_ENOTSUP: return "not supported";
    }
    return "unknown error";
}

int whf_version_flags(const char *name, int *vflags) {
    // game versions (biome_trees.h:70499-70526) and lysh's aliases (main.c:104-125)
    static const struct { const char *n; int f; } V[] = {
        {"1.18", WHF_V_118}, {"1.18.1", WHF_V_118}, {"1.18.2", WHF_V_118}, {"1.18.x", WHF_V_118},
        {"1.19", 0}, {"1.19.2", 0}, {"1.19.3", 0}, {"1.19.4", 0}, {"1.19.x", 0}, {"1.20", 0}, {"1.20.1", 0},
        {"1.20.2", 0}, {"1.20.4", 0}, {"1.20.5", 0}, {"1.20.6", 0}, {"1.20.x", 0}, {"1.21", 0}, {"1.21.1", 0},
        {"1.21.2", 0}, {"1.21.3", 0}, {"1.21.4", 0}, {"1.21.5", 0}, {"1.21.8", 0}, {"1.21.11", 0},
        {"1.21.x~26.1", 0}, {"1.21.5~26.1", 0}, {"1.19.x~26.1", 0}, {"26.1", 0}, {"26.2", WHF_V_262},
        {"26.3", WHF_V_262}, {"26.2~26.3", WHF_V_262},
    };
    if (!name || !vflags) return WHF_EINVAL;
    for (const auto &v : V)
        if (!std::strcmp(v.n, name)) { *vflags = v.f; return WHF_OK; }
    return WHF_EINVAL;
}

int whf_open(const whf_config *cfg, whf_ctx **out) {
    if (!out) return WHF_EINVAL;
    *out = nullptr;
    g_last_error[0] = 0;
    int dev = cfg ? cfg->device : 0;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (dev < 0 || dev >= n) {
        std::snprintf(g_last_error, sizeof g_last_error, "device %d not present (%d visible)", dev, n);
        return WHF_EDEVICE;
    }
    HIP_TRY(hipSetDevice(dev));
    whf_ctx *c = new (std::nothrow) whf_ctx();
    if (!c) return WHF_ENOMEM;
    c->device = dev;
    c->profile = cfg && (cfg->flags & WHF_CFG_PROFILE);
    c->one_stream = cfg && (cfg->flags & WHF_CFG_ONE_STREAM);
    hipDeviceProp_t prop;
    for (int l = 0; l < kLayouts; l++) tune_defaults(c->tune[l]);
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        // the device's key in tuned.json: its marketing name, or "AMD <arch>" where the runtime reports none
        std::string arch = prop.gcnArchName;
        arch = arch.substr(0, arch.find(':'));
        const std::string key = prop.name[0] ? std::string(prop.name) : "AMD " + arch;
        std::snprintf(c->name, sizeof c->name, "%s (%s)", key.c_str(), prop.gcnArchName);
        c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        tune_load(c->tune, tuned_path(), key.c_str());
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return WHF_EDEVICE;
    }
    int rc = ctx_alloc(c);
    if (rc != WHF_OK) {
        whf_close(c);
        return rc == WHF_EDEVICE ? WHF_ENOMEM : rc;
    }
    if ((rc = dd_scratch_init(c)) != WHF_OK) {
        whf_close(c);
        return rc;
    }
    *out = c;
    return WHF_OK;
}

void whf_close(whf_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->d_seeds, c->d_tab[0], c->d_tab[1], c->d_tab[2], c->d_par[0], c->d_par[1], c->d_par[2],
                    c->d_q1, c->d_q2, c->d_qx, c->d_needC, c->d_needL, c->d_slotC, c->d_slotL, c->d_seedC, c->d_seedL,
                    c->d_ctl, c->d_cnt, c->d_hits, c->d_gt, c->d_bd, c->d_dd, c->d_bio, c->d_ter, c->d_solve};
    for (hipStream_t s : c->stage_stream)
        if (s) (void)hipStreamSynchronize(s);
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (int k = 1; k < 3; k++) {
        const BatchSet &b = c->batch_set[k];
        void *own[] = {b.seeds, b.tabC, b.parC, b.q1, b.q2, b.needC, b.slotC, b.seedC, b.needL, b.ctl};
        for (void *p : own)
            if (p) (void)hipFree(p);
    }
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int s = 0; s < 3; s++)
        for (int k = 0; k < 3; k++)
            if (c->ev_stage[s][k]) (void)hipEventDestroy(c->ev_stage[s][k]);
    for (hipStream_t s : c->stage_stream)
        if (s) (void)hipStreamDestroy(s);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *whf_device_name(whf_ctx *c) { return c ? c->name : ""; }

const char *whf_build_id(void) { return WHF_BUILD_ID; }

int whf_get_tuning(whf_ctx *c, whf_tuning *out) { return whf_get_tuning_for(c, 0, out); }

int whf_set_tuning(whf_ctx *c, const whf_tuning *t) {
    if (!c || !t) return WHF_EINVAL;
    whf_tuning v = *t;
    if (!tune_complete(v)) return WHF_EINVAL;
    for (int l = 0; l < kLayouts; l++) c->tune[l] = v;
    return WHF_OK;
}

int whf_get_tuning_for(whf_ctx *c, int vflags, whf_tuning *out) {
    if (!c || !out || vflags < 0 || vflags >= kLayouts) return WHF_EINVAL;
    *out = c->tune[vflags];
    return WHF_OK;
}

int whf_set_tuning_for(whf_ctx *c, int vflags, const whf_tuning *t) {
    if (!c || !t || vflags < 0 || vflags >= kLayouts) return WHF_EINVAL;
    whf_tuning v = *t;
    if (!tune_complete(v)) return WHF_EINVAL;
    c->tune[vflags] = v;
    return WHF_OK;
}

int whf_load_tuning(const char *path, const char *device, int vflags, whf_tuning *out) {
    if (!device || !out || vflags < 0 || vflags >= kLayouts) return WHF_EINVAL;
    whf_tuning all[kLayouts];
    const bool ok = tune_load(all, path ? std::string(path) : tuned_path(), device);
    *out = all[vflags];
    return ok ? WHF_OK : WHF_EINVAL;
}

int whf_scan(whf_ctx *c, const whf_request *req, whf_sink *sink, whf_stats *stats) {
    if (!c || !req) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (req->rx1 <= req->rx0 || req->rz1 <= req->rz0) return WHF_EINVAL;
    if (req->seed_mode < WHF_SEEDS_LIST || req->seed_mode > WHF_SEEDS_COUNTER) return WHF_EINVAL;
    if (req->seed_mode == WHF_SEEDS_LIST && req->n_seeds && !req->seeds) return WHF_EINVAL;
    if (req->precision != WHF_PREC_FP64 && req->precision != WHF_PREC_FP32_GATE) return WHF_EINVAL;
    if (req->vflags < 0 || req->vflags >= kLayouts) return WHF_EINVAL;
    if (req->phase2 < WHF_P2_NONE || req->phase2 > WHF_P2_EXACTLY) return WHF_EINVAL;
    if (req->biome_tree < 0 || req->biome_tree > 8) return WHF_EINVAL;
    // gate 3 keeps one bit per check: the floodedness heights max_height..60 must fit beside the ladders
    if (req->max_height < -140 || req->max_height > 100000) return WHF_EINVAL;
    // chunk coordinates travel biased in 23 bits through the erosion kernel
    const int64_t lim = 65536;
    if (req->rx0 < -lim || req->rx1 > lim || req->rz0 < -lim || req->rz1 > lim) return WHF_EINVAL;
    HIP_TRY(hipSetDevice(c->device));

    const uint32_t W = (uint32_t)(req->rx1 - req->rx0);
    const uint64_t R = (uint64_t)W * (uint64_t)(req->rz1 - req->rz0);
    const int layout = req->vflags;
    const hipStream_t st = c->stream;
    HIP_TRY(hipMemsetAsync(c->d_cnt, 0, CNT_N * sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(c->d_ctl, 0, CTL_N * sizeof(uint32_t), st));
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    c->prof_ev.clear();
    c->prof_fam.clear();
    HIP_TRY(hipEventRecord(c->ev0, st));

    // batch plan: seeds per batch x regions per batch <= kMaxItems
    uint32_t seeds_per, regions_per;
    if (R <= kMaxItems) {
        regions_per = (uint32_t)R;
        seeds_per = (uint32_t)std::min<uint64_t>(kMaxSeeds, kMaxItems / R);
    } else {
        regions_per = kMaxItems;
        seeds_per = 1;
    }
    // batch k runs in three stages (Pipeline): seeds, set E and the erosion gate; set C and gate 2; set L and gate 3
    const uint64_t n_batches = ((req->n_seeds + seeds_per - 1) / seeds_per) * ((R + regions_per - 1) / regions_per);
    Pipeline pl(c);
    if (int rc = pl.begin(n_batches)) return rc;
    for (uint64_t s0 = 0; s0 < req->n_seeds; s0 += seeds_per) {
        const uint32_t ns = (uint32_t)std::min<uint64_t>(seeds_per, req->n_seeds - s0);
        for (uint64_t f0 = 0; f0 < R; f0 += regions_per) {
            const uint32_t fn = (uint32_t)std::min<uint64_t>(regions_per, R - f0);
            const BatchSet *bp = nullptr;
            if (int rc = pl.acquire(bp)) return rc;
            const BatchSet &b = *bp;
            if (req->seed_mode == WHF_SEEDS_LIST) {
                HIP_TRY(hipMemcpyAsync(b.seeds, req->seeds + s0, ns * sizeof(uint64_t), hipMemcpyHostToDevice, st));
            } else {
                Prof p(c, F_SEEDS);
                hipLaunchKernelGGL(k_seeds, dim3((ns + 255) / 256), dim3(256), 0, st, req->seed_mode, req->stream,
                                   req->offset + s0, ns, b.seeds);
            }
            int rc = stage_e(c, req, layout, ns, f0, fn, W, b);
            if (rc) return rc;
            if ((rc = pl.tail(b, layout, ns, req->vflags, req->max_height)) != WHF_OK) return rc;
        }
    }
    if (int rc = pl.end()) return rc;
    HIP_TRY(hipEventRecord(c->ev1, st));
    unsigned long long cnt[CNT_N];
    uint32_t ctl[CTL_N];
    HIP_TRY(hipMemcpyAsync(cnt, c->d_cnt, sizeof cnt, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(ctl, c->d_ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (cnt[CNT_QX_OVER]) {
        // the FP32 gate's re-evaluation queue overflowed in some batch: redo the whole request in FP64
        whf_request r64 = *req;
        r64.precision = WHF_PREC_FP64;
        return whf_scan(c, &r64, sink, stats);
    }
    uint64_t hits_total = ctl[CTL_HITS];
    std::vector<DevHit> host_hits(std::min<uint64_t>(hits_total, kHitCap));
    if (!host_hits.empty())
        HIP_TRY(hipMemcpy(host_hits.data(), c->d_hits, host_hits.size() * sizeof(DevHit), hipMemcpyDeviceToHost));
    int rc = WHF_OK;
    uint64_t p2[5] = {0, 0, 0, 0, 0};
    double p2_ms = 0.0;
    if (req->phase2 != WHF_P2_NONE) {
        if (hits_total > host_hits.size()) return WHF_EOVERFLOW;
        uint64_t kept = 0;
        const int rc2 = phase2_filter(c, req, reinterpret_cast<whf_hit *>(host_hits.data()), host_hits.size(), &kept,
                                      p2, &p2_ms);
        if (rc2 != WHF_OK) return rc2;
        host_hits.resize(kept);
        hits_total = kept;
    }
    if (sink) {
        sink->count = hits_total;
        const uint64_t m = std::min<uint64_t>(sink->capacity, host_hits.size());
        if (m && sink->hits) std::memcpy(sink->hits, host_hits.data(), m * sizeof(whf_hit));
        if (hits_total > sink->capacity || host_hits.size() < hits_total) rc = WHF_EOVERFLOW;
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->seeds = req->n_seeds;
        stats->checks = req->n_seeds * R;
        uint64_t ero = 0;
        for (int t = 0; t < 7; t++) { stats->tiers[t] = cnt[CNT_TIER0 + t]; ero += cnt[CNT_TIER0 + t]; }
        stats->funnel[0] = ero;
        for (int k = 1; k < WHF_KIND_N; k++) stats->funnel[k] = cnt[CNT_KIND0 + k];
        stats->accepted = cnt[CNT_KIND0 + 10];
        stats->reached_ladder = cnt[CNT_LADDER];
        for (int s = 0; s < SET_COUNT; s++) stats->samplers_built[s] = cnt[CNT_BUILT + s];
        stats->fp64_rechecks = cnt[CNT_QX];
        for (int k = 0; k < 5; k++) stats->phase2[k] = p2[k];
        stats->phase2_ms = p2_ms;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        stats->device_ms = ms;
        if (c->profile) {
            for (size_t i = 0; i + 1 < c->prof_ev.size(); i += 2) {
                float m2 = 0.f;
                (void)hipEventElapsedTime(&m2, c->prof_ev[i], c->prof_ev[i + 1]);
                const int f = c->prof_fam[i / 2];
                if (f >= 0 && f < 8) stats->kernel_ms[f] += m2;
            }
        }
    }
    return rc;
}

int whf_probe(whf_ctx *c, const int64_t *seeds, const int32_t *hut_x, const int32_t *hut_z, uint64_t n,
              int max_height, int vflags, double *out) {
    if (!c || (n && (!seeds || !hut_x || !hut_z || !out))) return WHF_EINVAL;
    if (vflags < 0 || vflags >= kLayouts) return WHF_EINVAL;
    if (n > kMaxSeeds) return WHF_EINVAL;
    if (!n) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t st = c->stream;
    const uint32_t nn = (uint32_t)n;
    int32_t *d_xz = nullptr;
    double *d_out = nullptr;
    HIP_TRY(hipMalloc(&d_xz, 2 * n * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&d_out, 38 * n * sizeof(double)));
    HIP_TRY(hipMemcpyAsync(c->d_seeds, seeds, n * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_xz, hut_x, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_xz + n, hut_z, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    for (int s = 0; s < SET_COUNT; s++) {
        const uint32_t jobs = nn * kSetSize[s];
        hipLaunchKernelGGL(k_build, dim3((jobs + 63) / 64), dim3(64), 0, st, c->d_seeds, bd_of(c, vflags, s),
                           kSetSize[s], (const uint32_t *)nullptr, (const uint32_t *)nullptr, nn, c->d_tab[s],
                           c->d_par[s], c->d_cnt + CNT_BUILT + s, (const uint8_t *)nullptr);
    }
    hipLaunchKernelGGL(k_probe, dim3((nn + 63) / 64), dim3(64), 0, st, c->d_gt + vflags, c->d_tab[SET_E],
                       c->d_par[SET_E], c->d_tab[SET_C], c->d_par[SET_C], c->d_tab[SET_L], c->d_par[SET_L], d_xz,
                       d_xz + n, nn, max_height, d_out);
    int rc = check_launch();
    if (rc == WHF_OK) {
        HIP_TRY(hipMemcpyAsync(out, d_out, 38 * n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    (void)hipFree(d_xz);
    (void)hipFree(d_out);
    return rc;
}

int whf_noise_error(whf_ctx *c, const int64_t *seeds, uint64_t n_seeds, uint64_t points_per_seed, int vflags,
                    double max_err[8]) {
    if (!c || !max_err || (n_seeds && !seeds)) return WHF_EINVAL;
    if (n_seeds > kMaxSeeds || vflags < 0 || vflags >= kLayouts || points_per_seed > (1u << 20)) return WHF_EINVAL;
    for (int k = 0; k < 8; k++) max_err[k] = 0.0;
    if (!n_seeds || !points_per_seed) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t st = c->stream;
    const uint32_t ns = (uint32_t)n_seeds, np = (uint32_t)points_per_seed;
    unsigned long long *d_max = nullptr;
    HIP_TRY(hipMalloc(&d_max, 8 * sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(d_max, 0, 8 * sizeof(unsigned long long), st));
    HIP_TRY(hipMemcpyAsync(c->d_seeds, seeds, n_seeds * sizeof(int64_t), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_build, dim3((ns * 8 + 63) / 64), dim3(64), 0, st, c->d_seeds, bd_of(c, vflags, SET_E), 8,
                       (const uint32_t *)nullptr, (const uint32_t *)nullptr, ns, c->d_tab[SET_E], c->d_par[SET_E],
                       c->d_cnt + CNT_BUILT + SET_E, (const uint8_t *)nullptr);
    const uint64_t total = (uint64_t)ns * np;
    hipLaunchKernelGGL(k_noise_error, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, c->d_tab[SET_E],
                       c->d_par[SET_E], c->d_gt + vflags, ns, np, d_max);
    int rc = check_launch();
    unsigned long long bits[8];
    if (rc == WHF_OK) {
        HIP_TRY(hipMemcpyAsync(bits, d_max, sizeof bits, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int k = 0; k < 8; k++) std::memcpy(&max_err[k], &bits[k], 8);
    }
    (void)hipFree(d_max);
    return rc;
}

double whf_fp32_bound(void) { return kFp32PerlinBound; }

}  // extern "C"
