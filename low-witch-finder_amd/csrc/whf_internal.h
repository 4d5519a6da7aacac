// whf_internal.h — what the library's two translation units share: the context, the device control words and
// counters, the wave-aggregated queue append, error and profiling helpers, and the Seed Check stages that the
// Deep Dive expansion reuses (whf_scan.hip defines them, whf_deepdive.hip calls them).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/whf.h"
#include "whf_kernels.h"
#include "whf_tables.h"

namespace whf {

constexpr int kWave = 64;

// Device control words of a batch.
// CTL_QM: gate-2 survivors of temperature and ridge, before the cave ladder. Words below CTL_HITS are per batch.
enum Ctl { CTL_Q1 = 0, CTL_Q2, CTL_NC, CTL_NL, CTL_QX, CTL_QM, CTL_HITS, CTL_DD_STARTS, CTL_DD_SPOTS, CTL_DD_WORK, CTL_N };

// Device counters of a scan (u64).
enum Cnt {
    CNT_TIER0 = 0,                       // 0..6 tiers
    CNT_KIND0 = 8,                       // 8 + kind, kinds 1..10 (kind 0 = sum of tiers)
    CNT_LADDER = 20,
    CNT_BUILT = 21,                      // 21 + set
    CNT_QX = 24,                         // FP32 gate: items re-evaluated in FP64
    CNT_QX_OVER = 25,                    // FP32 gate: re-evaluation queue overflow (the scan is redone in FP64)
    CNT_SLOT_OVER = 26,                  // a gate found a seed without a sampler-table slot: cannot happen in
                                         // whf_scan (a batch holds at most kMaxSeeds seeds); the Deep Dive
                                         // expansion, whose batches hold more, returns WHF_EOVERFLOW
    CNT_DD_DONE = 27,                    // Deep Dive layout walks: complete, cut below the line, capped
    CNT_DD_CUT = 28,
    CNT_DD_CAPPED = 29,
    CNT_DD_DRAWS = 30,
    CNT_DD_PIECES = 31,
    CNT_BIO_BUILT = 32,                  // biome stage: Perlin samplers built
    CNT_BIO_CHECKS = 33,                 // biome stage: (query, quart height) tree searches
    CNT_N = 40
};

struct DevHit {
    int64_t seed;
    int32_t rx, rz, cx, cz, floor_y, flags;
};
static_assert(sizeof(DevHit) == sizeof(whf_hit), "hit layout");

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated append of `pred` lanes to a queue: returns this lane's slot (valid where pred).
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t *counter) {
    unsigned long long mask = __ballot(pred);
    uint32_t base = 0;
    if (mask) {
        int leader = __ffsll((long long)mask) - 1;
        if ((int)lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
    }
    unsigned long long below = mask & ((1ull << lane_id()) - 1ull);
    return base + (uint32_t)__popcll(below);
}

constexpr uint32_t kMaxSeeds = 1u << 17;      // sampler-table slots per set (seeds with tables at once)
constexpr uint32_t kMaxBatchSeeds = 1u << 20; // seeds per batch (the Deep Dive expansion runs 16 spots at once)
constexpr uint32_t kMaxItems = 1u << 25;      // (seed, region) items per batch
constexpr uint32_t kHitCap = 1u << 20;
constexpr uint32_t kQxCap = 1u << 22;         // FP32 gate re-evaluation queue
constexpr int kLayouts = 16;                  // every combination of the four rule flags
constexpr int kBdPerLayout = 8 + 28 + 36;

extern thread_local char g_last_error[512];

#define HIP_TRY(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) {                                                                             \
            std::snprintf(::whf::g_last_error, sizeof ::whf::g_last_error, "%s: %s", #expr,                \
                          hipGetErrorString(e_));                                                           \
            return WHF_EDEVICE;                                                                             \
        }                                                                                                   \
    } while (0)

// Kernel families for the profile breakdown (Seed Check 0..7, Deep Dive 8..10).
enum Fam {
    F_SEEDS = 0, F_BUILD_E, F_EROSION, F_COMPACT, F_BUILD_C, F_GATE2, F_BUILD_L, F_GATE3,
    F_DD_SWEEP, F_DD_FLOOR, F_DD_EXPAND, F_COUNT
};

}  // namespace whf

namespace whf {

// ---- device-resident serving (whf_serve.hip): the request table and the packed batches ----
// A served request lives in one entry of a device table. Every batch of the pipeline is packed on the device: a
// claim kernel takes seeds from the table's requests (least remaining work first) into the batch's seed slots, so
// one batch holds tiles of many requests, rule layouts and rects. The kernels then find each slot's request, rule
// layout and rect through the tables below, count into per-request counters and append survivors to per-request
// queues.
constexpr uint32_t kSrvLive = 1024;                     // request-table entries
constexpr int kSrvClasses = 4;                          // tile classes: FP32 gate or FP64, x 4 or 8 seeds per tile
constexpr uint32_t kSrvTileCap = kMaxSeeds / 4 + kSrvLive;   // tiles per class list (a partial tile per claim)
constexpr uint32_t kSrvMaxRegions = 1u << 16;           // regions of a served request (one erosion tile covers all)
constexpr int kSrvWide = 64;                            // rects of at least this many regions take 4 seeds per tile

// A request as the device reads it (written by the host when the request is admitted to its entry).
struct SReq {
    uint64_t stream, offset;   // WHF_SEEDS_SPLITMIX / _COUNTER
    const uint64_t *list;      // WHF_SEEDS_LIST: the request's seeds, copied into the server's device seed pool
    uint32_t n_seeds, R;       // seeds; regions W x H
    int32_t mode, rx0, rz0, W;
    int32_t max_height, vflags, salt, f32;
    uint32_t hit_base, hit_cap;   // the request's survivor queue in the server's hit pool
    uint32_t seq;                 // admission order (ties of the claim order)
    uint32_t admit_batch;         // the first batch that could claim it (its age counts from there)
    FastDiv Rdiv, Wdiv;           // division by R and by W, prepared on the host
    uint32_t pad[2];
};

// Per-request progress, written by the device: seeds claimed into batches, seeds whose batch completed, survivors.
struct SState {
    uint64_t claimed, done;
    uint32_t hits, pad;
};

// A run of at most 4 or 8 consecutive slots of one request: one erosion workgroup's tile.
struct STile {
    uint32_t slot0;
    uint16_t ns, req;
};

// Seeds [first, first + n) of request `req` went to slots [slot0, slot0 + n) of the batch; their tiles are entries
// [tile0, ...) of the class list cls.
struct SClaim {
    uint32_t req, slot0, n, tile0;
    uint64_t first;
    uint32_t cls, pad;
};

// The device side of one batch of the served pipeline (one per batch set).
struct SBatch {
    uint32_t n_slots, n_claims;
    uint32_t n_tiles[kSrvClasses];
    uint32_t tile_next[kSrvClasses];   // work-stealing cursors of the erosion workgroups over each class's tiles
    uint64_t claimed_items;
    SClaim claims[kSrvLive];
    STile tiles[kSrvClasses][kSrvTileCap];
    uint16_t rq_of[kMaxSeeds];   // request entry of each slot
    uint8_t lay_of[kMaxSeeds];   // rule layout (whf_request.vflags) of each slot
};

// What the served kernels read besides a batch's own buffers (whf_scan.hip's kernels take it by value). Their
// gate-table argument is then the whole per-layout array (whf_ctx.d_gt), indexed by each request's rule layout.
struct SrvView {
    const SReq *req;             // [kSrvLive]
    const uint16_t *rq_of;       // [kMaxSeeds]
    const uint8_t *lay_of;       // [kMaxSeeds]
    unsigned long long *rcnt;    // [kSrvLive][CNT_N]: the Cnt counters per request
    SState *state;               // [kSrvLive]
    DevHit *hit_pool;
};

// Counts 1 into counter idx (a Cnt index; < 0: none) of the lane's request, with one atomic per distinct
// (request, counter) pair of the wave.
__device__ __forceinline__ void wave_count_req(int idx, uint32_t req, unsigned long long *rcnt) {
    unsigned long long todo = __ballot(idx >= 0);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t lr = __shfl(req, leader);
        const int li = __shfl(idx, leader);
        const unsigned long long same = __ballot(idx == li && req == lr) & todo;
        if ((int)lane_id() == leader) atomicAdd(&rcnt[(size_t)lr * CNT_N + li], (unsigned long long)__popcll(same));
        todo &= ~same;
    }
}

// The buffers of one batch that its later stages read: its seeds, erosion survivors (Q1) and their seeds' flags,
// set-C slots and tables, gate-2 survivors (Q2) and their seeds' flags, and its control words.
struct BatchSet {
    uint64_t *seeds;
    uint8_t *tabC;
    SPar *parC;
    uint4 *q1, *q2;
    uint32_t *needC, *slotC, *seedC, *needL, *ctl;
};
}  // namespace whf

struct whf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char name[256] = {0};
    bool profile = false;
    uint64_t *d_seeds = nullptr;     // [kMaxBatchSeeds]
    uint8_t *d_tab[whf::SET_COUNT] = {nullptr, nullptr, nullptr};
    whf::SPar *d_par[whf::SET_COUNT] = {nullptr, nullptr, nullptr};
    uint4 *d_q1 = nullptr, *d_q2 = nullptr, *d_qx = nullptr;
    uint32_t *d_needC = nullptr, *d_needL = nullptr, *d_slotC = nullptr, *d_slotL = nullptr;   // [kMaxBatchSeeds]
    uint32_t *d_seedC = nullptr, *d_seedL = nullptr, *d_ctl = nullptr;
    unsigned long long *d_cnt = nullptr;
    whf::DevHit *d_hits = nullptr;
    whf::GateTables *d_gt = nullptr;      // [kLayouts]
    whf::BuildDesc *d_bd = nullptr;       // [kLayouts][8 + 28 + 36]
    void *d_dd = nullptr;                 // Deep Dive buffers (whf_deepdive.hip), allocated on first use
    void *dd_scratch = nullptr;           // device address of whf_deepdive_noalloc's static scratch on this device
    size_t dd_bytes = 0;
    void *d_bio = nullptr;                // biome stage buffers and tables (whf_biome.hip), allocated on first use
    size_t bio_bytes = 0;
    void *d_ter = nullptr;                // terrain stage buffers and tables (whf_terrain.hip), allocated on first use
    size_t ter_bytes = 0;
    void *d_solve = nullptr;              // Deep Dive solver buffers and tables (whf_solve.hip), allocated on first use
    size_t solve_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Seed Check pipelining: a batch runs in three stages on three streams (the library's stream: seeds, set E and
    // the erosion gate; stage_stream[0]: set C and gate 2; stage_stream[1]: set L and gate 3, both created by the
    // first scan of more than one batch), so three batches are on the device at once. What a batch's later stages read exists three times (batch_set[0] aliases the buffers
    // above, 1 and 2 are their own) and batch k uses set k % 3; ev_stage[s][set] marks the end of stage s of the
    // batch on that set
    hipStream_t stage_stream[2] = {nullptr, nullptr};
    bool one_stream = false;              // WHF_CFG_ONE_STREAM: stages stay on the library's stream
    whf::BatchSet batch_set[3] = {};
    hipEvent_t ev_stage[3][3] = {};
    whf_tuning tune[16] = {};         // launch shapes per rule layout (defaults, then tuned.json, at whf_open)
    std::vector<hipEvent_t> prof_ev;  // pairs per kernel launch when profiling
    std::vector<int> prof_fam;
    int n_cu = 256;
};

namespace whf {

// Brackets the launches of one kernel family with HIP events on the stream they go to (the library's stream
// unless given) when profiling.
struct Prof {
    whf_ctx *c;
    int fam;
    hipStream_t st;
    Prof(whf_ctx *c_, int f, hipStream_t s = nullptr) : c(c_), fam(f), st(s ? s : c_->stream) {
        if (c->profile) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            (void)hipEventRecord(e, st);
            c->prof_ev.push_back(e);
            c->prof_fam.push_back(fam);
        }
    }
    ~Prof() {
        if (c->profile) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            (void)hipEventRecord(e, st);
            c->prof_ev.push_back(e);
        }
    }
};

// One Perlin sampler of a set: noise parameter id, sub-sampler (0/1) and octave slot.
struct SamplerRef {
    int noise, sub, slot;
};
// The MD5-keyed derivation descriptor of a sampler, and the descriptor of a double-Perlin noise over a set
// (its table indices, lacunarities, amplitudes, persistences and output amplitude, noise.c:80-156).
BuildDesc make_build_desc(const SamplerRef &r);
// MD5 of an identifier as two big-endian longs (the createRandom(name) key, rng.c:164-170).
void md5_longs(const char *name, uint64_t &hi, uint64_t &lo);
void fill_noise_desc(NoiseDesc &d, int pid, const std::vector<SamplerRef> &set, int set_id);

int check_launch();
const BuildDesc *bd_of(whf_ctx *c, int layout, int set);
// Perlin sampler tables of n_slots seeds (seeds[slot]) for the set_n samplers of bd, on the context's stream.
void launch_build(whf_ctx *c, const uint64_t *seeds, const BuildDesc *bd, int set_n, uint32_t n_slots,
                  uint8_t *tab, SPar *par, unsigned long long *built);
// Zero the batch's queue counts and the gate flags of its first ns seeds.
void launch_reset(whf_ctx *c, uint32_t ns);
// lysh's phase 2 on phase-1 survivors of a request (whf_scan.hip): keeps the accepted huts in place with their floor;
// adds to p2 (evaluated, accepted, rejected biome, height, flood) and to *ms.
int phase2_filter(whf_ctx *c, const whf_request *req, whf_hit *hits, uint64_t n, uint64_t *kept, uint64_t p2[5],
                  double *ms);
// The biome tree of a request's hut verdict (its biome_tree, or its rule flags' default as lysh picks it).
int phase2_tree(const whf_request *req, int *tree);
// A request's threshold applied to hut verdicts hr[0..n) of its survivors (evaluated with any max_y: the floor test
// is the verdict's last step): keeps the accepted huts in place with their floor and adds to p2 as phase2_filter.
void phase2_apply(const whf_request *req, whf_hit *hits, uint64_t n, const whf_hut_result *hr, uint64_t *kept,
                  uint64_t p2[5]);
// Pre-carver surface heights of n columns (seed, x, z) and the state of each column's top block (0 solid,
// 2 water, 3 lava; 1 for a column of air); whf_terrain.hip.
int terrain_heights(whf_ctx *c, const whf_column *cols, uint64_t n, int vflags, int32_t *height, int32_t *state);
// Gates 2 and 3 of a batch whose erosion survivors are already in Q1 (with needC set): compact, build set C,
// gate 2, compact, build set L, gate 3 -> hits.
int run_tail(whf_ctx *c, int layout, uint32_t ns, int vflags, int max_height);
// Looks up (once per device) the device address of whf_deepdive_noalloc's static scratch array and keeps it in the
// context (whf_deepdive.hip).
int dd_scratch_init(whf_ctx *c);
// Resets a batch set's control words and the seed flags of its first ns seeds, on the library's stream.
void reset_set(whf_ctx *c, const BatchSet &b, uint32_t ns);

// A serving lane's own buffers besides its batch set: set-E tables, the FP32 gate's re-evaluation queue, set-L tables
// and set-L slot maps (a lane runs whole batches on its own stream, beside the other lanes).
struct SrvLane {
    uint8_t *tabE;
    SPar *parE;
    uint4 *qx;
    uint8_t *tabL;
    SPar *parL;
    uint32_t *slotL, *seedL;
};
// The stages of a served batch (whf_scan.hip), in order on the lane's stream: reset, then (after the claim) set E and
// the erosion gate, set C and gate 2, set L and gate 3. The batch's slots, tiles and their requests are in sb and v
// (packed by whf_serve.hip's claim kernel); every launch has a fixed grid.
void serve_reset(const BatchSet &b, hipStream_t st);
void serve_stage_e(whf_ctx *c, const BatchSet &b, const SrvLane &ln, SBatch *sb, const SrvView &v, hipStream_t st,
                   int class_mask);
void serve_stage_c(whf_ctx *c, const BatchSet &b, const SrvView &v, hipStream_t st);
void serve_stage_l(whf_ctx *c, const BatchSet &b, const SrvLane &ln, const SrvView &v, hipStream_t st);

// The three-stage batch pipeline (whf_scan.hip). Batch k uses buffer set k % 3: the caller takes the set with
// acquire(), issues the batch's first stage into it on the library's stream, and hands it to tail(), which issues
// set C + gate 2 and set L + gate 3 on the two stage streams. end() makes the library's stream wait for all of it.
// With one batch, or a context opened with WHF_CFG_ONE_STREAM, every stage goes to the library's stream.
struct Pipeline {
    whf_ctx *c;
    bool piped = false;
    hipStream_t ss[3] = {nullptr, nullptr, nullptr};
    bool used[3] = {false, false, false};
    int set = 0;
    explicit Pipeline(whf_ctx *c_) : c(c_) {}
    int begin(uint64_t n_batches);
    int acquire(const BatchSet *&b);
    int tail(const BatchSet &b, int layout, uint32_t ns, int vflags, int max_height);
    int end();
};

}  // namespace whf
