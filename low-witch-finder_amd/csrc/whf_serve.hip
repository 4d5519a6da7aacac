// whf_serve.hip — the device-resident serving loop behind whf_server_* (C ABI in whf.h): a request table in device
// memory, batches packed on the device from every admitted request, per-request counters and survivor queues on the
// device, and admission against the device pools that an HBM budget sizes.
//
// Replaces the reference's one-request-at-a-time worker loop (engine/gpu/low54/phase1_scan.cu:1365-1387), in which
// each request line runs a whole scan before the next one is read.
//
// Per batch, on the three streams of the context's Seed Check pipeline:
//   stream 1: k_reset, k_srv_claim (one workgroup: sorts the live requests by remaining work and takes seeds from
//             them into the batch's slots, up to the batch's slot and item caps), k_srv_fill (the slots' seeds,
//             requests, rule layouts and erosion tiles), set E and the erosion gate over the tile lists;
//   stream 2: set C and gate 2;
//   stream 3: set L and gate 3 (survivors into their requests' queues), k_srv_account (samplers built per request),
//             k_srv_finish (seeds done per request; requests whose every seed is done are listed for the host).
// A batch thus holds tiles of many requests of every rect, rule layout and precision, and a request admitted while
// others run is claimed by the next batch: its work overlaps theirs on the device.
//
// The host side has two threads. The launcher admits queued requests into free table entries (copying their
// descriptors, and seed lists, to the device), issues batches while the device reports unclaimed work, and reads
// back only the claim kernel's summary. The completer waits for each batch's end, reads the requests it completed
// (counters and survivors), runs lysh's phase 2 on their survivors when asked, fills the sinks, posts the events
// and frees the entries and pool space, which admits the next requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "whf_internal.h"
#include "whf_serve.h"

// Device bytes one context holds (whf_scan.hip).
__attribute__((visibility("hidden"))) uint64_t whf_ctx_device_bytes(const whf_ctx *c);

namespace whf {

namespace {

constexpr uint32_t kSrvRing = 64;               // batch records the host can lag behind the device
constexpr int kSrvMaxLanes = 4;             // lanes: streams that each run whole batches, side by side (the device
                                            // has four hardware queues per process)
constexpr int kSrvMinLanes = 3;
constexpr uint32_t kSrvAge = 8;             // batches after which a live request is claimed first (at most half a
                                            // batch of it), so big requests are not starved by a stream of small ones
constexpr uint32_t kSrvHitMin = 4096;           // survivor entries reserved per request at least
constexpr uint64_t kSrvHitPool = 1ull << 22;    // survivor entries of the pool (128 MiB of host memory)
constexpr uint64_t kSrvSeedPool = 1ull << 26;   // listed seeds of the pool by default (512 MiB)

// The claim kernel's summary of a batch, written to host memory.
struct SClaimOut {
    uint64_t claimed_seeds, claimed_items, unclaimed_checks, n_claims;
};

// A request a batch completed: its entry, survivors and counters, as k_srv_finish copies them to host memory.
struct SDoneRec {
    uint32_t entry, hits;
    unsigned long long cnt[CNT_N];
};

// The requests a batch completed, written to host memory by k_srv_finish.
struct SDone {
    uint64_t seq;
    uint32_t n, pad;
    SDoneRec rec[kSrvLive];
};

// ---------------------------------------------------------------------------------------------- kernels

// One workgroup of kSrvLive lanes, lane t on table entry t. Live requests (seeds left to claim) are sorted by
// remaining checks, least first (ties: the earlier admission), and take seeds in that order: whole requests while
// the batch's slot and item caps hold, then part of the first one that does not fit. A request live for kSrvAge
// batches sorts before all others (oldest first) and takes at most half the batch's items. Each claim's slots follow
// the previous claims', and its tiles are a range of its class list. Claims run one after another (the host chains
// them with events), so `claimed` needs no atomics.
__global__ __launch_bounds__(kSrvLive) void k_srv_claim(const SReq *__restrict__ req, SState *__restrict__ state,
                                                       SBatch *__restrict__ sb, unsigned long long *__restrict__ rcnt,
                                                       uint64_t item_cap, uint32_t batch, SClaimOut *__restrict__ out) {
    __shared__ uint64_t s_key[kSrvLive];
    __shared__ uint16_t s_idx[kSrvLive];
    __shared__ uint64_t s_a[kSrvLive], s_b[kSrvLive];
    __shared__ uint32_t s_tiles[kSrvClasses];
    __shared__ unsigned long long s_sum[3];
    const uint32_t t = threadIdx.x;
    if (t < kSrvClasses) s_tiles[t] = 0;
    if (t < 3) s_sum[t] = 0;
    {
        const SReq &r = req[t];
        const uint64_t claimed = state[t].claimed;
        const uint64_t rem = r.n_seeds > claimed ? r.n_seeds - claimed : 0;
        const bool aged = batch - r.admit_batch >= kSrvAge;
        // remaining checks (< 2^48: seeds < 2^32, regions <= 2^16), then the admission order's low 16 bits; an aged
        // request's key has no remaining-checks part
        s_key[t] = rem ? ((aged ? 0ull : rem * r.R) << 16) | (uint64_t)(r.seq & 0xFFFFu) : ~0ull;
        s_idx[t] = (uint16_t)t;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (uint32_t k = 2; k <= kSrvLive; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t p = t ^ j;
            if (p > t) {
                const bool up = (t & k) == 0;
                const uint64_t a = s_key[t], b = s_key[p];
                if ((a > b) == up) {
                    s_key[t] = b;
                    s_key[p] = a;
                    const uint16_t x = s_idx[t];
                    s_idx[t] = s_idx[p];
                    s_idx[p] = x;
                }
            }
            __syncthreads();
        }
    }
    const bool live = s_key[t] != ~0ull;
    const uint32_t e = s_idx[t];
    const SReq &r = req[e];
    const uint64_t claimed = live ? state[e].claimed : 0;
    const uint64_t R = live ? r.R : 1;
    uint64_t rem = live ? r.n_seeds - claimed : 0;
    const uint64_t left = rem;
    if (live && batch - r.admit_batch >= kSrvAge) rem = min(rem, max((uint64_t)1, (item_cap / 2) / R));
    // inclusive prefix sums of seeds and items in the sorted order
    s_a[t] = rem;
    s_b[t] = rem * R;
    __syncthreads();
    for (uint32_t o = 1; o < kSrvLive; o <<= 1) {
        const uint64_t a = t >= o ? s_a[t - o] : 0, b = t >= o ? s_b[t - o] : 0;
        __syncthreads();
        s_a[t] += a;
        s_b[t] += b;
        __syncthreads();
    }
    const uint64_t a_in = s_a[t], b_in = s_b[t], a_ex = a_in - rem, b_ex = b_in - rem * R;
    const uint64_t S = kMaxSeeds;
    uint64_t take = 0;
    if (rem) {
        if (a_in <= S && b_in <= item_cap) take = rem;
        else if (a_ex < S && b_ex < item_cap) take = min(min(S - a_ex, (item_cap - b_ex) / R), rem);
    }
    // the claimed lanes are a prefix of the sorted order: lane t is claim t
    const uint32_t n_claims = (uint32_t)__syncthreads_count(take > 0);
    if (take) {
        const int cls = (r.f32 ? 2 : 0) + (R >= (uint64_t)kSrvWide ? 0 : 1);
        const uint32_t st = (cls & 1) ? 8u : 4u;
        const uint32_t n_tiles = (uint32_t)((take + st - 1) / st);
        const uint32_t tile0 = atomicAdd(&s_tiles[cls], n_tiles);
        SClaim cl;
        cl.req = e;
        cl.slot0 = (uint32_t)a_ex;
        cl.n = (uint32_t)take;
        cl.tile0 = tile0;
        cl.first = claimed;
        cl.cls = (uint32_t)cls;
        cl.pad = 0;
        sb->claims[t] = cl;
        state[e].claimed = claimed + take;
        atomicAdd(&rcnt[(size_t)e * CNT_N + CNT_BUILT + SET_E], 8ull * take);   // every claimed seed gets its set E
        atomicAdd(&s_sum[0], (unsigned long long)take);
        atomicAdd(&s_sum[1], (unsigned long long)(take * R));
    }
    if (left > take) atomicAdd(&s_sum[2], (unsigned long long)((left - take) * R));
    __syncthreads();
    if (t == 0) {
        sb->n_slots = (uint32_t)s_sum[0];
        sb->n_claims = n_claims;
        sb->claimed_items = s_sum[1];
        for (int k = 0; k < kSrvClasses; k++) {
            sb->n_tiles[k] = s_tiles[k];
            sb->tile_next[k] = 0;
        }
        out->claimed_seeds = s_sum[0];
        out->claimed_items = s_sum[1];
        out->unclaimed_checks = s_sum[2];
        out->n_claims = n_claims;
        __threadfence_system();
    }
}

// One lane per batch slot: its claim (claims are in slot order), then the seed, request, rule layout, and the tile
// that starts at it.
__global__ __launch_bounds__(256) void k_srv_fill(const SReq *__restrict__ req, SBatch *__restrict__ sb,
                                                  uint64_t *__restrict__ seeds) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= sb->n_slots) return;
    uint32_t lo = 0, hi = sb->n_claims - 1;   // the last claim with slot0 <= j
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (sb->claims[mid].slot0 <= j) lo = mid;
        else hi = mid - 1;
    }
    const SClaim cl = sb->claims[lo];
    const SReq &r = req[cl.req];
    const uint32_t i = j - cl.slot0;
    const uint64_t idx = cl.first + i;
    uint64_t seed;
    if (r.mode == WHF_SEEDS_LIST) seed = r.list[idx];
    else if (r.mode == WHF_SEEDS_SPLITMIX) seed = mix64(r.stream + (r.offset + idx + 1) * kGolden);
    else seed = mix64(r.stream + r.offset + idx + kGolden);
    seeds[j] = seed;
    sb->rq_of[j] = (uint16_t)cl.req;
    sb->lay_of[j] = (uint8_t)r.vflags;
    const uint32_t st = (cl.cls & 1) ? 8u : 4u;
    if (i % st == 0) sb->tiles[cl.cls][cl.tile0 + i / st] = STile{j, (uint16_t)min(st, cl.n - i), (uint16_t)cl.req};
}

// Samplers built per request: 28 set-C samplers for each slot with erosion survivors, 36 set-L samplers for each slot
// with gate-2 survivors (the builds of stages 2 and 3 make exactly these).
__global__ __launch_bounds__(256) void k_srv_account(const SBatch *__restrict__ sb, const uint32_t *__restrict__ needC,
                                                     const uint32_t *__restrict__ needL,
                                                     unsigned long long *__restrict__ rcnt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = sb->n_slots;
    const bool in = j < n;
    const uint32_t rq = in ? sb->rq_of[j] : 0;
    const bool c = in && needC[j], l = in && needL[j];
    // one atomic per distinct request of the wave and set, weighted by the set's sampler count
    for (int s = 0; s < 2; s++) {
        const bool f = s == 0 ? c : l;
        unsigned long long todo = __ballot(f);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const uint32_t lr = __shfl(rq, leader);
            const unsigned long long same = __ballot(f && rq == lr) & todo;
            if ((int)lane_id() == leader)
                atomicAdd(&rcnt[(size_t)lr * CNT_N + CNT_BUILT + (s == 0 ? SET_C : SET_L)],
                          (unsigned long long)__popcll(same) * (unsigned long long)kSetSize[s == 0 ? SET_C : SET_L]);
            todo &= ~same;
        }
    }
}

// One lane per claim: the claim's seeds are done; a request whose every seed is done goes to the batch's record in host
// memory with its survivor count and counters (every batch that holds its seeds has run its last kernel by now), so
// the host reads completions without a copy.
__global__ __launch_bounds__(kSrvLive) void k_srv_finish(const SReq *__restrict__ req, SState *__restrict__ state,
                                                        const SBatch *__restrict__ sb,
                                                        unsigned long long *__restrict__ rcnt,
                                                        SDone *__restrict__ out, uint64_t seq) {
    __shared__ uint32_t s_n;
    __shared__ uint16_t s_done[kSrvLive];
    const uint32_t t = threadIdx.x;
    if (t == 0) s_n = 0;
    __syncthreads();
    if (t < sb->n_claims) {
        // lanes finish batches in any order: the batch whose count completes the request reports it, after every
        // other batch holding its seeds has added its own count (and so run its gates)
        const SClaim cl = sb->claims[t];
        const uint64_t done =
            atomicAdd(reinterpret_cast<unsigned long long *>(&state[cl.req].done), (unsigned long long)cl.n) + cl.n;
        if (done == req[cl.req].n_seeds) s_done[atomicAdd(&s_n, 1u)] = (uint16_t)cl.req;
    }
    __syncthreads();
    // the records, one word per lane at a time: host-memory stores are slow one after another, fast side by side
    const uint32_t n_done = s_n;
    constexpr uint32_t kWords = 1 + CNT_N;
    for (uint32_t w = t; w < n_done * kWords; w += kSrvLive) {
        const uint32_t i = w / kWords, k = w - i * kWords;
        const uint32_t e = s_done[i];
        SDoneRec &o = out->rec[i];
        // other lanes' kernels counted into these with device-wide atomics: read them the same way
        if (k == 0) {
            o.entry = e;
            o.hits = __hip_atomic_load(&state[e].hits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            o.cnt[k - 1] = __hip_atomic_load(&rcnt[(size_t)e * CNT_N + (k - 1)], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    if (t == 0) {
        out->seq = seq;
        out->n = s_n;
        __threadfence_system();
    }
}

// ---------------------------------------------------------------------------------------------- host

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) (offsets in entries).
class Pool {
    std::map<uint64_t, uint64_t> free_;   // offset -> length
    uint64_t size_ = 0;

  public:
    void init(uint64_t size) {
        size_ = size;
        free_.clear();
        if (size) free_[0] = size;
    }
    uint64_t size() const { return size_; }
    bool can(uint64_t n) const {
        if (n == 0) return true;
        for (const auto &f : free_)
            if (f.second >= n) return true;
        return false;
    }
    bool alloc(uint64_t n, uint64_t &off) {
        off = 0;
        if (n == 0) return true;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const uint64_t rest = it->second - n, o2 = it->first + n;
            free_.erase(it);
            if (rest) free_[o2] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        if (!n) return;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// The erosion tile class of a request (whf_internal.h kSrvClasses): FP32 gate or not, 4 or 8 seeds per tile.
int tile_class(const whf_request &q, uint32_t R) {
    return (q.precision == WHF_PREC_FP32_GATE ? 2 : 0) + (R >= (uint32_t)kSrvWide ? 0 : 1);
}

struct DqReq {
    uint64_t id = 0;
    whf_request req{};
    whf_sink *sink = nullptr;
    uint32_t R = 0, hit_cap = 0;
    uint64_t seed_words = 0;   // listed seeds held in the device seed pool
    int entry = -1;
    uint64_t hit_off = 0, seed_off = 0;
    double submit_s = 0, start_s = 0;
};

}  // namespace

// A lane: a stream that runs whole batches, with its batch set, its own tables and queue (SrvLane) and its batch
// record on the device.
struct LaneBufs {
    hipStream_t st = nullptr;
    bool own_st = false;
    BatchSet b{};
    bool own_b = false;
    SrvLane ln{};
    bool own_ln = false;
    SBatch *sb = nullptr;
};

struct DqServer {
    int device = 0;
    whf_ctx *c = nullptr;     // the lanes' context: lane 0 uses its buffers and stream, lanes 1-2 its other batch sets
    whf_ctx *p2 = nullptr;    // the completer's context: phase 2, and whf_scan for a request the queue cannot finish
    hipStream_t copy_st = nullptr;
    DqHooks hooks{};
    Clock::time_point epoch;
    uint64_t item_cap = kMaxItems;
    // device
    SReq *d_req = nullptr;
    SState *d_state = nullptr;
    unsigned long long *d_rcnt = nullptr;
    whf_hit *h_hits = nullptr;   // survivor pool in host memory (hipHostMalloc): gate 3 appends to it directly
    uint64_t *d_seeds = nullptr;
    LaneBufs lanes[kSrvMaxLanes];
    int n_lanes = 3;
    // host memory the device writes or the copies read (hipHostMalloc)
    SReq *h_req = nullptr;
    unsigned char *h_zero = nullptr;
    SClaimOut *h_claim = nullptr;   // [kSrvRing]
    SDone *h_done = nullptr;        // [kSrvRing]
    hipEvent_t ev_claim[kSrvRing] = {};
    hipEvent_t ev_done[kSrvRing] = {};
    // host state, under mu
    std::mutex mu;
    std::condition_variable cv_launch, cv_complete;
    std::deque<std::shared_ptr<DqReq>> inbox;
    std::vector<std::shared_ptr<DqReq>> table;
    std::vector<int> free_entries;
    Pool hit_pool, seed_pool;
    uint32_t next_seq = 0;
    bool stop = false, launcher_done = false;
    uint64_t admitted_live = 0;
    bool unclaimed = false;         // the last claim read left work on the device
    uint64_t claims_read = 0;       // batches whose claim summary the launcher has read
    int class_live[kSrvClasses] = {};   // admitted, not completed requests per tile class
    uint64_t issued = 0, completed_batches = 0;
    std::deque<uint64_t> issued_q;
    std::thread launcher, completer;

    double now() const { return std::chrono::duration<double>(Clock::now() - epoch).count(); }

    SrvView view(const LaneBufs &L) const {
        SrvView v;
        v.req = d_req;
        v.rq_of = L.sb->rq_of;
        v.lay_of = L.sb->lay_of;
        v.rcnt = d_rcnt;
        v.state = d_state;
        v.hit_pool = reinterpret_cast<DevHit *>(h_hits);
        return v;
    }

    // mu held: whether the oldest queued request fits an entry and the pools now
    bool fits(const DqReq &r) const {
        if (free_entries.empty()) return false;
        if (r.hit_cap > hit_pool.size() || r.seed_words > seed_pool.size()) return true;   // refused at admission
        return hit_pool.can(r.hit_cap) && seed_pool.can(r.seed_words);
    }

    // Publishes an admitted request on stream st, the stream of the batch that claims next.
    int upload(DqReq &r, hipStream_t st) {
        if (r.seed_words) {
            // a pageable copy on the idle copy stream, finished before the descriptor is published
            HIP_TRY(hipMemcpyAsync(d_seeds + r.seed_off, r.req.seeds, r.seed_words * sizeof(uint64_t),
                                   hipMemcpyHostToDevice, copy_st));
            HIP_TRY(hipStreamSynchronize(copy_st));
        }
        SReq &d = h_req[r.entry];
        std::memset(&d, 0, sizeof d);
        d.stream = r.req.stream;
        d.offset = r.req.offset;
        d.list = r.seed_words ? d_seeds + r.seed_off : nullptr;
        d.n_seeds = (uint32_t)r.req.n_seeds;
        d.R = r.R;
        d.mode = r.req.seed_mode;
        d.rx0 = r.req.rx0;
        d.rz0 = r.req.rz0;
        d.W = r.req.rx1 - r.req.rx0;
        d.max_height = r.req.max_height;
        d.vflags = r.req.vflags;
        d.salt = r.req.salt ? r.req.salt : 14357620;
        d.f32 = r.req.precision == WHF_PREC_FP32_GATE ? 1 : 0;
        d.hit_base = (uint32_t)r.hit_off;
        d.hit_cap = r.hit_cap;
        d.seq = next_seq++;
        d.admit_batch = (uint32_t)issued;
        d.Rdiv = make_fastdiv(d.R);
        d.Wdiv = make_fastdiv((uint32_t)d.W);
        // the state and counters are zeroed before the descriptor makes the entry live; all three copies precede the
        // next claim kernel on the same stream
        HIP_TRY(hipMemcpyAsync(d_state + r.entry, h_zero, sizeof(SState), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_rcnt + (size_t)r.entry * CNT_N, h_zero, CNT_N * sizeof(unsigned long long),
                               hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_req + r.entry, &d, sizeof(SReq), hipMemcpyHostToDevice, st));
        return WHF_OK;
    }

    // Issues batch `issued` on lane issued % n_lanes: reset, claim (after the previous batch's claim, on whatever lane
    // it ran), fill, and all three stages; then waits for the claim's summary.
    int issue_batch(bool admitted_before) {
        const uint64_t seq = issued;
        LaneBufs &L = lanes[seq % n_lanes];
        const SrvView v = view(L);
        const hipStream_t st = L.st;
        serve_reset(L.b, st);
        if (seq > 0) HIP_TRY(hipStreamWaitEvent(st, ev_claim[(seq - 1) % kSrvRing], 0));
        hipLaunchKernelGGL(k_srv_claim, dim3(1), dim3(kSrvLive), 0, st, d_req, d_state, L.sb, d_rcnt, item_cap,
                           (uint32_t)seq, &h_claim[seq % kSrvRing]);
        HIP_TRY(hipEventRecord(ev_claim[seq % kSrvRing], st));
        hipLaunchKernelGGL(k_srv_fill, dim3(kMaxSeeds / 256), dim3(256), 0, st, d_req, L.sb, L.b.seeds);
        int class_mask = 0;
        {
            // a class with no live request has no tiles in this batch
            std::lock_guard<std::mutex> lk(mu);
            for (int k = 0; k < kSrvClasses; k++)
                if (class_live[k]) class_mask |= 1 << k;
        }
        serve_stage_e(c, L.b, L.ln, L.sb, v, st, class_mask);
        serve_stage_c(c, L.b, v, st);
        serve_stage_l(c, L.b, L.ln, v, st);
        hipLaunchKernelGGL(k_srv_account, dim3(kMaxSeeds / 256), dim3(256), 0, st, L.sb, L.b.needC, L.b.needL, d_rcnt);
        hipLaunchKernelGGL(k_srv_finish, dim3(1), dim3(kSrvLive), 0, st, d_req, d_state, L.sb, d_rcnt,
                           &h_done[seq % kSrvRing], seq);
        if (int rc = check_launch()) return rc;
        HIP_TRY(hipEventRecord(ev_done[seq % kSrvRing], st));
        {
            std::lock_guard<std::mutex> lk(mu);
            issued = seq + 1;
            issued_q.push_back(seq);
        }
        cv_complete.notify_all();
        // the launcher runs one batch ahead of the claims it has read, so that a lane whose batch ends finds the next
        // one queued: it reads the previous batch's claim now, and this one's before it stops issuing
        return seq > 0 ? read_claim(seq - 1, admitted_before) : WHF_OK;
    }

    // Reads the summary of batch seq's claim (waiting for it); requests admitted after that claim (fresh) leave work
    // it did not see.
    int read_claim(uint64_t seq, bool fresh) {
        HIP_TRY(hipEventSynchronize(ev_claim[seq % kSrvRing]));
        std::lock_guard<std::mutex> lk(mu);
        unclaimed = h_claim[seq % kSrvRing].unclaimed_checks > 0 || fresh;
        claims_read = seq + 1;
        return WHF_OK;
    }

    // Fails every request still queued or admitted (a device error stops the loop).
    void fail_all(int rc) {
        std::vector<std::shared_ptr<DqReq>> dead;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &r : inbox) dead.push_back(r);
            inbox.clear();
            for (auto &r : table)
                if (r) {
                    dead.push_back(r);
                    r.reset();
                }
            admitted_live = 0;
            for (int &k : class_live) k = 0;
        }
        for (auto &r : dead) post(*r, rc, whf_stats{}, {}, 0);
    }

    // Completes a request: the sink gets the hits (up to its capacity) and the total count, as whf_scan fills it.
    void post(const DqReq &r, int status, const whf_stats &st, const std::vector<whf_hit> &hits, uint64_t total) {
        whf_event e{};
        e.id = r.id;
        e.status = status;
        e.stats = st;
        e.submit_s = r.submit_s;
        e.start_s = std::max(r.start_s, r.submit_s);
        e.done_s = std::max(now(), e.start_s);
        if (r.sink) {
            r.sink->count = total;
            const uint64_t m = std::min<uint64_t>(r.sink->capacity, hits.size());
            if (m && r.sink->hits) std::memcpy(r.sink->hits, hits.data(), m * sizeof(whf_hit));
            if (total > r.sink->capacity && e.status == WHF_OK) e.status = WHF_EOVERFLOW;
        }
        hooks.post(hooks.ctx, e);
    }

    void launch_loop() {
        (void)hipSetDevice(device);
        for (;;) {
            std::vector<std::shared_ptr<DqReq>> admit, refused;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_launch.wait(lk, [&] {
                    const bool room = issued - completed_batches < kSrvRing - 4;
                    return (stop && inbox.empty() && admitted_live == 0) || claims_read < issued ||
                           (room && ((!inbox.empty() && fits(*inbox.front())) || unclaimed));
                });
                if (stop && inbox.empty() && admitted_live == 0) break;
                while (!inbox.empty() && fits(*inbox.front())) {
                    std::shared_ptr<DqReq> r = inbox.front();
                    inbox.pop_front();
                    if (r->hit_cap > hit_pool.size() || r->seed_words > seed_pool.size()) {
                        refused.push_back(r);
                        continue;
                    }
                    hit_pool.alloc(r->hit_cap, r->hit_off);
                    seed_pool.alloc(r->seed_words, r->seed_off);
                    r->entry = free_entries.back();
                    free_entries.pop_back();
                    table[r->entry] = r;
                    r->start_s = now();
                    admitted_live++;
                    class_live[tile_class(r->req, r->R)]++;
                    admit.push_back(r);
                }
            }
            for (auto &r : refused) post(*r, WHF_ENOMEM, whf_stats{}, {}, 0);
            int rc = WHF_OK;
            if (!admit.empty() && issued > 0) {
                // an entry is rewritten by three copies (state, counters, descriptor); no claim may run between them,
                // so they follow the previous batch's claim (on whatever lane it ran) and precede the next one
                const hipError_t e = hipStreamWaitEvent(lanes[issued % n_lanes].st, ev_claim[(issued - 1) % kSrvRing], 0);
                if (e != hipSuccess) rc = WHF_EDEVICE;
            }
            for (auto &r : admit)
                if (rc == WHF_OK && (rc = upload(*r, lanes[issued % n_lanes].st)) != WHF_OK) break;
            if (rc == WHF_OK && (!admit.empty() || unclaimed)) {
                rc = issue_batch(!admit.empty());
            } else if (rc == WHF_OK && claims_read < issued) {
                // about to wait: the newest claim decides whether work is left
                rc = read_claim(issued - 1, false);
            }
            if (rc != WHF_OK) {
                fail_all(rc);
                break;
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        launcher_done = true;
        cv_complete.notify_all();
    }

    int ensure_p2() {
        if (p2) return WHF_OK;
        whf_config cc{device, WHF_CFG_ONE_STREAM, 0};
        return whf_open(&cc, &p2);
    }

    // Completes the requests of one batch's record: their counters and survivors are in host memory already (the
    // record and the survivor pool); lysh's phase 2 runs over all their survivors at once per biome tree.
    void finish_batch(const SDone &d) {
        const uint32_t n_done = std::min<uint32_t>(d.n, kSrvLive);
        if (n_done == 0) return;
        struct Fin {
            std::shared_ptr<DqReq> r;
            int e;
            int status = WHF_OK;
            whf_stats st{};
            std::vector<whf_hit> hits;
            uint64_t total = 0;
            bool fallback = false;
        };
        std::vector<Fin> fin;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (uint32_t i = 0; i < n_done; i++) {
                const uint32_t e = d.rec[i].entry;
                if (e >= kSrvLive || !table[e]) continue;
                Fin f;
                f.r = table[e];
                f.e = (int)i;   // the record's index until the stats are read
                fin.push_back(std::move(f));
            }
        }
        for (Fin &f : fin) {
            const SDoneRec &rec = d.rec[f.e];
            f.e = (int)rec.entry;
            const DqReq &r = *f.r;
            const unsigned long long *cnt = rec.cnt;
            f.total = rec.hits;
            const uint64_t n = std::min<uint64_t>(f.total, r.hit_cap);
            f.hits.assign(h_hits + r.hit_off, h_hits + r.hit_off + n);
            whf_stats &st = f.st;
            st.seeds = r.req.n_seeds;
            st.checks = r.req.n_seeds * (uint64_t)r.R;
            uint64_t ero = 0;
            for (int t = 0; t < 7; t++) {
                st.tiers[t] = cnt[CNT_TIER0 + t];
                ero += cnt[CNT_TIER0 + t];
            }
            st.funnel[0] = ero;
            for (int k = 1; k < WHF_KIND_N; k++) st.funnel[k] = cnt[CNT_KIND0 + k];
            st.accepted = cnt[CNT_KIND0 + 10];
            st.reached_ladder = cnt[CNT_LADDER];
            for (int k = 0; k < SET_COUNT; k++) st.samplers_built[k] = cnt[CNT_BUILT + k];
            st.fp64_rechecks = cnt[CNT_QX];
            // The queue cannot give whf_scan's answer when the FP32 gate's re-evaluation queue overflowed in a batch,
            // or when the request's survivor queue overflowed and its survivors are all needed (phase 2) or more of
            // them fit its sink than the queue kept. Then the request is scanned alone, as whf_scan would.
            const uint64_t want = std::min<uint64_t>(r.sink ? r.sink->capacity : 0, kHitCap);
            f.fallback = cnt[CNT_QX_OVER] != 0 ||
                         (f.total > r.hit_cap && (r.req.phase2 != WHF_P2_NONE || r.hit_cap < want));
        }
        // phase 2: one hut evaluation per (biome tree, LARGE_BIOMES) over the survivors of every request that asks
        // for it, then each request's own threshold
        std::map<std::pair<int, int>, std::vector<Fin *>> groups;
        for (Fin &f : fin) {
            if (f.status != WHF_OK || f.fallback || f.r->req.phase2 == WHF_P2_NONE) continue;
            int tree;
            if (phase2_tree(&f.r->req, &tree) != WHF_OK) {
                f.status = WHF_EINVAL;
                continue;
            }
            groups[{tree, f.r->req.vflags & WHF_V_LARGE}].push_back(&f);
        }
        for (auto &g : groups) {
            const auto t0 = Clock::now();
            std::vector<whf_hut_query> hq;
            for (Fin *f : g.second)
                for (const whf_hit &h : f->hits) hq.push_back(whf_hut_query{h.seed, h.cx, h.cz});
            std::vector<whf_hut_result> hr(hq.size());
            int prc = hq.empty() ? WHF_OK : ensure_p2();
            if (prc == WHF_OK && !hq.empty())
                prc = whf_hut_eval(p2, hq.data(), hq.size(), g.first.first, g.first.second, INT32_MAX, hr.data());
            const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
            size_t o = 0;
            for (Fin *f : g.second) {
                if (prc != WHF_OK) {
                    f->status = prc;
                    continue;
                }
                uint64_t kept = 0, p2c[5] = {0, 0, 0, 0, 0};
                phase2_apply(&f->r->req, f->hits.data(), f->hits.size(), hr.data() + o, &kept, p2c);
                o += f->hits.size();
                f->hits.resize(kept);
                f->total = kept;
                for (int k = 0; k < 5; k++) f->st.phase2[k] = p2c[k];
                f->st.phase2_ms = ms;   // the shared evaluation's wall time
            }
        }
        for (Fin &f : fin) {
            if (f.status == WHF_OK && f.fallback) {
                f.status = ensure_p2();
                if (f.status == WHF_OK) {
                    const uint64_t cap = f.r->sink ? f.r->sink->capacity : 0;
                    std::vector<whf_hit> buf(std::max<uint64_t>(1, cap));
                    whf_sink sink{buf.data(), cap, 0};
                    f.status = whf_scan(p2, &f.r->req, &sink, &f.st);
                    buf.resize(std::min<uint64_t>(sink.count, cap));
                    f.hits = std::move(buf);
                    f.total = sink.count;
                    if (f.status == WHF_EOVERFLOW) f.status = WHF_OK;   // post() sets it again from the sink
                }
            }
            std::sort(f.hits.begin(), f.hits.end(), [](const whf_hit &x, const whf_hit &y) {
                if (x.seed != y.seed) return x.seed < y.seed;
                if (x.rx != y.rx) return x.rx < y.rx;
                return x.rz < y.rz;
            });
            if (f.status != WHF_OK) {
                f.hits.clear();
                f.total = 0;
            }
            // the sink's count is the total, which may exceed the hits the queue kept
            post(*f.r, f.status, f.st, f.hits, f.total);
            std::lock_guard<std::mutex> lk(mu);
            hit_pool.release(f.r->hit_off, f.r->hit_cap);
            seed_pool.release(f.r->seed_off, f.r->seed_words);
            table[f.e].reset();
            free_entries.push_back(f.e);
            admitted_live--;
            class_live[tile_class(f.r->req, f.r->R)]--;
            cv_launch.notify_all();
        }
    }

    void complete_loop() {
        (void)hipSetDevice(device);
        for (;;) {
            uint64_t seq;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_complete.wait(lk, [&] { return !issued_q.empty() || launcher_done; });
                if (issued_q.empty()) break;   // launcher_done
                seq = issued_q.front();
                issued_q.pop_front();
            }
            if (hipEventSynchronize(ev_done[seq % kSrvRing]) != hipSuccess) {
                fail_all(WHF_EDEVICE);
                std::lock_guard<std::mutex> lk(mu);
                completed_batches = seq + 1;
                cv_launch.notify_all();
                continue;
            }
            finish_batch(h_done[seq % kSrvRing]);
            std::lock_guard<std::mutex> lk(mu);
            completed_batches = seq + 1;
            cv_launch.notify_all();
        }
    }

    // A lane's buffers: lane 0 uses the context's primary buffers and stream, lanes 1 and 2 its second and third
    // batch sets and the stage streams, and every lane beyond 0 gets its own set-E and set-L tables, re-evaluation
    // queue and set-L slots; lane 3 also its own batch set and stream.
    int make_lane(int k) {
        LaneBufs &L = lanes[k];
        const size_t S = kMaxSeeds;
        HIP_TRY(hipMalloc(&L.sb, sizeof(SBatch)));
        HIP_TRY(hipMemset(L.sb, 0, sizeof(SBatch)));   // rq_of / lay_of of unused slots stay in range
        if (k == 0) {
            L.st = c->stream;
            L.b = c->batch_set[0];
            L.ln = SrvLane{c->d_tab[SET_E], c->d_par[SET_E], c->d_qx, c->d_tab[SET_L], c->d_par[SET_L], c->d_slotL,
                           c->d_seedL};
            return WHF_OK;
        }
        if (k < 3) {
            hipStream_t &s = c->stage_stream[k - 1];
            if (!s) HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            L.st = s;
            L.b = c->batch_set[k];
        } else {
            HIP_TRY(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
            L.own_st = true;
            BatchSet &b = L.b;
            L.own_b = true;
            HIP_TRY(hipMalloc(&b.seeds, S * sizeof(uint64_t)));
            HIP_TRY(hipMalloc(&b.tabC, S * kSetSize[SET_C] * 256));
            HIP_TRY(hipMalloc(&b.parC, S * kSetSize[SET_C] * sizeof(SPar)));
            HIP_TRY(hipMalloc(&b.q1, (size_t)kMaxItems * sizeof(uint4)));
            HIP_TRY(hipMalloc(&b.q2, (size_t)kMaxItems * sizeof(uint4)));
            HIP_TRY(hipMalloc(&b.needC, S * 4));
            HIP_TRY(hipMalloc(&b.slotC, S * 4));
            HIP_TRY(hipMalloc(&b.seedC, S * 4));
            HIP_TRY(hipMalloc(&b.needL, S * 4));
            HIP_TRY(hipMalloc(&b.ctl, 64));
        }
        SrvLane &ln = L.ln;
        L.own_ln = true;
        HIP_TRY(hipMalloc(&ln.tabE, S * kSetSize[SET_E] * 256));
        HIP_TRY(hipMalloc(&ln.parE, S * kSetSize[SET_E] * sizeof(SPar)));
        HIP_TRY(hipMalloc(&ln.qx, (size_t)kQxCap * sizeof(uint4)));
        HIP_TRY(hipMalloc(&ln.tabL, S * kSetSize[SET_L] * 256));
        HIP_TRY(hipMalloc(&ln.parL, S * kSetSize[SET_L] * sizeof(SPar)));
        HIP_TRY(hipMalloc(&ln.slotL, S * 4));
        HIP_TRY(hipMalloc(&ln.seedL, S * 4));
        return WHF_OK;
    }

    int open(int dev, int lanes_n, uint64_t cap, uint64_t hbm_budget) {
        device = dev;
        // 3 or 4 lanes: fewer are not run (a saturated 2-lane loop once stopped making progress, cause not found)
        n_lanes = lanes_n > 0 ? std::min(std::max(lanes_n, kSrvMinLanes), kSrvMaxLanes) : kSrvMinLanes;
        // at least one seed of the widest request fits a batch
        item_cap = cap ? std::min<uint64_t>(std::max<uint64_t>(cap, kSrvMaxRegions), kMaxItems) : kMaxItems;
        whf_config cc{device, 0, 0};
        if (int rc = whf_open(&cc, &c)) return rc;
        HIP_TRY(hipSetDevice(device));
        // the seed-list pool (device memory): the default, cut to what the budget leaves after the contexts (the
        // lanes' and the completer's), the lanes' own buffers and the request table; the survivor pool is host memory
        const uint64_t hit_n = kSrvHitPool;
        uint64_t seed_n = kSrvSeedPool;
        const uint64_t S = kMaxSeeds;
        const uint64_t set_bytes = S * (8 + kSetSize[SET_C] * (256 + sizeof(SPar)) + 4 * 4) +
                                   2ull * kMaxItems * sizeof(uint4) + 64;
        const uint64_t lane_bytes = S * ((kSetSize[SET_E] + kSetSize[SET_L]) * (256 + sizeof(SPar)) + 8) +
                                    (uint64_t)kQxCap * sizeof(uint4);
        const uint64_t fixed = 2 * whf_ctx_device_bytes(c) + n_lanes * sizeof(SBatch) +
                               (n_lanes - 1) * lane_bytes + (n_lanes > 3 ? set_bytes : 0) +
                               kSrvLive * (sizeof(SReq) + sizeof(SState) + CNT_N * sizeof(unsigned long long));
        if (hbm_budget) {
            const uint64_t min_pool = kSrvLive * sizeof(uint64_t);
            if (hbm_budget < fixed + min_pool) {
                std::snprintf(g_last_error, sizeof g_last_error, "hbm_budget %llu below the server's %llu bytes",
                              (unsigned long long)hbm_budget, (unsigned long long)(fixed + min_pool));
                return WHF_ENOMEM;
            }
            seed_n = std::min<uint64_t>(seed_n, (hbm_budget - fixed) / sizeof(uint64_t));
        }
        HIP_TRY(hipStreamCreateWithFlags(&copy_st, hipStreamNonBlocking));
        HIP_TRY(hipMalloc(&d_req, kSrvLive * sizeof(SReq)));
        HIP_TRY(hipMalloc(&d_state, kSrvLive * sizeof(SState)));
        HIP_TRY(hipMalloc(&d_rcnt, (size_t)kSrvLive * CNT_N * sizeof(unsigned long long)));
        HIP_TRY(hipHostMalloc((void **)&h_hits, hit_n * sizeof(whf_hit), hipHostMallocCoherent));
        HIP_TRY(hipMalloc(&d_seeds, seed_n * sizeof(uint64_t)));
        HIP_TRY(hipMemset(d_req, 0, kSrvLive * sizeof(SReq)));
        HIP_TRY(hipMemset(d_state, 0, kSrvLive * sizeof(SState)));
        HIP_TRY(hipMemset(d_rcnt, 0, (size_t)kSrvLive * CNT_N * sizeof(unsigned long long)));
        for (int k = 0; k < n_lanes; k++)
            if (int rc = make_lane(k)) return rc;
        HIP_TRY(hipHostMalloc((void **)&h_req, kSrvLive * sizeof(SReq), hipHostMallocCoherent));
        HIP_TRY(hipHostMalloc((void **)&h_zero, 4096, hipHostMallocCoherent));
        HIP_TRY(hipHostMalloc((void **)&h_claim, kSrvRing * sizeof(SClaimOut), hipHostMallocCoherent));
        HIP_TRY(hipHostMalloc((void **)&h_done, kSrvRing * sizeof(SDone), hipHostMallocCoherent));
        std::memset(h_zero, 0, 4096);
        std::memset(h_claim, 0, kSrvRing * sizeof(SClaimOut));
        std::memset(h_done, 0, kSrvRing * sizeof(SDone));
        for (hipEvent_t &e : ev_claim) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (hipEvent_t &e : ev_done) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIP_TRY(hipDeviceSynchronize());
        // the completer's context, with its terrain and biome tables set up by one hut verdict, so that no request
        // waits for them
        if (int rc = ensure_p2()) return rc;
        {
            const whf_hut_query q{0, 0, 0};
            whf_hut_result hr;
            if (int rc = whf_hut_eval(p2, &q, 1, 0, 0, INT32_MAX, &hr)) return rc;
        }
        hit_pool.init(hit_n);
        seed_pool.init(seed_n);
        table.assign(kSrvLive, nullptr);
        for (int e = (int)kSrvLive - 1; e >= 0; e--) free_entries.push_back(e);
        launcher = std::thread(&DqServer::launch_loop, this);
        completer = std::thread(&DqServer::complete_loop, this);
        return WHF_OK;
    }

    void release() {
        (void)hipSetDevice(device);
        for (hipEvent_t e : ev_claim)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : ev_done)
            if (e) (void)hipEventDestroy(e);
        if (copy_st) (void)hipStreamDestroy(copy_st);
        for (LaneBufs &L : lanes) {
            if (L.sb) (void)hipFree(L.sb);
            if (L.own_b)
                for (void *p : {(void *)L.b.seeds, (void *)L.b.tabC, (void *)L.b.parC, (void *)L.b.q1, (void *)L.b.q2,
                                (void *)L.b.needC, (void *)L.b.slotC, (void *)L.b.seedC, (void *)L.b.needL,
                                (void *)L.b.ctl})
                    if (p) (void)hipFree(p);
            if (L.own_ln)
                for (void *p : {(void *)L.ln.tabE, (void *)L.ln.parE, (void *)L.ln.qx, (void *)L.ln.tabL,
                                (void *)L.ln.parL, (void *)L.ln.slotL, (void *)L.ln.seedL})
                    if (p) (void)hipFree(p);
            if (L.own_st && L.st) (void)hipStreamDestroy(L.st);
        }
        (void)hipFree(d_req);
        (void)hipFree(d_state);
        (void)hipFree(d_rcnt);
        (void)hipFree(d_seeds);
        if (h_hits) (void)hipHostFree(h_hits);
        for (void *p : {(void *)h_req, (void *)h_zero, (void *)h_claim, (void *)h_done})
            if (p) (void)hipHostFree(p);
        if (p2) whf_close(p2);
        if (c) whf_close(c);
    }
};

int dq_open(int device, int lanes, uint64_t item_cap, uint64_t hbm_budget, const DqHooks &hooks,
            std::chrono::steady_clock::time_point epoch, DqServer **out) {
    *out = nullptr;
    DqServer *s = new (std::nothrow) DqServer();
    if (!s) return WHF_ENOMEM;
    s->hooks = hooks;
    s->epoch = epoch;
    const int rc = s->open(device, lanes, item_cap, hbm_budget);
    if (rc != WHF_OK) {
        s->release();
        delete s;
        return rc;
    }
    *out = s;
    return WHF_OK;
}

int dq_check(const whf_request *req) {
    const uint64_t R = (uint64_t)(req->rx1 - req->rx0) * (uint64_t)(req->rz1 - req->rz0);
    if (R == 0 || R > kSrvMaxRegions) return WHF_EINVAL;
    if (req->n_seeds >= (1ull << 32)) return WHF_EINVAL;
    if (req->seed_mode == WHF_SEEDS_LIST && req->n_seeds > kSrvSeedPool) return WHF_EINVAL;
    if (req->seed_mode < WHF_SEEDS_LIST || req->seed_mode > WHF_SEEDS_COUNTER) return WHF_EINVAL;
    if (req->precision != WHF_PREC_FP64 && req->precision != WHF_PREC_FP32_GATE) return WHF_EINVAL;
    if (req->vflags < 0 || req->vflags >= kLayouts) return WHF_EINVAL;
    if (req->phase2 < WHF_P2_NONE || req->phase2 > WHF_P2_EXACTLY) return WHF_EINVAL;
    if (req->biome_tree < 0 || req->biome_tree > 8) return WHF_EINVAL;
    if (req->max_height < -140 || req->max_height > 100000) return WHF_EINVAL;
    const int64_t lim = 65536;
    if (req->rx0 < -lim || req->rx1 > lim || req->rz0 < -lim || req->rz1 > lim) return WHF_EINVAL;
    return WHF_OK;
}

void dq_submit(DqServer *s, uint64_t id, const whf_request *req, whf_sink *sink, double submit_s) {
    auto r = std::make_shared<DqReq>();
    r->id = id;
    r->req = *req;
    r->sink = sink;
    r->R = (uint32_t)((uint64_t)(req->rx1 - req->rx0) * (uint64_t)(req->rz1 - req->rz0));
    const uint64_t checks = req->n_seeds * (uint64_t)r->R;
    const uint64_t want = req->phase2 != WHF_P2_NONE ? 0 : (sink ? sink->capacity : 0);
    r->hit_cap = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(checks, kHitCap), std::max<uint64_t>(kSrvHitMin, want));
    r->seed_words = req->seed_mode == WHF_SEEDS_LIST ? req->n_seeds : 0;
    r->submit_s = submit_s;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->inbox.push_back(r);
    }
    s->cv_launch.notify_all();
}

bool dq_busy(DqServer *s) {
    std::lock_guard<std::mutex> lk(s->mu);
    return !s->inbox.empty() || s->unclaimed;
}

void dq_close(DqServer *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_launch.notify_all();
    if (s->launcher.joinable()) s->launcher.join();
    if (s->completer.joinable()) s->completer.join();
    s->release();
    delete s;
}

}  // namespace whf
