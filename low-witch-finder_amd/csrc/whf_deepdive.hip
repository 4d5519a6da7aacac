// whf_deepdive.hip — the MI355X Deep Dive: mineshaft rooms under swamp-hut chunks, over 48-bit structure seeds.
//
// Per sweep batch of structure seeds S (one HIP stream):
//   k_dd_sweep     one lane per S: Random(S) -> a, b; per hut region the hut chunk (structure.c:4-24), the chunk's
//                  carver seed X = (cx*a ^ cz*b ^ S) mod 2^48 and the start test nextDouble() < 0.004
//                  (room_core.h:55-72, mstree.cpp:10-15); starts -> queue.
//   k_dd_floor     a group of G lanes per start walks the mineshaft layout (whf_mineshaft.h) with its pieces in
//                  LDS; a room floor <= floor_max makes the start a spot. A piece below -12 - floor_max ends
//                  the walk early: the room floor is r - minY - 12 with r >= 0 the final nextInt of the move
//                  below sea level, so such a start can no longer qualify.
// Expansion, 16 spots (2^20 world seeds u << 48 | S) per batch:
//   k_dd_xseeds    the batch's world seeds.
//   k_dd_erosion   the tiered erosion gate (noise.c:432-459) at the hut, one lane per world seed at a time,
//                  building each erosion octave sampler in LDS only when the seed is still alive at that
//                  evaluation (the reference's staged kernels stageA0a..A2b, hut_lookup_v4_source.cu:523-619,
//                  also build octave by octave).
//   then gates 2 and 3 of the Seed Check (whf_scan.hip) on the erosion survivors -> finds.
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "whf_internal.h"
#include "whf_mineshaft.h"

using namespace whf;

namespace {

struct DdStart {
    uint64_t s48, carver;
    int32_t cx, cz;
};

constexpr uint32_t kDdStartCap = 1u << 21;      // starts per sweep batch
constexpr uint32_t kDdSpotCap = 1u << 20;       // spots per sweep batch
constexpr uint64_t kDdSweepMax = 1ull << 24;    // structure seeds per sweep batch
constexpr uint32_t kDdSpotsPerExpand = kMaxBatchSeeds >> 16;
constexpr int kDdGroup = 32;                    // lanes per layout walk (measured: 16 -> 1.35x slower, 64 -> 1.2x)

// ---------------------------------------------------------------------------------------------- sweep

__global__ __launch_bounds__(256) void k_dd_sweep(uint64_t s_first, uint32_t n, int rx0, int nrx, int rz0, int nrz,
                                                   int salt, DdStart *__restrict__ starts, uint32_t cap,
                                                   uint32_t *__restrict__ ctl) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;
    const uint64_t s48 = (s_first + i) & kLcgM48;
    uint64_t st = (s48 ^ kLcgMul) & kLcgM48;
    const uint64_t a = lcg_next_long(st);
    const uint64_t b = lcg_next_long(st);
    for (int rz = 0; rz < nrz; rz++) {
        for (int rx = 0; rx < nrx; rx++) {
            bool start = false;
            int cx = 0, cz = 0;
            uint64_t X = 0;
            if (live) {
                hut_chunk(s48, rx0 + rx, rz0 + rz, salt, cx, cz);
                X = carver_seed(s48, a, b, cx, cz);
                start = ms_is_start(X);
            }
            const uint32_t pos = wave_append(start, &ctl[CTL_DD_STARTS]);
            if (start && pos < cap) starts[pos] = DdStart{s48, X, cx, cz};
        }
    }
}

// ---------------------------------------------------------------------------------------------- layout walks

template <int G>
__global__ __launch_bounds__(64) void k_dd_floor(const DdStart *__restrict__ starts, const uint32_t *__restrict__ ctl_in,
                                                  uint32_t cap, int cut_rel, int floor_max,
                                                  whf_dd_spot *__restrict__ spots, uint32_t spot_cap,
                                                  uint32_t *__restrict__ ctl, unsigned long long *__restrict__ cnt) {
    constexpr int NG = 64 / G;
    __shared__ uint2 s_box[NG][kMsMaxPieces];
    __shared__ uint32_t s_frm[NG][kMsMaxFrames];
    const int lane = (int)threadIdx.x, grp = lane / G, lig = lane % G;
    const uint32_t n = min(ctl_in[CTL_DD_STARTS], cap);
    unsigned long long c[5] = {0, 0, 0, 0, 0};   // done, cut, capped, draws, pieces
    for (uint32_t q = blockIdx.x * NG + (uint32_t)grp; q < n; q += gridDim.x * NG) {
        const DdStart s = starts[q];
        const MsOut o = ms_room_floor<G>(s.carver, cut_rel, s_box[grp], s_frm[grp], lig, grp * G);
        if (lig == 0) {
            c[o.status]++;
            c[3] += o.draws;
            c[4] += (unsigned long long)o.pieces;
            if (o.status == MS_DONE && o.floor <= floor_max) {
                const uint32_t pos = atomicAdd(&ctl[CTL_DD_SPOTS], 1u);
                if (pos < spot_cap) {
                    whf_dd_spot sp;
                    sp.s48 = s.s48;
                    sp.rx = s.cx >> 5;
                    sp.rz = s.cz >> 5;
                    sp.cx = s.cx;
                    sp.cz = s.cz;
                    sp.carver = s.carver;
                    sp.room_floor = o.floor;
                    sp.pieces = o.pieces;
                    spots[pos] = sp;
                }
            }
        }
    }
    for (int k = 0; k < 5; k++) {
        unsigned long long v = c[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
        if (lane == 0 && v) atomicAdd(&cnt[CNT_DD_DONE + k], v);
    }
}

// Full walks of given carver seeds (no cut): status, room floor, pieces.
template <int G>
__global__ __launch_bounds__(64) void k_dd_walks(const uint64_t *__restrict__ carvers, uint32_t n,
                                                  int4 *__restrict__ out) {
    constexpr int NG = 64 / G;
    __shared__ uint2 s_box[NG][kMsMaxPieces];
    __shared__ uint32_t s_frm[NG][kMsMaxFrames];
    const int lane = (int)threadIdx.x, grp = lane / G, lig = lane % G;
    for (uint32_t q = blockIdx.x * NG + (uint32_t)grp; q < n; q += gridDim.x * NG) {
        const MsOut o = ms_room_floor<G>(carvers[q], INT_MIN, s_box[grp], s_frm[grp], lig, grp * G);
        if (lig == 0) out[q] = make_int4(o.status, o.floor, o.pieces, (int)o.draws);
    }
}

// Spots named by the caller (structure seed and chunk, the reference's --spots lines): per spot, a group of G
// lanes computes the chunk's carver seed and the start test and, for a start, walks the whole layout; fills the
// spot's region, carver seed, room floor (INT32_MIN unless a start whose walk completed) and piece count in place.
// start[q] = 1 for a start.
template <int G>
__global__ __launch_bounds__(64) void k_dd_spot_walks(whf_dd_spot *__restrict__ spots, uint32_t n,
                                                       uint32_t *__restrict__ start,
                                                       unsigned long long *__restrict__ cnt) {
    constexpr int NG = 64 / G;
    __shared__ uint2 s_box[NG][kMsMaxPieces];
    __shared__ uint32_t s_frm[NG][kMsMaxFrames];
    const int lane = (int)threadIdx.x, grp = lane / G, lig = lane % G;
    unsigned long long c[5] = {0, 0, 0, 0, 0};   // done, cut, capped, draws, pieces
    for (uint32_t q = blockIdx.x * NG + (uint32_t)grp; q < n; q += gridDim.x * NG) {
        whf_dd_spot sp = spots[q];
        const uint64_t s48 = sp.s48 & kLcgM48;
        uint64_t st = (s48 ^ kLcgMul) & kLcgM48;
        const uint64_t a = lcg_next_long(st);
        const uint64_t b = lcg_next_long(st);
        const uint64_t X = carver_seed(s48, a, b, sp.cx, sp.cz);
        const bool is_start = ms_is_start(X);   // the same for the whole group
        int floor = INT_MIN, pieces = 0;
        if (is_start) {
            const MsOut o = ms_room_floor<G>(X, INT_MIN, s_box[grp], s_frm[grp], lig, grp * G);
            if (o.status == MS_DONE) floor = o.floor;
            pieces = o.pieces;
            c[o.status]++;
            c[3] += o.draws;
            c[4] += (unsigned long long)o.pieces;
        }
        if (lig == 0) {
            sp.s48 = s48;
            sp.rx = sp.cx >> 5;
            sp.rz = sp.cz >> 5;
            sp.carver = X;
            sp.room_floor = floor;
            sp.pieces = pieces;
            spots[q] = sp;
            start[q] = is_start ? 1u : 0u;
        }
    }
    for (int k = 0; k < 5; k++) {
        unsigned long long v = lig == 0 ? c[k] : 0ull;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
        if (lane == 0 && v) atomicAdd(&cnt[CNT_DD_DONE + k], v);
    }
}

// ---------------------------------------------------------------------------------------------- expansion

__global__ void k_dd_xseeds(const whf_dd_spot *__restrict__ spots, uint32_t ns, uint64_t *__restrict__ seeds) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ns) return;
    seeds[i] = ((uint64_t)(i & 0xFFFFu) << 48) | spots[i >> 16].s48;
}

// Byte i of this lane's 256-byte table in the lane-interleaved LDS layout (word w of lane l at w * 64 + l).
struct LanePerm {
    const uint8_t *bytes;
    uint32_t lane4;
    __device__ __forceinline__ int operator()(int i) const {
        i &= 255;
        return bytes[((uint32_t)(i >> 2) << 8) + lane4 + (uint32_t)(i & 3)];
    }
};

// One lane carries one world seed through the tiered gate: at level k it builds erosion sampler k (set-E order A0 B0
// A1 B1 A3 B3 A4 B4) in its LDS column, evaluates it at the hut, adds w[k] * v and, from k = 1 on, applies tier
// k - 1; a seed that fails a tier, or passes all eight, hands its lane the next seed from a batch-wide counter
// (fetched 64 at a time per wave). Every lane runs the same build loop each iteration, whatever level its seed is
// at, so no lane waits on another seed's level and no per-level lists are kept.
__global__ __launch_bounds__(64) void k_dd_erosion(const uint64_t *__restrict__ seeds, uint32_t ns,
                                                    const whf_dd_spot *__restrict__ spots,
                                                    const BuildDesc *__restrict__ bdE,
                                                    const GateTables *__restrict__ gt, int single,
                                                    uint4 *__restrict__ q1, uint32_t *__restrict__ ctl,
                                                    uint32_t *__restrict__ needC,
                                                    unsigned long long *__restrict__ cnt) {
    __shared__ uint32_t words[64 * 64];
    __shared__ uint32_t s_cnt[8];
    uint8_t *bytes = reinterpret_cast<uint8_t *>(words);
    const uint32_t lane = threadIdx.x;
    if (lane < 8) s_cnt[lane] = 0;
    __syncthreads();
    const EroDesc &ed = gt->ero;
    const double amplitude = ed.amplitude;
    unsigned long long built = 0;

    // fetch: every lane whose seed is finished takes the next index of the batch
    auto fetch = [&](bool need) -> uint32_t {
        const unsigned long long m = __ballot(need);
        uint32_t base = 0;
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            if ((int)lane == leader) base = atomicAdd(&ctl[CTL_DD_WORK], (uint32_t)__popcll(m));
            base = __shfl(base, leader);
        }
        return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    };

    if (single) {
        // SINGLE_BIOME: the climate gates, erosion included, are skipped
        for (uint32_t i = blockIdx.x * 64 + lane; i < ns; i += gridDim.x * 64) {
            const whf_dd_spot sp = spots[i >> 16];
            const uint32_t pos = atomicAdd(&ctl[CTL_Q1], 1u);
            q1[pos] = make_uint4(i, (uint32_t)sp.cx, (uint32_t)sp.cz, 0u);
            needC[i] = 1u;
        }
        return;
    }

    uint32_t idx = fetch(true);
    int k = 0;
    double sum = 0.0, x = 0.0, z = 0.0;
    bool raw = true;
    int cx = 0, cz = 0;
    if (idx < ns) {
        const whf_dd_spot sp = spots[idx >> 16];
        cx = sp.cx;
        cz = sp.cz;
        x = (double)(cx * 16 + 8) / 4.0;
        z = (double)(cz * 16 + 8) / 4.0;
        raw = fabs(x) < ed.coord_lim && fabs(z) < ed.coord_lim;
    }
    while (__ballot(idx < ns)) {
        bool done = false, pass = false;
        if (idx < ns) {
            // worldSeed -> deriver -> createRandom(erosion) -> octave deriver -> createRandom("octave_<n>")
            // (noise.c:80-143, rng.c:157-170), then PerlinNoiseSampler(random) (noise.c:61-78)
            const BuildDesc d = bdE[k];
            Xr r = xr_seed(seeds[idx]);
            Xr der = xr_fork(r);
            Xr nr = xr_make(d.na ^ der.lo, d.nb ^ der.hi);
            if (d.sub) { xr_next(nr); xr_next(nr); }
            Xr od = xr_fork(nr);
            Xr pr = xr_make(d.oa ^ od.lo, d.ob ^ od.hi);
            SPar s;
            s.ox = xr_next_double(pr) * 256.0;
            s.oy = xr_next_double(pr) * 256.0;
            s.oz = xr_next_double(pr) * 256.0;
            shuffle_perm(pr, words, lane);
            const double fj = floor(s.oy);
            s.jy = (int)fj;
            s.hy = s.oy - fj;
            s.hy1 = s.hy - 1.0;
            s.fy = fade(s.hy);
            s.pad = 0;
            built++;
            double xs = x * ed.sc[k] * ed.lac[k], zs = z * ed.sc[k] * ed.lac[k];
            if (!raw) { xs = keep_precision(xs); zs = keep_precision(zs); }
            const double v = perlin_y0(LanePerm{bytes, lane * 4}, s, xs, zs);
            sum += ed.w[k] * v;
            if (k >= 1 && sum * amplitude < ed.thr[k - 1]) {
                atomicAdd(&s_cnt[k - 1], 1u);
                done = true;
            } else if (k == 7) {
                pass = true;
                done = true;
            } else {
                k++;
            }
        }
        const uint32_t pos = wave_append(pass, &ctl[CTL_Q1]);
        if (pass) {
            q1[pos] = make_uint4(idx, (uint32_t)cx, (uint32_t)cz, 0u);
            needC[idx] = 1u;
        }
        const uint32_t nidx = fetch(done);
        if (done) {
            idx = nidx;
            k = 0;
            sum = 0.0;
            if (idx < ns) {
                const whf_dd_spot sp = spots[idx >> 16];
                cx = sp.cx;
                cz = sp.cz;
                x = (double)(cx * 16 + 8) / 4.0;
                z = (double)(cz * 16 + 8) / 4.0;
                raw = fabs(x) < ed.coord_lim && fabs(z) < ed.coord_lim;
            }
        }
    }
    __syncthreads();
    if (lane < 7 && s_cnt[lane]) atomicAdd(&cnt[CNT_TIER0 + lane], (unsigned long long)s_cnt[lane]);
    for (int o = 32; o > 0; o >>= 1) built += __shfl_down(built, o);
    if (lane == 0 && built) atomicAdd(&cnt[CNT_BUILT + SET_E], built);
}

// ---------------------------------------------------------------------------------------------- no-alloc sweep

// whf_deepdive_noalloc's working memory: the control words, the counters, one sweep batch's start queue and its
// spots. A file-scope device array, so the entry point allocates nothing; every device the library runs on holds its
// own copy (about 8.4 MB) until the process ends, and every call on a device shares it.
constexpr uint32_t kNaStartCap = 1u << 17;      // starts per sweep batch; batches expect half of it
constexpr uint32_t kNaSpotCap = kNaStartCap;    // a spot is a start, so the spot queue cannot overflow
constexpr size_t kNaCtlOff = 0;                 // uint32_t[CTL_N]
constexpr size_t kNaCntOff = 256;               // unsigned long long[CNT_N]
constexpr size_t kNaStartsOff = 1024;           // DdStart[kNaStartCap]
constexpr size_t kNaSpotsOff = kNaStartsOff + (((size_t)kNaStartCap * sizeof(DdStart) + 255) & ~(size_t)255);
constexpr size_t kNaScratchBytes = kNaSpotsOff + (size_t)kNaSpotCap * sizeof(whf_dd_spot);
static_assert(CTL_N * sizeof(uint32_t) <= kNaCntOff - kNaCtlOff, "control words fit");
static_assert(CNT_N * sizeof(unsigned long long) <= kNaStartsOff - kNaCntOff, "counters fit");

static __device__ __align__(256) uint8_t scratch[kNaScratchBytes];

// Zeroes the control words and, with all != 0, the counters (a kernel, not hipMemsetAsync: callers may capture the
// stream into a graph).
__global__ void k_na_clear(uint32_t *__restrict__ ctl, unsigned long long *__restrict__ cnt, int all) {
    const int t = (int)threadIdx.x;
    if (t < CTL_N) ctl[t] = 0u;
    if (all && t < CNT_N) cnt[t] = 0ull;
}

// ---------------------------------------------------------------------------------------------- host

struct DdBufs {
    DdStart *starts;
    whf_dd_spot *spots;
};

int dd_alloc(whf_ctx *c, DdBufs &b) {
    const size_t bytes = (size_t)kDdStartCap * sizeof(DdStart) + (size_t)kDdSpotCap * sizeof(whf_dd_spot);
    if (!c->d_dd) {
        HIP_TRY(hipMalloc(&c->d_dd, bytes));
        c->dd_bytes = bytes;
    }
    b.starts = reinterpret_cast<DdStart *>(c->d_dd);
    b.spots = reinterpret_cast<whf_dd_spot *>(reinterpret_cast<uint8_t *>(c->d_dd) +
                                              (size_t)kDdStartCap * sizeof(DdStart));
    return WHF_OK;
}

// One sweep batch: start tests and layout walks of S in [s_first, s_first + n) (mod 2^48); the batch's spots are
// left at the front of b.spots.
int dd_sweep(whf_ctx *c, const whf_dd_request *req, const DdBufs &b, uint64_t s_first, uint32_t n,
             uint32_t *n_spots) {
    const hipStream_t st = c->stream;
    const int nrx = req->rx1 - req->rx0, nrz = req->rz1 - req->rz0;
    const int salt = req->salt ? req->salt : 14357620;
    HIP_TRY(hipMemsetAsync(c->d_ctl + CTL_DD_STARTS, 0, 2 * sizeof(uint32_t), st));
    {
        Prof p(c, F_DD_SWEEP);
        hipLaunchKernelGGL(k_dd_sweep, dim3((n + 255) / 256), dim3(256), 0, st, s_first, n, req->rx0, nrx, req->rz0,
                           nrz, salt, b.starts, kDdStartCap, c->d_ctl);
    }
    uint32_t n_starts = 0;
    HIP_TRY(hipMemcpyAsync(&n_starts, c->d_ctl + CTL_DD_STARTS, sizeof n_starts, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (n_starts > kDdStartCap) {
        // batches are sized for half the queue on average (the start count is binomial, p = 0.004)
        std::snprintf(g_last_error, sizeof g_last_error, "start queue overflow (%u starts in one batch)", n_starts);
        return WHF_EOVERFLOW;
    }
    {
        Prof p(c, F_DD_FLOOR);
        const int cut_rel = -62 - req->floor_max;
        const uint32_t groups = 64 / kDdGroup;
        const uint32_t grid = std::max(1u, std::min((n_starts + groups - 1) / groups, (uint32_t)c->n_cu * 16));
        hipLaunchKernelGGL(k_dd_floor<kDdGroup>, dim3(grid), dim3(64), 0, st, b.starts, c->d_ctl, kDdStartCap, cut_rel,
                           req->floor_max, b.spots, kDdSpotCap, c->d_ctl, c->d_cnt);
    }
    HIP_TRY(hipMemcpyAsync(n_spots, c->d_ctl + CTL_DD_SPOTS, sizeof *n_spots, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return check_launch();
}

int dd_expand(whf_ctx *c, const whf_dd_request *req, const whf_dd_spot *d_spots, uint32_t n_spots) {
    const hipStream_t st = c->stream;
    const int layout = req->vflags;
    // each batch of 16 spots: its world seeds and the lazy erosion gate on the library's stream, then the Seed
    // Check's gates 2 and 3 on the pipeline's stage streams beside the next batch
    Pipeline pl(c);
    if (int rc = pl.begin((n_spots + kDdSpotsPerExpand - 1) / kDdSpotsPerExpand)) return rc;
    for (uint32_t s0 = 0; s0 < n_spots; s0 += kDdSpotsPerExpand) {
        const uint32_t k = std::min(kDdSpotsPerExpand, n_spots - s0);
        const uint32_t ns = k << 16;
        const BatchSet *bp = nullptr;
        if (int rc = pl.acquire(bp)) return rc;
        const BatchSet &b = *bp;
        reset_set(c, b, ns);
        {
            Prof p(c, F_DD_EXPAND);
            hipLaunchKernelGGL(k_dd_xseeds, dim3((ns + 255) / 256), dim3(256), 0, st, d_spots + s0, ns, b.seeds);
            HIP_TRY(hipMemsetAsync(b.ctl + CTL_DD_WORK, 0, sizeof(uint32_t), st));
            // persistent lanes: about one workgroup per LDS slot of the device (16 KB each)
            const uint32_t grid = std::min((ns + 63) / 64, (uint32_t)c->n_cu * 10);
            hipLaunchKernelGGL(k_dd_erosion, dim3(grid), dim3(64), 0, st, b.seeds, ns, d_spots + s0,
                               bd_of(c, layout, SET_E), c->d_gt + layout, (req->vflags & WHF_V_SINGLE) ? 1 : 0,
                               b.q1, b.ctl, b.needC, c->d_cnt);
        }
        if (int rc = pl.tail(b, layout, ns, req->vflags, req->max_height)) return rc;
    }
    return pl.end();
}

// After the last batch of a Deep Dive call: the finds (floor_y = the room floor of their spot) and the counters
// that the sweep and the spots entry share.
// The swamp check of the expansion's survivors (lysh's biome test at the chunk middle's pre-carver surface,
// biome.c:94-156): keeps, in place and in order, the survivors whose chunk middle is swamp with the tree `tree`.
int dd_biome_filter(whf_ctx *c, std::vector<DevHit> &hh, int tree, int vflags, uint64_t *rejected) {
    *rejected = 0;
    if (hh.empty()) return WHF_OK;
    std::vector<whf_biome_query> q(hh.size());
    for (size_t i = 0; i < hh.size(); i++) q[i] = whf_biome_query{hh[i].seed, hh[i].cx * 16, hh[i].cz * 16, WHF_OCC_SURFACE, 0};
    std::vector<whf_biome_result> r(hh.size());
    const int rc = whf_biome(c, q.data(), q.size(), tree, (vflags & WHF_V_LARGE) | WHF_BIO_OCC_ONLY, r.data(), nullptr);
    if (rc != WHF_OK) return rc;
    size_t k = 0;
    for (size_t i = 0; i < hh.size(); i++)
        if (r[i].ok) hh[k++] = hh[i];
    *rejected = hh.size() - k;
    hh.resize(k);
    return WHF_OK;
}

// The tree of a Deep Dive request's swamp check: -1 for none (req->biome_tree < 0).
int dd_tree(const whf_dd_request *req, int *tree) {
    if (req->biome_tree < 0) { *tree = -1; return WHF_OK; }
    if (req->biome_tree > 0) { *tree = req->biome_tree - 1; return WHF_OK; }
    // the rule flags' default tree, as lysh picks it (eval.c:104-106)
    const char *dflt = (req->vflags & WHF_V_118) ? "1.18.2" : ((req->vflags & WHF_V_262) ? "26.2" : "1.21.5");
    return whf_biome_tree(dflt, tree);
}

int dd_finish(whf_ctx *c, bool expand, int tree, int vflags, uint64_t spots_total,
              const std::map<std::tuple<uint64_t, int, int>, int> &floor_of, whf_sink *finds, whf_dd_stats *stats) {
    const hipStream_t st = c->stream;
    HIP_TRY(hipEventRecord(c->ev1, st));
    unsigned long long cnt[CNT_N];
    uint32_t ctl[CTL_N];
    HIP_TRY(hipMemcpyAsync(cnt, c->d_cnt, sizeof cnt, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(ctl, c->d_ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (cnt[CNT_SLOT_OVER]) {
        std::snprintf(g_last_error, sizeof g_last_error, "%llu gate survivors beyond the sampler-table slots",
                      (unsigned long long)cnt[CNT_SLOT_OVER]);
        return WHF_EOVERFLOW;
    }
    int rc = WHF_OK;
    uint64_t hits_total = expand ? ctl[CTL_HITS] : 0;
    uint64_t checked = 0, rejected = 0;
    if (finds || (expand && tree >= 0)) {
        std::vector<DevHit> hh(std::min<uint64_t>(hits_total, kHitCap));
        if (!hh.empty()) HIP_TRY(hipMemcpy(hh.data(), c->d_hits, hh.size() * sizeof(DevHit), hipMemcpyDeviceToHost));
        if (expand && tree >= 0) {
            if (hh.size() < hits_total) {
                std::snprintf(g_last_error, sizeof g_last_error, "%llu expansion survivors beyond the %u-survivor buffer",
                              (unsigned long long)hits_total, kHitCap);
                return WHF_EOVERFLOW;
            }
            checked = hh.size();
            if ((rc = dd_biome_filter(c, hh, tree, vflags, &rejected)) != WHF_OK) return rc;
            hits_total = hh.size();
        }
        if (finds) finds->count = hits_total;
        const uint64_t m = finds ? std::min<uint64_t>(finds->capacity, hh.size()) : 0;
        for (uint64_t i = 0; i < m; i++) {
            whf_hit h;
            std::memcpy(&h, &hh[i], sizeof h);
            auto it = floor_of.find(std::make_tuple((uint64_t)h.seed & kLcgM48, h.cx, h.cz));
            h.floor_y = it == floor_of.end() ? INT32_MIN : it->second;
            h.flags = WHF_HIT_DEEPDIVE;
            finds->hits[i] = h;
        }
        if (finds && (hits_total > finds->capacity || hh.size() < hits_total)) rc = WHF_EOVERFLOW;
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->walks[0] = cnt[CNT_DD_DONE];
        stats->walks[1] = cnt[CNT_DD_CUT];
        stats->walks[2] = cnt[CNT_DD_CAPPED];
        stats->starts = stats->walks[0] + stats->walks[1] + stats->walks[2];
        stats->spots = spots_total;
        stats->draws = cnt[CNT_DD_DRAWS];
        stats->pieces = cnt[CNT_DD_PIECES];
        if (expand) {
            stats->biome_checked = checked;
            stats->biome_rejected = rejected;
            stats->expanded_seeds = spots_total << 16;
            uint64_t ero = 0;
            for (int t = 0; t < 7; t++) { stats->tiers[t] = cnt[CNT_TIER0 + t]; ero += cnt[CNT_TIER0 + t]; }
            stats->funnel[0] = ero;
            for (int k = 1; k < WHF_KIND_N; k++) stats->funnel[k] = cnt[CNT_KIND0 + k];
            stats->accepted = cnt[CNT_KIND0 + 10];
            for (int s = 0; s < SET_COUNT; s++) stats->samplers_built[s] = cnt[CNT_BUILT + s];
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        stats->device_ms = ms;
        if (c->profile) {
            for (size_t i = 0; i + 1 < c->prof_ev.size(); i += 2) {
                float m2 = 0.f;
                (void)hipEventElapsedTime(&m2, c->prof_ev[i], c->prof_ev[i + 1]);
                const int f = c->prof_fam[i / 2];
                if (f == F_DD_SWEEP) stats->stage_ms[0] += m2;
                else if (f == F_DD_FLOOR) stats->stage_ms[1] += m2;
                else stats->stage_ms[2] += m2;   // the expansion and the Seed Check gates it runs
            }
        }
    }
    return rc;
}

}  // namespace

namespace whf {

// The device address of `scratch` on the context's device, looked up once per device (with that device current) and
// kept for every later context on it: in host code the array's name is only a host-side stand-in.
int dd_scratch_init(whf_ctx *c) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    if (c->device < 0 || c->device >= kMaxDevices) return WHF_EDEVICE;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[c->device]) {
        HIP_TRY(hipSetDevice(c->device));
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return WHF_EDEVICE;
        addr[c->device] = p;
    }
    c->dd_scratch = addr[c->device];
    return WHF_OK;
}

}  // namespace whf

extern "C" {

int whf_deepdive(whf_ctx *c, const whf_dd_request *req, whf_dd_sink *spots, whf_sink *finds, whf_dd_stats *stats) {
    if (!c || !req) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (req->rx1 <= req->rx0 || req->rz1 <= req->rz0) return WHF_EINVAL;
    const int64_t lim = 65536;
    if (req->rx0 < -lim || req->rx1 > lim || req->rz0 < -lim || req->rz1 > lim) return WHF_EINVAL;
    if ((uint64_t)(req->rx1 - req->rx0) * (uint64_t)(req->rz1 - req->rz0) > 4096) return WHF_EINVAL;
    if (req->count > (1ull << 48)) return WHF_EINVAL;
    if (req->floor_max < -1000 || req->floor_max > 1000) return WHF_EINVAL;
    if (req->expand != 0 && req->expand != 1) return WHF_EINVAL;
    if (req->vflags < 0 || req->vflags >= kLayouts) return WHF_EINVAL;
    if (req->expand && (req->max_height < -140 || req->max_height > 100000)) return WHF_EINVAL;
    if (req->expand && (req->biome_tree < -1 || req->biome_tree > 8)) return WHF_EINVAL;
    if (spots && spots->capacity && !spots->spots) return WHF_EINVAL;
    if (finds && finds->capacity && !finds->hits) return WHF_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    DdBufs b;
    int rc = dd_alloc(c, b);
    if (rc) return rc;
    const hipStream_t st = c->stream;
    HIP_TRY(hipMemsetAsync(c->d_cnt, 0, CNT_N * sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(c->d_ctl, 0, CTL_N * sizeof(uint32_t), st));
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    c->prof_ev.clear();
    c->prof_fam.clear();
    HIP_TRY(hipEventRecord(c->ev0, st));

    // structure seeds per batch: about kDdStartCap / 2 expected starts (start probability 0.004 = 1/250)
    const uint64_t regions = (uint64_t)(req->rx1 - req->rx0) * (uint64_t)(req->rz1 - req->rz0);
    const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(kDdSweepMax, (uint64_t)(kDdStartCap / 2) * 250 / regions));
    uint64_t spots_total = 0;
    std::map<std::tuple<uint64_t, int, int>, int> floor_of;   // (S, cx, cz) -> room floor, to label the finds
    std::vector<whf_dd_spot> host_spots;
    for (uint64_t done = 0; done < req->count; done += per) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(per, req->count - done);
        uint32_t ns = 0;
        rc = dd_sweep(c, req, b, (req->s_first + done) & kLcgM48, n, &ns);
        if (rc) return rc;
        if (ns > kDdSpotCap) {
            std::snprintf(g_last_error, sizeof g_last_error, "spot buffer overflow (%u spots in one batch)", ns);
            return WHF_EOVERFLOW;
        }
        if (ns) {
            host_spots.resize(ns);
            HIP_TRY(hipMemcpyAsync(host_spots.data(), b.spots, ns * sizeof(whf_dd_spot), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            for (const whf_dd_spot &s : host_spots) {
                if (spots && spots_total < spots->capacity) spots->spots[spots_total] = s;
                floor_of[std::make_tuple(s.s48, s.cx, s.cz)] = s.room_floor;
                spots_total++;
            }
            if (req->expand) {
                rc = dd_expand(c, req, b.spots, ns);
                if (rc) return rc;
            }
        }
    }
    int tree = -1;
    if (req->expand && (rc = dd_tree(req, &tree)) != WHF_OK) return rc;
    rc = dd_finish(c, req->expand != 0, tree, req->vflags, spots_total, floor_of, finds, stats);
    if (rc != WHF_OK && rc != WHF_EOVERFLOW) return rc;
    if (spots) {
        spots->count = spots_total;
        if (spots_total > spots->capacity) rc = WHF_EOVERFLOW;
    }
    if (stats) {
        stats->structure_seeds = req->count;
        stats->tests = req->count * regions;
    }
    return rc;
}

int whf_deepdive_noalloc(whf_ctx *c, const whf_dd_request *req, whf_dd_sink *spots, whf_dd_stats *stats) {
    if (!c || !req) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (req->expand != 0) {
        std::snprintf(g_last_error, sizeof g_last_error, "whf_deepdive_noalloc sweeps only (expand must be 0)");
        return WHF_EINVAL;
    }
    if (req->rx1 <= req->rx0 || req->rz1 <= req->rz0) return WHF_EINVAL;
    const int64_t lim = 65536;
    if (req->rx0 < -lim || req->rx1 > lim || req->rz0 < -lim || req->rz1 > lim) return WHF_EINVAL;
    const uint64_t regions = (uint64_t)(req->rx1 - req->rx0) * (uint64_t)(req->rz1 - req->rz0);
    if (regions > 4096) {
        std::snprintf(g_last_error, sizeof g_last_error, "%llu hut regions: at most 4096",
                      (unsigned long long)regions);
        return WHF_EINVAL;
    }
    if (req->count > (1ull << 48)) return WHF_EINVAL;
    if (req->floor_max < -1000 || req->floor_max > 1000) return WHF_EINVAL;
    if (spots && spots->capacity && !spots->spots) return WHF_EINVAL;
    if (!c->dd_scratch) return WHF_EDEVICE;
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t st = c->stream;
    uint8_t *base = static_cast<uint8_t *>(c->dd_scratch);
    uint32_t *ctl = reinterpret_cast<uint32_t *>(base + kNaCtlOff);
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(base + kNaCntOff);
    DdStart *starts = reinterpret_cast<DdStart *>(base + kNaStartsOff);
    whf_dd_spot *d_spots = reinterpret_cast<whf_dd_spot *>(base + kNaSpotsOff);
    HIP_TRY(hipEventRecord(c->ev0, st));
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, st, ctl, cnt, 1);

    // structure seeds per batch: about kNaStartCap / 2 expected starts (start probability 0.004 = 1/250)
    const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(kDdSweepMax, (uint64_t)(kNaStartCap / 2) * 250 / regions));
    const int nrx = req->rx1 - req->rx0, nrz = req->rz1 - req->rz0;
    const int salt = req->salt ? req->salt : 14357620;
    const int cut_rel = -62 - req->floor_max;
    uint64_t spots_total = 0;
    for (uint64_t done = 0; done < req->count; done += per) {
        const uint32_t n = (uint32_t)std::min<uint64_t>(per, req->count - done);
        if (done) hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, st, ctl, cnt, 0);
        hipLaunchKernelGGL(k_dd_sweep, dim3((n + 255) / 256), dim3(256), 0, st, (req->s_first + done) & kLcgM48, n,
                           req->rx0, nrx, req->rz0, nrz, salt, starts, kNaStartCap, ctl);
        uint32_t n_starts = 0;
        HIP_TRY(hipMemcpyAsync(&n_starts, ctl + CTL_DD_STARTS, sizeof n_starts, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (int rc = check_launch()) return rc;
        if (n_starts > kNaStartCap) {
            std::snprintf(g_last_error, sizeof g_last_error, "start queue overflow (%u starts in one batch, at most %u)",
                          n_starts, kNaStartCap);
            return WHF_EOVERFLOW;
        }
        const uint32_t groups = 64 / kDdGroup;
        const uint32_t grid = std::max(1u, std::min((n_starts + groups - 1) / groups, (uint32_t)c->n_cu * 16));
        hipLaunchKernelGGL(k_dd_floor<kDdGroup>, dim3(grid), dim3(64), 0, st, starts, ctl, kNaStartCap, cut_rel,
                           req->floor_max, d_spots, kNaSpotCap, ctl, cnt);
        uint32_t ns = 0;
        HIP_TRY(hipMemcpyAsync(&ns, ctl + CTL_DD_SPOTS, sizeof ns, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (int rc = check_launch()) return rc;
        // straight into the caller's sink, as far as it holds them
        if (spots && spots_total < spots->capacity && ns) {
            const uint64_t m = std::min<uint64_t>(ns, spots->capacity - spots_total);
            HIP_TRY(hipMemcpyAsync(spots->spots + spots_total, d_spots, m * sizeof(whf_dd_spot), hipMemcpyDeviceToHost,
                                   st));
            HIP_TRY(hipStreamSynchronize(st));
        }
        spots_total += ns;
    }
    HIP_TRY(hipEventRecord(c->ev1, st));
    unsigned long long h[CNT_N];
    HIP_TRY(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int rc = WHF_OK;
    if (spots) {
        spots->count = spots_total;
        if (spots_total > spots->capacity) rc = WHF_EOVERFLOW;
    }
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->structure_seeds = req->count;
        stats->tests = req->count * regions;
        stats->walks[0] = h[CNT_DD_DONE];
        stats->walks[1] = h[CNT_DD_CUT];
        stats->walks[2] = h[CNT_DD_CAPPED];
        stats->starts = stats->walks[0] + stats->walks[1] + stats->walks[2];
        stats->spots = spots_total;
        stats->draws = h[CNT_DD_DRAWS];
        stats->pieces = h[CNT_DD_PIECES];
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        stats->device_ms = ms;
    }
    return rc;
}

int whf_dd_terrain(whf_ctx *c, const whf_dd_request *req, whf_dd_spot *spots, uint64_t n, whf_sink *finds,
                   whf_dd_stats *stats) {
    if (!c || !req || (n && !spots)) return WHF_EINVAL;
    g_last_error[0] = 0;
    if (n > kDdSpotCap) return WHF_EINVAL;
    if (req->vflags < 0 || req->vflags >= kLayouts) return WHF_EINVAL;
    if (req->max_height < -140 || req->max_height > 100000) return WHF_EINVAL;
    if (finds && finds->capacity && !finds->hits) return WHF_EINVAL;
    HIP_TRY(hipSetDevice(c->device));
    DdBufs b;
    int rc = dd_alloc(c, b);
    if (rc) return rc;
    const hipStream_t st = c->stream;
    HIP_TRY(hipMemsetAsync(c->d_cnt, 0, CNT_N * sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(c->d_ctl, 0, CTL_N * sizeof(uint32_t), st));
    for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
    c->prof_ev.clear();
    c->prof_fam.clear();
    HIP_TRY(hipEventRecord(c->ev0, st));
    std::map<std::tuple<uint64_t, int, int>, int> floor_of;
    uint64_t n_starts = 0;
    if (n) {
        // the spots in b.spots and their start flags in the start queue's memory (kDdStartCap * 24 bytes >= n * 4)
        uint32_t *d_start = reinterpret_cast<uint32_t *>(b.starts);
        std::vector<uint32_t> is_start(n);
        HIP_TRY(hipMemcpyAsync(b.spots, spots, n * sizeof(whf_dd_spot), hipMemcpyHostToDevice, st));
        {
            Prof p(c, F_DD_FLOOR);
            const uint32_t groups = 64 / kDdGroup;
            const uint32_t grid = std::max(1u, std::min(((uint32_t)n + groups - 1) / groups, (uint32_t)c->n_cu * 16));
            hipLaunchKernelGGL(k_dd_spot_walks<kDdGroup>, dim3(grid), dim3(64), 0, st, b.spots, (uint32_t)n, d_start,
                               c->d_cnt);
        }
        HIP_TRY(hipMemcpyAsync(spots, b.spots, n * sizeof(whf_dd_spot), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(is_start.data(), d_start, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if ((rc = check_launch())) return rc;
        // every start is expanded, whatever its room floor (the reference's --spots mode tests the start only)
        std::vector<whf_dd_spot> starts;
        for (uint64_t i = 0; i < n; i++) {
            if (!is_start[i]) continue;
            starts.push_back(spots[i]);
            floor_of[std::make_tuple(spots[i].s48, spots[i].cx, spots[i].cz)] = spots[i].room_floor;
        }
        n_starts = starts.size();
        if (n_starts) {
            HIP_TRY(hipMemcpyAsync(b.spots, starts.data(), n_starts * sizeof(whf_dd_spot), hipMemcpyHostToDevice, st));
            if ((rc = dd_expand(c, req, b.spots, (uint32_t)n_starts))) return rc;
        }
    }
    // the terrain stage's survivors are the phase-1 chain's, as the reference's --spots mode reports them (hunt.py
    // confirms them with lysh hut, which includes the biome test)
    rc = dd_finish(c, true, -1, req->vflags, n_starts, floor_of, finds, stats);
    if (stats) {
        stats->structure_seeds = n;
        stats->tests = n;
    }
    return rc;
}

int whf_room_floors(whf_ctx *c, const uint64_t *carvers, uint64_t n, int32_t *status, int32_t *room_floor,
                    int32_t *pieces) {
    if (!c || (n && (!carvers || !status || !room_floor || !pieces))) return WHF_EINVAL;
    if (n > (1u << 24)) return WHF_EINVAL;
    if (!n) return WHF_OK;
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t st = c->stream;
    uint64_t *d_in = nullptr;
    int4 *d_out = nullptr;
    HIP_TRY(hipMalloc(&d_in, n * sizeof(uint64_t)));
    if (hipMalloc(&d_out, n * sizeof(int4)) != hipSuccess) {
        (void)hipFree(d_in);
        return WHF_ENOMEM;
    }
    std::vector<int4> out(n);
    int rc = WHF_OK;
    if (hipMemcpyAsync(d_in, carvers, n * sizeof(uint64_t), hipMemcpyHostToDevice, st) != hipSuccess) rc = WHF_EDEVICE;
    if (rc == WHF_OK) {
        const uint32_t groups = 64 / kDdGroup;
        const uint32_t grid = std::min<uint32_t>(((uint32_t)n + groups - 1) / groups, (uint32_t)c->n_cu * 16);
        hipLaunchKernelGGL(k_dd_walks<kDdGroup>, dim3(grid), dim3(64), 0, st, d_in, (uint32_t)n, d_out);
        rc = check_launch();
    }
    if (rc == WHF_OK && (hipMemcpyAsync(out.data(), d_out, n * sizeof(int4), hipMemcpyDeviceToHost, st) != hipSuccess ||
                         hipStreamSynchronize(st) != hipSuccess))
        rc = WHF_EDEVICE;
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (rc) return rc;
    for (uint64_t i = 0; i < n; i++) {
        status[i] = out[i].x;
        room_floor[i] = out[i].y;
        pieces[i] = out[i].z;
    }
    return WHF_OK;
}

}  // extern "C"
