// whf_solve.hip — the Deep Dive solver on gfx950: the structure seeds S whose swamp-hut chunk carries a catalogued
// mineshaft layout (carver seed X), over the reference solver's task space (hut_lookup_v4 --solve5,
// hut_lookup_v4_source.cu:1293-1403): bundles of 2^20 groups (taskspace.h), the hut regions of a group, the chunk
// offsets of a tier.
//
// A hit of task (X, chunk (cx, cz), region (rx, rz) = (cx >> 5, cz >> 5), offset (ox, oz)) is an S with
//   carver(S, cx, cz) = (cx * a ^ cz * b ^ S) mod 2^48 = X, where a, b are the first two nextLong() of
//                      java.util.Random(S) (room_core.h:55-58), and
//   the region's hut at (ox, oz): two nextInt(24) of Random(S + Creg(rx, rz)) (solver_kernels.cuh:78-82),
// among the S whose first two hut draws agree with (ox, oz) modulo 8 in the low 20 state bits, the candidate set
// the reference's pools are built from (solver_kernels.cuh:241-255).
//
// With t = S ^ 0x5DEECE66D (Random's state), bit k of the carver equation's low 32 bits depends on t's bits
// <= k - m + 16 only, m being the smaller 2-adic valuation of cx and cz (their products with the draws have m
// trailing zeros). So the solver works from the low end:
//   k_sv_regions  one thread per (group, rx): the rz in [-58593, 58593] whose region constant has the group's
//                 20-bit residue (taskspace.h:27-37), kept if the z regime holds it.
//   k_sv_pool     one workgroup per (group, low-bit class): a class is (m, (cx >> m) mod 16, (cz >> m) mod 16),
//                 which fixes the hut filter's key (cx mod 8, cz mod 8) and the equation's low m + 4 bits as
//                 functions of t mod 2^20. The workgroup walks the 20-bit region-seed residues of that key (a table
//                 sorted by the first two draws modulo 8, built once) and keeps each t20 whose equation bits
//                 0 .. m + 3 are zero: about 1024 / 2^m of 16384.
//   k_sv_lift     one wave per (region, offset) task, one lane per pool entry: the lane extends t20 to 48 bits.
//                 Where the valuations differ, bit k flips with t's bit k - m + 16 and with nothing below, so each
//                 level has exactly one choice: 28 - m levels of a few adds and xors on the running products
//                 cx * h2 and cz * h4 (h2, h4 = bits 16..47 of the second and fourth LCG states), then the m top
//                 bits, which the low 32 do not see, take every value. Where they are equal the two products'
//                 flips cancel and the level branches: a depth-first search with a bit stack of pending branches.
//                 Every full state gets the exact carver check (all 48 bits) and then the exact hut check.
// Hits come back sorted by (bundle, S, cx, cz); the reference sorts and de-duplicates per job
// (hut_lookup_v4_source.cu:1333-1334), and a task cannot produce an S twice.
#include "whf_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace {

using namespace whf;

constexpr uint64_t kM48 = 0xFFFFFFFFFFFFull;
constexpr uint64_t kLA = 0x5DEECE66Dull;   // java.util.Random
constexpr uint64_t kLC = 0xBull;
constexpr uint64_t kLA2 = (kLA * kLA) & kM48;                                  // two steps: t -> a2 t + k2
constexpr uint64_t kLK2 = (kLC * (kLA + 1)) & kM48;
constexpr uint64_t kLA4 = (kLA2 * kLA2) & kM48;                                // four steps
constexpr uint64_t kLK4 = (kLC * (kLA * kLA * kLA + kLA * kLA + kLA + 1)) & kM48;
constexpr uint64_t kRegC0 = 14357620ull, kRegC1 = 341873128712ull, kRegC2 = 132897987541ull;   // taskspace.h:10
constexpr int kRadius = 58593;                                                                  // taskspace.h:9
constexpr uint32_t kPoolCap = 4096;      // pool entries per (group, class): about 1024 expected, 4x headroom
constexpr uint32_t kMaxSlots = 1u << 14; // (group, class) pools per launch
constexpr uint32_t kMaxGroups = 256;     // groups per launch
constexpr uint32_t kHitCapSolve = 1u << 22;
constexpr uint32_t kMaxChunks = 8192;    // chunk tasks per launch of whf_solve_chunks (one pool each)
constexpr int kSvWG = 256;

struct SvRegion {
    int32_t rx, rz;
    uint32_t g;     // group (pool block) of the region
    uint32_t pad;
};

struct SvHit {
    uint64_t s48;
    int32_t cx, cz;
    uint32_t idx;   // group or chunk index in the launch
    uint32_t pad;
};

enum SvCnt { SV_CAND = 0, SV_LEAVES, SV_TASKS, SV_POOL_OVER, SV_BUDGET, SV_N };
// Nodes one branching search may visit. Each level halves the survivors and doubles them again, so a candidate's
// search visits about 2 x 28 nodes; a search that reaches the budget stops and makes the call fail (never a silent
// partial answer).
constexpr uint32_t kNodeBudget = 1u << 16;

struct SolveBufs {
    uint32_t *keyq;      // [2^20] region-seed residues q sorted by key (first two draws mod 8)
    uint32_t *key_off;   // [65]
    uint32_t *c20;       // [kMaxChunks] residue of each group (or chunk)
    SvRegion *regions;   // [kMaxGroups * 51650]
    uint32_t *n_regions; // [1]
    uint32_t *pool;      // [kMaxSlots * kPoolCap]
    uint32_t *pool_n;    // [kMaxSlots]
    uint16_t *kappa;     // [kMaxChunks] class descriptors (m << 8 | cxr << 4 | czr): of the tier, by pool index
                         // within a group (whf_solve), or of each chunk (whf_solve_chunks)
    int16_t *kap_idx;    // [1280] class -> pool index, -1 if not in the launch's list
    uint8_t *offs;       // [2 * 288] offsets of the tier (ox then oz)
    int32_t *chunk_xz;   // [2 * kMaxChunks]
    SvHit *hits;         // [kHitCapSolve]
    uint32_t *hit_ctr;   // [1]
    unsigned long long *cnt;   // [SV_N]
};

constexpr size_t kRegionsCap = (size_t)kMaxGroups * 51650;

size_t solve_layout(SolveBufs *b, uint8_t *base) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> uint8_t * {
        uint8_t *p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    b->keyq = reinterpret_cast<uint32_t *>(take((size_t)4 << 20));
    b->key_off = reinterpret_cast<uint32_t *>(take(65 * 4));
    b->c20 = reinterpret_cast<uint32_t *>(take(kMaxChunks * 4));
    b->regions = reinterpret_cast<SvRegion *>(take(kRegionsCap * sizeof(SvRegion)));
    b->n_regions = reinterpret_cast<uint32_t *>(take(4));
    b->pool = reinterpret_cast<uint32_t *>(take((size_t)kMaxSlots * kPoolCap * 4));
    b->pool_n = reinterpret_cast<uint32_t *>(take(kMaxSlots * 4));
    b->kappa = reinterpret_cast<uint16_t *>(take(kMaxChunks * 2));
    b->kap_idx = reinterpret_cast<int16_t *>(take(1280 * 2));
    b->offs = reinterpret_cast<uint8_t *>(take(2 * 288));
    b->chunk_xz = reinterpret_cast<int32_t *>(take(2 * kMaxChunks * 4));
    b->hits = reinterpret_cast<SvHit *>(take((size_t)kHitCapSolve * sizeof(SvHit)));
    b->hit_ctr = reinterpret_cast<uint32_t *>(take(4));
    b->cnt = reinterpret_cast<unsigned long long *>(take(SV_N * 8));
    return off;
}

// The hut filter's key of a 20-bit region-seed residue q: the first two draws' low 3 bits (bits 17..19 of the
// state after one and two steps), which are the hut offsets modulo 8 (24 = 8 x 3).
inline uint32_t draw_key(uint32_t q) {
    const uint32_t a = (uint32_t)kLA, c = (uint32_t)kLC;
    const uint32_t r0 = (q ^ a) & 0xFFFFFu;
    const uint32_t r1 = (r0 * a + c) & 0xFFFFFu;
    const uint32_t r2 = (r1 * a + c) & 0xFFFFFu;
    return (((r1 >> 17) & 7u) << 3) | ((r2 >> 17) & 7u);
}

int solve_init(whf_ctx *c, SolveBufs &b) {
    if (!c->d_solve) {
        SolveBufs probe;
        const size_t bytes = solve_layout(&probe, nullptr);
        HIP_TRY(hipMalloc(&c->d_solve, bytes));
        c->solve_bytes = bytes;
        solve_layout(&b, static_cast<uint8_t *>(c->d_solve));
        // the residue table: counting sort of all 2^20 residues by key
        std::vector<uint32_t> off(65, 0), list(1u << 20);
        for (uint32_t q = 0; q < (1u << 20); q++) off[draw_key(q) + 1]++;
        for (int k = 0; k < 64; k++) off[k + 1] += off[k];
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        for (uint32_t q = 0; q < (1u << 20); q++) list[fill[draw_key(q)]++] = q;
        HIP_TRY(hipMemcpy(b.keyq, list.data(), list.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b.key_off, off.data(), 65 * 4, hipMemcpyHostToDevice));
        return WHF_OK;
    }
    solve_layout(&b, static_cast<uint8_t *>(c->d_solve));
    return WHF_OK;
}

// ------------------------------------------------------------------------------------------------ device side

__device__ __forceinline__ int val32(int32_t v) { return v == 0 ? 40 : __builtin_ctz((uint32_t)v); }

__device__ __forceinline__ uint32_t h_bits(uint64_t mul, uint64_t add, uint64_t t) {
    return (uint32_t)(((mul * t + add) & kM48) >> 16);
}

// The swamp hut of a region seed R: two nextInt(24) with Java's rejection (rcHutExact, room_core.h:86-103).
__device__ __forceinline__ bool hut_at(uint64_t R, int ox, int oz) {
    uint64_t s = (R ^ kLA) & kM48;
    int v[2];
    for (int k = 0; k < 2; k++) {
        for (;;) {
            s = (s * kLA + kLC) & kM48;
            const int32_t bits = (int32_t)(s >> 17);
            const int32_t val = bits % 24;
            if ((int32_t)((uint32_t)bits - (uint32_t)val + 23u) >= 0) { v[k] = val; break; }
        }
    }
    return v[0] == ox && v[1] == oz;
}

struct TaskK {
    uint64_t X, creg;
    int32_t cx, cz;
    int ox, oz, m;
    uint32_t idx;
};

// A full 48-bit state: the exact carver equation, then the exact hut.
__device__ __forceinline__ void leaf(const TaskK &T, uint64_t t, SvHit *__restrict__ hits,
                                     uint32_t *__restrict__ hit_ctr, uint32_t hit_cap) {
    const uint64_t s1 = (kLA * t + kLC) & kM48;
    const uint64_t s2 = (kLA * s1 + kLC) & kM48;
    const uint64_t s3 = (kLA * s2 + kLC) & kM48;
    const uint64_t s4 = (kLA * s3 + kLC) & kM48;
    const uint64_t a = ((uint64_t)(uint32_t)(s1 >> 16) << 32) + (uint64_t)(int64_t)(int32_t)(uint32_t)(s2 >> 16);
    const uint64_t b = ((uint64_t)(uint32_t)(s3 >> 16) << 32) + (uint64_t)(int64_t)(int32_t)(uint32_t)(s4 >> 16);
    const uint64_t S = (t ^ kLA) & kM48;
    const uint64_t X = ((uint64_t)((int64_t)T.cx * (int64_t)a) ^ (uint64_t)((int64_t)T.cz * (int64_t)b) ^ S) & kM48;
    if (X != T.X) return;
    if (!hut_at((S + T.creg) & kM48, T.ox, T.oz)) return;
    const uint32_t pos = atomicAdd(hit_ctr, 1u);
    if (pos < hit_cap) {
        SvHit h;
        h.s48 = S;
        h.cx = T.cx;
        h.cz = T.cz;
        h.idx = T.idx;
        h.pad = 0;
        hits[pos] = h;
    }
}

// A full 48-bit state t = th:tlo whose carver equation is zero in its low 32 bits (the lift makes it so): the bits
// 32..47 from the draws' halves, then, for the 2^-16 that pass, the exact checks. With a = h1 * 2^32 + sext(h2),
// bits 32..47 of cx * a are mulhi(cx, h2) + cx * (h1 mod 2^16); h2, h4 are bits 16..47 of the second and fourth
// states (48-bit products), h1, h3 mod 2^16 bits 16..31 of the first and third (32-bit products).
__device__ __forceinline__ uint32_t bits16_47(uint64_t mul, uint64_t add, uint32_t tlo, uint32_t th) {
    const uint64_t p = (uint64_t)tlo * (uint32_t)mul + add;   // wraps modulo 2^64, exact modulo 2^48
    const uint32_t hi = (uint32_t)(p >> 32) + th * (uint32_t)mul + tlo * (uint32_t)(mul >> 32);
    return (hi << 16) | ((uint32_t)p >> 16);
}

constexpr uint64_t kLA3 = (kLA2 * kLA) & kM48;
constexpr uint64_t kLK3 = (kLC * (kLA * kLA + kLA + 1)) & kM48;

__device__ __forceinline__ void leaf_fast(const TaskK &T, uint32_t tlo, uint32_t th, SvHit *__restrict__ hits,
                                          uint32_t *__restrict__ hit_ctr, uint32_t hit_cap) {
    const int32_t h2 = (int32_t)bits16_47(kLA2, kLK2, tlo, th), h4 = (int32_t)bits16_47(kLA4, kLK4, tlo, th);
    const int32_t h1lo = (int32_t)((tlo * (uint32_t)kLA + (uint32_t)kLC) >> 16);
    const int32_t h3lo = (int32_t)((tlo * (uint32_t)kLA3 + (uint32_t)kLK3) >> 16);
    const uint32_t ha = (uint32_t)(__mulhi(T.cx, h2) + T.cx * h1lo);
    const uint32_t hb = (uint32_t)(__mulhi(T.cz, h4) + T.cz * h3lo);
    const uint32_t e = ha ^ hb ^ th ^ (uint32_t)(kLA >> 32) ^ (uint32_t)(T.X >> 32);
    if (e & 0xFFFFu) return;
    leaf(T, ((uint64_t)th << 32) | tlo, hits, hit_ctr, hit_cap);
}

// Lift where the valuations differ (m = M): one choice per level, then the M free top bits. The running values are
// SX = (S xor X) mod 2^32, P = cx * h2 and Q = cz * h4 mod 2^32; setting t's bit j adds cx * a2 * 2^(j-16) to P
// and cz * a4 * 2^(j-16) to Q (h2 grows by a2 * 2^(j-16): the LCG is affine and j >= 16) and flips S's bit j.
// Returns the leaves checked.
template <int M>
__device__ __forceinline__ uint32_t lift_unique_m(const TaskK &T, uint32_t t20, SvHit *__restrict__ hits,
                                                  uint32_t *__restrict__ hit_ctr, uint32_t hit_cap) {
    const uint32_t cx = (uint32_t)T.cx, cz = (uint32_t)T.cz, x32 = (uint32_t)T.X;
    const uint32_t A2 = cx * (uint32_t)kLA2, A4 = cz * (uint32_t)kLA4;
    uint32_t SX = t20 ^ (uint32_t)kLA ^ x32;
    uint32_t P = cx * h_bits(kLA2, kLK2, t20);
    uint32_t Q = cz * h_bits(kLA4, kLK4, t20);
    uint32_t th = 0;   // t bits 32..47
#pragma unroll
    for (int j = 20; j <= 47 - M; j++) {
        const uint32_t bit = ((SX ^ P ^ Q) >> (j + M - 16)) & 1u;
        const uint32_t msk = 0u - bit;
        P += (A2 << (j - 16)) & msk;
        Q += (A4 << (j - 16)) & msk;
        if (j < 32) SX ^= (1u << j) & msk;
        else th |= bit << (j - 32);
    }
    const uint32_t tlo = t20 | ((SX ^ x32 ^ (uint32_t)kLA) & 0xFFF00000u);
#pragma unroll
    for (uint32_t tl = 0; tl < (1u << M); tl++) leaf_fast(T, tlo, th | (tl << (16 - M)), hits, hit_ctr, hit_cap);
    return 1u << M;
}

__device__ __forceinline__ uint32_t lift_unique(const TaskK &T, uint32_t t20, SvHit *__restrict__ hits,
                                                uint32_t *__restrict__ hit_ctr, uint32_t hit_cap) {
    switch (T.m) {   // the same for the whole wave (one task per wave)
    case 0: return lift_unique_m<0>(T, t20, hits, hit_ctr, hit_cap);
    case 1: return lift_unique_m<1>(T, t20, hits, hit_ctr, hit_cap);
    case 2: return lift_unique_m<2>(T, t20, hits, hit_ctr, hit_cap);
    case 3: return lift_unique_m<3>(T, t20, hits, hit_ctr, hit_cap);
    default: return lift_unique_m<4>(T, t20, hits, hit_ctr, hit_cap);
    }
}

// Lift where the valuations are equal: a level's bit does not decide the equation bit it reaches, so both values
// survive wherever that bit is zero. Depth-first, the pending 1-branches in a bit stack.
__device__ __forceinline__ uint32_t lift_branch(const TaskK &T, uint32_t t20, SvHit *__restrict__ hits,
                                                uint32_t *__restrict__ hit_ctr, uint32_t hit_cap,
                                                unsigned long long *__restrict__ cnt) {
    const uint32_t cx = (uint32_t)T.cx, cz = (uint32_t)T.cz, x32 = (uint32_t)T.X;
    const int m = T.m;
    const uint32_t A2 = cx * (uint32_t)kLA2, A4 = cz * (uint32_t)kLA4;
    uint64_t t = t20;
    uint32_t S = t20 ^ (uint32_t)kLA;
    uint32_t P = cx * h_bits(kLA2, kLK2, t);
    uint32_t Q = cz * h_bits(kLA4, kLK4, t);
    uint32_t pend = 0, leaves = 0, nodes = 0;
    int j = 20;
    for (;;) {
        if (++nodes > kNodeBudget) {
            atomicAdd(&cnt[SV_BUDGET], 1ull);
            break;
        }
        bool dead = false;
        if (j == 48) {
            leaf_fast(T, (uint32_t)t, (uint32_t)(t >> 32), hits, hit_ctr, hit_cap);
            leaves++;
            dead = true;
        } else {
            const int k = j + m - 16;
            const uint32_t dP = A2 << (j - 16), dQ = A4 << (j - 16), dS = j < 32 ? (1u << j) : 0u;
            const uint32_t E0 = S ^ P ^ Q ^ x32;
            const uint32_t E1 = (S ^ dS) ^ (P + dP) ^ (Q + dQ) ^ x32;
            const bool ok0 = k > 31 || !((E0 >> k) & 1u);
            const bool ok1 = k > 31 || !((E1 >> k) & 1u);
            if (ok0) {
                if (ok1) pend |= 1u << (j - 20);
                j++;
            } else if (ok1) {
                t |= 1ull << j;
                S ^= dS;
                P += dP;
                Q += dQ;
                j++;
            } else {
                dead = true;
            }
        }
        if (dead) {
            if (!pend) break;
            const int d = 31 - __builtin_clz(pend);
            pend &= ~(1u << d);
            j = d + 20;
            t = (t & ((1ull << j) - 1ull)) | (1ull << j);
            S = (uint32_t)t ^ (uint32_t)kLA;
            P = cx * h_bits(kLA2, kLK2, t);
            Q = cz * h_bits(kLA4, kLK4, t);
            j++;
        }
    }
    return leaves;
}

// One wave per task: lanes take the task's pool entries 64 at a time.
__device__ __forceinline__ void solve_task(const TaskK &T, const uint32_t *__restrict__ list, uint32_t n, bool eqv,
                                           SvHit *__restrict__ hits, uint32_t *__restrict__ hit_ctr,
                                           uint32_t hit_cap, uint32_t &leaves, unsigned long long *__restrict__ cnt) {
    for (uint32_t i = lane_id(); i < n; i += kWave) {
        const uint32_t t20 = list[i];
        leaves += eqv ? lift_branch(T, t20, hits, hit_ctr, hit_cap, cnt) : lift_unique(T, t20, hits, hit_ctr, hit_cap);
    }
}

__device__ __forceinline__ uint64_t creg_of(int32_t rx, int32_t rz) {
    return (kRegC0 + (uint64_t)(int64_t)rx * kRegC1 + (uint64_t)(int64_t)rz * kRegC2) & kM48;
}

// The regions of each group: thread (g, i) takes the i-th rx of the x regime and solves Creg(rx, rz) = c20 (mod
// 2^20) for rz with the inverse of C2 (taskspace.h:27-37).
__global__ __launch_bounds__(kSvWG) void k_sv_regions(const uint32_t *__restrict__ c20s, uint32_t n_groups,
                                                      int32_t xlo0, uint32_t xn0, int32_t xlo1, uint32_t xn1,
                                                      int32_t zlo0, int32_t zhi0, int32_t zlo1, int32_t zhi1,
                                                      uint32_t inv_c2, SvRegion *__restrict__ out,
                                                      uint32_t *__restrict__ n_out) {
    const uint32_t nrx = xn0 + xn1;
    const uint64_t i = (uint64_t)blockIdx.x * kSvWG + threadIdx.x;
    bool keep = false;
    SvRegion r{0, 0, 0, 0};
    if (i < (uint64_t)n_groups * nrx) {
        const uint32_t g = (uint32_t)(i / nrx), k = (uint32_t)(i - (uint64_t)g * nrx);
        const int32_t rx = k < xn0 ? xlo0 + (int32_t)k : xlo1 + (int32_t)(k - xn0);
        const uint32_t rem = (c20s[g] - (uint32_t)((kRegC0 + (uint64_t)(int64_t)rx * kRegC1) & 0xFFFFFu)) & 0xFFFFFu;
        const uint32_t res = (rem * inv_c2) & 0xFFFFFu;
        int32_t rz = 0;
        bool ok = true;
        if (res <= (uint32_t)kRadius) rz = (int32_t)res;
        else if (res >= (1u << 20) - (uint32_t)kRadius) rz = (int32_t)res - (1 << 20);
        else ok = false;
        if (ok && ((rz >= zlo0 && rz <= zhi0) || (rz >= zlo1 && rz <= zhi1))) {
            keep = true;
            r = SvRegion{rx, rz, g, 0};
        }
    }
    const uint32_t pos = wave_append(keep, n_out);
    if (keep) out[pos] = r;
}

// The pool of one (group, class), slot = group x n_kappa + class; with n_kappa = 0 one pool per chunk, slot =
// chunk, with the chunk's own residue and class: t20 of the key's residues whose equation bits 0 .. m + 3 are zero.
__global__ __launch_bounds__(kSvWG) void k_sv_pool(const uint32_t *__restrict__ c20s, const uint16_t *__restrict__ kappa,
                                                   uint32_t n_kappa, const uint32_t *__restrict__ keyq,
                                                   const uint32_t *__restrict__ key_off, uint32_t x32,
                                                   uint32_t *__restrict__ pool, uint32_t *__restrict__ pool_n,
                                                   unsigned long long *__restrict__ cnt) {
    __shared__ uint32_t s_n;
    const uint32_t slot = blockIdx.x;
    const uint32_t g = n_kappa ? slot / n_kappa : slot, kp = kappa[n_kappa ? slot - g * n_kappa : slot];
    const int m = kp >> 8;
    const uint32_t cxk = ((kp >> 4) & 15u) << m, czk = (kp & 15u) << m;   // cx, cz modulo 2^(m+4)
    const uint32_t key = ((cxk & 7u) << 3) | (czk & 7u);
    const uint32_t c20 = c20s[g];
    const uint32_t mask = (1u << (m + 4)) - 1u;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const uint32_t lo = key_off[key], hi = key_off[key + 1];
    uint32_t *dst = pool + (size_t)slot * kPoolCap;
    for (uint32_t i0 = lo; i0 < hi; i0 += kSvWG) {
        const uint32_t i = i0 + threadIdx.x;
        bool keep = false;
        uint32_t t20 = 0;
        if (i < hi) {
            t20 = ((keyq[i] - c20) & 0xFFFFFu) ^ ((uint32_t)kLA & 0xFFFFFu);
            const uint32_t e = (t20 ^ (uint32_t)kLA) ^ (cxk * h_bits(kLA2, kLK2, t20)) ^
                               (czk * h_bits(kLA4, kLK4, t20)) ^ x32;
            keep = (e & mask) == 0;
        }
        const uint32_t pos = wave_append(keep, &s_n);
        if (keep && pos < kPoolCap) dst[pos] = t20;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        pool_n[slot] = s_n;
        if (s_n > kPoolCap) atomicAdd(&cnt[SV_POOL_OVER], 1ull);
    }
}

__device__ __forceinline__ void add_counts(unsigned long long *s_acc, uint32_t cand, uint32_t leaves, uint32_t tasks) {
    if (cand) atomicAdd(&s_acc[SV_CAND], (unsigned long long)cand);
    if (leaves) atomicAdd(&s_acc[SV_LEAVES], (unsigned long long)leaves);
    if (tasks) atomicAdd(&s_acc[SV_TASKS], (unsigned long long)tasks);
}

// Tasks of a batch of groups: (region, offset) = (task / nJ, task % nJ), waves striding over them.
__global__ __launch_bounds__(kSvWG) void k_sv_lift_groups(const SvRegion *__restrict__ regions,
                                                          const uint32_t *__restrict__ n_regions,
                                                          const uint8_t *__restrict__ offs, uint32_t nJ,
                                                          const int16_t *__restrict__ kap_idx, uint32_t n_kappa,
                                                          const uint32_t *__restrict__ pool,
                                                          const uint32_t *__restrict__ pool_n, uint64_t X,
                                                          SvHit *__restrict__ hits, uint32_t *__restrict__ hit_ctr,
                                                          uint32_t hit_cap, unsigned long long *__restrict__ cnt) {
    __shared__ unsigned long long s_acc[3];
    __shared__ uint8_t s_off[2 * 288];
    if (threadIdx.x < 3) s_acc[threadIdx.x] = 0;
    for (uint32_t i = threadIdx.x; i < 2 * nJ; i += kSvWG) s_off[i] = offs[i];
    __syncthreads();
    const uint64_t n_tasks = (uint64_t)(*n_regions) * nJ;
    const uint32_t waves = kSvWG / kWave, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    uint32_t cand = 0, leaves = 0, tasks = 0;
    for (uint64_t task = (uint64_t)blockIdx.x * waves + wave; task < n_tasks; task += (uint64_t)gridDim.x * waves) {
        const uint32_t r = (uint32_t)(task / nJ), j = (uint32_t)(task - (uint64_t)r * nJ);
        const SvRegion R = regions[r];
        TaskK T;
        T.ox = __builtin_amdgcn_readfirstlane(s_off[j]);
        T.oz = __builtin_amdgcn_readfirstlane(s_off[nJ + j]);
        T.cx = 32 * R.rx + T.ox;
        T.cz = 32 * R.rz + T.oz;
        const int vx = val32(T.cx), vz = val32(T.cz);
        T.m = vx < vz ? vx : vz;
        T.X = X;
        T.creg = creg_of(R.rx, R.rz);
        T.idx = R.g;
        const uint32_t cls = ((uint32_t)T.m << 8) | ((((uint32_t)T.cx >> T.m) & 15u) << 4) | (((uint32_t)T.cz >> T.m) & 15u);
        const int16_t ki = kap_idx[cls];
        if (ki < 0) continue;   // not a class of this tier (cannot happen for the tier's own offsets)
        const uint32_t slot = R.g * n_kappa + (uint32_t)ki;
        const uint32_t n = min(pool_n[slot], kPoolCap);
        solve_task(T, pool + (size_t)slot * kPoolCap, n, vx == vz, hits, hit_ctr, hit_cap, leaves, cnt);
        if (lane_id() == 0) {
            cand += n;
            tasks++;
        }
    }
    add_counts(s_acc, cand, leaves, tasks);
    __syncthreads();
    if (threadIdx.x < 3 && s_acc[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], s_acc[threadIdx.x]);
}

// Given chunks: task i uses pool slot i.
__global__ __launch_bounds__(kSvWG) void k_sv_lift_chunks(const int32_t *__restrict__ xz, uint32_t n, uint32_t idx0,
                                                          const uint32_t *__restrict__ pool,
                                                          const uint32_t *__restrict__ pool_n, uint64_t X,
                                                          SvHit *__restrict__ hits, uint32_t *__restrict__ hit_ctr,
                                                          uint32_t hit_cap, unsigned long long *__restrict__ cnt) {
    __shared__ unsigned long long s_acc[3];
    if (threadIdx.x < 3) s_acc[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t waves = kSvWG / kWave, wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    uint32_t cand = 0, leaves = 0, tasks = 0;
    for (uint32_t task = blockIdx.x * waves + wave; task < n; task += gridDim.x * waves) {
        TaskK T;
        T.cx = xz[task];
        T.cz = xz[kMaxChunks + task];
        T.ox = T.cx & 31;
        T.oz = T.cz & 31;
        const int vx = val32(T.cx), vz = val32(T.cz);
        T.m = vx < vz ? vx : vz;
        T.X = X;
        T.creg = creg_of(T.cx >> 5, T.cz >> 5);
        T.idx = idx0 + task;
        const uint32_t nn = min(pool_n[task], kPoolCap);
        solve_task(T, pool + (size_t)task * kPoolCap, nn, vx == vz, hits, hit_ctr, hit_cap, leaves, cnt);
        if (lane_id() == 0) {
            cand += nn;
            tasks++;
        }
    }
    add_counts(s_acc, cand, leaves, tasks);
    __syncthreads();
    if (threadIdx.x < 3 && s_acc[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], s_acc[threadIdx.x]);
}

__global__ void k_sv_zero(uint32_t *a, uint32_t *b, unsigned long long *cnt) {
    if (threadIdx.x == 0) {
        *a = 0;
        *b = 0;
    }
    if (threadIdx.x < SV_N) cnt[threadIdx.x] = 0;
}

// ------------------------------------------------------------------------------------------------ host side

uint32_t perm_group(uint32_t p) {   // taskspace.h:24
    uint32_t x = (p * 0x9E3779B1u) & 0xFFFFFu;
    x ^= x >> 9;
    x = (x * 0x85EBCA6Bu) & 0xFFFFFu;
    x ^= x >> 7;
    return x & 0xFFFFFu;
}

struct Axis {
    int n;
    int32_t lo[2], hi[2];
};

Axis axis_of(int regime) {   // taskspace.h:13-19
    Axis a{};
    if (regime == 0) { a.n = 1; a.lo[0] = -16384; a.hi[0] = 16383; a.lo[1] = 1; a.hi[1] = 0; }
    else if (regime == 1) { a.n = 2; a.lo[0] = -32768; a.hi[0] = -16385; a.lo[1] = 16384; a.hi[1] = 32767; }
    else { a.n = 2; a.lo[0] = -kRadius; a.hi[0] = -32769; a.lo[1] = 32768; a.hi[1] = kRadius; }
    return a;
}

int valuation(int v) {   // of an offset; 99 for 0 (solver_fused.cuh:24)
    if (v == 0) return 99;
    int k = 0;
    while (!(v & 1)) { v >>= 1; k++; }
    return k;
}

int tier_of(int a, int b) {   // solver_fused.cuh:27-31
    const int va = valuation(a), vb = valuation(b), m = std::min(va, vb);
    if (va != vb) return m == 0 ? 1 : 2;
    return va < 99 ? 3 : 0;
}

// The tier's offsets (a-major, solver_fused.cuh:25-33) and its low-bit classes: (m, (cx >> m) mod 16,
// (cz >> m) mod 16) for every m its chunks can have (tier 1: 0; tier 2: 1..4; tier 3: 0..4) and every residue
// pair whose parities fit (the coordinate of valuation m odd after the shift; for tier 3 both).
struct TierPlan {
    std::vector<uint8_t> offs;   // ox[nJ] then oz[nJ]
    int nJ = 0;
    std::vector<uint16_t> kappa;
    std::vector<int16_t> kap_idx;
};

const TierPlan &tier_plan(int tier) {
    static TierPlan plans[4];
    static std::once_flag once[4];
    std::call_once(once[tier], [tier] {
        TierPlan &p = plans[tier];
        std::vector<uint8_t> ox, oz;
        for (int a = 0; a < 24; a++)
            for (int b = 0; b < 24; b++)
                if (tier_of(a, b) == tier) { ox.push_back((uint8_t)a); oz.push_back((uint8_t)b); }
        p.nJ = (int)ox.size();
        p.offs = ox;
        p.offs.insert(p.offs.end(), oz.begin(), oz.end());
        p.kap_idx.assign(1280, -1);
        const int mlo = tier == 2 ? 1 : 0, mhi = tier == 1 ? 0 : 4;
        for (int m = mlo; m <= mhi; m++)
            for (int xr = 0; xr < 16; xr++)
                for (int zr = 0; zr < 16; zr++) {
                    const bool xo = xr & 1, zo = zr & 1;
                    if (tier == 3 ? !(xo && zo) : (xo == zo)) continue;
                    p.kap_idx[m * 256 + xr * 16 + zr] = (int16_t)p.kappa.size();
                    p.kappa.push_back((uint16_t)(m << 8 | xr << 4 | zr));
                }
    });
    return plans[tier];
}

uint32_t inv20(uint64_t v) {   // v odd: inverse modulo 2^20 by Newton's iteration
    uint64_t x = v;
    for (int i = 0; i < 6; i++) x *= 2 - v * x;
    return (uint32_t)(x & 0xFFFFFu);
}

struct Collected {
    std::vector<SvHit> hits;
    whf_solve_stats st{};
};

// Copies the launch's hits and counters back and adds them to `col`; idx_base maps the launch-local index.
int harvest(whf_ctx *c, SolveBufs &b, Collected &col, uint32_t idx_base, uint32_t cap) {
    uint32_t nh = 0;
    unsigned long long cnt[SV_N];
    HIP_TRY(hipMemcpyAsync(&nh, b.hit_ctr, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(cnt, b.cnt, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (cnt[SV_BUDGET]) {
        std::snprintf(g_last_error, sizeof g_last_error, "solver: a branching search passed %u nodes", kNodeBudget);
        return WHF_EOVERFLOW;
    }
    if (cnt[SV_POOL_OVER]) {
        std::snprintf(g_last_error, sizeof g_last_error, "solver: a candidate pool exceeded %u entries", kPoolCap);
        return WHF_EOVERFLOW;
    }
    if (nh > cap) {
        std::snprintf(g_last_error, sizeof g_last_error, "solver: %u hits in one launch (more than %u)", nh, cap);
        return WHF_EOVERFLOW;
    }
    const size_t base = col.hits.size();
    col.hits.resize(base + nh);
    if (nh) HIP_TRY(hipMemcpy(col.hits.data() + base, b.hits, (size_t)nh * sizeof(SvHit), hipMemcpyDeviceToHost));
    for (size_t i = base; i < col.hits.size(); i++) col.hits[i].idx += idx_base;
    col.st.candidates += cnt[SV_CAND];
    col.st.leaves += cnt[SV_LEAVES];
    col.st.tasks += cnt[SV_TASKS];
    return WHF_OK;
}

int deliver(Collected &col, const std::vector<uint64_t> &layout_of_idx, whf_solve_sink *sink,
            whf_solve_stats *stats) {
    std::sort(col.hits.begin(), col.hits.end(), [](const SvHit &a, const SvHit &b) {
        if (a.idx != b.idx) return a.idx < b.idx;
        if (a.s48 != b.s48) return a.s48 < b.s48;
        if (a.cx != b.cx) return a.cx < b.cx;
        return a.cz < b.cz;
    });
    col.st.hits = col.hits.size();
    sink->count = col.hits.size();
    const uint64_t n = std::min<uint64_t>(sink->capacity, col.hits.size());
    for (uint64_t i = 0; i < n; i++) {
        const SvHit &h = col.hits[i];
        whf_solve_hit o;
        o.s48 = h.s48;
        o.layout = layout_of_idx[h.idx];
        o.cx = h.cx;
        o.cz = h.cz;
        o.bundle = h.idx;
        o.reserved = 0;
        sink->hits[i] = o;
    }
    if (stats) *stats = col.st;
    return col.hits.size() > sink->capacity ? WHF_EOVERFLOW : WHF_OK;
}

}  // namespace

extern "C" {

int whf_solve(whf_ctx *c, const whf_solve_bundle *bundles, uint64_t n_bundles, uint64_t first_group,
              uint64_t n_groups, whf_solve_sink *sink, whf_solve_stats *stats) {
    if (!c || !bundles || !sink || (!sink->hits && sink->capacity) || n_bundles == 0 || n_bundles > (1u << 20))
        return WHF_EINVAL;
    if (first_group + n_groups > (n_bundles << 20) || first_group + n_groups < first_group) {
        std::snprintf(g_last_error, sizeof g_last_error, "group range outside the %llu bundles x 2^20 groups",
                      (unsigned long long)n_bundles);
        return WHF_EINVAL;
    }
    for (uint64_t i = 0; i < n_bundles; i++) {
        const whf_solve_bundle &B = bundles[i];
        if (B.layout > kM48 || B.regime_x < 0 || B.regime_x > 2 || B.regime_z < 0 || B.regime_z > 2 || B.tier < 1 ||
            B.tier > 3)
            return WHF_EINVAL;
    }
    HIP_TRY(hipSetDevice(c->device));
    SolveBufs b;
    if (int rc = solve_init(c, b)) return rc;
    Collected col;
    const uint32_t inv_c2 = inv20(kRegC2);
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    // hits carry the bundle index (groups of one launch are in one bundle)
    std::vector<uint64_t> layout_of(n_bundles);
    for (uint64_t i = 0; i < n_bundles; i++) layout_of[i] = bundles[i].layout;
    uint64_t gi = first_group;
    const uint64_t g_end = first_group + n_groups;
    int last_tier = 0;
    std::vector<uint32_t> c20;
    while (gi < g_end) {
        const uint64_t bi = gi >> 20;
        const whf_solve_bundle &B = bundles[bi];
        const TierPlan &tp = tier_plan(B.tier);
        const Axis ax = axis_of(B.regime_x), az = axis_of(B.regime_z);
        const uint32_t xn0 = (uint32_t)(ax.hi[0] - ax.lo[0] + 1), xn1 = ax.n > 1 ? (uint32_t)(ax.hi[1] - ax.lo[1] + 1) : 0;
        const uint32_t nk = (uint32_t)tp.kappa.size();
        // groups per launch: as many as the pools allow, and about 2^23 tasks
        const double zfrac = (double)((az.hi[0] - az.lo[0] + 1) + (az.n > 1 ? az.hi[1] - az.lo[1] + 1 : 0)) /
                             (double)(1u << 20);
        const double tasks_per_group = std::max(1.0, (double)(xn0 + xn1) * zfrac * tp.nJ);
        uint64_t gpb = (uint64_t)std::max(1.0, std::min<double>(kMaxGroups, (double)(1u << 23) / tasks_per_group));
        gpb = std::min<uint64_t>(gpb, kMaxSlots / nk);
        const uint64_t batch_end = std::min<uint64_t>(std::min<uint64_t>(g_end, (bi + 1) << 20), gi + gpb);
        const uint32_t nb = (uint32_t)(batch_end - gi);
        c20.resize(nb);
        for (uint32_t k = 0; k < nb; k++) c20[k] = perm_group((uint32_t)((gi + k) & 0xFFFFFu));
        HIP_TRY(hipMemcpyAsync(b.c20, c20.data(), (size_t)nb * 4, hipMemcpyHostToDevice, c->stream));
        if (B.tier != last_tier) {
            HIP_TRY(hipMemcpyAsync(b.kappa, tp.kappa.data(), tp.kappa.size() * 2, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(b.kap_idx, tp.kap_idx.data(), 1280 * 2, hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(b.offs, tp.offs.data(), tp.offs.size(), hipMemcpyHostToDevice, c->stream));
            last_tier = B.tier;
        }
        hipLaunchKernelGGL(k_sv_zero, dim3(1), dim3(64), 0, c->stream, b.n_regions, b.hit_ctr, b.cnt);
        const uint64_t rthreads = (uint64_t)nb * (xn0 + xn1);
        hipLaunchKernelGGL(k_sv_regions, dim3((uint32_t)((rthreads + kSvWG - 1) / kSvWG)), dim3(kSvWG), 0, c->stream,
                           b.c20, nb, ax.lo[0], xn0, ax.lo[1], xn1, az.lo[0], az.hi[0], az.n > 1 ? az.lo[1] : 1,
                           az.n > 1 ? az.hi[1] : 0, inv_c2, b.regions, b.n_regions);
        hipLaunchKernelGGL(k_sv_pool, dim3(nb * nk), dim3(kSvWG), 0, c->stream, b.c20, b.kappa, nk, b.keyq, b.key_off,
                           (uint32_t)B.layout, b.pool, b.pool_n, b.cnt);
        hipLaunchKernelGGL(k_sv_lift_groups, dim3((uint32_t)c->n_cu * 8), dim3(kSvWG), 0, c->stream, b.regions,
                           b.n_regions, b.offs, (uint32_t)tp.nJ, b.kap_idx, nk, b.pool, b.pool_n, B.layout, b.hits,
                           b.hit_ctr, kHitCapSolve, b.cnt);
        if (int rc = check_launch()) return rc;
        uint32_t nreg = 0;
        HIP_TRY(hipMemcpyAsync(&nreg, b.n_regions, 4, hipMemcpyDeviceToHost, c->stream));
        const size_t before = col.hits.size();
        if (int rc = harvest(c, b, col, 0, kHitCapSolve)) return rc;
        for (size_t i = before; i < col.hits.size(); i++) col.hits[i].idx = (uint32_t)bi;
        col.st.regions += nreg;
        col.st.groups += nb;
        gi = batch_end;
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    col.st.device_ms = ms;
    return deliver(col, layout_of, sink, stats);
}

int whf_solve_chunks(whf_ctx *c, uint64_t layout, const int32_t *cx, const int32_t *cz, uint64_t n,
                     whf_solve_sink *sink, whf_solve_stats *stats) {
    if (!c || (!cx && n) || (!cz && n) || !sink || (!sink->hits && sink->capacity) || n > (1u << 20) ||
        layout > kM48)
        return WHF_EINVAL;
    for (uint64_t i = 0; i < n; i++) {
        const int ox = cx[i] & 31, oz = cz[i] & 31;
        if (ox >= 24 || oz >= 24 || (ox == 0 && oz == 0)) {
            std::snprintf(g_last_error, sizeof g_last_error, "chunk (%d, %d): offset (%d, %d) is in no tier", cx[i],
                          cz[i], ox, oz);
            return WHF_EINVAL;
        }
    }
    HIP_TRY(hipSetDevice(c->device));
    SolveBufs b;
    if (int rc = solve_init(c, b)) return rc;
    Collected col;
    std::vector<uint64_t> layout_of(n ? n : 1, layout);
    HIP_TRY(hipEventRecord(c->ev0, c->stream));
    std::vector<uint32_t> c20;
    std::vector<uint16_t> kap;
    std::vector<int32_t> xz(2 * kMaxChunks);
    for (uint64_t i0 = 0; i0 < n; i0 += kMaxChunks) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(kMaxChunks, n - i0);
        c20.resize(nb);
        kap.resize(nb);
        for (uint32_t k = 0; k < nb; k++) {
            const int32_t x = cx[i0 + k], z = cz[i0 + k];
            const int vx = x == 0 ? 40 : __builtin_ctz((uint32_t)x), vz = z == 0 ? 40 : __builtin_ctz((uint32_t)z);
            const int m = std::min(vx, vz);
            c20[k] = (uint32_t)((kRegC0 + (uint64_t)(int64_t)(x >> 5) * kRegC1 + (uint64_t)(int64_t)(z >> 5) * kRegC2) &
                                0xFFFFFu);
            kap[k] = (uint16_t)(m << 8 | ((((uint32_t)x >> m) & 15u) << 4) | (((uint32_t)z >> m) & 15u));
            xz[k] = x;
            xz[kMaxChunks + k] = z;
        }
        HIP_TRY(hipMemcpyAsync(b.c20, c20.data(), (size_t)nb * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(b.kappa, kap.data(), (size_t)nb * 2, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(b.chunk_xz, xz.data(), xz.size() * 4, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_sv_zero, dim3(1), dim3(64), 0, c->stream, b.n_regions, b.hit_ctr, b.cnt);
        // one pool per chunk: its own residue and class
        hipLaunchKernelGGL(k_sv_pool, dim3(nb), dim3(kSvWG), 0, c->stream, b.c20, b.kappa, 0u, b.keyq, b.key_off,
                           (uint32_t)layout, b.pool, b.pool_n, b.cnt);
        hipLaunchKernelGGL(k_sv_lift_chunks, dim3((uint32_t)c->n_cu * 8), dim3(kSvWG), 0, c->stream, b.chunk_xz, nb,
                           (uint32_t)i0, b.pool, b.pool_n, layout, b.hits, b.hit_ctr, kHitCapSolve, b.cnt);
        if (int rc = check_launch()) return rc;
        if (int rc = harvest(c, b, col, 0, kHitCapSolve)) return rc;
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    col.st.device_ms = ms;
    col.st.regions = n;
    return deliver(col, layout_of, sink, stats);
}

}  // extern "C"
