/* whf_oracle_solve.c — CPU oracle of the Deep Dive solver (hut_lookup_v4 --solve-only --solve5).
 *
 * TEST INFRASTRUCTURE ONLY (see whf_oracle.h). A restatement of what the reference's solver lists, written from its
 * definition rather than its method: for a layout X (a mineshaft carver seed) and a chunk task (cx, cz) in region
 * (rx, rz) = (cx >> 5, cz >> 5), every 48-bit structure seed S with
 *   (1) carver(S, cx, cz) == X                    (the chunk's carver seed, room_core.h:55-58; the final check
 *                                                   of solver_fused.cuh:117-122)
 *   (2) the region's hut at offset (cx & 31, cz & 31)  (hutExact, solver_kernels.cuh:78-82)
 *   (3) the first two nextInt(24) draws of the region's Random agree with that offset modulo 8 in their low 20
 *       state bits (hutPrune20, solver_kernels.cuh:21-28). The reference builds its candidate pools from this
 *       (g_hutQ, solver_kernels.cuh:241-255; fusedBuildKernel, solver_fused.cuh:37-69), so it is part of what the
 *       solver outputs. It differs from (2) only where a first draw is rejected (8 in 2^31), and then the
 *       reference misses the seed; the oracle does the same.
 * The task space (bundles x 2^20 groups, the regions of a group, the offsets of a tier) follows taskspace.h:9-43 and
 * solver_fused.cuh:24-35; the run's order (hut_lookup_v4_source.cu:1361-1397) does not matter to the hit set.
 *
 * Method: scan all 2^20 low state bits t20 of t = S ^ 0x5DEECE66D for (3) and for the low m + 4 bits of the
 * carver equation (m = the smaller 2-adic valuation of cx and cz; those bits depend on t20 only), then extend t
 * one bit at a time from bit 20 to bit 47, trying both values and keeping a prefix while the equation's bits that
 * the prefix fixes are zero (bit k of the low 32 depends on t bits <= k - m + 16), recomputing the equation from
 * scratch each time; at full width check (1) and (2) exactly.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "whf_oracle.h"

#define M48 0xFFFFFFFFFFFFull
#define LCG_A 0x5DEECE66Dull
#define LCG_C 0xBull

static const uint64_t kC0 = 14357620ull, kC1 = 341873128712ull, kC2 = 132897987541ull;   /* taskspace.h:10 */
static const int kRadius = 58593;                                                        /* taskspace.h:9 */

/* taskspace.h:24 permGroup: the group index -> the low 20 bits of the region constant. */
uint32_t wo_solve_perm_group(uint32_t p) {
    uint32_t x = (p * 0x9E3779B1u) & 0xFFFFFu;
    x ^= x >> 9;
    x = (x * 0x85EBCA6Bu) & 0xFFFFFu;
    x ^= x >> 7;
    return x & 0xFFFFFu;
}

/* taskspace.h:25 cregOf: the region constant added to the structure seed. */
uint64_t wo_solve_creg(int32_t rx, int32_t rz) {
    return (kC0 + (uint64_t)(int64_t)rx * kC1 + (uint64_t)(int64_t)rz * kC2) & M48;
}

/* taskspace.h:13-19 axisOf: the region coordinates of a regime digit, as up to two closed intervals. */
static int axis_of(int regime, int64_t lo[2], int64_t hi[2]) {
    if (regime == 0) { lo[0] = -16384; hi[0] = 16383; return 1; }
    if (regime == 1) { lo[0] = -32768; hi[0] = -16385; lo[1] = 16384; hi[1] = 32767; return 2; }
    lo[0] = -kRadius; hi[0] = -32769; lo[1] = 32768; hi[1] = kRadius;
    return 2;
}

/* taskspace.h:27-37 regionsOfGroup: the regions whose constant is c20 modulo 2^20, with rx in the x regime and the
 * one rz in [-58593, 58593] that fits (2^20 > 2 x 58593 + 1), kept if it is in the z regime. Brute force over rz,
 * not the modular inverse the reference uses. Returns the count (rx ascending within the x intervals). */
int wo_solve_regions(uint32_t c20, int regime_x, int regime_z, int32_t *rx_out, int32_t *rz_out, int cap) {
    int64_t xl[2], xh[2], zl[2], zh[2];
    const int nx = axis_of(regime_x, xl, xh), nz = axis_of(regime_z, zl, zh);
    /* rz -> low 20 bits of rz * C2, inverted by a table over the whole radius */
    static uint32_t *inv = NULL;   /* [2^20]: rz + kRadius + 1 of the rz with that residue, 0 for none */
    static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    pthread_mutex_lock(&mu);
    if (!inv) {
        uint32_t *t = (uint32_t *)calloc(1u << 20, sizeof *t);
        for (int64_t rz = -kRadius; rz <= kRadius; rz++) t[((uint64_t)rz * kC2) & 0xFFFFFu] = (uint32_t)(rz + kRadius + 1);
        inv = t;
    }
    pthread_mutex_unlock(&mu);
    int n = 0;
    for (int i = 0; i < nx; i++)
        for (int64_t rx = xl[i]; rx <= xh[i]; rx++) {
            const uint32_t need = (c20 - (uint32_t)((kC0 + (uint64_t)rx * kC1) & 0xFFFFFu)) & 0xFFFFFu;
            const uint32_t e = inv[need];
            if (!e) continue;
            const int64_t rz = (int64_t)e - kRadius - 1;
            int in = 0;
            for (int k = 0; k < nz; k++) in |= rz >= zl[k] && rz <= zh[k];
            if (!in) continue;
            if (n < cap) { rx_out[n] = (int32_t)rx; rz_out[n] = (int32_t)rz; }
            n++;
        }
    return n;
}

static int val_offset(int v) {   /* 2-adic valuation of an offset, 99 for 0 (solver_fused.cuh:24) */
    if (v == 0) return 99;
    int k = 0;
    while (!(v & 1)) { v >>= 1; k++; }
    return k;
}

/* solver_fused.cuh:25-33 fusedOffsetList: the chunk offsets (a, b) in [0,24)^2 of a tier, a-major. Tier 1 "odd":
 * different valuations, the smaller 0; tier 2 "u2": different, the smaller >= 1; tier 3 "eq": equal and finite. */
int wo_solve_offsets(int tier, uint8_t *ox, uint8_t *oz) {
    int n = 0;
    for (int a = 0; a < 24; a++)
        for (int b = 0; b < 24; b++) {
            const int va = val_offset(a), vb = val_offset(b), m = va < vb ? va : vb;
            const int t = va != vb ? (m == 0 ? 1 : 2) : (va < 99 ? 3 : 0);
            if (t == tier) { ox[n] = (uint8_t)a; oz[n] = (uint8_t)b; n++; }
        }
    return n;
}

static int val_coord(int32_t c) {   /* 2-adic valuation of a chunk coordinate, 40 for 0 */
    if (c == 0) return 40;
    return __builtin_ctz((uint32_t)c);
}

/* The low 32 bits of the carver equation's left side xor X for a (possibly partial) state t: bit k is zero for
 * every k the prefix decides when t is a prefix of a solution. */
static uint32_t eq_low32(uint64_t t, uint32_t cx, uint32_t cz, uint32_t x32) {
    const uint64_t a2 = (LCG_A * LCG_A) & M48, k2 = (LCG_C * (LCG_A + 1)) & M48;
    const uint64_t a4 = (a2 * a2) & M48, k4 = (LCG_C * (LCG_A * LCG_A * LCG_A + LCG_A * LCG_A + LCG_A + 1)) & M48;
    const uint32_t h2 = (uint32_t)((((a2 * t) + k2) & M48) >> 16);
    const uint32_t h4 = (uint32_t)((((a4 * t) + k4) & M48) >> 16);
    return (uint32_t)(t ^ LCG_A) ^ (cx * h2) ^ (cz * h4) ^ x32;
}

/* The first two draws modulo 8 from the low 20 state bits of the region's Random (solver_kernels.cuh:21-28). */
static int prune20(uint64_t t20, uint32_t creg20, int ox, int oz) {
    const uint32_t q = (uint32_t)(((t20 ^ LCG_A) + creg20) & 0xFFFFFu);
    const uint32_t r0 = q ^ (uint32_t)(LCG_A & 0xFFFFFu);
    const uint32_t r1 = (uint32_t)((r0 * (LCG_A & 0xFFFFFFFFu) + LCG_C) & 0xFFFFFu);
    const uint32_t r2 = (uint32_t)((r1 * (LCG_A & 0xFFFFFFFFu) + LCG_C) & 0xFFFFFu);
    return (int)((r1 >> 17) & 7u) == (ox & 7) && (int)((r2 >> 17) & 7u) == (oz & 7);
}

typedef struct {
    uint64_t X;
    int32_t cx, cz, rx, rz, ox, oz, m;
    uint32_t x32;
    uint64_t *out;
    int cap, n;
    int stop;   /* 48, or a smaller width at which the prefixes are recorded instead */
} task_ctx;

static void lift(task_ctx *tc, uint64_t t, int j) {
    if (j == tc->stop && j < 48) {
        if (tc->n < tc->cap) tc->out[tc->n] = t;
        tc->n++;
        return;
    }
    if (j == 48) {
        const uint64_t S = (t ^ LCG_A) & M48;
        if (wo_ms_carver_seed(S, tc->cx, tc->cz) != tc->X) return;
        int hx, hz;
        wo_hut_chunk(S, tc->rx, tc->rz, 14357620, &hx, &hz);
        if (hx != tc->cx || hz != tc->cz) return;
        if (tc->n < tc->cap) tc->out[tc->n] = S;
        tc->n++;
        return;
    }
    const int k = j + tc->m - 16;   /* the equation bit this level's prefix fixes */
    for (uint64_t b = 0; b < 2; b++) {
        const uint64_t t1 = t | (b << j);
        if (k <= 31) {
            const uint32_t e = eq_low32(t1, (uint32_t)tc->cx, (uint32_t)tc->cz, tc->x32);
            const uint32_t mask = k == 31 ? 0xFFFFFFFFu : ((1u << (k + 1)) - 1u);
            if (e & mask) continue;
        }
        lift(tc, t1, j + 1);
    }
}

static int cmp_u64(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static int solve_to(uint64_t X, int32_t cx, int32_t cz, int stop, uint64_t *out, int cap);

/* Every structure seed of one chunk task, ascending. Returns the count (seeds beyond cap are counted, not
 * stored); -1 for an offset that no tier holds ((0, 0), or >= 24). */
int wo_solve_task(uint64_t X, int32_t cx, int32_t cz, uint64_t *out, int cap) {
    return solve_to(X, cx, cz, 48, out, cap);
}

/* The search stopped at `width` bits (20 <= width < 48): the states t mod 2^width it keeps, ascending, which are
 * the t that pass the hut filter and whose equation bits below width + m - 16 are zero. For checking the search
 * against brute force at small widths. */
int wo_solve_prefixes(uint64_t X, int32_t cx, int32_t cz, int width, uint64_t *out, int cap) {
    if (width < 20 || width >= 48) return -1;
    return solve_to(X, cx, cz, width, out, cap);
}

static int solve_to(uint64_t X, int32_t cx, int32_t cz, int stop, uint64_t *out, int cap) {
    task_ctx tc;
    tc.stop = stop;
    tc.X = X & M48;
    tc.cx = cx;
    tc.cz = cz;
    tc.rx = cx >> 5;
    tc.rz = cz >> 5;
    tc.ox = cx & 31;
    tc.oz = cz & 31;
    if (tc.ox >= 24 || tc.oz >= 24 || (tc.ox == 0 && tc.oz == 0)) return -1;
    const int vx = val_coord(cx), vz = val_coord(cz);
    tc.m = vx < vz ? vx : vz;
    tc.x32 = (uint32_t)tc.X;
    tc.out = out;
    tc.cap = cap;
    tc.n = 0;
    const uint32_t creg20 = (uint32_t)(wo_solve_creg(tc.rx, tc.rz) & 0xFFFFFu);
    const uint32_t low = (1u << (tc.m + 4)) - 1u;
    for (uint64_t t20 = 0; t20 < (1u << 20); t20++) {
        if (!prune20(t20, creg20, tc.ox, tc.oz)) continue;
        if (eq_low32(t20, (uint32_t)cx, (uint32_t)cz, tc.x32) & low) continue;
        lift(&tc, t20, 20);
    }
    if (out && tc.n > 1) qsort(out, (size_t)(tc.n < cap ? tc.n : cap), sizeof *out, cmp_u64);
    return tc.n;
}

/* ---- many tasks on host threads ---- */
typedef struct {
    uint64_t X;
    const int32_t *cx, *cz;
    int n_tasks, per_task_cap;
    uint64_t *seeds;   /* [n_tasks * per_task_cap] */
    int *counts;       /* [n_tasks] */
    int next;
    pthread_mutex_t mu;
} mt_ctx;

static void *mt_worker(void *arg) {
    mt_ctx *m = (mt_ctx *)arg;
    for (;;) {
        pthread_mutex_lock(&m->mu);
        const int i = m->next++;
        pthread_mutex_unlock(&m->mu);
        if (i >= m->n_tasks) return NULL;
        m->counts[i] = wo_solve_task(m->X, m->cx[i], m->cz[i], m->seeds + (size_t)i * m->per_task_cap,
                                     m->per_task_cap);
    }
}

/* wo_solve_task over n chunk tasks of one layout on `threads` host threads: counts[i] and the first per_task_cap
 * seeds of task i at seeds[i * per_task_cap]. Returns the total count. */
int64_t wo_solve_tasks(uint64_t X, const int32_t *cx, const int32_t *cz, int n, int per_task_cap, int threads,
                       uint64_t *seeds, int *counts) {
    mt_ctx m;
    m.X = X;
    m.cx = cx;
    m.cz = cz;
    m.n_tasks = n;
    m.per_task_cap = per_task_cap;
    m.seeds = seeds;
    m.counts = counts;
    m.next = 0;
    pthread_mutex_init(&m.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, mt_worker, &m);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&m.mu);
    int64_t tot = 0;
    for (int i = 0; i < n; i++) tot += counts[i] > 0 ? counts[i] : 0;
    return tot;
}
