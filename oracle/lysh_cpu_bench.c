/* lysh_cpu_bench — CPU baseline harness around the reference scanner (test infrastructure).
 *
 * Links the reference's own lysh objects (built by oracle/Makefile from engine/tools/lysh/src) and runs
 * its Seed Check as `lysh scan` does for one seed: lysh_scan_rect (search.c:198) with threads = 1, which is
 * the first pass of cmd_scan (main.c:139-216); with --phase2-max-y, every phase-1 survivor also gets lysh's hut
 * verdict through the scanner's own phase-2 hook (search.c:167-192 calling lysh_eval_hut, eval.c:121-226), the
 * work of lysh's grading pass (eval.c:381-477). Seeds come from the sequential SplitMix64 stream (Vigna) and are
 * handed out one at a time to T worker threads, so a T-core host runs T seeds at once (lysh caps threads per seed
 * at the number of rz rows, search.c:205-207).
 *
 *   lysh_cpu_bench --stream S --count N [--offset O] --rx0 A --rx1 B --rz0 C --rz1 D
 *                  [--max-y M] [--version V] [--threads T] [--time-limit SEC] [--phase2-max-y Y [--exactly 1]]
 *                  [--list 1]
 *
 * Seeds are handed out in stream order and a handed-out seed is always finished, so the seeds scanned are exactly
 * the first `seeds` of the stream even when the time limit stops the run.
 *
 * Prints one JSON line: seeds and region checks done, seconds, checks/s, and the summed funnel
 * (reject kinds, main.c:96-100 order) and erosion tier histogram, which must equal the GPU's counters; with
 * phase 2, the huts evaluated and accepted (avg_y <= Y, or == Y with --exactly 1), the rejections by reason
 * (biome, floor, flood: the order of lysh's grading counters, main.c:333-345) and the phase-2 seconds; with
 * --list 1 also every accepted hut as [seed, chunk x, chunk z, avg_y]. `cflags` and `compiler` name the build.
 */
#include <inttypes.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "search.h"
#include "biome.h"
#include "structure.h"

static uint64_t g_stream, g_offset;
static long long g_count;
static int g_rx0, g_rx1, g_rz0, g_rz1, g_maxy, g_threads = 1;
static int g_p2 = 0, g_p2_maxy = 0, g_exactly = 0;
static double g_limit = 0.0;
static lysh_phase1_opts g_opts;
static long long g_next;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static double g_t0;
static int g_list = 0;

#ifndef LYSH_BENCH_CFLAGS
#define LYSH_BENCH_CFLAGS "unknown"
#endif

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec * 1e-9;
}

/* Sequential SplitMix64: the i-th output (i from 0) of the generator seeded with s. */
static uint64_t splitmix64_at(uint64_t s, uint64_t i) {
    uint64_t z = s + (i + 1) * UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

typedef struct {
    long long seeds, scanned, accepted, reached;
    long long reject[LYSH_P1_REJECT_KIND_N];
    long long tier[LYSH_TIER_MAX];
    long long p2_eval, p2_ok;
    long long p2_rej[3];          /* biome, floor, flood */
    double p2_seconds;
    long long (*finds)[4];        /* --list: seed, chunk x, chunk z, avg_y */
    long long n_finds, cap_finds;
} acc_t;

/* phase-2 hook: the worker's accumulator rides on the scanner's per-thread context (search.h:36-45) */
static void p2_ctx_init(void *user, lysh_search_ctx *ctx, int thread_index) {
    (void)thread_index;
    ctx->p2_slot = user;
}
static int p2_hook(void *user, lysh_search_ctx *ctx, int rx, int rz, int hut_x, int hut_z) {
    (void)user; (void)rx; (void)rz;
    acc_t *a = (acc_t *)ctx->p2_slot;
    lysh_hut_result r = lysh_eval_hut(ctx, hut_x, hut_z, g_p2_maxy);
    const int ok = r.ok && (!g_exactly || r.avg_y == g_p2_maxy);
    a->p2_eval++;
    a->p2_ok += ok;
    if (!ok) {
        /* the reject order of lysh's verdict (eval.c:221-224); 'exactly' misses count as floor rejections */
        if (r.reject == LYSH_HUT_REJECT_BIOME) a->p2_rej[0]++;
        else if (r.reject == LYSH_HUT_REJECT_FLOOD) a->p2_rej[2]++;
        else a->p2_rej[1]++;
    } else if (g_list) {
        if (a->n_finds == a->cap_finds) {
            a->cap_finds = a->cap_finds ? 2 * a->cap_finds : 64;
            a->finds = realloc(a->finds, (size_t)a->cap_finds * sizeof *a->finds);
            if (!a->finds) { fprintf(stderr, "out of memory\n"); exit(1); }
        }
        long long *f = a->finds[a->n_finds++];
        f[0] = (long long)ctx->seed;
        f[1] = hut_x >> 4;
        f[2] = hut_z >> 4;
        f[3] = r.avg_y;
    }
    return ok;
}

static void *worker(void *arg) {
    acc_t *a = (acc_t *)arg;
    for (;;) {
        long long i;
        pthread_mutex_lock(&g_mu);
        i = g_next;
        if (i < g_count && (g_limit <= 0.0 || now_s() - g_t0 < g_limit)) g_next++;
        else i = -1;
        pthread_mutex_unlock(&g_mu);
        if (i < 0) break;

        lysh_scan_opts o;
        memset(&o, 0, sizeof o);
        o.seed = splitmix64_at(g_stream, g_offset + (uint64_t)i);
        o.opts = g_opts;
        o.max_height = g_maxy;
        o.salt = LYSH_SWAMP_HUT_SALT;
        o.threads = 1;
        if (g_p2) {
            o.phase2_hook = p2_hook;
            o.phase2_user = a;
            o.phase2_ctx_init = p2_ctx_init;
        }
        lysh_scan_result r;
        if (lysh_scan_rect(&o, g_rx0, g_rx1, g_rz0, g_rz1, &r)) continue;
        a->seeds++;
        a->scanned += r.scanned;
        a->accepted += r.accepted;
        a->reached += r.reached_ladder;
        for (int k = 0; k < LYSH_P1_REJECT_KIND_N; k++) a->reject[k] += r.reject_hist[k];
        for (int k = 0; k < LYSH_TIER_MAX; k++) a->tier[k] += r.tier_hist[k];
        a->p2_seconds += r.p2_seconds;
    }
    return NULL;
}

/* Version name -> phase-1 flags: the game versions lysh_biome_version knows, then main.c:111-125's aliases. */
static int set_version(const char *s) {
    int tr, is118, pre;
    if (!lysh_biome_version(s, &tr, &is118, &pre)) { g_opts.mc_1_18_2 = is118; g_opts.pre_26_2 = pre; g_opts.biome_tree = tr + 1; return 0; }
    if (!strcmp(s, "1.18.x")) { g_opts.mc_1_18_2 = 1; g_opts.pre_26_2 = 1; return 0; }
    if (!strcmp(s, "1.19.x") || !strcmp(s, "1.20.x") || !strcmp(s, "1.21.x~26.1") || !strcmp(s, "1.21.5~26.1")
        || !strcmp(s, "1.19.x~26.1")) { g_opts.mc_1_18_2 = 0; g_opts.pre_26_2 = 1; return 0; }
    if (!strcmp(s, "26.2~26.3")) { g_opts.mc_1_18_2 = 0; g_opts.pre_26_2 = 0; return 0; }
    return 1;
}

int main(int argc, char **argv) {
    g_maxy = -50;
    g_rx0 = g_rz0 = -2; g_rx1 = g_rz1 = 2;
    g_count = 1;
    memset(&g_opts, 0, sizeof g_opts);
    set_version("1.20");
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!v) { fprintf(stderr, "missing value for %s\n", a); return 2; }
        i++;
        if (!strcmp(a, "--stream")) g_stream = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--offset")) g_offset = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--count")) g_count = strtoll(v, NULL, 10);
        else if (!strcmp(a, "--rx0")) g_rx0 = atoi(v);
        else if (!strcmp(a, "--rx1")) g_rx1 = atoi(v);
        else if (!strcmp(a, "--rz0")) g_rz0 = atoi(v);
        else if (!strcmp(a, "--rz1")) g_rz1 = atoi(v);
        else if (!strcmp(a, "--max-y")) g_maxy = atoi(v);
        else if (!strcmp(a, "--threads")) g_threads = atoi(v);
        else if (!strcmp(a, "--time-limit")) g_limit = atof(v);
        else if (!strcmp(a, "--phase2-max-y")) { g_p2 = 1; g_p2_maxy = atoi(v); }
        else if (!strcmp(a, "--exactly")) g_exactly = atoi(v);
        else if (!strcmp(a, "--list")) g_list = atoi(v);
        else if (!strcmp(a, "--version")) { if (set_version(v)) { fprintf(stderr, "bad --version %s\n", v); return 2; } }
        else { fprintf(stderr, "unknown option %s\n", a); return 2; }
    }
    if (g_maxy < -50) g_maxy = -50;   /* phase1_check_height, main.c:135-137 */
    if (g_threads < 1) g_threads = 1;

    acc_t *acc = calloc((size_t)g_threads, sizeof(acc_t));
    pthread_t *th = calloc((size_t)g_threads, sizeof(pthread_t));
    g_t0 = now_s();
    for (int t = 0; t < g_threads; t++) pthread_create(&th[t], NULL, worker, &acc[t]);
    acc_t s;
    memset(&s, 0, sizeof s);
    for (int t = 0; t < g_threads; t++) {
        pthread_join(th[t], NULL);
        s.seeds += acc[t].seeds; s.scanned += acc[t].scanned; s.accepted += acc[t].accepted; s.reached += acc[t].reached;
        for (int k = 0; k < LYSH_P1_REJECT_KIND_N; k++) s.reject[k] += acc[t].reject[k];
        for (int k = 0; k < LYSH_TIER_MAX; k++) s.tier[k] += acc[t].tier[k];
        s.p2_eval += acc[t].p2_eval; s.p2_ok += acc[t].p2_ok; s.p2_seconds += acc[t].p2_seconds;
        for (int k = 0; k < 3; k++) s.p2_rej[k] += acc[t].p2_rej[k];
    }
    double secs = now_s() - g_t0;
    printf("{\"seeds\": %lld, \"checks\": %lld, \"seconds\": %.6f, \"checks_per_s\": %.1f, \"threads\": %d, "
           "\"accepted\": %lld, \"reached_ladder\": %lld, \"funnel\": [",
           s.seeds, s.scanned, secs, secs > 0 ? (double)s.scanned / secs : 0.0, g_threads, s.accepted, s.reached);
    for (int k = 0; k < LYSH_P1_REJECT_KIND_N; k++) printf("%s%lld", k ? ", " : "", s.reject[k]);
    printf("], \"tiers\": [");
    for (int k = 0; k < 7; k++) printf("%s%lld", k ? ", " : "", s.tier[k]);
    printf("], \"phase2\": %d, \"phase2_evaluated\": %lld, \"phase2_accepted\": %lld, "
           "\"phase2_rejected\": [%lld, %lld, %lld], \"phase2_seconds\": %.6f, \"cflags\": \"%s\", "
           "\"compiler\": \"%s\"",
           g_p2, s.p2_eval, s.p2_ok, s.p2_rej[0], s.p2_rej[1], s.p2_rej[2], s.p2_seconds, LYSH_BENCH_CFLAGS,
           __VERSION__);
    if (g_list) {
        printf(", \"finds\": [");
        int first = 1;
        for (int t = 0; t < g_threads; t++) {
            for (long long i = 0; i < acc[t].n_finds; i++) {
                const long long *f = acc[t].finds[i];
                printf("%s[%lld, %lld, %lld, %lld]", first ? "" : ", ", f[0], f[1], f[2], f[3]);
                first = 0;
            }
            free(acc[t].finds);
        }
        printf("]");
    }
    printf("}\n");
    free(acc);
    free(th);
    return 0;
}
