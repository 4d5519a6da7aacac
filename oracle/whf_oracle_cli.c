/* whf_oracle_cli — command-line front end of the CPU oracle (test infrastructure only).
 *
 *   whf_oracle_cli scan --stream S --count N [--offset O] --rx0 A --rx1 B --rz0 C --rz1 D
 *                       [--max-y M] [--version V] [--threads T]
 *       phase-1 scan of N SplitMix64(S) seeds; prints one JSON line with the funnel, tiers and hits.
 *   whf_oracle_cli probe --seed S --x X --z Z [--max-y M] [--version V]
 *       the 38 probe values (hex bit patterns) of one hut position.
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "whf_oracle.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec * 1e-9;
}

int main(int argc, char **argv) {
    if (argc < 2) { fprintf(stderr, "usage: whf_oracle_cli scan|probe ...\n"); return 2; }
    const char *cmd = argv[1];
    uint64_t stream = 0, offset = 0, seed = 0;
    int64_t count = 1;
    int rx0 = -2, rx1 = 2, rz0 = -2, rz1 = 2, maxy = -50, threads = 1, x = 0, z = 0, vf = 0;
    wo_version_flags("1.20", &vf);
    for (int i = 2; i < argc; i++) {
        const char *a = argv[i], *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!v) { fprintf(stderr, "missing value for %s\n", a); return 2; }
        i++;
        if (!strcmp(a, "--stream")) stream = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--offset")) offset = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--count")) count = strtoll(v, NULL, 10);
        else if (!strcmp(a, "--seed")) seed = strtoull(v, NULL, 10);
        else if (!strcmp(a, "--x")) x = atoi(v);
        else if (!strcmp(a, "--z")) z = atoi(v);
        else if (!strcmp(a, "--rx0")) rx0 = atoi(v);
        else if (!strcmp(a, "--rx1")) rx1 = atoi(v);
        else if (!strcmp(a, "--rz0")) rz0 = atoi(v);
        else if (!strcmp(a, "--rz1")) rz1 = atoi(v);
        else if (!strcmp(a, "--max-y")) maxy = atoi(v);
        else if (!strcmp(a, "--threads")) threads = atoi(v);
        else if (!strcmp(a, "--version")) { if (wo_version_flags(v, &vf)) { fprintf(stderr, "bad --version\n"); return 2; } }
        else { fprintf(stderr, "unknown option %s\n", a); return 2; }
    }
    if (!strcmp(cmd, "probe")) {
        wo_phase1 *p = malloc(sizeof *p);
        wo_phase1_init(p, seed, vf);
        double out[38];
        wo_phase1_probe(p, x, z, maxy, out);
        for (int k = 0; k < 38; k++) {
            uint64_t b;
            memcpy(&b, &out[k], 8);
            printf("%016" PRIx64 "\n", b);
        }
        free(p);
        return 0;
    }
    if (strcmp(cmd, "scan")) { fprintf(stderr, "unknown command %s\n", cmd); return 2; }
    int64_t *seeds = malloc((size_t)(count > 0 ? count : 1) * sizeof *seeds);
    for (int64_t i = 0; i < count; i++) seeds[i] = (int64_t)wo_splitmix64_at(stream, offset + (uint64_t)i);
    int64_t cap = 1 << 16, nh = 0;
    wo_hit *hits = malloc((size_t)cap * sizeof *hits);
    wo_funnel f;
    double t0 = now_s();
    if (wo_scan(seeds, count, rx0, rx1, rz0, rz1, maxy, vf, threads, &f, hits, cap, &nh)) return 1;
    double secs = now_s() - t0;
    printf("{\"seeds\": %" PRId64 ", \"checks\": %" PRIu64 ", \"seconds\": %.6f, \"threads\": %d, \"accepted\": %" PRIu64
           ", \"reached_ladder\": %" PRIu64 ", \"funnel\": [",
           count, f.checks, secs, threads, f.accepted, f.reached_ladder);
    for (int k = 0; k < WO_KIND_N; k++) printf("%s%" PRIu64, k ? ", " : "", f.funnel[k]);
    printf("], \"tiers\": [");
    for (int k = 0; k < 7; k++) printf("%s%" PRIu64, k ? ", " : "", f.tiers[k]);
    printf("], \"hits\": [");
    for (int64_t h = 0; h < nh && h < cap; h++)
        printf("%s[%" PRId64 ", %d, %d, %d, %d]", h ? ", " : "", hits[h].seed, hits[h].rx, hits[h].rz, hits[h].cx,
               hits[h].cz);
    printf("]}\n");
    free(seeds);
    free(hits);
    return 0;
}
