/* gen_hut_golden.c — writes golden vectors of the carvers and the hut evaluation from the reference CPU scanner.
 *
 * Built by tests/golden/make_hut_golden.py against oracle/_ref/liblysh.a; prints JSON on stdout, stored as
 * tests/golden/hut_vectors.json. The evaluation context is lysh's own (lysh_search_ctx_init, eval.c:72-110).
 *
 * Sections (data only: inputs and the reference's outputs):
 *   eval   lysh_eval_hut (eval.c:121-226) for hut chunks: orientation, footprint, biome (ok, winner, occupied
 *          height), wet columns, flooded, the height sum and average, ok and the reject reason
 *   post   post-carver heights (lysh_phase2_height, phase2.c:438-456) and pre-carver heights of columns, 60 of
 *          them lowered by the carvers
 *   eval2  lysh_eval_hut for huts placed on chunks with carved columns (tree 3)
 * Hut chunks are phase-1 survivors (1.20 rules, maxHeight 60) of a set of seeds, the reference's known-answer hut,
 * and random chunks; versions 1.20 (tree 3), 1.18.2 (tree 0) and 26.2 (tree 6).
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "biome.h"
#include "eval.h"
#include "phase1.h"
#include "phase2.h"
#include "structure.h"

static uint64_t sm64(uint64_t s, uint64_t i) {
    uint64_t z = s + (i + 1) * UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}
static uint64_t rng_state = UINT64_C(0x5EED5EED5EED5EED);
static uint64_t rnd(void) { return sm64(rng_state++, 0); }
static int rnd_int(int lo, int hi) { return lo + (int)(rnd() % (uint64_t)(hi - lo + 1)); }

typedef struct { uint64_t seed; int hx, hz, ver; } cand_t;
#define MAXC 256
static cand_t C[MAXC];
static int NC, NC0;
static void add(uint64_t seed, int hx, int hz, int ver) {
    if (NC < MAXC) { C[NC].seed = seed; C[NC].hx = hx; C[NC].hz = hz; C[NC].ver = ver; NC++; }
}

static void survivors(uint64_t seed, int r, int cap) {
    lysh_phase1_opts o;
    memset(&o, 0, sizeof o);
    o.pre_26_2 = 1;
    lysh_phase1 *p1 = malloc(sizeof *p1);
    lysh_phase1_init(p1, seed, &o);
    int got = 0;
    for (int rz = -r; rz < r && got < cap; rz++)
        for (int rx = -r; rx < r && got < cap; rx++) {
            lysh_chunkpos c = lysh_swamp_hut_in_region(seed, rx, rz, LYSH_SWAMP_HUT_SALT);
            lysh_phase1_result res;
            lysh_phase1_check_ex(p1, &o, c.chunkX * 16, c.chunkZ * 16, 60, &res, NULL);
            if (res.reject == LYSH_P1_ACCEPT) { add(seed, c.chunkX * 16, c.chunkZ * 16, 0); got++; }
        }
    free(p1);
}

static const char *VER[3] = {"1.20", "1.18.2", "26.2"};

int main(void) {
    add(UINT64_C(1473884015873122059), -15792, -17168, 0);   /* the reference's known answer (hunt.py:86) */
    for (int s = 0; s < 3000 && NC < 40; s++) survivors(sm64(31, (uint64_t)s), 24, 1);
    for (int s = 0; s < 24; s++) add(sm64(32, (uint64_t)s), rnd_int(-4000, 4000) * 16, rnd_int(-4000, 4000) * 16, s % 3);
    for (int i = 1; i < NC; i += 4) C[i].ver = 1 + (i / 4) % 2;

    NC0 = NC;
    printf("{\n\"eval\": [");
    lysh_search_ctx *ctx = malloc(sizeof *ctx);
    int first = 1;
    for (int i = 0; i < NC; i++) {
        lysh_phase1_opts o;
        memset(&o, 0, sizeof o);
        int tree = 0, is118 = 0, pre = 1;
        lysh_biome_version(VER[C[i].ver], &tree, &is118, &pre);
        o.mc_1_18_2 = is118;
        o.pre_26_2 = pre;
        o.biome_tree = tree + 1;
        lysh_search_ctx_init(ctx, C[i].seed, &o, -50);
        lysh_hut_result r = lysh_eval_hut(ctx, C[i].hx, C[i].hz, -50);
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%lld,%d,%d,%d]", first ? "" : ",", C[i].seed, C[i].hx,
               C[i].hz, tree, r.dir, r.size_x, r.size_z, r.biome_ok, r.biome_winner, r.biome_occ_y, r.wet_columns,
               r.flooded, r.sum_h, r.avg_y, r.ok, r.reject);
        first = 0;
        lysh_search_ctx_free(ctx);
    }
    printf("],\n\"post\": [");
    /* columns that the carvers lower, searched for over random chunks, and as many that they leave alone */
    first = 1;
    int lowered = 0, kept = 0;
    for (int t = 0; t < 400 && (lowered < 60 || kept < 30); t++) {
        const uint64_t seed = sm64(33, (uint64_t)t);
        lysh_phase1_opts o;
        memset(&o, 0, sizeof o);
        o.pre_26_2 = 1;
        lysh_search_ctx_init(ctx, seed, &o, -50);
        const int cx = rnd_int(-3000, 3000), cz = rnd_int(-3000, 3000);
        int here = 0;
        for (int k = 0; k < 64 && here < 4; k++) {
            const int x = cx * 16 + (k & 7) * 2, z = cz * 16 + (k >> 3) * 2;
            const int post = lysh_phase2_height(&ctx->p2, x, z);
            const int pre = lysh_phase2_height_pre_carver(&ctx->p2, x, z);
            if (post != pre ? lowered < 60 : (kept < 30 && k == 0)) {
                printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d]", first ? "" : ",", seed, x, z, post, pre);
                first = 0;
                if (post != pre) { lowered++; here++; }
                else kept++;
            }
        }
        if (here) add(seed, cx * 16, cz * 16, 0);   /* a hut on a carved chunk for the evaluation */
        lysh_search_ctx_free(ctx);
    }
    printf("],\n\"eval2\": [");
    first = 1;
    for (int i = 0; i < NC; i++) {
        if (i < NC0) continue;
        lysh_phase1_opts o;
        memset(&o, 0, sizeof o);
        o.pre_26_2 = 1;
        o.biome_tree = 3 + 1;
        lysh_search_ctx_init(ctx, C[i].seed, &o, -50);
        lysh_hut_result r = lysh_eval_hut(ctx, C[i].hx, C[i].hz, -50);
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%lld,%d,%d,%d]", first ? "" : ",", C[i].seed, C[i].hx,
               C[i].hz, 3, r.dir, r.size_x, r.size_z, r.biome_ok, r.biome_winner, r.biome_occ_y, r.wet_columns,
               r.flooded, r.sum_h, r.avg_y, r.ok, r.reject);
        first = 0;
        lysh_search_ctx_free(ctx);
    }
    printf("]\n}\n");
    free(ctx);
    return 0;
}
