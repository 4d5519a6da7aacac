/* gen_biome_golden.c — writes golden vectors of the biome stage from the reference CPU scanner (lysh).
 *
 * Built by tests/golden/make_biome_golden.py against oracle/_ref/liblysh.a. It compiles the reference's
 * engine/tools/lysh/src/biome.c into this program (by #include, from where it lies) to reach the tree search,
 * which that file keeps static. Prints JSON on stdout; the script stores it as tests/golden/biome_vectors.json.
 *
 * Sections (data only: inputs and the reference's outputs):
 *   spline   offset / factor / peaks of lysh_terrain_offset/factor/peak for float (c, e, w)   (spline.c:83-111)
 *   search   tree search winner and leaf for 6-parameter targets, per tree                    (biome.c:34-53)
 *   point    terrain info, temperature, humidity and overworld offset at quart points         (terrain.c:22-44,
 *                                                                                              column_top.c:124)
 *   check    lysh_swamp_hut_biome_ok with a constant height callback: targets, quart y, winner (biome.c:94-156)
 *   heights  the winner at every quart height -17..79 for a hut position and tree
 * Hut positions are phase-1 survivors (lysh_phase1_check_ex, 1.20 rules, maxHeight -50) of a few seeds and
 * random chunks, some far out; a fifth of them use the LARGE_BIOMES noises.
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "biome.c"
#include "column_top.h"
#include "phase1.h"
#include "structure.h"
#include "spline.h"
#include "terrain.h"
#include "terrain_table.h"

static uint64_t sm64(uint64_t s, uint64_t i) {
    uint64_t z = s + (i + 1) * UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

static uint64_t rng_state = UINT64_C(0x2545F4914F6CDD1D);
static uint64_t rnd(void) { return sm64(rng_state++, 0); }
static int rnd_int(int lo, int hi) { return lo + (int)(rnd() % (uint64_t)(hi - lo + 1)); }
static float rnd_float(float lo, float hi) { return lo + (hi - lo) * (float)((rnd() >> 40) * (1.0 / 16777216.0)); }

static uint32_t fbits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static uint64_t dbits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

static int g_height;
static int height_cb(void *user, int x, int z) { (void)user; (void)x; (void)z; return g_height; }

#define MAXPOS 4096
typedef struct { uint64_t seed; int hut_x, hut_z, large; } pos_t;
static pos_t P[MAXPOS];
static int NP;

static void add_pos(uint64_t seed, int hx, int hz, int large) {
    if (NP < MAXPOS) { P[NP].seed = seed; P[NP].hut_x = hx; P[NP].hut_z = hz; P[NP].large = large; NP++; }
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
            lysh_phase1_check_ex(p1, &o, c.chunkX * 16, c.chunkZ * 16, -50, &res, NULL);
            if (res.reject == LYSH_P1_ACCEPT) { add_pos(seed, c.chunkX * 16, c.chunkZ * 16, 0); got++; }
        }
    free(p1);
}

int main(void) {
    printf("{\n");

    /* ---- spline ---- */
    printf("\"spline\": [");
    for (int i = 0; i < 600; i++) {
        float v[3];
        const float lo[3] = {-1.3f, -1.2f, -1.2f}, hi[3] = {1.3f, 1.2f, 1.2f};
        for (int k = 0; k < 3; k++) {
            if (i % 2) v[k] = LYSH_TERRAIN_FLOATS[rnd() % (sizeof LYSH_TERRAIN_FLOATS / sizeof(float))];
            else v[k] = rnd_float(lo[k], hi[k]);
        }
        printf("%s[\"%08x\",\"%08x\",\"%08x\",\"%08x\",\"%08x\",\"%08x\"]", i ? "," : "", fbits(v[0]), fbits(v[1]),
               fbits(v[2]), fbits(lysh_terrain_offset(v[0], v[1], v[2])), fbits(lysh_terrain_factor(v[0], v[1], v[2])),
               fbits(lysh_terrain_peak(v[0], v[1], v[2])));
    }
    printf("],\n");

    /* ---- tree search ---- */
    printf("\"search\": [");
    int first = 1;
    for (int t = 0; t < LYSH_TREE_COUNT; t++) {
        const lysh_gtree *g = &LYSH_TREES[t];
        const int n = g->end[0];
        for (int i = 0; i < 300; i++) {
            int64_t s[6];
            if (i % 2 == 0) {
                for (int k = 0; k < 6; k++) s[k] = rnd_int(k == 4 ? -25000 : -15000, k == 4 ? 25000 : 15000);
            } else {
                int leaf;
                do leaf = rnd_int(0, n - 1); while (g->kids[leaf]);
                for (int k = 0; k < 6; k++) {
                    const int edge = g->range[leaf][2 * k + (int)(rnd() & 1)];
                    s[k] = edge + rnd_int(-1, 1);
                }
            }
            const int leaf = gt_search(g, 0, s, -1);
            printf("%s[%d,%" PRId64 ",%" PRId64 ",%" PRId64 ",%" PRId64 ",%" PRId64 ",%" PRId64 ",%d,%d]",
                   first ? "" : ",", t, s[0], s[1], s[2], s[3], s[4], s[5], g->biome[leaf], leaf);
            first = 0;
        }
    }
    printf("],\n");

    /* ---- hut positions ---- */
    survivors(UINT64_C(1473884015873122059), 40, 40);
    for (int s = 0; s < 24; s++) survivors(sm64(5, (uint64_t)s), 10, 3);
    for (int s = 0; s < 24; s++) {
        const uint64_t seed = sm64(6, (uint64_t)s);
        add_pos(seed, rnd_int(-2000, 2000) * 16, rnd_int(-2000, 2000) * 16, s % 5 == 0);
        add_pos(seed, rnd_int(-1870000, 1870000) * 16, rnd_int(-1870000, 1870000) * 16, s % 5 == 1);
        add_pos(seed, (s & 1 ? 1 : -1) * 29999984 + (s & 2 ? 0 : 16 * rnd_int(-40, 0)), rnd_int(-100, 100) * 16, 0);
    }
    for (int i = 0; i < NP; i++)
        if (i % 5 == 4) P[i].large = 1;

    lysh_terrain *tr = malloc(sizeof *tr);
    lysh_biome *bi = malloc(sizeof *bi);

    /* ---- point ---- */
    printf("\"point\": [");
    for (int i = 0; i < NP; i++) {
        lysh_terrain_init(tr, P[i].seed, P[i].large);
        lysh_biome_init(bi, P[i].seed, P[i].large, 0);
        const int qx = ((P[i].hut_x >> 4) * 16 + 8) >> 2, qz = ((P[i].hut_z >> 4) * 16 + 8) >> 2;
        lysh_terrain_info ti;
        lysh_terrain_info_at(tr, qx, qz, &ti);
        const double te = lysh_dblnoise_sample_y0(&bi->temperature, ti.shifted_x, ti.shifted_z);
        const double hu = lysh_dblnoise_sample_y0(&bi->vegetation, ti.shifted_x, ti.shifted_z);
        const double off = lysh_overworld_offset(tr, qx * 4, qz * 4);
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64
               "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\",\"%016" PRIx64 "\"]",
               i ? "," : "", P[i].seed, qx, qz, P[i].large, dbits(ti.shifted_x), dbits(ti.shifted_z),
               dbits(ti.continentalness), dbits(ti.weirdness), dbits(ti.erosion), dbits(te), dbits(hu), dbits(off));
    }
    printf("],\n");

    /* ---- check ---- */
    static const int HEIGHTS[] = {-65, -64, -61, -60, -57, -52, -40, -1, 0, 3, 62, 63, 64, 70, 100, 200, 319};
    const int NH = (int)(sizeof HEIGHTS / sizeof HEIGHTS[0]);
    printf("\"check\": [");
    first = 1;
    for (int i = 0; i < NP; i++) {
        lysh_terrain_init(tr, P[i].seed, P[i].large);
        for (int k = 0; k < 4; k++) {
            const int tree = k == 0 ? 3 : (k == 1 ? 5 : rnd_int(0, LYSH_TREE_COUNT - 1));
            const int occ = HEIGHTS[rnd_int(0, NH - 1)];
            lysh_biome_init(bi, P[i].seed, P[i].large, tree);
            g_height = occ + 1;
            lysh_biome_result br;
            lysh_swamp_hut_biome_ok(bi, tr, P[i].hut_x, P[i].hut_z, height_cb, NULL, &br);
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,%d,%lld,%lld,%lld,%lld,%lld,%lld,%d,%d,%d,%d]", first ? "" : ",",
                   P[i].seed, P[i].hut_x, P[i].hut_z, P[i].large, tree, occ, br.target.t[0], br.target.t[1],
                   br.target.t[2], br.target.t[3], br.target.t[4], br.target.t[5], br.target.qx, br.target.qy,
                   br.target.qz, br.winner);
            first = 0;
        }
    }
    printf("],\n");

    /* ---- heights ---- */
    printf("\"heights\": [");
    first = 1;
    for (int i = 0; i < NP; i += 3) {
        const int tree = i % LYSH_TREE_COUNT;
        lysh_terrain_init(tr, P[i].seed, P[i].large);
        lysh_biome_init(bi, P[i].seed, P[i].large, tree);
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,[", first ? "" : ",", P[i].seed, P[i].hut_x, P[i].hut_z, P[i].large,
               tree);
        first = 0;
        for (int qy = -17; qy <= 79; qy++) {
            g_height = qy * 4 + 1;
            lysh_biome_result br;
            lysh_swamp_hut_biome_ok(bi, tr, P[i].hut_x, P[i].hut_z, height_cb, NULL, &br);
            printf("%s%d", qy > -17 ? "," : "", br.winner);
        }
        printf("]]");
    }
    printf("]\n}\n");
    free(tr);
    free(bi);
    return 0;
}
