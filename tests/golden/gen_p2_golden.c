/* gen_p2_golden.c — writes golden vectors of the terrain stage from the reference CPU scanner (lysh).
 *
 * Built by tests/golden/make_p2_golden.py against oracle/_ref/liblysh.a; prints JSON on stdout, stored as
 * tests/golden/p2_vectors.json. The phase-2 state is set up as lysh_search_ctx_init does (eval.c:77-83): 26.1.2
 * terrain semantics and the preliminary surface level as the aquifer's column top.
 *
 * Sections (data only: inputs and the reference's outputs; doubles and floats as hex bit patterns):
 *   point    terrain point offset/factor/peaks at quart positions            (terrain.c:22-44, spline.c:116-131)
 *   blended  blended base noise at blocks                                    (interp_noise.c:91-139)
 *   column   the noise column at corner blocks (4 cx, -64 + 8 i, 4 cz)      (phase2.c:152-159, density.c:228-281)
 *   noodle   the noodle and its three children                              (density.c:186-209)
 *   density  the interpolated final density                                 (phase2.c:161-194)
 *   psl      the preliminary surface level                                  (column_top.c:85-122)
 *   skip     the aquifer's skipSamplingAboveY for a chunk                    (aquifer.c:160-196)
 *   fluid    computeFluid level and type                                    (aquifer.c:286-331)
 *   block    the block state of density + aquifer                            (phase2.c:311-316)
 *   height   the pre-carver surface height                                   (phase2.c:318-334)
 * Columns are chunk middles of phase-1 survivors (1.20 rules, maxHeight 60) and random chunks near and far from
 * the origin, a fifth of the seeds with the LARGE_BIOMES noises.
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aquifer.h"
#include "column_top.h"
#include "density.h"
#include "interp_noise.h"
#include "phase1.h"
#include "phase2.h"
#include "structure.h"
#include "terrain.h"

static uint64_t sm64(uint64_t s, uint64_t i) {
    uint64_t z = s + (i + 1) * UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}
static uint64_t rng_state = UINT64_C(0x1234567887654321);
static uint64_t rnd(void) { return sm64(rng_state++, 0); }
static int rnd_int(int lo, int hi) { return lo + (int)(rnd() % (uint64_t)(hi - lo + 1)); }
static uint64_t dbits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static uint32_t fbits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

static int column_top_cb(void *user, int x, int z) {
    return lysh_preliminary_surface_level((const lysh_terrain *)user, x, z);
}

typedef struct { uint64_t seed; int x, z, large; } col_t;
#define MAXC 512
static col_t C[MAXC];
static int NC;
static void add(uint64_t seed, int x, int z, int large) {
    if (NC < MAXC) { C[NC].seed = seed; C[NC].x = x; C[NC].z = z; C[NC].large = large; NC++; }
}

static lysh_phase2 *P2;
static uint64_t p2_seed = 0;
static int p2_large = -1;
static lysh_phase2 *p2_for(uint64_t seed, int large) {
    if (P2 && p2_seed == seed && p2_large == large) return P2;
    if (P2) { lysh_phase2_free(P2); free(P2); }
    P2 = (lysh_phase2 *)malloc(sizeof *P2);
    lysh_phase2_init(P2, seed, large);
    lysh_phase2_set_column_top(P2, column_top_cb, &P2->terrain);
    p2_seed = seed;
    p2_large = large;
    return P2;
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
            if (res.reject == LYSH_P1_ACCEPT) { add(seed, c.chunkX * 16 + 8, c.chunkZ * 16 + 8, 0); got++; }
        }
    free(p1);
}

int main(void) {
    for (int s = 0; s < 40; s++) survivors(sm64(21, (uint64_t)s), 16, 2);
    for (int s = 0; s < 30; s++) {
        const uint64_t seed = sm64(22, (uint64_t)s);
        add(seed, rnd_int(-3000, 3000) * 16 + 8, rnd_int(-3000, 3000) * 16 + 8, s % 5 == 0);
        add(seed, rnd_int(-1870000, 1870000) * 16 + 8, rnd_int(-1870000, 1870000) * 16 + 8, 0);
        add(seed, rnd_int(-500, 500), rnd_int(-500, 500), s % 5 == 1);
    }
    int first;
    printf("{\n");

    printf("\"point\": [");
    for (int i = 0; i < NC; i++) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        lysh_terrain_info ti;
        const int qx = C[i].x >> 2, qz = C[i].z >> 2;
        lysh_terrain_info_at(&p2->terrain, qx, qz, &ti);
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,\"%08x\",\"%08x\",\"%08x\"]", i ? "," : "", C[i].seed, C[i].large, qx, qz,
               fbits(ti.point.offset), fbits(ti.point.factor), fbits(ti.point.peaks));
    }
    printf("],\n\"blended\": [");
    first = 1;
    for (int i = 0; i < NC; i += 2) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        for (int k = 0; k < 4; k++) {
            const int x = C[i].x + rnd_int(-20, 20), y = rnd_int(-64, 320), z = C[i].z + rnd_int(-20, 20);
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,\"%016" PRIx64 "\"]", first ? "" : ",", C[i].seed, C[i].large, x, y, z,
                   dbits(lysh_interp_calculate_noise_261(&p2->interp, x, y, z)));
            first = 0;
        }
    }
    printf("],\n\"column\": [");
    first = 1;
    for (int i = 0; i < NC; i += 3) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        const int cx = C[i].x >> 2, cz = C[i].z >> 2;
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,[", first ? "" : ",", C[i].seed, C[i].large, cx, cz);
        first = 0;
        for (int k = 0; k <= 48; k++)
            printf("%s\"%016" PRIx64 "\"", k ? "," : "", dbits(lysh_phase2_density_raw_at(p2, cx * 4, -64 + 8 * k, cz * 4)));
        printf("]]");
    }
    printf("],\n\"noodle\": [");
    first = 1;
    for (int i = 0; i < NC; i += 3) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        for (int k = 0; k < 4; k++) {
            const int x = C[i].x + rnd_int(-40, 40), y = rnd_int(-70, 330), z = C[i].z + rnd_int(-40, 40);
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d", first ? "" : ",", C[i].seed, C[i].large, x, y, z);
            for (int w = 0; w < 4; w++) printf(",\"%016" PRIx64 "\"", dbits(lysh_noodle_child(&p2->density, w, x, y, z)));
            printf("]");
            first = 0;
        }
    }
    printf("],\n\"density\": [");
    first = 1;
    for (int i = 0; i < NC; i++) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        for (int k = 0; k < 4; k++) {
            const int x = C[i].x + rnd_int(-3, 3), y = rnd_int(-64, 319), z = C[i].z + rnd_int(-3, 3);
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,\"%016" PRIx64 "\"]", first ? "" : ",", C[i].seed, C[i].large, x, y, z,
                   dbits(lysh_phase2_density_at(p2, x, y, z)));
            first = 0;
        }
    }
    printf("],\n\"psl\": [");
    first = 1;
    for (int i = 0; i < NC; i++) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        for (int k = 0; k < 3; k++) {
            const int x = C[i].x + rnd_int(-64, 64), z = C[i].z + rnd_int(-64, 64);
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d]", first ? "" : ",", C[i].seed, C[i].large, x, z,
                   lysh_preliminary_surface_level(&p2->terrain, x, z));
            first = 0;
        }
    }
    printf("],\n\"skip\": [");
    for (int i = 0; i < NC; i++) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        lysh_phase2_ensure_chunk(p2, C[i].x, C[i].z);
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d]", i ? "," : "", C[i].seed, C[i].large, C[i].x >> 4, C[i].z >> 4,
               p2->aquifer.skip_sampling_above_y);
    }
    printf("],\n\"fluid\": [");
    first = 1;
    for (int i = 0; i < NC; i += 2) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        for (int k = 0; k < 3; k++) {
            const int x = C[i].x + rnd_int(-30, 30), y = rnd_int(-64, 200), z = C[i].z + rnd_int(-30, 30);
            int level;
            lysh_block_state type;
            lysh_aquifer_compute_fluid(&p2->aquifer, x, y, z, &level, &type);
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,%d,%d]", first ? "" : ",", C[i].seed, C[i].large, x, y, z, level,
                   (int)type);
            first = 0;
        }
    }
    printf("],\n\"block\": [");
    first = 1;
    for (int i = 0; i < NC; i++) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        const int h = lysh_phase2_height_pre_carver(p2, C[i].x, C[i].z);
        for (int k = 0; k < 6; k++) {
            const int y = k < 3 ? h - 1 - k * 3 : rnd_int(-64, 100);
            if (y < -64 || y > 319) continue;
            printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,%d]", first ? "" : ",", C[i].seed, C[i].large, C[i].x, y, C[i].z,
                   (int)lysh_phase2_block_at(p2, C[i].x, y, C[i].z));
            first = 0;
        }
    }
    printf("],\n\"height\": [");
    for (int i = 0; i < NC; i++) {
        lysh_phase2 *p2 = p2_for(C[i].seed, C[i].large);
        const int h = lysh_phase2_height_pre_carver(p2, C[i].x, C[i].z);
        const int top = h > -64 ? (int)lysh_phase2_block_at(p2, C[i].x, h - 1, C[i].z) : 1;
        printf("%s[\"%" PRIu64 "\",%d,%d,%d,%d,%d]", i ? "," : "", C[i].seed, C[i].large, C[i].x, C[i].z, h, top);
    }
    printf("]\n}\n");
    return 0;
}
