/* gen_golden.c — writes golden vectors from the reference CPU scanner (lysh) for the oracle tests.
 *
 * Built by tests/golden/make_golden.py against oracle/_ref/liblysh.a (the reference's own sources compiled
 * by oracle/Makefile). Prints JSON on stdout; the script stores it as tests/golden/lysh_vectors.json.
 *
 * Sections (each is data only: inputs and the reference's outputs):
 *   md5        MD5 digests of the noise identifiers and octave names            (rng.c:52-73)
 *   xr         Xoroshiro128++ streams, nextInt(bound), nextDouble for seeds      (rng.c:75-155)
 *   lcg        swamp-hut chunk for (seed, rx, rz)                                (structure.c:14-24)
 *   sampler    FNV-1a fingerprints of perm tables + origins per (seed, noise, sub, slot)   (noise.c:61-156)
 *   probe      the 38 raw doubles of lysh_phase1_probe as hex bit patterns       (phase1.c:47-80)
 *   regions    per-region reject kind and tier of lysh_phase1_check_ex           (phase1.c:130-200)
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "noise.h"
#include "phase1.h"
#include "rng.h"
#include "structure.h"

static uint64_t sm64(uint64_t s, uint64_t i) {
    uint64_t z = s + (i + 1) * UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

static uint64_t fnv(const lysh_perlin *p) {
    uint64_t h = UINT64_C(0xCBF29CE484222325);
    uint8_t buf[256 + 24];
    memcpy(buf, p->perm, 256);
    memcpy(buf + 256, &p->ox, 8);
    memcpy(buf + 264, &p->oy, 8);
    memcpy(buf + 272, &p->oz, 8);
    for (size_t i = 0; i < sizeof buf; i++) { h ^= buf[i]; h *= UINT64_C(0x100000001B3); }
    return h;
}

static const char *NAMES[] = {
    "minecraft:temperature", "minecraft:vegetation", "minecraft:continentalness", "minecraft:erosion",
    "minecraft:temperature_large", "minecraft:vegetation_large", "minecraft:continentalness_large",
    "minecraft:erosion_large", "minecraft:ridge", "minecraft:aquifer_fluid_level_floodedness",
    "minecraft:cave_entrance", "minecraft:spaghetti_3d_rarity", "minecraft:spaghetti_3d_thickness",
    "minecraft:spaghetti_3d_1", "minecraft:spaghetti_3d_2", "minecraft:spaghetti_roughness_modulator",
    "minecraft:spaghetti_roughness", "minecraft:cave_layer", "minecraft:cave_cheese",
};
#define NN ((int)(sizeof NAMES / sizeof NAMES[0]))

static void hex16(const uint8_t *h) { for (int i = 0; i < 16; i++) printf("%02x", h[i]); }

int main(void) {
    printf("{\n");
    /* md5 */
    printf("\"md5\": [");
    for (int i = 0; i < NN; i++) {
        uint8_t h[16];
        lysh_md5((const uint8_t *)NAMES[i], strlen(NAMES[i]), h);
        printf("%s[\"%s\", \"", i ? ", " : "", NAMES[i]);
        hex16(h);
        printf("\"]");
    }
    for (int o = -16; o <= 4; o++) {
        char nm[32];
        snprintf(nm, sizeof nm, "octave_%d", o);
        uint8_t h[16];
        lysh_md5((const uint8_t *)nm, strlen(nm), h);
        printf(", [\"%s\", \"", nm);
        hex16(h);
        printf("\"]");
    }
    printf("],\n");

    /* xoroshiro */
    printf("\"xr\": [");
    for (int s = 0; s < 8; s++) {
        uint64_t seed = s < 4 ? (uint64_t)s : sm64(7, (uint64_t)s);
        xoroshiro_t r;
        lysh_xr_init(&r, seed);
        printf("%s{\"seed\": \"%" PRIu64 "\", \"longs\": [", s ? ", " : "", seed);
        for (int k = 0; k < 4; k++) printf("%s\"%" PRIu64 "\"", k ? ", " : "", lysh_xr_next_long(&r));
        printf("], \"ints\": [");
        for (int k = 0; k < 8; k++) printf("%s%d", k ? ", " : "", lysh_xr_next_int_bound(&r, 256 - 31 * k));
        printf("], \"double\": \"%016" PRIx64 "\"}", ({ double d = lysh_xr_next_double(&r); uint64_t b; memcpy(&b, &d, 8); b; }));
    }
    printf("],\n");

    /* lcg placement */
    printf("\"lcg\": [");
    int first = 1;
    for (int s = 0; s < 16; s++) {
        uint64_t seed = sm64(11, (uint64_t)s);
        for (int rx = -3; rx <= 3; rx += 3)
            for (int rz = -58594; rz <= 58593; rz += 58593) {
                lysh_chunkpos c = lysh_swamp_hut_in_region(seed, rx, rz, LYSH_SWAMP_HUT_SALT);
                printf("%s[\"%" PRIu64 "\", %d, %d, %d, %d]", first ? "" : ", ", seed, rx, rz, c.chunkX, c.chunkZ);
                first = 0;
            }
    }
    printf("],\n");

    /* sampler fingerprints */
    printf("\"sampler\": [");
    first = 1;
    for (int s = 0; s < 3; s++) {
        uint64_t seed = sm64(2, (uint64_t)s);
        for (int n = 0; n < NN; n++) {
            lysh_dblnoise *dn = malloc(sizeof *dn);
            lysh_dblnoise_from_seed_id(dn, seed, NAMES[n]);
            for (int sub = 0; sub < 2; sub++)
                for (int slot = 0; slot < dn->sub[sub].n; slot++) {
                    if (!dn->sub[sub].oct[slot].present) continue;
                    printf("%s[\"%" PRIu64 "\", %d, %d, %d, \"%016" PRIx64 "\"]", first ? "" : ", ", seed, n, sub, slot,
                           fnv(&dn->sub[sub].oct[slot]));
                    first = 0;
                }
            uint64_t ab;
            memcpy(&ab, &dn->amplitude, 8);
            printf(", [\"%" PRIu64 "\", %d, -1, -1, \"%016" PRIx64 "\"]", seed, n, ab);
            free(dn);
        }
    }
    printf("],\n");

    /* probes: near the known answer and at seeded random positions, for two versions' max heights */
    printf("\"probe\": [");
    first = 1;
    lysh_phase1 *p1 = malloc(sizeof *p1);
    lysh_phase1_opts opts;
    memset(&opts, 0, sizeof opts);
    for (int s = 0; s < 6; s++) {
        uint64_t seed = s == 0 ? UINT64_C(1473884015873122059) : sm64(5, (uint64_t)s);
        lysh_phase1_init(p1, seed, &opts);
        for (int q = 0; q < 6; q++) {
            int hx, hz, maxh = q % 2 ? -50 : -40;
            if (s == 0 && q == 0) { hx = -15792; hz = -17168; }
            else {
                uint64_t r = sm64(seed, (uint64_t)q);
                hx = (int)((int64_t)(r % 60000001ULL) - 30000000) & ~15;
                hz = (int)((int64_t)((r >> 20) % 60000001ULL) - 30000000) & ~15;
            }
            double out[LYSH_P1_PROBE_N];
            lysh_phase1_probe(p1, hx, hz, maxh, out);
            printf("%s{\"seed\": \"%" PRIu64 "\", \"x\": %d, \"z\": %d, \"max_y\": %d, \"v\": [", first ? "" : ", ", seed,
                   hx, hz, maxh);
            for (int k = 0; k < LYSH_P1_PROBE_N; k++) {
                uint64_t b;
                memcpy(&b, &out[k], 8);
                printf("%s\"%016" PRIx64 "\"", k ? ", " : "", b);
            }
            printf("]}");
            first = 0;
        }
    }
    printf("],\n");

    /* per-region kinds and tiers: the known-answer seed's +-79 block and C1-shaped seeds, three rule sets */
    printf("\"regions\": [");
    first = 1;
    for (int v = 0; v < 3; v++) {
        lysh_phase1_opts o;
        memset(&o, 0, sizeof o);
        o.mc_1_18_2 = v == 1;
        o.pre_26_2 = v != 2;
        const char *vn = v == 0 ? "1.20" : (v == 1 ? "1.18" : "26.2");
        for (int s = 0; s < 4; s++) {
            uint64_t seed = s == 0 ? UINT64_C(1473884015873122059) : sm64(1, (uint64_t)s);
            int r = s == 0 ? 79 : 12;
            lysh_phase1_init(p1, seed, &o);
            printf("%s{\"seed\": \"%" PRIu64 "\", \"version\": \"%s\", \"rx0\": %d, \"rx1\": %d, \"rz0\": %d, \"rz1\": %d, "
                   "\"max_y\": -50, \"kinds\": \"",
                   first ? "" : ", ", seed, vn, -r, r, -r, r);
            first = 0;
            char *tiers = malloc((size_t)(4 * r * r) + 1);
            int n = 0;
            for (int rz = -r; rz < r; rz++)
                for (int rx = -r; rx < r; rx++) {
                    lysh_chunkpos c = lysh_swamp_hut_in_region(seed, rx, rz, LYSH_SWAMP_HUT_SALT);
                    long long th[LYSH_TIER_MAX] = {0};
                    lysh_phase1_result res;
                    lysh_phase1_check_ex(p1, &o, c.chunkX * 16, c.chunkZ * 16, -50, &res, th);
                    int t = -1;
                    for (int k = 0; k < LYSH_TIER_MAX; k++) if (th[k]) t = k;
                    putchar("0123456789A"[res.reject]);
                    tiers[n++] = t < 0 ? '-' : (char)('0' + t);
                }
            tiers[n] = 0;
            printf("\", \"tiers\": \"%s\"}", tiers);
            free(tiers);
        }
    }
    printf("]\n}\n");
    free(p1);
    return 0;
}
