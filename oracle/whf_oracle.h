/* whf_oracle.h — CPU oracle for the witch-hut phase-1 Seed Check.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference CPU scanner's phase-1 algorithm
 * (lysh, engine/tools/lysh/src). Only tests/, the smoke check and the benchmark's CPU-baseline leg load it,
 * and only as the checker. The HIP product path never links, loads or calls it.
 *
 * Every function names the reference lines it restates. How each part is pinned (golden vectors from the
 * reference build, closed forms, brute force) is listed in tests/test_oracle.py.
 */
#ifndef WHF_ORACLE_H
#define WHF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeds and RNGs ---- */
uint64_t wo_splitmix64_at(uint64_t stream, uint64_t index);          /* i-th output of sequential SplitMix64 */
uint64_t wo_scramble_index(uint64_t index);                          /* hunt.py:121-124 seed_of_index */
void     wo_md5(const uint8_t *data, size_t len, uint8_t out[16]);  /* RFC 1321 */

typedef struct { uint64_t lo, hi; } wo_xr;
void     wo_xr_seed(wo_xr *s, uint64_t seed);                        /* rng.c:97-102 */
uint64_t wo_xr_next(wo_xr *s);                                       /* rng.c:115-122 */
int      wo_xr_next_int(wo_xr *s, int bound);                        /* rng.c:131-146 */
double   wo_xr_next_double(wo_xr *s);                                /* rng.c:153-155 */

/* java.util.Random as used by structure placement (rng.c:187-227) */
int      wo_lcg_next_int_seeded(uint64_t *state, int bound);
uint64_t wo_lcg_next_long(uint64_t *state);

/* Swamp-hut chunk in region (rx, rz) (structure.c:4-24). */
void wo_hut_chunk(uint64_t world_seed, int rx, int rz, int salt, int *cx, int *cz);

/* ---- noise ---- */
typedef struct {
    double ox, oy, oz;
    uint8_t perm[256];
} wo_perlin;

typedef struct {
    int n;                   /* amplitude slots */
    int first_octave;
    double amp[16];
    uint8_t present[16];
    wo_perlin oct[16];
    double lac0, per0;       /* lacunarity and persistence of slot 0 */
} wo_octave;

typedef struct {
    wo_octave sub[2];
    double amplitude;
} wo_noise;

void   wo_noise_init(wo_noise *n, uint64_t world_seed, int noise_id);
double wo_noise_sample(const wo_noise *n, double x, double y, double z);
double wo_noise_sample_y0(const wo_noise *n, double x, double z);
/* Tiered erosion gate (noise.c:432-459). Returns 1 when a tier fired (tier in *tier), else 0 with the
 * interleaved sum in *value. */
int    wo_erosion_tiered(const wo_noise *erosion, double x, double z, double *value, int *tier);

/* ---- phase 1 ---- */
enum {
    WO_REJ_EROSION = 0, WO_REJ_TEMPERATURE, WO_REJ_RIDGE, WO_REJ_ENTRANCE50, WO_REJ_ENTRANCE60,
    WO_REJ_CAVE_MAXH, WO_REJ_LADDER_NEG, WO_REJ_LADDER_POS, WO_REJ_CONTINENTALNESS, WO_REJ_FLOODEDNESS,
    WO_ACCEPT, WO_KIND_N
};

/* Phase-1 rule flags (phase1.h:36-48). */
#define WO_V_118     1   /* 1.18.x: temperature checked against its lower bound only */
#define WO_V_262     2   /* 26.2+: ridge <= -0.91 also rejects */
#define WO_V_LARGE   4   /* LARGE_BIOMES noise variants */
#define WO_V_SINGLE  8   /* SINGLE_BIOME: climate gates skipped */

typedef struct {
    wo_noise erosion, temperature, ridge, continentalness, floodedness;
    wo_noise cave_entrance, spag_rarity, spag_thickness, spag_1, spag_2, spag_rough_mod, spag_rough;
    wo_noise cave_layer, cave_cheese;
} wo_phase1;

void wo_phase1_init(wo_phase1 *p, uint64_t world_seed, int vflags);
/* Returns the reject kind (WO_ACCEPT when the region survives); *tier = fired erosion tier or -1. */
int  wo_phase1_check(const wo_phase1 *p, int vflags, int hut_x, int hut_z, int max_height, int *tier);
/* The 38 raw values of lysh_phase1_probe (phase1.c:39-80), in its order. */
void wo_phase1_probe(const wo_phase1 *p, int hut_x, int hut_z, int max_height, double out[38]);

double wo_entrance(const wo_phase1 *p, int x, int y, int z);
double wo_entrance2(const wo_phase1 *p, int x, int y, int z);
double wo_cheese(const wo_phase1 *p, int x, int y, int z);

/* Version name -> rule flags, following lysh's names (main.c:104-125; biome_trees.h:70499-70526).
 * Returns 0 on success. */
int wo_version_flags(const char *name, int *vflags);

/* ---- scan ---- */
typedef struct {
    uint64_t checks, accepted, reached_ladder;
    uint64_t funnel[WO_KIND_N];
    uint64_t tiers[8];
} wo_funnel;

typedef struct { int64_t seed; int32_t rx, rz, cx, cz; } wo_hit;

/* Phase-1 scan of each seed over regions [rx0,rx1) x [rz0,rz1) (search.c:126-160 per seed), using `threads`
 * worker threads over seeds. Survivors go to hits (up to cap; *n_hits counts all). Returns 0 on success. */
int wo_scan(const int64_t *seeds, int64_t n_seeds, int rx0, int rx1, int rz0, int rz1, int max_height,
            int vflags, int threads, wo_funnel *out, wo_hit *hits, int64_t cap, int64_t *n_hits);

/* Per-region reject kinds and tiers for one seed (row-major rz outer, rx inner). */
int wo_scan_regions(int64_t seed, int rx0, int rx1, int rz0, int rz1, int max_height, int vflags,
                    int8_t *kind, int8_t *tier);

/* FNV-1a over a Perlin sampler's perm table and origin bits: a compact fingerprint for fixtures. */
uint64_t wo_perlin_fingerprint(const wo_perlin *p);
/* Fingerprint of (noise id, sub, slot) for a world seed; 0 when the slot is empty. */
uint64_t wo_sampler_fingerprint(uint64_t world_seed, int noise_id, int sub, int slot);

/* Heap helpers for bindings: a phase-1 noise set or a single noise for a world seed (free with wo_probe_free). */
void  *wo_probe_alloc(uint64_t world_seed, int vflags);
void  *wo_noise_alloc(uint64_t world_seed, int noise_id);
void   wo_probe_free(void *p);
double wo_noise_amplitude(const void *noise);
void   wo_noise_perm(const void *noise, int sub, int slot, uint8_t out[256]);

#ifdef __cplusplus
}
#endif
#endif /* WHF_ORACLE_H */
