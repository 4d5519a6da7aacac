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

/* ---- Deep Dive mineshaft stage (whf_oracle_ms.c) ---- */
enum { WO_MS_ROOM = 0, WO_MS_CORRIDOR = 1, WO_MS_CROSSING = 2, WO_MS_STAIRS = 3 };
#define WO_MS_MAXP 512   /* piece cap of the reference generator (mineshaft_tree.h:11-13) */

typedef struct {
    int32_t x0, y0, z0, x1, y1, z1;
    int8_t type, facing, depth, rails;   /* facing: 0 N, 1 S, 2 W, 3 E, -1 for the room */
    int16_t parent;                      /* index of the piece whose placement tried this one; -1 for the room */
} wo_ms_piece;

typedef struct {
    int status;        /* 0 complete; 1 a piece went below the abort cut; 2 piece cap or int8 box range exceeded */
    int pieces, span_y, min_y, max_y, room_floor;   /* min_y/max_y before the move; room_floor after it */
    uint64_t draws;          /* LCG steps taken, the two start-test draws included */
    uint64_t overlap_tests;  /* box-collision queries */
} wo_ms_result;

uint64_t wo_ms_carver_seed(uint64_t s48, int32_t cx, int32_t cz);   /* room_core.h:55-58 */
double   wo_ms_start_double(uint64_t carver);                        /* new Random(X).nextDouble() */
int      wo_ms_is_start(uint64_t carver);                            /* mstree.cpp:10-15 */
/* Layout of the mineshaft started by `carver` in chunk (cx, cz) (mineshaft_tree.h:529-565). Boxes are
 * absolute block coordinates after the move below sea level (status 0). Pieces stop at `abort_below`
 * (INT_MIN: never), as the reference's abortBelow cut. Returns the status. */
int wo_ms_layout(uint64_t carver, int32_t cx, int32_t cz, int abort_below, wo_ms_piece *pieces, int cap,
                 wo_ms_result *out);
int wo_ms_room_floor(uint64_t carver, int abort_below, wo_ms_result *out);

typedef struct {
    uint64_t structure_seeds, tests, starts, survivors;
    uint64_t status[3];      /* per start: complete, cut below the abort line, cap/range */
    uint64_t draws, pieces;  /* summed over starts */
} wo_dd_stats;

typedef struct {
    uint64_t s48;
    int32_t rx, rz, cx, cz;
    uint64_t carver;
    int32_t room_floor, pieces;
} wo_dd_hit;

/* Forward sweep: S = (s_first + i) mod 2^48 for i < count, hut regions [rx0,rx1) x [rz0,rz1); a hit is a
 * mineshaft start in the hut chunk whose room floor is <= floor_max. */
int wo_deepdive_sweep(uint64_t s_first, uint64_t count, int rx0, int rx1, int rz0, int rz1, int floor_max,
                      int abort_below, int threads, wo_dd_stats *out, wo_dd_hit *hits, int64_t cap,
                      int64_t *n_hits);

/* ---- biome stage (whf_oracle_biome.c) ---- */
typedef struct {
    wo_noise offset, continentalness, erosion, ridge, temperature, vegetation;
} wo_biome;

typedef struct {
    double shifted_x, shifted_z;
    double continentalness, weirdness, erosion;
} wo_terrain_info;

typedef struct {
    int64_t t[6];          /* quantized temperature, humidity, continentalness, erosion, depth, weirdness */
    int qx, qy, qz, occ_y;
    int winner, ok;
} wo_biome_result;

/* Quart heights an occupied height can fall in: the surface height is -64..320 (phase2.c:318-334), the
 * occupied height one less, qy = occ_y >> 2 in [-17, 79]. */
#define WO_BIOME_QY0  (-17)
#define WO_BIOME_QY_N 97

void    wo_biome_init(wo_biome *b, uint64_t world_seed, int vflags);
void    wo_terrain_info_at(const wo_biome *b, int qx, int qz, wo_terrain_info *out);   /* terrain.c:22-44 */
float   wo_spline_apply(int node, const float p[4]);                                   /* spline.c:28-62 */
float   wo_normalized_weirdness(float w);                                              /* spline.c:63-70 */
void    wo_terrain_spline(float c, float e, float w, float out[3]);                     /* spline.c:83-111 */
double  wo_overworld_offset(const wo_biome *b, int block_x, int block_z);               /* column_top.c:124-130 */
double  wo_depth_gradient(double y);                                                   /* biome.c:77-80 */
int64_t wo_tree_dist(int tree, int node, const int64_t np[6]);                         /* biome.c:23-31 */
int     wo_tree_leaf(int tree, const int64_t np[6]);                                   /* biome.c:34-47 */
int     wo_tree_biome(int tree, const int64_t np[6]);                                  /* biome.c:49-53 */
int64_t wo_tree_min_dist(int tree, const int64_t np[6], int *n_min);                   /* brute force */
/* The swamp-hut biome check at the hut chunk's middle for an occupied height (biome.c:94-156). */
int     wo_biome_check(const wo_biome *b, int tree, int hut_x, int hut_z, int occ_y, wo_biome_result *out);
/* The tree's winner at every quart height qy = WO_BIOME_QY0 + i. */
void    wo_biome_heights(const wo_biome *b, int tree, int hut_x, int hut_z, uint8_t winners[WO_BIOME_QY_N]);
/* Version name -> biome tree (main.c:104-125, eval.c:104-106). Returns 0 when known. */
int     wo_biome_tree(const char *version, int *tree);
/* shifted x, z, continentalness, weirdness, erosion, temperature, humidity, overworld offset at a quart point */
void    wo_biome_point(const wo_biome *b, int qx, int qz, double out[8]);
void   *wo_biome_alloc(uint64_t world_seed, int vflags);   /* free with wo_probe_free */

/* ---- terrain stage (whf_oracle_p2.c) ---- */
#define WO_P2_NOISES 25
typedef struct {
    wo_biome bio;                 /* offset, continentalness, erosion, ridge (+ temperature, vegetation) */
    wo_noise n[WO_P2_NOISES];     /* density, noodle and aquifer noises, in the order wo_p2_init lists them */
    wo_perlin lower[16], upper[16], main[8];   /* blended noise stacks, by octave (0 = the coarsest) */
    uint64_t aq_lo, aq_hi;        /* the aquifer's positional random factory */
} wo_phase2;

void   wo_p2_init(wo_phase2 *p, uint64_t seed, int vflags);
void  *wo_p2_alloc(uint64_t seed, int vflags);                            /* free with wo_probe_free */
/* offset, factor, peaks of the terrain point at a quart position, 26.1.2 semantics (terrain.c:33-39) */
void   wo_p2_terrain_point(const wo_phase2 *p, int qx, int qz, float out[3]);
double wo_p2_blended(const wo_phase2 *p, int x, int y, int z);            /* interp_noise.c:91-139 */
double wo_p2_noise_column(const wo_phase2 *p, int x, int y, int z, const float tp[3], double base); /* density.c:228-281 */
double wo_p2_noodle_child(const wo_phase2 *p, int which, int x, int y, int z);   /* density.c:186-209 */
/* the 49 corner samples of main density and the three interpolated noodle children (phase2.c:71-106) */
void   wo_p2_corner_column(const wo_phase2 *p, int cell_x, int cell_z, double main[49], double noodle[3][49]);
double wo_p2_density(const wo_phase2 *p, int x, int y, int z);            /* phase2.c:161-194 */
int    wo_p2_prelim_surface(const wo_phase2 *p, int x, int z);            /* column_top.c:85-122 */
int    wo_p2_skip_above(const wo_phase2 *p, int chunk_x, int chunk_z);    /* aquifer.c:160-196 */
void   wo_p2_fluid(const wo_phase2 *p, int x, int y, int z, int *level, int *type);   /* aquifer.c:286-331 */
/* block state 0 solid, 1 air, 2 water, 3 lava for a block of the density (aquifer.c:416-563) */
int    wo_p2_substance(const wo_phase2 *p, int x, int y, int z, double density);
/* pre-carver surface height (phase2.c:318-334); *top_state = state of the block below it */
int    wo_p2_height(const wo_phase2 *p, int x, int z, int *top_state);
/* The pillar cube: 0 = pow(t, 3.0) as the reference (default), 1 = the device's compensated cube; wo_cube_stats reads
 * and clears (cubes evaluated in mode 1, cubes differing from pow). Process-wide. */
void   wo_set_cube_mode(int mode);
void   wo_cube_stats(unsigned long long out[2]);

/* ---- carvers and hut evaluation (whf_oracle_hut.c) ---- */
enum { WO_HUT_OK = 0, WO_HUT_REJECT_Y = 1, WO_HUT_REJECT_FLOOD = 2, WO_HUT_REJECT_BIOME = 3 };   /* eval.h:88-93 */
typedef struct {
    int dir, size_x, size_z;       /* orientation (0 N, 1 E, 2 S, 3 W) and footprint */
    int biome_ok, winner, occ_y;   /* the biome check at the chunk middle's pre-carver surface */
    int wet, flooded;              /* footprint columns with water above y = 62; all 63 */
    int carve_skipped;             /* flooded with every pre-carver top block at or below 63: no carving */
    long long sum_h;
    int avg_y;                     /* sum_h / 63, truncated */
    int ok, reject;
} wo_hut_result;

int wo_hut_orientation(uint64_t seed, int hut_x, int hut_z);                               /* eval.c:41-50 */
/* post-carver surface height of a column (phase2.c:438-456, carvers of carver.c:498-523) */
int wo_post_carver_height(const wo_phase2 *p, uint64_t seed, int x, int z);
/* lysh_eval_hut (eval.c:121-226) with biome tree `tree` */
int wo_hut_eval(const wo_phase2 *p, int tree, uint64_t seed, int hut_x, int hut_z, int max_y, wo_hut_result *out);

/* Heap helpers for bindings: a phase-1 noise set or a single noise for a world seed (free with wo_probe_free). */
void  *wo_probe_alloc(uint64_t world_seed, int vflags);
void  *wo_noise_alloc(uint64_t world_seed, int noise_id);
void   wo_probe_free(void *p);
double wo_noise_amplitude(const void *noise);
void   wo_noise_perm(const void *noise, int sub, int slot, uint8_t out[256]);

/* ---- Deep Dive solver (whf_oracle_solve.c; hut_lookup_v4 --solve5, taskspace.h, solver_fused.cuh) ---- */
uint32_t wo_solve_perm_group(uint32_t p);                                       /* taskspace.h:24 */
uint64_t wo_solve_creg(int32_t rx, int32_t rz);                                 /* taskspace.h:25 */
int      wo_solve_regions(uint32_t c20, int regime_x, int regime_z, int32_t *rx, int32_t *rz, int cap);
int      wo_solve_offsets(int tier, uint8_t *ox, uint8_t *oz);                  /* solver_fused.cuh:25-33 */
int      wo_solve_task(uint64_t X, int32_t cx, int32_t cz, uint64_t *out, int cap);
int      wo_solve_prefixes(uint64_t X, int32_t cx, int32_t cz, int width, uint64_t *out, int cap);
int64_t  wo_solve_tasks(uint64_t X, const int32_t *cx, const int32_t *cz, int n, int per_task_cap, int threads,
                        uint64_t *seeds, int *counts);

#ifdef __cplusplus
}
#endif
#endif /* WHF_ORACLE_H */
