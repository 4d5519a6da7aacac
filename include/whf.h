/* whf.h — C ABI of the MI355X witch-hut Seed Check library (libwhf_hip.so).
 *
 * Plain C types only: callers own every input array and output buffer, the library owns device memory,
 * streams and kernels. Errors are negative return codes; the library never exits the process.
 *
 * Each entry point replaces a piece of the reference's GPU engine (engine/gpu/low54/phase1_scan.cu) or of
 * its CPU scanner (engine/tools/lysh/src), cited per function. The reference has no C ABI for this path:
 * hunt.py drives a `phase1_scan` process over argv/stdin (phase1_scan.cu:561-601, 1365-1387); this library
 * is what that process (low-witch-finder_amd/host/phase1_scan.cpp) now calls.
 */
#ifndef WHF_H
#define WHF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WHF_ABI_VERSION 1

/* Return codes. */
#define WHF_OK          0
#define WHF_EINVAL     (-1)   /* bad argument: empty rect, null pointer, unknown mode */
#define WHF_ENOMEM     (-2)   /* device or host allocation failed, or the request exceeds the HBM budget */
#define WHF_EDEVICE    (-3)   /* a HIP call failed (no device, launch error, fault) */
#define WHF_EOVERFLOW  (-4)   /* more survivors than the sink holds; counters are still complete */
#define WHF_ENOTSUP    (-5)   /* option not supported by this build */

/* Seed sources of a request. */
#define WHF_SEEDS_LIST      0   /* seeds[0..n_seeds) as given */
#define WHF_SEEDS_SPLITMIX  1   /* i-th output (offset + i) of sequential SplitMix64 seeded with `stream` (Vigna) */
#define WHF_SEEDS_COUNTER   2   /* SplitMix64 finalizer of (stream + offset + i + 0x9E3779B97F4A7C15): the reference's
                                   --seed-gen base count form (device_construct.h:39-46, phase1_scan.cu:87) */

/* Phase-1 rule flags: which game-version rules apply (phase1.h:36-48). */
#define WHF_V_118      1   /* 1.18.x: temperature only has its lower bound -0.45 */
#define WHF_V_262      2   /* 26.2+: ridge <= -0.91 also rejects */
#define WHF_V_LARGE    4   /* LARGE_BIOMES preset: the _large climate noises */
#define WHF_V_SINGLE   8   /* SINGLE_BIOME preset: climate gates skipped */

/* Arithmetic of the noise bulk. Both give the same survivors and counters. */
#define WHF_PREC_FP64      0   /* plain FP64 path, bit-exact with the reference CPU scanner */
#define WHF_PREC_FP32_GATE 1   /* erosion stage: FP32 corner blend decides every tier check that is farther than
                                  the stated error bound from its threshold; the rest is re-evaluated in FP64 */

typedef struct whf_ctx whf_ctx;

#define WHF_CFG_PROFILE    1   /* whf_config.flags: time kernel families with HIP events (whf_stats.kernel_ms) */
#define WHF_CFG_ONE_STREAM 2   /* whf_config.flags: run every stage of a scan on one stream (no batch pipeline;
                                  for contexts that share the device's hardware queues with many others) */

typedef struct {
    int device;                 /* HIP device ordinal */
    int flags;                  /* WHF_CFG_* */
    uint64_t hbm_budget;        /* unused by whf_open (a context holds about 9 GB); whf_server_config caps slots */
} whf_config;

typedef struct {
    int seed_mode;              /* WHF_SEEDS_* */
    int precision;              /* WHF_PREC_* */
    const int64_t *seeds;       /* WHF_SEEDS_LIST: host array of n_seeds world seeds */
    uint64_t n_seeds;
    uint64_t stream, offset;    /* WHF_SEEDS_SPLITMIX / _COUNTER */
    int rx0, rx1, rz0, rz1;     /* hut regions, half-open [rx0,rx1) x [rz0,rz1) */
    int max_height;             /* phase-1 maxHeight (lysh_phase1_check_ex's max_height, used as given); -140..100000 */
    int vflags;                 /* WHF_V_* */
    int salt;                   /* structure salt; 0 = swamp hut 14357620 */
    int phase2;                 /* WHF_P2_*: phase-1 survivors only, or lysh's hut verdict on them (whf_hut_eval) */
    int phase2_max_y;           /* the floor threshold of the hut verdict (lysh --phase2-max-y) */
    int biome_tree;             /* biome tree of the hut verdict: 1 + a tree index (whf_biome_tree); 0 = the rule
                                   flags' default, as lysh picks it (eval.c:104-106) */
} whf_request;

/* Phase 2 of a request: what whf_scan returns. */
#define WHF_P2_NONE     0   /* the phase-1 survivors (the reference GPU engine's candidates) */
#define WHF_P2_OR_LOWER 1   /* the huts lysh accepts with avg_y <= phase2_max_y (lysh scan --list), floor_y = avg_y */
#define WHF_P2_EXACTLY  2   /* of those, the huts with avg_y == phase2_max_y (hunt.py's "exactly", hunt.py:1342-1349) */

/* One phase-1 survivor: a hut position that passed every phase-1 gate; with phase 2, a hut lysh accepts. */
#define WHF_HIT_DEEPDIVE 1
#define WHF_HIT_PHASE2   2   /* floor_y is the hut's lysh floor (avg_y) */
typedef struct {
    int64_t seed;
    int32_t rx, rz;             /* region */
    int32_t cx, cz;             /* hut chunk; the hut's block is (16*cx, 16*cz) */
    int32_t floor_y;            /* Seed Check: the hut's floor with phase 2, else INT32_MIN; Deep Dive: the room floor */
    int32_t flags;              /* WHF_HIT_PHASE2 for a Seed Check hit with phase 2, WHF_HIT_DEEPDIVE for a Deep Dive
                                   find, else 0 */
} whf_hit;

typedef struct {
    whf_hit *hits;              /* caller-owned array */
    uint64_t capacity;          /* entries in hits */
    uint64_t count;             /* out: survivors found (may exceed capacity -> WHF_EOVERFLOW) */
} whf_sink;

/* Counters of one scan; the funnel follows lysh's reject kinds (phase1.h:53-66, main.c:96-100). */
#define WHF_KIND_N 11
typedef struct {
    uint64_t seeds;
    uint64_t checks;            /* hut-region checks (seeds x regions) */
    uint64_t accepted;          /* phase-1 survivors */
    uint64_t reached_ladder;    /* regions past the climate gates (lysh scan's "reached cave ladder") */
    uint64_t funnel[WHF_KIND_N];/* per reject kind; [10] = accepted */
    uint64_t tiers[8];          /* erosion tier where the tiered gate fired (noise.c:362-377); [7] unused */
    uint64_t samplers_built[3]; /* Perlin samplers constructed on the device per set: erosion (8 per seed),
                                   climate + Entrance (28 per seed past erosion), ladder tail (36 per seed past
                                   Entrance@50/60) */
    double device_ms;           /* device time of the scan (HIP events around the kernels) */
    double kernel_ms[8];        /* with whf_config.flags & 1: device time per kernel family (seeds, erosion build,
                                   erosion gate, compact, climate build, gate 2, ladder build, gate 3) */
    uint64_t fp64_rechecks;     /* WHF_PREC_FP32_GATE: items the FP32 gate left to the FP64 re-evaluation */
    uint64_t phase2[5];         /* with phase 2 (lysh's p2grade, main.c:333-345): evaluated, accepted, rejected for
                                   biome, for the floor (avg_y, or != threshold with WHF_P2_EXACTLY), for flood */
    double phase2_ms;           /* wall time of the hut evaluation */
} whf_stats;

/* Opens a context on a device and allocates its working buffers (replaces phase1_scan's allocateBuffers,
 * phase1_scan.cu:820-868). */
int whf_open(const whf_config *cfg, whf_ctx **out);

/* Releases everything the context holds (replaces freeBuffers, phase1_scan.cu:870-879). */
void whf_close(whf_ctx *ctx);

/* Synchronous Seed Check, one request per call: the plain FP64 drop-in. Scans every (seed, region) of the
 * request through the full phase-1 chain of the reference CPU scanner (lysh_scan_rect search.c:198 ->
 * lysh_phase1_check_ex phase1.c:130-200) and returns the survivors and the funnel. Replaces runScan
 * (phase1_scan.cu:881-1345), whose candidates were a superset of these. Not re-entrant per context. */
int whf_scan(whf_ctx *ctx, const whf_request *req, whf_sink *sink, whf_stats *stats);

/* Raw phase-1 values at hut positions: the 38 doubles of lysh_phase1_probe (phase1.c:39-80) for each
 * (seeds[i], hut_x[i], hut_z[i]), computed by the device kernels; out has 38*n doubles. For parity checks. */
int whf_probe(whf_ctx *ctx, const int64_t *seeds, const int32_t *hut_x, const int32_t *hut_z, uint64_t n,
              int max_height, int vflags, double *out);

/* Measured error of the FP32-gate corner blend against FP64 for the 8 erosion samplers (the noise the FP32
 * gate evaluates): max |v32 - v64| over points_per_seed points spread over the +-30M world for each seed.
 * No counterpart in the reference; it reports the per-octave error beside whf_fp32_bound(). */
int whf_noise_error(whf_ctx *ctx, const int64_t *seeds, uint64_t n_seeds, uint64_t points_per_seed, int vflags,
                    double max_err[8]);

/* The per-evaluation error bound the FP32 gate uses as its margin. */
double whf_fp32_bound(void);

/* ---- Deep Dive: mineshaft rooms under hut chunks, over 48-bit structure seeds ---- */

typedef struct {
    uint64_t s_first;           /* first structure seed; S_i = (s_first + i) mod 2^48, i < count */
    uint64_t count;             /* structure seeds of the sweep; 0..2^48 */
    int rx0, rx1, rz0, rz1;     /* hut regions, half-open */
    int floor_max;              /* keep mineshaft starts whose room floor is <= floor_max (-62..63) */
    int expand;                 /* 1: expand every kept start to its 2^16 world seeds (u << 48 | S) and run the
                                   phase-1 swamp check on them at the hut's region; 0: sweep only */
    int max_height;             /* phase-1 maxHeight of the expansion's check (as in whf_request) */
    int vflags;                 /* WHF_V_* rules of the expansion's check */
    int salt;                   /* structure salt; 0 = swamp hut 14357620 */
    int biome_tree;             /* expand = 1: the swamp check of the expansion's finds, lysh's biome test at the
                                   chunk middle's pre-carver surface (biome.c:94-156) with tree biome_tree - 1
                                   (whf_biome_tree); 0 = the rule flags' default tree (eval.c:104-106); -1 = none,
                                   the finds are the phase-1 survivors */
} whf_dd_request;

/* A mineshaft start in a hut chunk with a deep room: the solver's hit (S, cx, cz, X) plus the room floor. */
typedef struct {
    uint64_t s48;               /* structure seed (the world seed's low 48 bits) */
    int32_t rx, rz, cx, cz;     /* hut region and chunk (the mineshaft starts in the same chunk) */
    uint64_t carver;            /* the chunk's carver seed X */
    int32_t room_floor;         /* the room's floor y after the move below sea level */
    int32_t pieces;             /* pieces of the layout */
} whf_dd_spot;

typedef struct {
    whf_dd_spot *spots;         /* caller-owned array */
    uint64_t capacity;
    uint64_t count;             /* out: spots found (may exceed capacity -> WHF_EOVERFLOW) */
} whf_dd_sink;

typedef struct {
    uint64_t structure_seeds;   /* S values swept */
    uint64_t tests;             /* hut-chunk start tests (structure seeds x regions) */
    uint64_t starts;            /* mineshaft starts in hut chunks */
    uint64_t walks[3];          /* layout walks: complete, stopped by a piece too low for floor_max, capped
                                   (more than 512 pieces or a box beyond the int8 offsets the reference keeps) */
    uint64_t spots;             /* starts with room floor <= floor_max */
    uint64_t draws, pieces;     /* LCG steps and pieces over all walks */
    uint64_t expanded_seeds;    /* world seeds checked by the expansion (spots x 65536) */
    uint64_t accepted;          /* expansion survivors (finds) */
    uint64_t funnel[WHF_KIND_N];/* expansion: per phase-1 reject kind, as in whf_stats */
    uint64_t tiers[8];
    uint64_t samplers_built[3];
    double device_ms;
    double stage_ms[3];         /* sweep (start tests), layout walks, expansion (device time per stage) */
    uint64_t biome_checked;     /* expansion survivors given the swamp check (req->biome_tree >= 0) */
    uint64_t biome_rejected;    /* of those, not swamp at the chunk middle's surface (not among the finds) */
} whf_dd_stats;

/* Deep Dive over a range of structure seeds: for every S and hut region, the hut chunk's carver seed X
 * (room_core.h:55-58), the mineshaft start test nextDouble() < 0.004 (mstree.cpp:10-15) and, for starts, the
 * layout walk that fixes the room floor (mineshaft_tree.h:529-565). Starts with a room floor <= floor_max go
 * to `spots` (the hits of the reference solver, hut_lookup_v4_source.cu:1341 `S,cx,cz,X,1`, here with the
 * floor); with expand = 1 each is expanded to its 2^16 world seeds and checked by the phase-1 chain, and the
 * survivors by the swamp biome check (req->biome_tree), finds in `finds` (floor_y = the room floor, flags = 1). Replaces the solve + terrain stages hunt.py drives through
 * SolveRun (hunt.py:947-957). Not re-entrant per context. */
int whf_deepdive(whf_ctx *ctx, const whf_dd_request *req, whf_dd_sink *spots, whf_sink *finds, whf_dd_stats *stats);

/* The sweep of whf_deepdive (req->expand must be 0) without allocating: no hipMalloc, hipFree or other allocating
 * call, host or device, inside the call. It has no counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid allocation), which open the
 * context beforehand. Its start queue, spots and counters live in a static device array of about 8.4 MB per device,
 * set up by whf_open; a sweep batch holds at most 2^17 mineshaft starts (batches are sized to expect half of that,
 * more returns WHF_EOVERFLOW with a message) and at most 4096 hut regions are accepted. Spots are copied straight into
 * `spots` (count as in whf_deepdive, WHF_EOVERFLOW when more than its capacity); stats as whf_deepdive's with expand =
 * 0, stage times zero. Every call on a device shares that device's array: calls that overlap on one device, from any
 * context or thread, race and must not be made. */
int whf_deepdive_noalloc(whf_ctx *ctx, const whf_dd_request *req, whf_dd_sink *spots, whf_dd_stats *stats);

/* The terrain stage of the reference's --spots mode (hut_lookup_v4_source.cu:1262-1292, the run hunt.py starts
 * per chunk of solver hits, hunt.py:954-957): the caller fills s48, cx and cz of n <= 2^20 spots (the "S cx cz"
 * lines of its spots file); the library fills in place rx = cx >> 5, rz = cz >> 5, the chunk's carver seed X and,
 * where X is a mineshaft start, the room floor (INT32_MIN where the walk does not complete, and for a spot that is
 * not a start, which the reference reports as "not a mineshaft start" and skips) and piece count (0 exactly for a
 * spot that is not a start). Every start,
 * whatever its room floor, is expanded to its 2^16 world seeds and checked by the phase-1 chain at its chunk
 * (req->max_height, req->vflags; the other request fields are ignored); finds in `finds` as in whf_deepdive.
 * stats: structure_seeds = tests = n, starts, walks, spots = starts expanded, and the expansion's counters. */
int whf_dd_terrain(whf_ctx *ctx, const whf_dd_request *req, whf_dd_spot *spots, uint64_t n, whf_sink *finds,
                   whf_dd_stats *stats);

/* Layout walks of given carver seeds, each as if it were a start: status (0 complete, 2 capped), room floor
 * and piece count per seed (mstree's emit, mstree.cpp:17-34, without the piece list). For parity checks. */
int whf_room_floors(whf_ctx *ctx, const uint64_t *carvers, uint64_t n, int32_t *status, int32_t *room_floor,
                    int32_t *pieces);

/* ---- Deep Dive solver: structure seeds whose hut chunk carries a catalogued mineshaft layout ---- */

/* A plan line (hut_lookup_v4_source.cu:1297-1305, data/hunt/58/plan.txt): a layout's carver seed X, the regime of
 * the hut regions on each axis (0: |r| <= 16384, 1: 16384..32768, 2: 32768..58593, taskspace.h:13-19) and the tier,
 * which chunk offsets (a, b) in [0,24)^2 are searched by their 2-adic valuations: 1 "odd" (different, the smaller
 * 0), 2 "u2" (different, the smaller >= 1), 3 "eq" (equal) (solver_fused.cuh:25-33). A bundle is 2^20 groups. */
typedef struct {
    uint64_t layout;            /* X, 48 bits */
    int regime_x, regime_z;     /* 0..2 */
    int tier;                   /* 1..3 */
    int reserved;
} whf_solve_bundle;

/* A solver hit: the reference's dump line "S,cx,cz,X,1" (hut_lookup_v4_source.cu:1341). */
typedef struct {
    uint64_t s48;               /* structure seed */
    uint64_t layout;            /* X */
    int32_t cx, cz;             /* hut chunk = mineshaft chunk */
    uint32_t bundle;            /* index of the plan line (whf_solve), or of the chunk (whf_solve_chunks) */
    uint32_t reserved;
} whf_solve_hit;

typedef struct {
    whf_solve_hit *hits;        /* caller-owned array */
    uint64_t capacity;
    uint64_t count;             /* out: hits found (may exceed capacity -> WHF_EOVERFLOW; the first `capacity`
                                   are stored, in no particular order then) */
} whf_solve_sink;

typedef struct {
    uint64_t groups;            /* groups solved */
    uint64_t regions;           /* hut regions of those groups */
    uint64_t tasks;             /* (region, chunk offset) tasks */
    uint64_t candidates;        /* low-20-bit states that pass the hut and low-bit carver filters, summed per task */
    uint64_t leaves;            /* full 48-bit states the lift reaches (carver equation checked in full) */
    uint64_t hits;
    double device_ms;
} whf_solve_stats;

/* The solve5 solver over groups [first_group, first_group + n_groups) of the bundles (group g is bundle g >> 20,
 * residue permGroup(g & 0xFFFFF), taskspace.h:24; hut_lookup_v4_source.cu:1361-1397): for every hut region of the
 * group (taskspace.h:27-37) and chunk offset of the bundle's tier, every structure seed S whose chunk carver seed is
 * the layout (room_core.h:55-58) and whose region's hut is at that chunk (solver_kernels.cuh:78-82, with the
 * low-20-bit draw filter the reference's candidate pools apply, solver_kernels.cuh:241-255). The hits equal the
 * reference solver's (solver_fused.cuh:37-129); they come back sorted by (bundle, S, cx, cz). Replaces
 * `hut_lookup_v4 --solve-only --solve5` (hut_lookup_v4_source.cu:1293-1403), which hunt.py runs per chunk of groups
 * (hunt.py:947-952, 1031-1059). Not re-entrant per context. */
int whf_solve(whf_ctx *ctx, const whf_solve_bundle *bundles, uint64_t n_bundles, uint64_t first_group,
              uint64_t n_groups, whf_solve_sink *sink, whf_solve_stats *stats);

/* The same search for given chunks of one layout (the reference's --tasks-file mode of the solver,
 * hut_lookup_v4_source.cu:1417-1419): chunk i is (cx[i], cz[i]), region (cx >> 5, cz >> 5); an offset outside
 * [0,24)^2, or (0, 0), which no tier holds, is WHF_EINVAL. Hits carry the chunk index in `bundle`, sorted by
 * (chunk, S). n <= 2^20. */
int whf_solve_chunks(whf_ctx *ctx, uint64_t layout, const int32_t *cx, const int32_t *cz, uint64_t n,
                     whf_solve_sink *sink, whf_solve_stats *stats);

/* ---- Biome stage: the swamp-hut biome check of the multi-noise biome source ---- */

#define WHF_BIOME_SWAMP 6     /* the trees' biome id of minecraft:swamp (biome.h:77) */
#define WHF_BIOME_QY0   (-17) /* quart heights an occupied height can fall in: -17 .. 79 (surface -64 .. 320) */
#define WHF_BIOME_QY_N  97
#define WHF_OCC_SURFACE (-2147483647 - 1)   /* whf_biome_query.occ_y: use the pre-carver surface at the chunk middle */

typedef struct {
    int64_t seed;               /* world seed */
    int32_t hut_x, hut_z;       /* a block of the hut's chunk (16*cx, 16*cz); the check samples the chunk's middle */
    int32_t occ_y;              /* first occupied height at the chunk middle: its WORLD_SURFACE_WG height - 1;
                                   WHF_OCC_SURFACE: computed here from the pre-carver surface (whf_surface_heights) */
    int32_t reserved;
} whf_biome_query;

typedef struct {
    int64_t target[6];          /* quantized temperature, humidity, continentalness, erosion, depth, weirdness
                                   (x 10000, Climate.quantizeCoord) at the query's occupied height */
    int32_t qx, qy, qz;         /* quart position of the check */
    int32_t occ_y;
    int32_t winner;             /* biome id of the tree's nearest leaf at the query's height */
    int32_t ok;                 /* winner == WHF_BIOME_SWAMP */
    uint64_t swamp_qy[2];       /* bit i (word i / 64): the winner at quart height WHF_BIOME_QY0 + i is swamp */
} whf_biome_result;

/* The biome tree a game version uses (biome_trees.h:70501-70526; the rule-set aliases of the command line, such
 * as "1.19.x~26.1", take their rule set's default tree, eval.c:104-106). Returns WHF_EINVAL for unknown names. */
int whf_biome_tree(const char *version, int *tree);

/* The swamp-hut biome check of lysh_swamp_hut_biome_ok (biome.c:94-156) for each query, with the height callback's
 * answer given as occ_y: the six climate parameters at the chunk middle's quart point (terrain.c:22-44 shifted
 * coordinates, column_top.c:124-130 overworld offset for depth), quantized, and the nearest leaf of biome tree
 * `tree` (0..7, Climate.RTree.SubTree.search, biome.c:34-47). Besides the answer at occ_y, each result carries the
 * swamp verdict at every quart height the surface can have; `winners` (may be NULL) receives the winner at each
 * of them, WHF_BIOME_QY_N bytes per query. vflags: WHF_V_LARGE selects the LARGE_BIOMES climate noises;
 * WHF_BIO_OCC_ONLY computes only the answer at occ_y (one tree walk per query instead of 98, for maps; swamp_qy is
 * then 0 and `winners` must be NULL); other bits are ignored. Not re-entrant per context. */
#define WHF_BIO_OCC_ONLY 0x100
int whf_biome(whf_ctx *ctx, const whf_biome_query *queries, uint64_t n, int tree, int vflags, whf_biome_result *out,
              uint8_t *winners);

/* ---- Terrain stage: final density and the pre-carver surface height ---- */

/* Block states of a column's top block. */
#define WHF_BLOCK_SOLID 0
#define WHF_BLOCK_AIR   1
#define WHF_BLOCK_WATER 2
#define WHF_BLOCK_LAVA  3

typedef struct {
    int64_t seed;               /* world seed */
    int32_t x, z;               /* block column */
} whf_column;

/* Pre-carver surface height of each column: one above the first block from y = 319 down that is not air in the
 * overworld's final density and aquifer (lysh_phase2_height_pre_carver, phase2.c:318-334: the interpolated final
 * density of phase2.c:161-194 and computeSubstance of aquifer.c:416-563), -64 for a column of air; this is
 * WORLD_SURFACE_WG before carvers, the height the biome check uses (eval.c:62-67). `state` (may be NULL) receives
 * the WHF_BLOCK_* state of the top block (water for a sea surface). vflags: WHF_V_LARGE selects the LARGE_BIOMES
 * noises. Not re-entrant per context. */
int whf_surface_heights(whf_ctx *ctx, const whf_column *cols, uint64_t n, int vflags, int32_t *height,
                        int32_t *state);

/* Post-carver surface height of each column: the pre-carver surface after the cave and canyon carvers of the 17 x 17
 * chunks around the column's chunk (lysh_phase2_height, phase2.c:438-456, carvers carver.c:498-523); `pre` (may be
 * NULL) receives the pre-carver height. This is the height the hut's floor averages (eval.c:170-219). */
int whf_post_carver_heights(whf_ctx *ctx, const whf_column *cols, uint64_t n, int vflags, int32_t *post,
                            int32_t *pre);

/* ---- Hut evaluation ---- */

#define WHF_HUT_OK           0
#define WHF_HUT_REJECT_Y     1   /* avg_y > max_y */
#define WHF_HUT_REJECT_FLOOD 2   /* all 63 footprint columns hold water above y = 62 */
#define WHF_HUT_REJECT_BIOME 3   /* the chunk middle's biome is not swamp */

typedef struct {
    int64_t seed;
    int32_t cx, cz;             /* the hut's chunk; the hut stands at (16 cx, 16 cz) */
} whf_hut_query;

typedef struct {
    int32_t dir;                /* orientation: nextInt(4) after setLargeFeatureSeed, 0 N 1 E 2 S 3 W */
    int32_t size_x, size_z;     /* footprint, 7 x 9 or 9 x 7 */
    int32_t biome_ok, winner, occ_y;   /* the biome check at the chunk middle's pre-carver surface */
    int32_t wet;                /* footprint columns with water above y = 62 */
    int32_t flooded;            /* wet == 63 */
    int32_t carve_skipped;      /* flooded with all 63 pre-carver top blocks at or below 63: pre-carver sum used */
    int32_t avg_y;              /* sum_h / 63, truncated toward zero: the hut's floor */
    int64_t sum_h;              /* the 63 footprint heights summed */
    int32_t ok, reject;         /* WHF_HUT_OK or the WHF_HUT_REJECT_* reason, in lysh's order */
} whf_hut_result;

/* The reference CPU scanner's verdict on each hut (lysh_eval_hut, eval.c:121-226) with biome tree `tree` and
 * floor threshold max_y: orientation, the biome at the pre-carver surface, the flood test, the average post-carver
 * height of the 63 footprint columns, and the verdict biome_ok && !flooded && avg_y <= max_y. Replaces the `lysh
 * hut` processes hunt.py runs per candidate (hunt.py:1390-1392). vflags: WHF_V_LARGE for the LARGE_BIOMES noises.
 * Not re-entrant per context. */
int whf_hut_eval(whf_ctx *ctx, const whf_hut_query *huts, uint64_t n, int tree, int vflags, int max_y,
                 whf_hut_result *out);

/* ---- Serving: continuous batching of Seed Check requests ---- */

typedef struct whf_server whf_server;

#define WHF_SERVE_SLOTS 1   /* whf_server_config.flags: the host-driven slot scheduler instead of the device loop */

typedef struct {
    int device;
    int slots;                  /* device loop: lanes, streams that run whole batches side by side (3 or 4; 0, 1
                                   and 2 give 3);
                                   WHF_SERVE_SLOTS: concurrent device slots (a context, stream and buffers each);
                                   0 = 4 */
    uint64_t chunk_checks;      /* device loop: (seed, region) checks per batch at most (0 = 2^25, the most a batch
                                   holds; at least 2^16); WHF_SERVE_SLOTS: checks per scheduled chunk (0 = 2^24) */
    uint64_t hbm_budget;        /* device bytes the server may hold; 0 = no limit. Device loop: its survivor and
                                   seed-list pools are cut to fit (WHF_ENOMEM when the fixed part does not);
                                   WHF_SERVE_SLOTS: slots are cut to fit */
    int flags;                  /* WHF_SERVE_* */
    int reserved;
} whf_server_config;

/* Completion of one request. Times are seconds on the server's steady clock (0 = whf_server_open). */
typedef struct {
    uint64_t id;                /* as returned by whf_submit */
    int status;                 /* WHF_OK, WHF_EOVERFLOW (more hits than the request's sink) or an error */
    int reserved;
    whf_stats stats;            /* the request's own counters: the same as a whf_scan of it alone */
    double submit_s, start_s, done_s;
} whf_event;

/* Opens a server. By default it is a device-resident loop: admitted requests live in a request table in device
 * memory, and every batch of the Seed Check chain is packed on the device by a claim kernel that takes seeds from
 * the live requests, least remaining work first (a request live for 8 batches first, at most half a batch of it),
 * so one batch holds tiles of many requests, rects, game versions and precisions, and the lanes run batches side by
 * side. Counters and survivor queues are per request on the device; a host thread only
 * admits requests (when a table entry and space in the survivor and seed-list pools are free: admission against
 * the HBM budget) and keeps batches flowing while the device reports unclaimed work, and another reads completed
 * requests back, runs lysh's phase 2 on their survivors and posts their events. A request the loop cannot finish
 * exactly (its survivor queue or the FP32 gate's re-evaluation queue overflowed) is scanned again alone with
 * whf_scan. Served requests take at most 2^16 regions and fewer than 2^32 seeds (WHF_EINVAL otherwise).
 * With WHF_SERVE_SLOTS: `slots` contexts on the device, each driven by its own host thread. Requests are cut into
 * chunks of about chunk_checks checks; an idle slot takes the next chunk of the request with the least work left
 * (the oldest on ties). Either way a request that arrives while others run starts while they run, small requests
 * overtake big ones, and the work of several requests overlaps on the device. Replaces the one-request-at-a-time
 * `--worker` loop (phase1_scan.cu:1346-1387) that hunt.py's GpuScanWorker drives (hunt.py:1276-1332). */
int whf_server_open(const whf_server_config *cfg, whf_server **out);

/* Finishes the requests in flight, then releases the slots. Events not polled are dropped. */
void whf_server_close(whf_server *srv);

/* Queues a request (copied; for WHF_SEEDS_LIST the seeds array must stay valid until its event). The sink is
 * filled when the request completes, its hits sorted by (seed, rx, rz). With phase 2, a request whose phase-1
 * survivors overflow the library's survivor buffer completes with WHF_EOVERFLOW and an empty sink (count 0), as
 * whf_scan returns WHF_EOVERFLOW without filling the sink. Thread-safe. */
int whf_submit(whf_server *srv, const whf_request *req, whf_sink *sink, uint64_t *id);

/* Queues a background Deep Dive (request copied; see whf_deepdive). Slots run its chunks of 2^18 structure seeds
 * only when no Seed Check work is waiting (device loop: none left to claim; one slot context of its own), and at
 * most half the slots at once, so interactive requests keep their latency while the Deep Dive uses the rest of
 * the device. The sinks and *stats are filled when it
 * completes; its completion comes through whf_poll like a request's (with zero whf_stats). Thread-safe. No
 * counterpart in the reference, which runs Seed Check and Deep Dive as separate programs (hunt.py:58-63). */
int whf_submit_deepdive(whf_server *srv, const whf_dd_request *req, whf_dd_sink *spots, whf_sink *finds,
                        whf_dd_stats *stats, uint64_t *id);

/* Waits up to timeout_ms (-1: no limit) for completions; returns the number of events written (0..max_events)
 * or a negative error code. Thread-safe. */
int whf_poll(whf_server *srv, whf_event *events, int max_events, int timeout_ms);

/* ---- Tuning: launch shapes chosen per device ---- */

/* The launch shapes the library leaves open. whf_open starts from the built-in defaults, then applies the entry for
 * its device (whf_device_name up to " (", or "default") from the JSON file named by the WHF_TUNED
 * environment variable, or else from tuned.json beside libwhf_hip.so, if one exists; low-witch-finder_amd/tune.py
 * measures the candidates on the device and writes that file. No counterpart in the reference's engine, whose launch
 * shapes are constants; hunt.py keeps a tuned configuration per GPU the same way (hunt.py:927-945, 959-1021). Every
 * choice gives the same answers; only the speed changes. A field of 0 means the default. */
typedef struct {
    int ero_wg;                 /* erosion gate workgroup: 256 or 512 lanes (default 256) */
    int ero_ilp;                /* FP32-gated erosion gate: items evaluated at once per lane, 1 or 2 (default 2;
                                   512-lane workgroups take 1) */
    int ero_wg_per_cu;          /* erosion gate: persistent workgroups per compute unit, 1..64, or -1 for one
                                   workgroup per tile (default -1) */
    int gate_wg_per_cu;         /* workgroups per compute unit of the persistent gate-2 and gate-3 kernels, 1..32
                                   (default 16) */
    uint64_t chunk_checks;      /* serving: (seed, region) checks per scheduled chunk when whf_server_config leaves
                                   it 0 (default 2^24) */
    int reserved[8];
} whf_tuning;

/* The context's launch shapes for the version rules of vflags 0 (1.19 - 26.1, NORMAL preset). */
int whf_get_tuning(whf_ctx *ctx, whf_tuning *out);
/* Sets the context's launch shapes for every rule layout (0 fields take the defaults); WHF_EINVAL for a value out of
 * range. */
int whf_set_tuning(whf_ctx *ctx, const whf_tuning *t);
/* The same for the rule layout of one set of WHF_V_* flags (0..15): tuned.json may give each game version its own. */
int whf_get_tuning_for(whf_ctx *ctx, int vflags, whf_tuning *out);
int whf_set_tuning_for(whf_ctx *ctx, int vflags, const whf_tuning *t);
/* Reads a tuned-configuration file without a device: the shapes whf_open would apply on `device` for vflags, from
 * `path` (NULL: $WHF_TUNED, else tuned.json beside the library). File format: {"<device name>" or "default": {"build":
 * "<whf_build_id>", "ero_wg": ..., "ero_ilp": ..., "ero_wg_per_cu": ..., "gate_wg_per_cu": ..., "chunk_checks": ...,
 * "versions": {"<version name>": {<fields>}}, <other keys, ignored>}}. Returns WHF_OK when an entry measured with this
 * build applied, else WHF_EINVAL with the defaults in *out. */
int whf_load_tuning(const char *path, const char *device, int vflags, whf_tuning *out);

/* Version name -> WHF_V_* flags, accepting lysh's version names (main.c:104-125, biome_trees.h:70499-70526). */
int whf_version_flags(const char *name, int *vflags);

/* Text of a return code. */
const char *whf_strerror(int code);

/* Identifier of this library build (a digest of its sources): tuned.json entries record the build they were
 * measured with, and whf_open applies only an entry measured with its own build. */
const char *whf_build_id(void);

/* Name of the device the context runs on (for logs), or "" when unknown. */
const char *whf_device_name(whf_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* WHF_H */
