/* whf_game_data.h — game data of the biome and terrain stages: Minecraft's overworld biome search trees (one per
 * game version that changed them), the overworld terrain spline and the game's sine table.
 *
 * Both are data the game defines; the values must be the reference's node for node, so they are not retyped
 * here. gen_game_data.py reads them at build time from the reference's generated tables
 * (engine/tools/lysh/src/biome_trees.h: 8 trees of {int16 range[12], kids, biome, end} plus the version table
 * :70499-70526; engine/tools/lysh/src/terrain_table.h: 115 spline nodes and their float/child pools;
 * engine/tools/lysh/src/sin_table.h: 65536 floats) and writes
 * gen/whf_game_data.c in the layout below. The same object links into the HIP library and into the CPU oracle.
 */
#ifndef WHF_GAME_DATA_H
#define WHF_GAME_DATA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WHF_GD_TREES 8
#define WHF_GD_SWAMP 6            /* the tree's biome id of minecraft:swamp */

/* One R-tree node, 32 bytes: six (min, max) parameter ranges in quantized units (x 10000), the index just past
 * the node's subtree (tree-local), the child count (0 = leaf) and the leaf's biome id. */
typedef struct {
    int16_t range[12];
    uint16_t end;
    uint8_t kids;
    uint8_t biome;
    uint32_t pad;
} whf_gd_node;

/* Nodes of all trees, concatenated; tree t holds nodes [tree_base[t], tree_base[t + 1]). */
extern const whf_gd_node whf_gd_nodes[];
extern const uint32_t whf_gd_tree_base[WHF_GD_TREES + 1];
extern const int whf_gd_tree_depth;       /* nodes from the root to the deepest leaf, over all trees */

/* Version name -> tree and the phase-1 rule set that version uses. */
typedef struct {
    const char *name;
    int tree, is_118, pre_262;
} whf_gd_version;
extern const whf_gd_version whf_gd_versions[];
extern const int whf_gd_n_versions;

/* Spline node: kind 0 = piecewise cubic over `n` locations of coordinate `coord` (0 continentalness,
 * 1 erosion, 2 normalized weirdness, 3 weirdness) with locations, derivatives and children at the given pool
 * offsets; kind 1 = the constant. */
typedef struct {
    int32_t kind, coord, n, loc, der, child;
    float constant;
} whf_gd_spline_node;

extern const whf_gd_spline_node whf_gd_spline_nodes[];
extern const int whf_gd_n_spline_nodes;
extern const float whf_gd_spline_floats[];
extern const int whf_gd_n_spline_floats;
extern const int32_t whf_gd_spline_child[];
extern const int whf_gd_n_spline_child;
extern const int whf_gd_spline_root[3];   /* offset, factor, peaks */
extern const int whf_gd_spline_depth;     /* nesting depth of the deepest root (1 = a constant) */

/* Mth.SIN as the game's JVM computed it (sin_table.h): sin(a) = SIN[(long)(a * 10430.378350470453) & 65535],
 * cos(a) the same at + 16384.0. The carvers' only trigonometry. */
extern const float whf_gd_mc_sin[65536];

#ifdef __cplusplus
}
#endif
#endif /* WHF_GAME_DATA_H */
