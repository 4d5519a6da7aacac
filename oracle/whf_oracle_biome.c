/* whf_oracle_biome.c — CPU oracle of the biome stage: the terrain point at a quart position, the overworld
 * offset spline, the six quantized climate parameters and the per-version biome tree search.
 *
 * TEST INFRASTRUCTURE ONLY (see whf_oracle.h). Restates, in plain C:
 *   terrain.c:22-44          shifted coordinates and the three terrain noises at a quart point
 *   spline.c:8-62, 63-70     the nested cubic spline (FP32) and the normalized weirdness
 *   spline.c:95-111          the offset / factor / peaks roots
 *   column_top.c:57-77,124-130  the overworld offset in double (blend alpha 1, blend offset 0)
 *   biome.c:23-53            node distance and the R-tree search (Climate.RTree.SubTree.search)
 *   biome.c:71-80            quantization and the depth gradient
 *   biome.c:94-156           the swamp-hut biome check at the chunk middle for a given occupied height
 * The trees and the spline nodes are the game's data (whf_game_data.h), shared with the HIP library.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "whf_oracle.h"
#include "whf_oracle_tables.h"
#include "../low-witch-finder_amd/csrc/whf_game_data.h"

void wo_biome_init(wo_biome *b, uint64_t world_seed, int vflags) {
    /* terrain.c:4-14 (shift, continentalness, erosion, ridge) and biome.c:82-90 (temperature, vegetation);
     * LARGE_BIOMES swaps in the _large climate noises, the shift and ridge noises stay */
    const int large = (vflags & WO_V_LARGE) != 0;
    wo_noise_init(&b->offset, world_seed, WO_N_OFFSET);
    wo_noise_init(&b->continentalness, world_seed, large ? WO_N_CONTINENTALNESS_LARGE : WO_N_CONTINENTALNESS);
    wo_noise_init(&b->erosion, world_seed, large ? WO_N_EROSION_LARGE : WO_N_EROSION);
    wo_noise_init(&b->ridge, world_seed, WO_N_RIDGE);
    wo_noise_init(&b->temperature, world_seed, large ? WO_N_TEMPERATURE_LARGE : WO_N_TEMPERATURE);
    wo_noise_init(&b->vegetation, world_seed, large ? WO_N_VEGETATION_LARGE : WO_N_VEGETATION);
}

void wo_terrain_info_at(const wo_biome *b, int qx, int qz, wo_terrain_info *out) {
    /* terrain.c:18-31: shift = offset noise x 4 at (qx, 0, qz) for x and at (qz, qx, 0) for z */
    const double sx = wo_noise_sample(&b->offset, (double)qx, 0.0, (double)qz) * 4.0;
    const double sz = wo_noise_sample(&b->offset, (double)qz, (double)qx, 0.0) * 4.0;
    out->shifted_x = (double)qx + sx;
    out->shifted_z = (double)qz + sz;
    out->continentalness = wo_noise_sample_y0(&b->continentalness, out->shifted_x, out->shifted_z);
    out->weirdness = wo_noise_sample_y0(&b->ridge, out->shifted_x, out->shifted_z);
    out->erosion = wo_noise_sample_y0(&b->erosion, out->shifted_x, out->shifted_z);
}

/* ---- spline (FP32 throughout) ---- */

static int first_above(const float *loc, int n, float v) {
    /* spline.c:8-23: first index whose location is above v, n when none; NaN finds none */
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (v < loc[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

float wo_spline_apply(int node, const float p[4]) {
    const whf_gd_spline_node *sn = &whf_gd_spline_nodes[node];
    if (sn->kind == 1) return sn->constant;
    const float *loc = &whf_gd_spline_floats[sn->loc];
    const float *der = &whf_gd_spline_floats[sn->der];
    const int32_t *kid = &whf_gd_spline_child[sn->child];
    const float f = p[sn->coord];
    const int last = sn->n - 1;
    const int i = first_above(loc, sn->n, f) - 1;
    /* outside the locations: the end child extended linearly with its derivative (spline.c:40-45) */
    if (i < 0) return wo_spline_apply(kid[0], p) + der[0] * (f - loc[0]);
    if (i == last) return wo_spline_apply(kid[last], p) + der[last] * (f - loc[last]);
    /* cubic Hermite between locations i and i + 1 (spline.c:47-61) */
    const float x0 = loc[i], x1 = loc[i + 1];
    const float t = (f - x0) / (x1 - x0);
    const float y0 = wo_spline_apply(kid[i], p);
    const float y1 = wo_spline_apply(kid[i + 1], p);
    const float a = der[i] * (x1 - x0) - (y1 - y0);
    const float bq = -der[i + 1] * (x1 - x0) + (y1 - y0);
    const float base = y0 + t * (y1 - y0);
    const float bend = a + t * (bq - a);
    return base + t * (1.0f - t) * bend;
}

float wo_normalized_weirdness(float w) {
    /* spline.c:63-70: -(||w| - 2/3| - 1/3) * 3 in float */
    const float a = w < 0.0f ? -w : w;
    const float b = a - 0.6666667f;
    const float c = b < 0.0f ? -b : b;
    const float d = c - 0.33333334f;
    return -d * 3.0f;
}

void wo_terrain_spline(float c, float e, float w, float out[3]) {
    /* spline.c:83-111: offset (with its -0.50375 shift in float), factor and peaks */
    const float p[4] = { c, e, wo_normalized_weirdness(w), w };
    out[0] = wo_spline_apply(whf_gd_spline_root[0], p) + (-0.50375f);
    out[1] = wo_spline_apply(whf_gd_spline_root[1], p);
    out[2] = wo_spline_apply(whf_gd_spline_root[2], p);
}

double wo_overworld_offset(const wo_biome *b, int block_x, int block_z) {
    /* column_top.c:124-130 over the terrain point at the quart position (column_top.c:80-83), then
     * ps_offset_factor (column_top.c:57-77): the raw spline in float, the blend in double */
    wo_terrain_info ti;
    wo_terrain_info_at(b, block_x >> 2, block_z >> 2, &ti);
    const float c = (float)ti.continentalness, e = (float)ti.erosion, w = (float)ti.weirdness;
    const float p[4] = { c, e, wo_normalized_weirdness(w), w };
    const double raw = (double)wo_spline_apply(whf_gd_spline_root[0], p);
    const double alpha = 1.0, blend = 0.0;
    return blend * (1.0 + -1.0 * alpha) + (-0.5037500262260437 + raw) * alpha;
}

/* ---- the biome tree ---- */

static const whf_gd_node *tree_nodes(int tree) { return &whf_gd_nodes[whf_gd_tree_base[tree]]; }

int64_t wo_tree_dist(int tree, int node, const int64_t np[6]) {
    /* biome.c:23-31: per parameter the distance to the node's range (0 inside), squared and summed */
    const whf_gd_node *g = &tree_nodes(tree)[node];
    int64_t sum = 0;
    for (int k = 0; k < 6; k++) {
        const int64_t above = np[k] - g->range[2 * k + 1];
        const int64_t below = (int64_t)g->range[2 * k] - np[k];
        const int64_t d = above > 0 ? above : (below > 0 ? below : 0);
        sum += d * d;
    }
    return sum;
}

static int tree_walk(int tree, int node, const int64_t np[6], int leaf) {
    /* biome.c:34-47: a child is entered only when its box is nearer than the best leaf so far; the best leaf
     * is handed down, and a child's answer replaces it only when strictly nearer */
    const whf_gd_node *g = tree_nodes(tree);
    if (g[node].kids == 0) return node;
    int64_t best = leaf < 0 ? INT64_MAX : wo_tree_dist(tree, leaf, np);
    int result = leaf;
    int c = node + 1;
    for (int k = 0; k < g[node].kids; k++, c = g[c].end) {
        const int64_t d = wo_tree_dist(tree, c, np);
        if (best > d) {
            const int got = tree_walk(tree, c, np, result);
            const int64_t dg = got == c ? d : wo_tree_dist(tree, got, np);
            if (best > dg) { best = dg; result = got; }
        }
    }
    return result;
}

int wo_tree_leaf(int tree, const int64_t np[6]) {
    if (tree < 0 || tree >= WHF_GD_TREES) return -1;
    return tree_walk(tree, 0, np, -1);
}

int wo_tree_biome(int tree, const int64_t np[6]) {
    const int leaf = wo_tree_leaf(tree, np);
    return leaf < 0 ? -1 : tree_nodes(tree)[leaf].biome;
}

int64_t wo_tree_min_dist(int tree, const int64_t np[6], int *n_min) {
    /* brute force over every leaf: the smallest distance and how many leaves share it */
    const whf_gd_node *g = tree_nodes(tree);
    const int n = (int)(whf_gd_tree_base[tree + 1] - whf_gd_tree_base[tree]);
    int64_t best = INT64_MAX;
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        if (g[i].kids) continue;
        const int64_t d = wo_tree_dist(tree, i, np);
        if (d < best) { best = d; cnt = 1; }
        else if (d == best) cnt++;
    }
    if (n_min) *n_min = cnt;
    return best;
}

/* ---- the swamp-hut biome check ---- */

static int64_t quantize(float f) { return (int64_t)(f * 10000.0f); }   /* biome.c:71-73, product in float */

double wo_depth_gradient(double y) {
    /* biome.c:77-80: clamp((y + 64) / 384, 0, 1) mapped linearly from 1.5 to -1.5 */
    double t = (y - -64.0) / (320.0 - -64.0);
    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
    return 1.5 + t * (-1.5 - 1.5);
}

static void targets_at(const wo_biome *b, int qx, int qz, int64_t t[6], double *offset) {
    wo_terrain_info ti;
    wo_terrain_info_at(b, qx, qz, &ti);
    const double temperature = wo_noise_sample_y0(&b->temperature, ti.shifted_x, ti.shifted_z);
    const double humidity = wo_noise_sample_y0(&b->vegetation, ti.shifted_x, ti.shifted_z);
    t[0] = quantize((float)temperature);
    t[1] = quantize((float)humidity);
    t[2] = quantize((float)ti.continentalness);
    t[3] = quantize((float)ti.erosion);
    t[4] = 0;
    t[5] = quantize((float)ti.weirdness);
    *offset = wo_overworld_offset(b, qx * 4, qz * 4);
}

int wo_biome_check(const wo_biome *b, int tree, int hut_x, int hut_z, int occ_y, wo_biome_result *out) {
    /* biome.c:94-156 with the height callback's answer given: occ_y = surface height - 1 at the chunk middle */
    memset(out, 0, sizeof *out);
    out->winner = -1;
    if (tree < 0 || tree >= WHF_GD_TREES) return 0;
    const int mid_x = (hut_x >> 4) * 16 + 8, mid_z = (hut_z >> 4) * 16 + 8;
    const int qx = mid_x >> 2, qy = occ_y >> 2, qz = mid_z >> 2;
    double offset;
    targets_at(b, qx, qz, out->t, &offset);
    const double depth = wo_depth_gradient((double)(qy * 4)) + offset;
    out->t[4] = quantize((float)depth);
    out->qx = qx;
    out->qy = qy;
    out->qz = qz;
    out->occ_y = occ_y;
    out->winner = wo_tree_biome(tree, out->t);
    out->ok = out->winner == WHF_GD_SWAMP;
    return out->ok;
}

void wo_biome_heights(const wo_biome *b, int tree, int hut_x, int hut_z, uint8_t winners[WO_BIOME_QY_N]) {
    /* the winner at every quart height an occupied height can fall in, qy = WO_BIOME_QY0 + i */
    const int mid_x = (hut_x >> 4) * 16 + 8, mid_z = (hut_z >> 4) * 16 + 8;
    int64_t t[6];
    double offset;
    targets_at(b, mid_x >> 2, mid_z >> 2, t, &offset);
    for (int i = 0; i < WO_BIOME_QY_N; i++) {
        const int qy = WO_BIOME_QY0 + i;
        t[4] = quantize((float)(wo_depth_gradient((double)(qy * 4)) + offset));
        winners[i] = (uint8_t)wo_tree_biome(tree, t);
    }
}

int wo_biome_tree(const char *version, int *tree) {
    /* a version with a tree of its own uses it (main.c:104-108); the rule-set aliases take the default of their
     * rule set, as lysh_search_ctx_init does (eval.c:104-106): 1.18.2, 1.21.5 or 26.2 */
    for (int i = 0; i < whf_gd_n_versions; i++)
        if (!strcmp(whf_gd_versions[i].name, version)) { *tree = whf_gd_versions[i].tree; return 0; }
    int vf;
    if (wo_version_flags(version, &vf)) return 1;
    const char *dflt = (vf & WO_V_118) ? "1.18.2" : ((vf & WO_V_262) ? "26.2" : "1.21.5");
    for (int i = 0; i < whf_gd_n_versions; i++)
        if (!strcmp(whf_gd_versions[i].name, dflt)) { *tree = whf_gd_versions[i].tree; return 0; }
    return 1;
}

void wo_biome_point(const wo_biome *b, int qx, int qz, double out[8]) {
    /* the doubles the check is built from, at a quart point: terrain info, temperature, humidity, offset */
    wo_terrain_info ti;
    wo_terrain_info_at(b, qx, qz, &ti);
    out[0] = ti.shifted_x;
    out[1] = ti.shifted_z;
    out[2] = ti.continentalness;
    out[3] = ti.weirdness;
    out[4] = ti.erosion;
    out[5] = wo_noise_sample_y0(&b->temperature, ti.shifted_x, ti.shifted_z);
    out[6] = wo_noise_sample_y0(&b->vegetation, ti.shifted_x, ti.shifted_z);
    out[7] = wo_overworld_offset(b, qx * 4, qz * 4);
}

void *wo_biome_alloc(uint64_t world_seed, int vflags) {
    wo_biome *b = (wo_biome *)malloc(sizeof *b);
    if (b) wo_biome_init(b, world_seed, vflags);
    return b;
}
