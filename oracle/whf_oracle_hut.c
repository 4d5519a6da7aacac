/* whf_oracle_hut.c — CPU oracle of the carvers and the hut evaluation: the cave and canyon carvers of the 17 x 17
 * chunks around a target chunk, the post-carver surface height, and lysh's hut verdict (biome, flood, average
 * floor).
 *
 * TEST INFRASTRUCTURE ONLY (see whf_oracle.h). Restates, in plain C (engine/tools/lysh/src):
 *   rng.c:186-231, structure.c:26-42  java.util.Random and setLargeFeatureSeed
 *   carver.c:53-523                   Mth.sin/cos, carveBlock, carveEllipsoid, the cave and canyon carvers, and
 *                                     applyCarvers over the 17 x 17 neighbourhood
 *   phase2.c:337-345, 438-456         the post-carver block state and surface height
 *   eval.c:41-50, 121-226             the hut's orientation and its evaluation
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "whf_oracle.h"
#include "whf_oracle_int.h"
#include "../low-witch-finder_amd/csrc/whf_game_data.h"

/* ---------------------------------------------------------------- java.util.Random */

#define JR_MUL UINT64_C(0x5DEECE66D)
#define JR_M48 UINT64_C(0xFFFFFFFFFFFF)

typedef struct { uint64_t s; } jr;

static void jr_seed(jr *r, uint64_t seed) { r->s = (seed ^ JR_MUL) & JR_M48; }
static int jr_next(jr *r, int bits) {
    r->s = (r->s * JR_MUL + 0xB) & JR_M48;
    return (int)(r->s >> (48 - bits));
}
static int jr_int(jr *r, int bound) {
    /* rng.c:214-231 */
    if (bound <= 0) return 0;
    if ((bound & (bound - 1)) == 0) return (int)(((int64_t)bound * (int64_t)jr_next(r, 31)) >> 31);
    int u = jr_next(r, 31), v;
    for (;;) {
        v = u % bound;
        if ((int32_t)((uint32_t)u - (uint32_t)v + (uint32_t)(bound - 1)) >= 0) break;
        u = jr_next(r, 31);
    }
    return v;
}
static uint64_t jr_long(jr *r) {
    /* ((long)next(32) << 32) + next(32), the low word sign-extended (rng.c:197-211) */
    const uint64_t hi = (uint64_t)(uint32_t)jr_next(r, 32);
    const int64_t lo = (int64_t)jr_next(r, 32);
    return (hi << 32) + (uint64_t)lo;
}
static float jr_float(jr *r) { return (float)jr_next(r, 24) * 5.9604645e-8f; }
static void jr_large_feature(jr *r, uint64_t seed, int cx, int cz) {
    /* setLargeFeatureSeed (structure.c:26-42) */
    jr_seed(r, seed);
    const uint64_t a = jr_long(r), b = jr_long(r);
    jr_seed(r, (uint64_t)((int64_t)cx * (int64_t)a) ^ (uint64_t)((int64_t)cz * (int64_t)b) ^ seed);
}

/* ---------------------------------------------------------------- Mth.sin / Mth.cos (carver.c:46-59) */

static float mc_sin(double a) { return whf_gd_mc_sin[(int)((int64_t)(a * 10430.378350470453) & 65535)]; }
static float mc_cos(double a) { return whf_gd_mc_sin[(int)((int64_t)(a * 10430.378350470453 + 16384.0) & 65535)]; }

/* ---------------------------------------------------------------- carving into one target chunk */

enum { C_UNCHANGED = 0, C_AIR = 1, C_WATER = 2, C_LAVA = 3 };
enum { S_SOLID = 0, S_AIR = 1, S_WATER = 2, S_LAVA = 3 };
#define CELL(lx, y, lz) ((((lz) & 15) << 4) | ((lx) & 15) | (((y) + 64) << 8))

typedef struct {
    wo_p2_chunk *ch;
    int min_x, min_z;
    double mid_x, mid_z;
    uint8_t *cells;          /* 16 x 16 x 384 carve results */
    uint64_t mask[1536];     /* CarvingMask: cells an ellipsoid has claimed */
} carve_ctx;

typedef struct {
    int canyon;
    double floor_level;      /* caves */
    const float *width;      /* canyons: width factors by y - minY */
} carve_skip;

static int skip_cell(const carve_skip *sk, double nx, double ny, double nz, int y) {
    /* CaveWorldCarver / CanyonWorldCarver.shouldSkip (carver.c:182-198) */
    if (!sk->canyon) {
        if (ny <= sk->floor_level) return 1;
        return (nx * nx + ny * ny + nz * nz) >= 1.0;
    }
    return (nx * nx + nz * nz) * (double)sk->width[y - -64 - 1] + ny * ny / 6.0 >= 1.0;
}

static int carve_block(carve_ctx *c, int lx, int y, int lz) {
    /* carveBlock (carver.c:145-175): only solid or water is replaced; below y = -56 with lava, elsewhere with
     * the aquifer's substance at density 0 (nothing when that is solid) */
    const int x = c->min_x + lx, z = c->min_z + lz;
    const uint8_t v = c->cells[CELL(lx, y, lz)];
    const int st = v == C_AIR ? S_AIR : (v == C_WATER ? S_WATER : (v == C_LAVA ? S_LAVA : wo_p2_chunk_state(c->ch, x, y, z)));
    if (!(st == S_SOLID || st == S_WATER)) return 0;
    int cs;
    if (y <= -56) {
        cs = S_LAVA;
    } else {
        cs = wo_p2_chunk_substance(c->ch, x, y, z, 0.0);
        if (cs == S_SOLID) return 0;
    }
    c->cells[CELL(lx, y, lz)] = cs == S_AIR ? C_AIR : (cs == S_WATER ? C_WATER : C_LAVA);
    return 1;
}

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

static void carve_ellipsoid(carve_ctx *c, const carve_skip *sk, double x, double y, double z, double hr, double vr) {
    /* carveEllipsoid (carver.c:202-245): an early out on the target chunk's middle, then the unit-ball cells of
     * the chunk, top-down per column; an ellipsoid claims a cell (mask) before it tries to carve it */
    const double reach = 16.0 + hr * 2.0;
    if (fabs(x - c->mid_x) > reach || fabs(z - c->mid_z) > reach) return;
    const int i0 = imax((int)floor(x - hr) - c->min_x - 1, 0);
    const int i1 = imin((int)floor(x + hr) - c->min_x, 15);
    const int k0 = imax((int)floor(y - vr) - 1, -64 + 1);
    const int k1 = imin((int)floor(y + vr) + 1, -64 + 384 - 1 - 7);
    const int n0 = imax((int)floor(z - hr) - c->min_z - 1, 0);
    const int n1 = imin((int)floor(z + hr) - c->min_z, 15);
    for (int p = i0; p <= i1; p++) {
        const double nx = ((double)(c->min_x + p) + 0.5 - x) / hr;
        for (int s = n0; s <= n1; s++) {
            const double nz = ((double)(c->min_z + s) + 0.5 - z) / hr;
            if (nx * nx + nz * nz >= 1.0) continue;
            for (int v = k1; v > k0; v--) {
                const double ny = ((double)v - 0.5 - y) / vr;
                if (skip_cell(sk, nx, ny, nz, v)) continue;
                const unsigned idx = (unsigned)CELL(p, v, s);
                if ((c->mask[idx >> 6] >> (idx & 63)) & 1u) continue;
                c->mask[idx >> 6] |= (uint64_t)1 << (idx & 63);
                carve_block(c, p, v, s);
            }
        }
    }
}

static float uniform_float(jr *r, float lo, float hi) { return jr_float(r) * (hi - lo) + lo; }

static void cave_tunnel(carve_ctx *c, const carve_skip *sk, uint64_t seed, double x, double y, double z, double hrm,
                        double vrm, float thickness, float yaw, float pitch, int first, int steps, double y_scale) {
    /* CaveWorldCarver.createTunnel (carver.c:268-329) */
    jr r;
    jr_seed(&r, seed);
    const int branch = jr_int(&r, steps / 2) + steps / 4;
    const int wide = jr_int(&r, 6) == 0;
    float dyaw = 0.0f, dpitch = 0.0f;
    for (int i = first; i < steps; i++) {
        const float ang = (float)i * 3.1415927f / (float)steps;
        const double radius = 1.5 + (double)(mc_sin((double)ang) * thickness);
        const double vert = radius * y_scale;
        const float cp = mc_cos((double)pitch);
        x += (double)(mc_cos((double)yaw) * cp);
        y += (double)mc_sin((double)pitch);
        z += (double)(mc_sin((double)yaw) * cp);
        pitch *= wide ? 0.92f : 0.7f;
        pitch += dpitch * 0.1f;
        yaw += dyaw * 0.1f;
        dpitch *= 0.9f;
        dyaw *= 0.75f;
        {
            const float a1 = jr_float(&r), b1 = jr_float(&r), c1 = jr_float(&r);
            dpitch += (a1 - b1) * c1 * 2.0f;
            const float a2 = jr_float(&r), b2 = jr_float(&r), c2 = jr_float(&r);
            dyaw += (a2 - b2) * c2 * 4.0f;
        }
        if (i == branch && thickness > 1.0f) {
            const uint64_t s1 = jr_long(&r);
            const float t1 = jr_float(&r) * 0.5f + 0.5f;
            cave_tunnel(c, sk, s1, x, y, z, hrm, vrm, t1, yaw - 1.5707964f, pitch / 3.0f, i, steps, 1.0);
            const uint64_t s2 = jr_long(&r);
            const float t2 = jr_float(&r) * 0.5f + 0.5f;
            cave_tunnel(c, sk, s2, x, y, z, hrm, vrm, t2, yaw + 1.5707964f, pitch / 3.0f, i, steps, 1.0);
            return;
        }
        if (jr_int(&r, 4) == 0) continue;
        {
            const double dx = x - c->mid_x, dz = z - c->mid_z, dy = (double)(steps - i);
            const double rr = (double)(thickness + 2.0f + 16.0f);
            if (!(dx * dx + dz * dz - dy * dy <= rr * rr)) return;
        }
        carve_ellipsoid(c, sk, x, y, z, radius * hrm, vert * vrm);
    }
}

static void cave_carve(carve_ctx *c, jr *r, int ox, int oz, int y_hi) {
    /* CaveWorldCarver.carve (carver.c:332-372) */
    const int caves = jr_int(r, jr_int(r, jr_int(r, 15) + 1) + 1);
    for (int k = 0; k < caves; k++) {
        const double x = (double)(ox * 16 + jr_int(r, 16));
        const int y = jr_int(r, y_hi - -56 + 1) + -56;
        const double z = (double)(oz * 16 + jr_int(r, 16));
        const double hrm = (double)uniform_float(r, 0.7f, 1.4f);
        const double vrm = (double)uniform_float(r, 0.8f, 1.3f);
        carve_skip sk;
        sk.canyon = 0;
        sk.floor_level = (double)uniform_float(r, -1.0f, -0.4f);
        sk.width = NULL;
        int tunnels = 1;
        if (jr_int(r, 4) == 0) {
            const double room_scale = (double)uniform_float(r, 0.1f, 0.9f);
            const float f = 1.0f + jr_float(r) * 6.0f;
            const double d17 = 1.5 + (double)(mc_sin(1.5707963705062866) * f);
            carve_ellipsoid(c, &sk, x + 1.0, (double)y, z, d17, d17 * room_scale);
            tunnels += jr_int(r, 4);
        }
        for (int t = 0; t < tunnels; t++) {
            const float yaw = jr_float(r) * 6.2831855f;
            const float pitch = (jr_float(r) - 0.5f) / 4.0f;
            /* getThickness (carver.c:250-258), its draws in Java's left-to-right order */
            const float t0 = jr_float(r), t1 = jr_float(r);
            float thick = t0 * 2.0f + t1;
            if (jr_int(r, 10) == 0) {
                const float a = jr_float(r), b = jr_float(r);
                thick = thick * (a * b * 3.0f + 1.0f);
            }
            const int steps = 112 - jr_int(r, 112 / 4);
            const uint64_t seed = jr_long(r);
            cave_tunnel(c, &sk, seed, x, (double)y, z, hrm, vrm, thick, yaw, pitch, 0, steps, 1.0);
        }
    }
}

static void canyon_carve(carve_ctx *c, jr *r, int ox, int oz) {
    /* CanyonWorldCarver.carve and doCarve (carver.c:376-494) */
    const double x0 = (double)(ox * 16 + jr_int(r, 16));
    const int y0 = jr_int(r, 67 - 10 + 1) + 10;
    const double z0 = (double)(oz * 16 + jr_int(r, 16));
    float yaw = jr_float(r) * 6.2831855f;
    float pitch = uniform_float(r, -0.125f, 0.125f);
    const double y_scale = 3.0;
    float thick;
    {
        /* TrapezoidFloat(0, 6, plateau 2) */
        const float d = 6.0f - 0.0f, e = (d - 2.0f) / 2.0f, f = d - e;
        const float a = jr_float(r) * f + 0.0f;
        thick = a + jr_float(r) * e;
    }
    const float dist = uniform_float(r, 0.75f, 1.0f);
    const int steps = (int)((float)112 * dist);
    const uint64_t seed = jr_long(r);

    jr q;
    jr_seed(&q, seed);
    float width[384];
    float fw = 1.0f;
    for (int i = 0; i < 384; i++) {
        if (i == 0 || jr_int(&q, 3) == 0) {
            const float a = jr_float(&q), b = jr_float(&q);
            fw = 1.0f + a * b;
        }
        width[i] = fw * fw;
    }
    carve_skip sk;
    sk.canyon = 1;
    sk.floor_level = 0.0;
    sk.width = width;
    double x = x0, y = (double)y0, z = z0;
    float dyaw = 0.0f, dpitch = 0.0f;
    for (int i = 0; i < steps; i++) {
        const float ang = (float)i * 3.1415927f / (float)steps;
        double radius = 1.5 + (double)(mc_sin((double)ang) * thick);
        double vert = radius * y_scale;
        radius *= (double)uniform_float(&q, 0.75f, 1.0f);
        {
            const float f7 = 1.0f - fabsf(0.5f - (float)i / (float)steps) * 2.0f;
            const float f8 = 0.0f + 1.0f * f7;
            vert = (double)f8 * vert * (double)uniform_float(&q, 0.75f, 1.0f);
        }
        const float cp = mc_cos((double)pitch), sp = mc_sin((double)pitch);
        x += (double)(mc_cos((double)yaw) * cp);
        y += (double)sp;
        z += (double)(mc_sin((double)yaw) * cp);
        pitch *= 0.7f;
        pitch += dpitch * 0.05f;
        yaw += dyaw * 0.05f;
        dpitch *= 0.8f;
        dyaw *= 0.5f;
        {
            const float a1 = jr_float(&q), b1 = jr_float(&q), c1 = jr_float(&q);
            dpitch += (a1 - b1) * c1 * 2.0f;
            const float a2 = jr_float(&q), b2 = jr_float(&q), c2 = jr_float(&q);
            dyaw += (a2 - b2) * c2 * 4.0f;
        }
        if (jr_int(&q, 4) == 0) continue;
        {
            const double dx = x - c->mid_x, dz = z - c->mid_z, dy = (double)(steps - i);
            const double rr = (double)(thick + 2.0f + 16.0f);
            if (!(dx * dx + dz * dz - dy * dy <= rr * rr)) return;
        }
        carve_ellipsoid(c, &sk, x, y, z, radius, vert);
    }
}

static void carve_chunk(wo_p2_chunk *ch, uint64_t seed, uint8_t *cells) {
    /* applyCarvers (carver.c:498-523): origins dx outer, dz inner over -8..8; per origin the cave (0.15), extra
     * underground cave (0.07) and canyon (0.01) carvers, each seeded setLargeFeatureSeed(seed + index, origin) */
    carve_ctx *c = (carve_ctx *)calloc(1, sizeof *c);
    c->ch = ch;
    c->min_x = ch->cx * 16;
    c->min_z = ch->cz * 16;
    c->mid_x = (double)(c->min_x + 8);
    c->mid_z = (double)(c->min_z + 8);
    c->cells = cells;
    memset(cells, C_UNCHANGED, 16 * 16 * 384);
    for (int dx = -8; dx <= 8; dx++)
        for (int dz = -8; dz <= 8; dz++) {
            const int ox = ch->cx + dx, oz = ch->cz + dz;
            for (int ci = 0; ci < 3; ci++) {
                const float prob = ci == 0 ? 0.15f : (ci == 1 ? 0.07f : 0.01f);
                jr r;
                jr_large_feature(&r, seed + (uint64_t)ci, ox, oz);
                const float f = jr_float(&r);
                if (!(f <= prob)) continue;
                if (ci == 2) canyon_carve(c, &r, ox, oz);
                else cave_carve(c, &r, ox, oz, ci == 0 ? 180 : 47);
            }
        }
    free(c);
}

static int post_height(wo_p2_chunk *ch, const uint8_t *cells, int x, int z) {
    /* phase2.c:337-345, 438-456: the carved state where a carver wrote, the block state elsewhere */
    for (int y = 319; y >= -64; y--) {
        const uint8_t v = cells[CELL(x & 15, y, z & 15)];
        const int st = v == C_AIR ? S_AIR : (v == C_WATER ? S_WATER : (v == C_LAVA ? S_LAVA : wo_p2_chunk_state(ch, x, y, z)));
        if (st != S_AIR) return y + 1;
    }
    return -64;
}

int wo_post_carver_height(const wo_phase2 *p, uint64_t seed, int x, int z) {
    wo_p2_chunk *ch = (wo_p2_chunk *)malloc(sizeof *ch);
    uint8_t *cells = (uint8_t *)malloc(16 * 16 * 384);
    wo_p2_chunk_init(ch, p, x >> 4, z >> 4);
    carve_chunk(ch, seed, cells);
    const int h = post_height(ch, cells, x, z);
    free(cells);
    free(ch);
    return h;
}

int wo_hut_orientation(uint64_t seed, int hut_x, int hut_z) {
    /* eval.c:41-50: nextInt(4) after setLargeFeatureSeed(seed, chunk) */
    jr r;
    jr_large_feature(&r, seed, hut_x >> 4, hut_z >> 4);
    return jr_int(&r, 4);
}

int wo_hut_eval(const wo_phase2 *p, int tree, uint64_t seed, int hut_x, int hut_z, int max_y, wo_hut_result *out) {
    /* lysh_eval_hut (eval.c:121-226) */
    memset(out, 0, sizeof *out);
    out->dir = wo_hut_orientation(seed, hut_x, hut_z);
    const int axis_z = out->dir == 0 || out->dir == 2;
    out->size_x = axis_z ? 7 : 9;
    out->size_z = axis_z ? 9 : 7;

    /* the biome at the chunk middle's pre-carver surface (eval.c:141-143) */
    int top;
    const int h_mid = wo_p2_height(p, (hut_x >> 4) * 16 + 8, (hut_z >> 4) * 16 + 8, &top);
    wo_biome_result br;
    wo_biome_check(&p->bio, tree, hut_x, hut_z, h_mid - 1, &br);
    out->biome_ok = br.ok;
    out->winner = br.winner;
    out->occ_y = h_mid - 1;

    /* flood: water above y = 62 in all 63 footprint columns (eval.c:151-168) */
    int wet = 0;
    for (int dx = 0; dx < out->size_x; dx++)
        for (int dz = 0; dz < out->size_z; dz++) {
            int level, type;
            wo_p2_fluid(p, hut_x + dx, 62, hut_z + dz, &level, &type);
            if (type == S_WATER && 62 < level) wet++;
        }
    out->wet = wet;
    out->flooded = wet == 63;

    /* the footprint's average surface (eval.c:170-219): post-carver heights, except a flooded hut whose 63
     * pre-carver top blocks are all at or below 63 keeps its pre-carver sum (no carving) */
    wo_p2_chunk *ch = (wo_p2_chunk *)malloc(sizeof *ch);
    wo_p2_chunk_init(ch, p, hut_x >> 4, hut_z >> 4);
    long long sum = 0;
    int skipped = 0;
    if (out->flooded) {
        int blocked = 0;
        long long sum_pre = 0;
        for (int dx = 0; dx < out->size_x && !blocked; dx++)
            for (int dz = 0; dz < out->size_z; dz++) {
                const int hp = wo_p2_chunk_height(ch, hut_x + dx, hut_z + dz);
                sum_pre += hp;
                if (hp - 1 >= 64) { blocked = 1; break; }
            }
        if (!blocked) { sum = sum_pre; skipped = 1; }
    }
    if (!skipped) {
        uint8_t *cells = (uint8_t *)malloc(16 * 16 * 384);
        carve_chunk(ch, seed, cells);
        for (int dx = 0; dx < out->size_x; dx++)
            for (int dz = 0; dz < out->size_z; dz++) sum += post_height(ch, cells, hut_x + dx, hut_z + dz);
        free(cells);
    }
    free(ch);
    out->carve_skipped = skipped;
    out->sum_h = sum;
    out->avg_y = (int)(sum / 63);
    if (!out->biome_ok) { out->ok = 0; out->reject = WO_HUT_REJECT_BIOME; }
    else if (out->flooded) { out->ok = 0; out->reject = WO_HUT_REJECT_FLOOD; }
    else if (out->avg_y > max_y) { out->ok = 0; out->reject = WO_HUT_REJECT_Y; }
    else { out->ok = 1; out->reject = WO_HUT_OK; }
    return out->ok;
}
