/* whf_oracle_p2.c — CPU oracle of the terrain stage: the overworld final density, the preliminary surface level,
 * the aquifer's block state and the pre-carver surface height of a column.
 *
 * TEST INFRASTRUCTURE ONLY (see whf_oracle.h). Restates, in plain C, the reference CPU scanner's 26.1.2-semantics
 * phase 2 (engine/tools/lysh/src):
 *   noise.c:160-221, 246-261  legacy octave stacks (sequential construction) and the Perlin sample with y smear
 *   interp_noise.c:19-139     the blended base noise (minecraft:old_blended_noise)
 *   spline.c:75-81, 116-131   the terrain point with double-evaluated folded ridges
 *   density.c:1-281           the noise column (jagged, sloped cheese, caves, pillars, slides) and the noodle
 *   phase2.c:47-194           the 4 x 8 x 4 cell interpolation of the final density
 *   column_top.c:1-122        the preliminary surface level
 *   aquifer.c:1-563           fluid levels, barriers and computeSubstance
 *   phase2.c:318-334          the pre-carver surface height (first non-air block from the top, + 1)
 * The anchor fluid statuses are memoized per grid cell within one column query, as the reference memoizes them
 * per chunk; they are pure functions of the cell, so the answers do not depend on it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "whf_oracle.h"
#include "whf_oracle_int.h"
#include "whf_oracle_tables.h"
#include "../low-witch-finder_amd/csrc/whf_game_data.h"

/* ---------------------------------------------------------------- Java-semantics helpers (mcmath.h) */

static double j_clamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static double j_lerp(double t, double a, double b) { return a + t * (b - a); }
static double j_clamped_lerp(double a, double b, double t) { return t < 0.0 ? a : (t > 1.0 ? b : j_lerp(t, a, b)); }
static double j_progress(double v, double lo, double hi) { return (v - lo) / (hi - lo); }
static double j_lerp_progress(double v, double lo, double hi, double a, double b) {
    return j_lerp(j_progress(v, lo, hi), a, b);
}
static int is_neg_zero(double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    return b == UINT64_C(0x8000000000000000);
}
/* Math.max / Math.min: NaN first argument wins, and -0.0 < +0.0 */
static double j_max(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && is_neg_zero(a)) return b;
    return a >= b ? a : b;
}
static double j_min(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && is_neg_zero(b)) return b;
    return a <= b ? a : b;
}
static int floor_div(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}
static double y_gradient(double y, double y0, double y1, double v0, double v1) {
    return j_lerp(j_clamp((y - y0) / (y1 - y0), 0.0, 1.0), v0, v1);
}

/* ---------------------------------------------------------------- construction */

static void legacy_stack(wo_perlin *slot, int n, int first_octave, wo_xr *r) {
    /* createLegacy with every amplitude 1 (noise.c:167-209): the sampler of slot -firstOctave is built first,
     * then slots -firstOctave-1 down to 0, all from the one random in that order */
    const int top = -first_octave;
    wo_perlin_build(&slot[top], r);
    for (int j = top - 1; j >= 0; j--) wo_perlin_build(&slot[j], r);
    (void)n;
}

void wo_p2_init(wo_phase2 *p, uint64_t seed, int vflags) {
    memset(p, 0, sizeof *p);
    wo_biome_init(&p->bio, seed, vflags);
    static const int ids[] = {
        WO_N_JAGGED, WO_N_CAVE_ENTRANCE, WO_N_SPAG_ROUGH, WO_N_SPAG_ROUGH_MOD, WO_N_SPAG3D_1, WO_N_SPAG3D_2,
        WO_N_SPAG3D_RARITY, WO_N_SPAG3D_THICKNESS, WO_N_SPAG2D, WO_N_SPAG2D_ELEVATION, WO_N_SPAG2D_MODULATOR,
        WO_N_SPAG2D_THICKNESS, WO_N_PILLAR, WO_N_PILLAR_RARENESS, WO_N_PILLAR_THICKNESS, WO_N_CAVE_LAYER,
        WO_N_CAVE_CHEESE, WO_N_NOODLE, WO_N_NOODLE_THICKNESS, WO_N_NOODLE_RIDGE_A, WO_N_NOODLE_RIDGE_B,
        WO_N_AQUIFER_BARRIER, WO_N_FLOODEDNESS, WO_N_AQUIFER_SPREAD, WO_N_AQUIFER_LAVA};
    for (int i = 0; i < WO_P2_NOISES; i++) wo_noise_init(&p->n[i], seed, ids[i]);

    /* blended noise: deriver of the world random, createRandom("minecraft:terrain"), then the lower, upper and
     * main stacks in that order (interp_noise.c:19-48) */
    wo_xr r, der, terrain;
    wo_xr_seed(&r, seed);
    wo_xr_fork(&der, &r);
    wo_xr_fork_named(&terrain, &der, "minecraft:terrain");
    wo_perlin lower[16], upper[16], mainn[8];
    legacy_stack(lower, 16, -15, &terrain);
    legacy_stack(upper, 16, -15, &terrain);
    legacy_stack(mainn, 8, -7, &terrain);
    /* octave o of a stack reads slot n - 1 - o (noise.c:211-216) */
    for (int o = 0; o < 16; o++) { p->lower[o] = lower[15 - o]; p->upper[o] = upper[15 - o]; }
    for (int o = 0; o < 8; o++) p->main[o] = mainn[7 - o];

    /* aquifer positional random: forkPositional of the world random, fromHashOf("minecraft:aquifer"),
     * forkPositional again (aquifer.c:123-135) */
    wo_xr w, aq;
    wo_xr_seed(&w, seed);
    wo_xr base;
    base.lo = wo_xr_next(&w);
    base.hi = wo_xr_next(&w);
    wo_xr_fork_named(&aq, &base, "minecraft:aquifer");
    p->aq_lo = wo_xr_next(&aq);
    p->aq_hi = wo_xr_next(&aq);
}

void *wo_p2_alloc(uint64_t seed, int vflags) {
    wo_phase2 *p = (wo_phase2 *)malloc(sizeof *p);
    if (p) wo_p2_init(p, seed, vflags);
    return p;
}

#define N(p, k) (&(p)->n[k])
enum {
    P_JAGGED = 0, P_CAVE_ENTRANCE, P_ROUGH, P_ROUGH_MOD, P_SPAG3D_1, P_SPAG3D_2, P_SPAG3D_RARITY, P_SPAG3D_THICK,
    P_SPAG2D, P_SPAG2D_ELEV, P_SPAG2D_MOD, P_SPAG2D_THICK, P_PILLAR, P_PILLAR_RARE, P_PILLAR_THICK, P_CAVE_LAYER,
    P_CAVE_CHEESE, P_NOODLE, P_NOODLE_THICK, P_RIDGE_A, P_RIDGE_B, P_AQ_BARRIER, P_AQ_FLOOD, P_AQ_SPREAD, P_AQ_LAVA
};

/* ---------------------------------------------------------------- terrain point (26.1.2 semantics) */

void wo_p2_terrain_point(const wo_phase2 *p, int qx, int qz, float out[3]) {
    /* terrain.c:33-39 with semantics_261: the folded ridges are evaluated in double from the raw weirdness
     * (spline.c:75-81), the spline itself in float (spline.c:116-131) */
    wo_terrain_info ti;
    wo_terrain_info_at(&p->bio, qx, qz, &ti);
    const double w = ti.weirdness;
    const double folded = -3.0 * (-0.3333333333333333 + fabs(-0.6666666666666666 + fabs(w)));
    const float pt[4] = {(float)ti.continentalness, (float)ti.erosion, (float)folded, (float)w};
    out[0] = wo_spline_apply(whf_gd_spline_root[0], pt) + (-0.50375f);
    out[1] = wo_spline_apply(whf_gd_spline_root[1], pt);
    out[2] = wo_spline_apply(whf_gd_spline_root[2], pt);
}

/* ---------------------------------------------------------------- blended base noise */

static double perlin_smeared(const wo_perlin *pr, double x, double y, double z, double ys, double ymax) {
    /* noise.c:246-261: the y offset is snapped down to a multiple of ys (at most ymax) for the gradients, while
     * the fade keeps the unsnapped fraction */
    const double d = x + pr->ox, e = y + pr->oy, f = z + pr->oz;
    const int i = (int)floor(d), j = (int)floor(e), k = (int)floor(f);
    const double g = d - i, h = e - j, l = f - k;
    double m = 0.0;
    if (ys != 0.0) {
        const double nn = (ymax >= 0.0 && ymax < h) ? ymax : h;
        m = floor(nn / ys + 1.0000000116860974E-7) * ys;
    }
    return wo_perlin_core(pr, i, j, k, g, h - m, l, h);
}

static double keep(double v) { return (v > -1.0e7 && v < 1.0e7) ? v : wo_keep_precision(v); }

double wo_p2_blended(const wo_phase2 *p, int x, int y, int z) {
    /* interp_noise.c:91-139 with xz_scale 0.25, y_scale 0.125, xz_factor 80, y_factor 160, smear 8
     * (phase2.c:262-271) */
    const double xzm = 684.412 * 0.25, ym = 684.412 * 0.125, xzf = 80.0, yf = 160.0, smear = 8.0;
    const double d = (double)x * xzm, e = (double)y * ym, f = (double)z * xzm;
    const double g = d / xzf, h = e / yf, i = f / xzf;
    const double j = ym * smear, k = j / yf;
    double n = 0.0, q = 1.0;
    for (int o = 0; o < 8; o++) {
        n += perlin_smeared(&p->main[o], keep(g * q), keep(h * q), keep(i * q), k * q, h * q) / q;
        q /= 2.0;
    }
    const double t = (n / 10.0 + 1.0) / 2.0;
    const int skip_lo = t >= 1.0, skip_hi = t <= 0.0;
    double lo = 0.0, hi = 0.0;
    q = 1.0;
    for (int o = 0; o < 16; o++) {
        const double xx = keep(d * q), yy = keep(e * q), zz = keep(f * q), ys = j * q, ymax = e * q;
        if (!skip_lo) lo += perlin_smeared(&p->lower[o], xx, yy, zz, ys, ymax) / q;
        if (!skip_hi) hi += perlin_smeared(&p->upper[o], xx, yy, zz, ys, ymax) / q;
        q /= 2.0;
    }
    return j_clamped_lerp(lo / 512.0, hi / 512.0, t) / 128.0;
}

/* ---------------------------------------------------------------- noise column (density.c) */

static double sample3(const wo_noise *n, double x, double y, double z) { return wo_noise_sample(n, x, y, z); }
static double from_noise(const wo_noise *n, double x, double y, double z, double a, double b) {
    return j_lerp_progress(sample3(n, x, y, z), -1.0, 1.0, a, b);
}
static double tunnel_scale(double v) { return v < -0.5 ? 0.75 : (v < 0.0 ? 1.0 : (v < 0.5 ? 1.5 : 2.0)); }
static double cave_scale(double v) {
    return v < -0.75 ? 0.5 : (v < -0.5 ? 0.75 : (v < 0.5 ? 1.0 : (v < 0.75 ? 2.0 : 3.0)));
}

static double cave_entrance(const wo_phase2 *p, int x, int y, int z) {
    /* density.c:75-81 */
    const double s = sample3(N(p, P_CAVE_ENTRANCE), (double)x * 0.75, (double)y * 0.5, (double)z * 0.75);
    return s + 0.37 + j_clamped_lerp(0.3, 0.0, (10.0 + (double)y) / 40.0);
}

static double roughness(const wo_phase2 *p, int x, int y, int z) {
    /* density.c:87-94 */
    const double m = from_noise(N(p, P_ROUGH_MOD), (double)x, (double)y, (double)z, 0.0, 0.1);
    const double r = fabs(sample3(N(p, P_ROUGH), (double)x, (double)y, (double)z));
    return (0.4 - r) * m;
}

static double spaghetti3d(const wo_phase2 *p, int x, int y, int z) {
    /* density.c:97-110 */
    const double rarity = sample3(N(p, P_SPAG3D_RARITY), (double)(x * 2), (double)y, (double)(z * 2));
    const double e = tunnel_scale(rarity);
    const double h = from_noise(N(p, P_SPAG3D_THICK), (double)x, (double)y, (double)z, 0.065, 0.088);
    const double l = sample3(N(p, P_SPAG3D_1), (double)x / e, (double)y / e, (double)z / e);
    const double m = fabs(e * l) - h;
    const double n = sample3(N(p, P_SPAG3D_2), (double)x / e, (double)y / e, (double)z / e);
    const double o = fabs(e * n) - h;
    return j_clamp(j_max(m, o), -1.0, 1.0);
}

static double spaghetti2d(const wo_phase2 *p, int x, int y, int z) {
    /* density.c:122-141; the elevation's lower end is the cell minimum -8 */
    const double mod = sample3(N(p, P_SPAG2D_MOD), (double)(x * 2), (double)y, (double)(z * 2));
    const double s = cave_scale(mod);
    const double h = from_noise(N(p, P_SPAG2D_THICK), (double)(x * 2), (double)y, (double)(z * 2), 0.6, 1.3);
    const double v = sample3(N(p, P_SPAG2D), (double)x / s, (double)y / s, (double)z / s);
    const double j = fabs(s * v) - 0.083 * h;
    const double l = from_noise(N(p, P_SPAG2D_ELEV), (double)x, 0.0, (double)z, -8.0, 8.0);
    double m = fabs(l - (double)y / 8.0) - h;
    m = m * m * m;
    return j_clamp(j_max(m, j), -1.0, 1.0);
}

/* The pillar cube. Mode 0: the C library's pow, as the reference computes it. Mode 1: the cube the device computes
 * (the exact square and cube error terms by FMA, summed before the final rounding), counting the cubes it evaluates
 * and those that differ from pow's, so that a test can measure what the difference does to the terrain. */
static int g_cube_mode = 0;
static unsigned long long g_cube_stats[2];   /* cubes evaluated in mode 1, cubes differing from pow */

void wo_set_cube_mode(int mode) { g_cube_mode = mode; }

void wo_cube_stats(unsigned long long out[2]) {
    out[0] = __atomic_exchange_n(&g_cube_stats[0], 0ull, __ATOMIC_RELAXED);
    out[1] = __atomic_exchange_n(&g_cube_stats[1], 0ull, __ATOMIC_RELAXED);
}

static double pillar_cube(double t) {
    const double pw = pow(t, 3.0);
    if (!g_cube_mode) return pw;
    const double h = t * t, e1 = fma(t, t, -h);
    const double q = h * t, e2 = fma(h, t, -q);
    const double c = q + fma(e1, t, e2);
    __atomic_add_fetch(&g_cube_stats[0], 1ull, __ATOMIC_RELAXED);
    if (c != pw) __atomic_add_fetch(&g_cube_stats[1], 1ull, __ATOMIC_RELAXED);
    return c;
}

static double pillar(const wo_phase2 *p, int x, int y, int z) {
    /* density.c:151-162 */
    const double rare = from_noise(N(p, P_PILLAR_RARE), (double)x, (double)y, (double)z, 0.0, 2.0);
    double thick = from_noise(N(p, P_PILLAR_THICK), (double)x, (double)y, (double)z, 0.0, 1.1);
    thick = pillar_cube(thick);
    const double f = sample3(N(p, P_PILLAR), (double)x * 25.0, (double)y * 0.3, (double)z * 25.0);
    const double g = thick * (f * 2.0 - rare);
    return g > 0.03 ? g : -INFINITY;
}

static double cave_layer(const wo_phase2 *p, int x, int y, int z) {
    /* density.c:165-168 (y * 8 in int) */
    const double v = sample3(N(p, P_CAVE_LAYER), (double)x, (double)(y * 8), (double)z);
    return v * v * 4.0;
}

static double slide(double target, int size, int offset, double v, int y) {
    /* density.c:214-219 */
    if (size <= 0) return v;
    return j_clamped_lerp(target, v, (double)(y - offset) / (double)size);
}

double wo_p2_noise_column(const wo_phase2 *p, int x, int y, int z, const float tp[3], double base) {
    /* density.c:228-281 with jagged on and noise caves on */
    const double peaks = (double)tp[2];
    double jag = 0.0;
    if (peaks != 0.0) {   /* density.c:66-70 */
        const double n = sample3(N(p, P_JAGGED), (double)x * 1500.0, 0.0, (double)z * 1500.0);
        jag = n > 0.0 ? peaks * n : (peaks / 2.0) * n;
    }
    const double depth = (1.0 - (double)y / 128.0) + (double)tp[0];
    double dens = (depth + jag) * (double)tp[1];
    dens = dens > 0.0 ? dens * 4.0 : dens;
    const double g = dens + base;
    double cheese, spag, pil;
    if (g < -64.0) {
        cheese = g;
        spag = 64.0;
        pil = -64.0;
    } else {
        const double i = g - 1.5625;
        const double ce = cave_entrance(p, x, y, z);
        const double rough = roughness(p, x, y, z);
        const double s3 = spaghetti3d(p, x, y, z);
        const double m = j_min(ce, s3 + rough);
        if (i < 0.0) {
            cheese = g;
            spag = m * 5.0;
            pil = -64.0;
        } else {
            const double layer = cave_layer(p, x, y, z);
            if (layer > 64.0) {
                cheese = 64.0;
            } else {
                const double raw = sample3(N(p, P_CAVE_CHEESE), (double)x, (double)y / 1.5, (double)z);
                const double pc = j_clamp(raw + 0.27, -1.0, 1.0);
                cheese = pc + j_clamped_lerp(0.5, 0.0, i * 1.28) + layer;
            }
            const double s2 = spaghetti2d(p, x, y, z);
            spag = j_min(m, s2 + rough);
            pil = pillar(p, x, y, z);
        }
    }
    double v = j_max(j_min(cheese, spag), pil);
    /* slides (density.c:221-226): top over the cell count above, bottom over the cells above the minimum; the
     * cell index is y / 8 truncated toward zero */
    const int cell = y / 8;
    const int k = cell - -8;
    v = slide(-0.078125, 2, 8, v, 48 - k);
    v = slide(0.1171875, 3, 0, v, k);
    return j_clamp(v, -64.0, 64.0);
}

double wo_p2_noodle_child(const wo_phase2 *p, int which, int x, int y, int z) {
    /* density.c:186-209: noodle, thickness, ridge a, ridge b; outside -60 <= y <= 320 their fixed values */
    const int in = y >= -60 && y <= 320;
    const double sc = 2.6666666666666665;
    switch (which) {
    case 0: return in ? sample3(N(p, P_NOODLE), (double)x, (double)y, (double)z) : -1.0;
    case 1: return in ? -0.07500000000000001 + -0.025 * sample3(N(p, P_NOODLE_THICK), (double)x, (double)y, (double)z)
                      : 0.0;
    case 2: return in ? sample3(N(p, P_RIDGE_A), (double)x * sc, (double)y * sc, (double)z * sc) : 0.0;
    case 3: return in ? sample3(N(p, P_RIDGE_B), (double)x * sc, (double)y * sc, (double)z * sc) : 0.0;
    }
    return 0.0;
}

static double noodle_combine(double n, double thick, double ra, double rb) {
    /* density.c:211-214 */
    if (n >= -1000000.0 && n < 0.0) return 64.0;
    return thick + 1.5 * j_max(fabs(ra), fabs(rb));
}

/* ---------------------------------------------------------------- interpolated final density (phase2.c) */

void wo_p2_corner_column(const wo_phase2 *p, int cell_x, int cell_z, double main[49], double noodle[3][49]) {
    /* phase2.c:71-106: 49 samples at y = -64 + 8 i of the corner (4 cell_x, 4 cell_z); the terrain point is the
     * quart point (cell_x, cell_z) */
    float tp[3];
    wo_p2_terrain_point(p, cell_x, cell_z, tp);
    const int bx = cell_x * 4, bz = cell_z * 4;
    for (int i = 0; i <= 48; i++) {
        const int y = -64 + i * 8;
        main[i] = wo_p2_noise_column(p, bx, y, bz, tp, wo_p2_blended(p, bx, y, bz));
        for (int c = 0; c < 3; c++) noodle[c][i] = wo_p2_noodle_child(p, c + 1, bx, y, bz);
    }
}

static double lerp3c(double dx, double dy, double dz, const double *c00, const double *c10, const double *c01,
                     const double *c11, int cy) {
    /* Mth.lerp3 over the cell's eight corners, x innermost and z outermost (phase2.c:55-65, 132-150) */
    const double a = j_lerp(dy, c00[cy], c00[cy + 1]);
    const double b = j_lerp(dy, c10[cy], c10[cy + 1]);
    const double c = j_lerp(dy, c01[cy], c01[cy + 1]);
    const double d = j_lerp(dy, c11[cy], c11[cy + 1]);
    return j_lerp(dz, j_lerp(dx, a, b), j_lerp(dx, c, d));
}

typedef struct {
    int cell_x, cell_z;
    double main[4][49], noodle[4][3][49];   /* corners (x0,z0), (x1,z0), (x0,z1), (x1,z1) */
} corner_set;

static void corners_for(const wo_phase2 *p, int x, int z, corner_set *cs) {
    cs->cell_x = x >> 2;
    cs->cell_z = z >> 2;
    for (int c = 0; c < 4; c++)
        wo_p2_corner_column(p, cs->cell_x + (c & 1), cs->cell_z + (c >> 1), cs->main[c], cs->noodle[c]);
}

static double density_core(const wo_phase2 *p, const double *const m[4], const double (*const nd[4])[49], int x, int y,
                           int z) {
    /* phase2.c:161-194 for -64 <= y < 320, over the corner columns (x0,z0), (x1,z0), (x0,z1), (x1,z1) */
    const int yo = y + 64;
    const int cy = yo / 8;
    const double dx = (double)(x & 3) / 4.0, dy = (double)(yo - cy * 8) / 8.0, dz = (double)(z & 3) / 4.0;
    const double v = lerp3c(dx, dy, dz, m[0], m[1], m[2], m[3], cy);
    const double a = j_clamp(v * 0.64, -1.0, 1.0);
    const double squeezed = a / 2.0 - a * a * a / 24.0;   /* density.c:172-175 */
    double ch[3];
    for (int c = 0; c < 3; c++) ch[c] = lerp3c(dx, dy, dz, nd[0][c], nd[1][c], nd[2][c], nd[3][c], cy);
    const double noodle = noodle_combine(wo_p2_noodle_child(p, 0, x, y, z), ch[0], ch[1], ch[2]);
    return j_min(squeezed, noodle);
}

static double density_in(const wo_phase2 *p, const corner_set *cs, int x, int y, int z) {
    const double *const m[4] = {cs->main[0], cs->main[1], cs->main[2], cs->main[3]};
    const double (*const nd[4])[49] = {cs->noodle[0], cs->noodle[1], cs->noodle[2], cs->noodle[3]};
    return density_core(p, m, nd, x, y, z);
}

double wo_p2_density(const wo_phase2 *p, int x, int y, int z) {
    corner_set *cs = (corner_set *)malloc(sizeof *cs);
    corners_for(p, x, z, cs);
    const double v = density_in(p, cs, x, y, z);
    free(cs);
    return v;
}

/* ---------------------------------------------------------------- preliminary surface level (column_top.c) */

static double ps_density(double y, double offset, double factor) {
    /* column_top.c:28-35 */
    const double depth = y_gradient(y, -64.0, 320.0, 1.5, -1.5) + offset;
    const double dq = depth * factor;
    const double cheese = 4.0 * (dq > 0.0 ? dq : dq * 0.25);
    const double c = j_clamp(-0.703125 + cheese, -64.0, 64.0);
    const double a1 = y_gradient(y, 240.0, 256.0, 1.0, 0.0);
    const double a2 = y_gradient(y, -64.0, -40.0, 0.0, 1.0);
    return -0.390625 + (0.1171875 + a2 * (-0.1171875 + (-0.078125 + a1 * (0.078125 + c))));
}

int wo_p2_prelim_surface(const wo_phase2 *p, int x, int z) {
    /* column_top.c:57-122: offset and factor in double over the float spline at the quart point of (x, z); an
     * upper bound from them, then the first y (step 8, down to -64) whose density is positive */
    wo_terrain_info ti;
    wo_terrain_info_at(&p->bio, x >> 2, z >> 2, &ti);
    const float c = (float)ti.continentalness, e = (float)ti.erosion, w = (float)ti.weirdness;
    const float pt[4] = {c, e, wo_normalized_weirdness(w), w};
    const double raw_off = (double)wo_spline_apply(whf_gd_spline_root[0], pt);
    const double raw_fac = (double)wo_spline_apply(whf_gd_spline_root[1], pt);
    const double alpha = 1.0, blend = 0.0;
    const double offset = blend * (1.0 + -1.0 * alpha) + (-0.5037500262260437 + raw_off) * alpha;
    const double factor = 10.0 + alpha * (-10.0 + raw_fac);
    const double inner = (1.0 / factor) * 0.2734375 + offset * -1.0;
    const double ub = j_clamp(128.0 + -128.0 * inner, -40.0, 320.0);
    const int y0 = (int)(ub / 8.0) * 8;
    if (y0 <= -64) return -64;   /* the lower bound is the world's minimum y (column_top.h:69) */
    for (int y = y0; y >= -64; y -= 8)
        if (ps_density((double)y, offset, factor) > 0.0) return y;
    return -64;
}

/* ---------------------------------------------------------------- aquifer (aquifer.c) */

enum { B_SOLID = 0, B_AIR = 1, B_WATER = 2, B_LAVA = 3 };
#define NO_FLUID (-32512)

static void global_fluid(int y, int *level, int *type) {
    /* aquifer.c:86-95: lava at -54 below y = -54, water at sea level 63 otherwise */
    if (y < -54) { *level = -54; *type = B_LAVA; }
    else { *level = 63; *type = B_WATER; }
}
static int level_block(int level, int type, int y) { return y < level ? type : B_AIR; }

static int deep_dark(const wo_phase2 *p, int x, int y, int z) {
    /* aquifer.c:218-228 */
    wo_terrain_info ti;
    wo_terrain_info_at(&p->bio, x >> 2, z >> 2, &ti);
    if (!(ti.erosion < -0.22499999403953552)) return 0;
    const double depth = y_gradient((double)y, -64.0, 320.0, 1.5, -1.5) + wo_overworld_offset(&p->bio, x, z);
    return depth > 0.8999999761581421;
}

static const int SURFACE_OFFSETS[13][2] = {{0, 0},   {-2, -1}, {-1, -1}, {0, -1}, {1, -1}, {-3, 0}, {-2, 0},
                                           {-1, 0},  {1, 0},   {-2, 1},  {-1, 1}, {0, 1},  {1, 1}};

static void compute_fluid(const wo_phase2 *p, int x, int y, int z, int *out_level, int *out_type) {
    /* aquifer.c:286-331 computeFluid, with computeSurfaceLevel (:248-273), the randomized level (:232-245) and
     * computeFluidType (:277-289) */
    int gl, gt;
    global_fluid(y, &gl, &gt);
    int min_top = INT32_MAX, has_above = 0;
    for (int i = 0; i < 13; i++) {
        const int top = wo_p2_prelim_surface(p, x + SURFACE_OFFSETS[i][0] * 16, z + SURFACE_OFFSETS[i][1] * 16);
        const int adj = top + 8;
        const int center = i == 0;
        if (center && y - 12 > adj) { *out_level = gl; *out_type = gt; return; }
        const int above = y + 12 > adj;
        if (above || center) {
            int lv, ty;
            global_fluid(adj, &lv, &ty);
            if (level_block(lv, ty, adj) != B_AIR) {
                if (center) has_above = 1;
                if (above) { *out_level = lv; *out_type = ty; return; }
            }
        }
        if (top < min_top) min_top = top;
    }
    int level;
    if (deep_dark(p, x, y, z)) {
        level = NO_FLUID;
    } else {
        const int d = (min_top + 8) - y;
        const double f = has_above ? j_clamped_lerp(1.0, 0.0, j_progress((double)d, 0.0, 64.0)) : 0.0;
        const double high = j_lerp_progress(f, 1.0, 0.0, -0.3, 0.8);
        const double low = j_lerp_progress(f, 1.0, 0.0, -0.8, 0.4);
        const double flood = j_clamp(sample3(N(p, P_AQ_FLOOD), (double)x, (double)y * 0.67, (double)z), -1.0, 1.0);
        if (flood - high > 0.0) {
            level = gl;
        } else if (flood - low > 0.0) {
            const int fx = floor_div(x, 16), fy = floor_div(y, 40), fz = floor_div(z, 16);
            const double n = sample3(N(p, P_AQ_SPREAD), (double)fx, (double)fy * 0.7142857142857143, (double)fz) * 10.0;
            const int lvl = fy * 40 + 20 + (int)floor(n / 3.0) * 3;
            level = min_top < lvl ? min_top : lvl;
        } else {
            level = NO_FLUID;
        }
    }
    int type = gt;
    if (level <= -10 && level != NO_FLUID && gt != B_LAVA) {
        const double n = sample3(N(p, P_AQ_LAVA), (double)floor_div(x, 64), (double)floor_div(y, 40),
                                 (double)floor_div(z, 64));
        if (fabs(n) > 0.3) type = B_LAVA;
    }
    *out_level = level;
    *out_type = type;
}

void wo_p2_fluid(const wo_phase2 *p, int x, int y, int z, int *level, int *type) { compute_fluid(p, x, y, z, level, type); }

static int64_t pos_hash(int x, int y, int z) {
    /* Mth.getSeed (rng.c:170-175) */
    const int64_t xm = (int64_t)(int32_t)((uint32_t)x * UINT32_C(3129871));
    int64_t i = xm ^ ((int64_t)z * INT64_C(116129781)) ^ (int64_t)y;
    i = i * i * INT64_C(42317861) + i * INT64_C(11);
    return i >> 16;
}

static void anchor_of(const wo_phase2 *p, int gx, int gy, int gz, int *ax, int *ay, int *az) {
    /* aquifer.c:452-459: the grid cell's random anchor */
    wo_xr r;
    wo_xr_set(&r, (uint64_t)pos_hash(gx, gy, gz) ^ p->aq_lo, p->aq_hi);
    *ax = gx * 16 + wo_xr_next_int(&r, 10);
    *ay = gy * 12 + wo_xr_next_int(&r, 9);
    *az = gz * 16 + wo_xr_next_int(&r, 10);
}

int wo_p2_skip_above(const wo_phase2 *p, int chunk_x, int chunk_z) {
    /* aquifer.c:160-196: skipSamplingAboveY of the chunk's aquifer grid */
    const int min_gx = (chunk_x * 16 - 5) >> 4, max_gx = ((chunk_x * 16 + 15 - 5) >> 4) + 1;
    const int min_gz = (chunk_z * 16 - 5) >> 4, max_gz = ((chunk_z * 16 + 15 - 5) >> 4) + 1;
    int max_psl = INT32_MIN;
    for (int z = min_gz * 16; z <= max_gz * 16 + 9; z += 4)
        for (int x = min_gx * 16; x <= max_gx * 16 + 9; x += 4) {
            const int v = wo_p2_prelim_surface(p, x, z);
            if (v > max_psl) max_psl = v;
        }
    const int g = floor_div((max_psl + 8) + 12, 12) + 1;
    return g * 12 + 11 - 1;
}

typedef wo_aq_memo aq_column;

static void status_of(const wo_phase2 *p, aq_column *ac, int gx, int gy, int gz, int ax, int ay, int az, int *level,
                      int *type) {
    /* fluid statuses of the anchors met, by open addressing on the grid cell */
    const unsigned mask = WO_AQ_SLOTS - 1u;
    unsigned h = (unsigned)(gx * 73856093 ^ gy * 19349663 ^ gz * 83492791) & mask;
    for (int probe = 0; probe < WO_AQ_SLOTS; probe++) {
        wo_aq_slot *s = &ac->st[(h + probe) & mask];
        if (s->valid && s->gx == gx && s->gy == gy && s->gz == gz) { *level = s->level; *type = s->type; return; }
        if (!s->valid) {
            compute_fluid(p, ax, ay, az, &s->level, &s->type);
            s->valid = 1;
            s->gx = gx; s->gy = gy; s->gz = gz;
            *level = s->level;
            *type = s->type;
            return;
        }
    }
    compute_fluid(p, ax, ay, az, level, type);
}

static double similarity(int a, int b) { return 1.0 - (double)(b - a) / 25.0; }   /* aquifer.c:49-52 */

static double pressure(const wo_phase2 *p, int x, int y, int z, double *barrier, int la, int ta, int lb, int tb) {
    /* aquifer.c:366-404 calculatePressure; the barrier noise is sampled at most once per block */
    const int sa = level_block(la, ta, y), sb = level_block(lb, tb, y);
    if ((sa == B_LAVA && sb == B_WATER) || (sa == B_WATER && sb == B_LAVA)) return 2.0;
    int dy = la - lb;
    if (dy < 0) dy = -dy;
    if (dy == 0) return 0.0;
    const double mid = 0.5 * ((double)la + (double)lb);
    const double f = ((double)y + 0.5) - mid;
    const double n = (double)dy / 2.0 - fabs(f);
    double o;
    if (f > 0.0) {
        const double q = 0.0 + n;
        o = q > 0.0 ? q / 1.5 : q / 2.5;
    } else {
        const double q = 3.0 + n;
        o = q > 0.0 ? q / 3.0 : q / 10.0;
    }
    double inner;
    if (o < -2.0 || o > 2.0) {
        inner = 0.0;
    } else if (*barrier != *barrier) {
        *barrier = sample3(N(p, P_AQ_BARRIER), (double)x, (double)y * 0.5, (double)z);
        inner = *barrier;
    } else {
        inner = *barrier;
    }
    return 2.0 * (inner + o);
}

static int substance(const wo_phase2 *p, aq_column *ac, int x, int y, int z, double density) {
    /* aquifer.c:416-563 computeSubstance; only the block state (the fluid-update flag is not needed) */
    if (density > 0.0) return B_SOLID;
    int gl, gt;
    global_fluid(y, &gl, &gt);
    if (y > ac->skip_above) return level_block(gl, gt, y);
    if (level_block(gl, gt, y) == B_LAVA) return B_LAVA;
    const int gx = floor_div(x - 5, 16), gy = floor_div(y + 1, 12), gz = floor_div(z - 5, 16);
    int d[4] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
    int cell[4][6];
    memset(cell, 0, sizeof cell);
    for (int i = 0; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = 0; k <= 1; k++) {
                const int cx = gx + i, cy = gy + j, cz = gz + k;
                int ax, ay, az;
                anchor_of(p, cx, cy, cz, &ax, &ay, &az);
                const int ddx = ax - x, ddy = ay - y, ddz = az - z;
                const int dist = ddx * ddx + ddy * ddy + ddz * ddz;
                const int rec[6] = {cx, cy, cz, ax, ay, az};
                /* keep the four nearest, a later cell winning ties (aquifer.c:478-491) */
                int at = -1;
                for (int s = 0; s < 4; s++)
                    if (dist <= d[s]) { at = s; break; }
                if (at < 0) continue;
                for (int s = 3; s > at; s--) { d[s] = d[s - 1]; memcpy(cell[s], cell[s - 1], sizeof cell[s]); }
                d[at] = dist;
                memcpy(cell[at], rec, sizeof rec);
            }
    int l[4], t[4];
#define STATUS(s) status_of(p, ac, cell[s][0], cell[s][1], cell[s][2], cell[s][3], cell[s][4], cell[s][5], &l[s], &t[s])
    STATUS(0);
    const double s01 = similarity(d[0], d[1]);
    const int result = level_block(l[0], t[0], y);
    if (!(s01 > 0.0)) return result;
    {
        int bl, bt;
        global_fluid(y - 1, &bl, &bt);
        if (result == B_WATER && level_block(bl, bt, y - 1) == B_LAVA) return result;
    }
    double barrier = NAN;
    STATUS(1);
    if (density + s01 * pressure(p, x, y, z, &barrier, l[0], t[0], l[1], t[1]) > 0.0) return B_SOLID;
    STATUS(2);
    const double s02 = similarity(d[0], d[2]);
    if (s02 > 0.0 && density + s01 * s02 * pressure(p, x, y, z, &barrier, l[0], t[0], l[2], t[2]) > 0.0)
        return B_SOLID;
    const double s12 = similarity(d[1], d[2]);
    if (s12 > 0.0 && density + s01 * s12 * pressure(p, x, y, z, &barrier, l[1], t[1], l[2], t[2]) > 0.0)
        return B_SOLID;
#undef STATUS
    return result;
}

int wo_p2_substance(const wo_phase2 *p, int x, int y, int z, double density) {
    aq_column *ac = (aq_column *)calloc(1, sizeof *ac);
    ac->skip_above = wo_p2_skip_above(p, x >> 4, z >> 4);
    const int s = substance(p, ac, x, y, z, density);
    free(ac);
    return s;
}

/* ---------------------------------------------------------------- pre-carver surface height */

int wo_p2_height(const wo_phase2 *p, int x, int z, int *top_state) {
    /* phase2.c:318-334: from y = 319 down, the first block that is not air (solid, water or lava) gives y + 1;
     * a column of air gives -64 */
    corner_set *cs = (corner_set *)malloc(sizeof *cs);
    aq_column *ac = (aq_column *)calloc(1, sizeof *ac);
    corners_for(p, x, z, cs);
    ac->skip_above = wo_p2_skip_above(p, x >> 4, z >> 4);
    int h = -64, st = B_AIR;
    for (int y = 319; y >= -64; y--) {
        const int s = substance(p, ac, x, y, z, density_in(p, cs, x, y, z));
        if (s != B_AIR) { h = y + 1; st = s; break; }
    }
    free(cs);
    free(ac);
    if (top_state) *top_state = st;
    return h;
}

/* ---------------------------------------------------------------- a whole chunk (for the carvers) */

void wo_p2_chunk_init(wo_p2_chunk *ch, const wo_phase2 *p, int cx, int cz) {
    /* the chunk's 5 x 5 corner columns, computed on first use (phase2.c:108-130), and its aquifer grid */
    memset(ch, 0, sizeof *ch);
    ch->p = p;
    ch->cx = cx;
    ch->cz = cz;
    ch->aq.skip_above = wo_p2_skip_above(p, cx, cz);
}

static void chunk_corner(wo_p2_chunk *ch, int gx, int gz) {
    const int idx = gx * 5 + gz;
    if (ch->have & (1u << idx)) return;
    wo_p2_corner_column(ch->p, ch->cx * 4 + gx, ch->cz * 4 + gz, ch->main[idx], ch->noodle[idx]);
    ch->have |= 1u << idx;
}

double wo_p2_chunk_density(wo_p2_chunk *ch, int x, int y, int z) {
    const int gx = (x >> 2) - ch->cx * 4, gz = (z >> 2) - ch->cz * 4;
    for (int c = 0; c < 4; c++) chunk_corner(ch, gx + (c & 1), gz + (c >> 1));
    const int i00 = gx * 5 + gz, i10 = (gx + 1) * 5 + gz, i01 = gx * 5 + gz + 1, i11 = (gx + 1) * 5 + gz + 1;
    const double *const m[4] = {ch->main[i00], ch->main[i10], ch->main[i01], ch->main[i11]};
    const double (*const nd[4])[49] = {ch->noodle[i00], ch->noodle[i10], ch->noodle[i01], ch->noodle[i11]};
    return density_core(ch->p, m, nd, x, y, z);
}

int wo_p2_chunk_substance(wo_p2_chunk *ch, int x, int y, int z, double density) {
    return substance(ch->p, &ch->aq, x, y, z, density);
}

int wo_p2_chunk_state(wo_p2_chunk *ch, int x, int y, int z) {
    /* lysh_phase2_block_at (phase2.c:311-316): the aquifer's verdict on the interpolated density */
    return substance(ch->p, &ch->aq, x, y, z, wo_p2_chunk_density(ch, x, y, z));
}

int wo_p2_chunk_height(wo_p2_chunk *ch, int x, int z) {
    /* phase2.c:318-334 within the chunk */
    for (int y = 319; y >= -64; y--)
        if (wo_p2_chunk_state(ch, x, y, z) != B_AIR) return y + 1;
    return -64;
}
