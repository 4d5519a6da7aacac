// whf_terrain.h — device functions of the terrain stage: the blended base noise, the noise column, the noodle, the
// interpolated final density, the preliminary surface level and the aquifer's block state, for one seed's sampler
// tables. Each follows the reference CPU scanner's 26.1.2-semantics phase 2 (engine/tools/lysh/src: interp_noise.c,
// density.c, phase2.c, column_top.c, aquifer.c) operation for operation, so every value is the reference's in FP64;
// the one exception is the pillar thickness cube (see cube()).
#pragma once

#include "whf_biome.h"
#include "whf_kernels.h"

namespace whf {

// The terrain stage's noises, in the order of their descriptors; the first six are the biome stage's.
enum TNoise {
    T_OFFSET = 0, T_CONT, T_EROSION, T_RIDGE, T_TEMPERATURE, T_VEGETATION,
    T_JAGGED, T_CAVE_ENTRANCE, T_ROUGH, T_ROUGH_MOD, T_SPAG3D_1, T_SPAG3D_2, T_SPAG3D_RARITY, T_SPAG3D_THICK,
    T_SPAG2D, T_SPAG2D_ELEV, T_SPAG2D_MOD, T_SPAG2D_THICK, T_PILLAR, T_PILLAR_RARE, T_PILLAR_THICK, T_CAVE_LAYER,
    T_CAVE_CHEESE, T_NOODLE, T_NOODLE_THICK, T_RIDGE_A, T_RIDGE_B, T_AQ_BARRIER, T_AQ_FLOOD, T_AQ_SPREAD, T_AQ_LAVA,
    T_COUNT
};

constexpr int kLegacyPerlins = 40;   // blended noise: lower 16, upper 16, main 8 octaves, in construction order

struct TerrainTables {
    NoiseDesc noise[T_COUNT];
};

enum BlockState { BS_SOLID = 0, BS_AIR = 1, BS_WATER = 2, BS_LAVA = 3 };
constexpr int kNoFluid = -32512;   // DimensionType.WAY_BELOW_MIN_Y: no fluid here (aquifer.c:17-18)

__device__ __forceinline__ double j_lerp(double t, double a, double b) { return a + t * (b - a); }
__device__ __forceinline__ double j_clamped_lerp(double a, double b, double t) {
    return t < 0.0 ? a : (t > 1.0 ? b : j_lerp(t, a, b));
}
__device__ __forceinline__ double j_progress(double v, double lo, double hi) { return (v - lo) / (hi - lo); }
__device__ __forceinline__ double j_lerp_progress(double v, double lo, double hi, double a, double b) {
    return j_lerp(j_progress(v, lo, hi), a, b);
}
__device__ __forceinline__ double y_gradient(double y, double y0, double y1, double v0, double v1) {
    return j_lerp(jclamp((y - y0) / (y1 - y0), 0.0, 1.0), v0, v1);
}
__device__ __forceinline__ int floor_div(int a, int b) {
    int q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
    return q;
}

// t^3 rounded once: the exact square and cube error terms by FMA, summed before the final rounding. The reference
// takes the C library's pow(t, 3.0), which is not correctly rounded (about 0.08% of cubes on [0, 1.1] differ from
// the exact cube's rounding by one ulp); no device function reproduces that library's tables, so this is the
// correctly rounded cube and the pillar term can differ from the reference by one ulp.
__device__ __forceinline__ double cube(double t) {
    const double h = t * t, e1 = fma(t, t, -h);
    const double q = h * t, e2 = fma(h, t, -q);
    return q + fma(e1, t, e2);
}

// The aquifer tables of one chunk, computed once per chunk and read by all of its columns (the game memoizes the same
// values per quart position and per aquifer grid cell). Quart table: the preliminary surface and the deep-dark offset
// (Terr::quart_surface) at quart (4 cx - 16 + i, 4 cz - 8 + j), entry i * kAqQz + j; that covers the 13 surface
// probes of computeFluid around every anchor and column of the chunk, and the 11 x 11 skipSamplingAboveY grid.
// Cell table: the anchor (x, y, z) and the fluid (level * 4 + type) of grid cell (cx - 1 + i, gy, cz - 1 + k), entry
// ((gy - kAqGy0) * 3 + k) * 3 + i, for gy in [-6, 27], the cells computeSubstance reads for y in [-54, 319].
constexpr int kAqQx = 27, kAqQz = 19, kAqQuarts = kAqQx * kAqQz;
constexpr int kAqGy0 = -6, kAqGyN = 34, kAqCells = 3 * 3 * kAqGyN;

struct ChunkAq {
    const int *surface;    // [kAqQuarts]
    const double *dd;      // [kAqQuarts]
    const int4 *cells;     // [kAqCells]
    int cx, cz;

    __device__ int quart(int x, int z) const { return ((x >> 2) - (4 * cx - 16)) * kAqQz + ((z >> 2) - (4 * cz - 8)); }
    __device__ int prelim(int x, int z) const { return surface[quart(x, z)]; }
    __device__ bool deep_dark(int x, int y, int z) const {
        // aquifer.c:218-228: erosion below -0.225 (folded into the offset) and depth above 0.9
        return y_gradient((double)y, -64.0, 320.0, 1.5, -1.5) + dd[quart(x, z)] > 0.8999999761581421;
    }
    __device__ int4 cell(int gx, int gy, int gz) const {
        return cells[((gy - kAqGy0) * 3 + (gz - (cz - 1))) * 3 + (gx - (cx - 1))];
    }
};

// One seed's view of the terrain stage's tables.
struct Terr {
    const TerrainTables *tt;
    SetView sv;                 // named-noise samplers of the seed slot
    uint32_t slot;
    const uint8_t *ltab;        // the seed's 40 blended-noise permutations
    const SPar *lpar;
    const SplineTab *S;
    uint64_t aq_lo, aq_hi;      // the aquifer's positional random factory

    __device__ double n3(int k, double x, double y, double z) const { return noise_3d(tt->noise[k], sv, slot, x, y, z); }
    __device__ double n0(int k, double x, double z) const { return noise_y0(tt->noise[k], sv, slot, x, z); }
    __device__ double from_noise(int k, double x, double y, double z, double a, double b) const {
        return j_lerp_progress(n3(k, x, y, z), -1.0, 1.0, a, b);
    }

    // terrain.c:22-44: shifted coordinates and continentalness, weirdness, erosion at a quart position
    __device__ void info(int qx, int qz, double &c, double &w, double &e) const {
        const double sx = n3(T_OFFSET, (double)qx, 0.0, (double)qz) * 4.0;
        const double sz = n3(T_OFFSET, (double)qz, (double)qx, 0.0) * 4.0;
        const double x = (double)qx + sx, z = (double)qz + sz;
        c = n0(T_CONT, x, z);
        w = n0(T_RIDGE, x, z);
        e = n0(T_EROSION, x, z);
    }

    // terrain point with the 26.1.2 folded ridges (terrain.c:33-39, spline.c:75-81, 116-131): offset, factor, peaks
    __device__ void point261(int qx, int qz, float tp[3]) const {
        double c, w, e;
        info(qx, qz, c, w, e);
        const float folded = (float)(-3.0 * (-0.3333333333333333 + fabs(-0.6666666666666666 + fabs(w))));
        const float cf = (float)c, ef = (float)e, wf = (float)w;
        tp[0] = spline_at<kSplineMaxDepth>(S, S->root[0], cf, ef, folded, wf) + (-0.50375f);
        tp[1] = spline_at<kSplineMaxDepth>(S, S->root[1], cf, ef, folded, wf);
        tp[2] = spline_at<kSplineMaxDepth>(S, S->root[2], cf, ef, folded, wf);
    }

    // offset and factor of the preliminary surface, in double over the float spline (column_top.c:57-77), from the
    // quart position's continentalness, weirdness and erosion
    __device__ void offset_factor(double c, double w, double e, double &offset, double &factor) const {
        const float cf = (float)c, ef = (float)e, wf = (float)w, nw = normalized_weirdness(wf);
        const double ro = (double)spline_at<kSplineMaxDepth>(S, S->root[0], cf, ef, nw, wf);
        const double rf = (double)spline_at<kSplineMaxDepth>(S, S->root[1], cf, ef, nw, wf);
        const double alpha = 1.0, blend = 0.0;
        offset = blend * (1.0 + -1.0 * alpha) + (-0.5037500262260437 + ro) * alpha;
        factor = 10.0 + alpha * (-10.0 + rf);
    }

    // ---- blended base noise (interp_noise.c:91-139) ----
    __device__ double smeared(int idx, double x, double y, double z, double ys, double ymax) const {
        // noise.c:246-261: the gradients use y snapped down to a multiple of ys (at most ymax); the fade does not
        const SPar sp = lpar[idx];
        const GPerm P{ltab + (size_t)idx * 256};
        const double d = x + sp.ox, e = y + sp.oy, f = z + sp.oz;
        const double fi = floor(d), fj = floor(e), fk = floor(f);
        const double g = d - fi, h = e - fj, l = f - fk;
        double m = 0.0;
        if (ys != 0.0) {
            const double nn = (ymax >= 0.0 && ymax < h) ? ymax : h;
            m = floor(nn / ys + 1.0000000116860974E-7) * ys;
        }
        return perlin_cell(P, (int)fi, (int)fj, (int)fk, g, h - m, l, fade(h));
    }
    __device__ double blended(int x, int y, int z) const {
        const double xzm = 684.412 * 0.25, ym = 684.412 * 0.125, xzf = 80.0, yf = 160.0, smear = 8.0;
        const double d = (double)x * xzm, e = (double)y * ym, f = (double)z * xzm;
        const double g = d / xzf, h = e / yf, i = f / xzf;
        const double j = ym * smear, k = j / yf;
        double n = 0.0, q = 1.0;
        for (int o = 0; o < 8; o++) {
            n += smeared(32 + o, keep_precision(g * q), keep_precision(h * q), keep_precision(i * q), k * q, h * q) / q;
            q /= 2.0;
        }
        const double t = (n / 10.0 + 1.0) / 2.0;
        const bool skip_lo = t >= 1.0, skip_hi = t <= 0.0;
        double lo = 0.0, hi = 0.0;
        q = 1.0;
        for (int o = 0; o < 16; o++) {
            const double xx = keep_precision(d * q), yy = keep_precision(e * q), zz = keep_precision(f * q);
            const double ys = j * q, ymax = e * q;
            if (!skip_lo) lo += smeared(o, xx, yy, zz, ys, ymax) / q;
            if (!skip_hi) hi += smeared(16 + o, xx, yy, zz, ys, ymax) / q;
            q /= 2.0;
        }
        return j_clamped_lerp(lo / 512.0, hi / 512.0, t) / 128.0;
    }

    // ---- noise column (density.c) ----
    __device__ double cave_entrance(int x, int y, int z) const {
        const double s = n3(T_CAVE_ENTRANCE, (double)x * 0.75, (double)y * 0.5, (double)z * 0.75);
        return s + 0.37 + j_clamped_lerp(0.3, 0.0, (10.0 + (double)y) / 40.0);
    }
    __device__ double roughness(int x, int y, int z) const {
        const double m = from_noise(T_ROUGH_MOD, (double)x, (double)y, (double)z, 0.0, 0.1);
        const double r = fabs(n3(T_ROUGH, (double)x, (double)y, (double)z));
        return (0.4 - r) * m;
    }
    __device__ double spaghetti3d(int x, int y, int z) const {
        const double rarity = n3(T_SPAG3D_RARITY, (double)(x * 2), (double)y, (double)(z * 2));
        const double e = rarity < -0.5 ? 0.75 : (rarity < 0.0 ? 1.0 : (rarity < 0.5 ? 1.5 : 2.0));
        const double h = from_noise(T_SPAG3D_THICK, (double)x, (double)y, (double)z, 0.065, 0.088);
        const double l = n3(T_SPAG3D_1, (double)x / e, (double)y / e, (double)z / e);
        const double m = fabs(e * l) - h;
        const double n = n3(T_SPAG3D_2, (double)x / e, (double)y / e, (double)z / e);
        const double o = fabs(e * n) - h;
        return jclamp(jmax(m, o), -1.0, 1.0);
    }
    __device__ double spaghetti2d(int x, int y, int z) const {
        const double mod = n3(T_SPAG2D_MOD, (double)(x * 2), (double)y, (double)(z * 2));
        const double s = mod < -0.75 ? 0.5 : (mod < -0.5 ? 0.75 : (mod < 0.5 ? 1.0 : (mod < 0.75 ? 2.0 : 3.0)));
        const double h = from_noise(T_SPAG2D_THICK, (double)(x * 2), (double)y, (double)(z * 2), 0.6, 1.3);
        const double v = n3(T_SPAG2D, (double)x / s, (double)y / s, (double)z / s);
        const double j = fabs(s * v) - 0.083 * h;
        const double l = from_noise(T_SPAG2D_ELEV, (double)x, 0.0, (double)z, -8.0, 8.0);
        double m = fabs(l - (double)y / 8.0) - h;
        m = m * m * m;
        return jclamp(jmax(m, j), -1.0, 1.0);
    }
    __device__ double pillar(int x, int y, int z) const {
        const double rare = from_noise(T_PILLAR_RARE, (double)x, (double)y, (double)z, 0.0, 2.0);
        const double thick = cube(from_noise(T_PILLAR_THICK, (double)x, (double)y, (double)z, 0.0, 1.1));
        const double f = n3(T_PILLAR, (double)x * 25.0, (double)y * 0.3, (double)z * 25.0);
        const double g = thick * (f * 2.0 - rare);
        return g > 0.03 ? g : -__builtin_huge_val();
    }
    __device__ double column(int x, int y, int z, const float tp[3], double base) const {
        // density.c:228-281 with jagged and noise caves on
        const double peaks = (double)tp[2];
        double jag = 0.0;
        if (peaks != 0.0) {
            const double n = n3(T_JAGGED, (double)x * 1500.0, 0.0, (double)z * 1500.0);
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
            const double rough = roughness(x, y, z);
            const double m = jmin(cave_entrance(x, y, z), spaghetti3d(x, y, z) + rough);
            if (i < 0.0) {
                cheese = g;
                spag = m * 5.0;
                pil = -64.0;
            } else {
                const double layer = [&] {
                    const double v = n3(T_CAVE_LAYER, (double)x, (double)(y * 8), (double)z);
                    return v * v * 4.0;
                }();
                if (layer > 64.0) {
                    cheese = 64.0;
                } else {
                    const double raw = n3(T_CAVE_CHEESE, (double)x, (double)y / 1.5, (double)z);
                    cheese = jclamp(raw + 0.27, -1.0, 1.0) + j_clamped_lerp(0.5, 0.0, i * 1.28) + layer;
                }
                spag = jmin(m, spaghetti2d(x, y, z) + rough);
                pil = pillar(x, y, z);
            }
        }
        double v = jmax(jmin(cheese, spag), pil);
        // slides over the cell index y / 8 (truncated), top then bottom (density.c:211-226)
        const int k = y / 8 - -8;
        v = j_clamped_lerp(-0.078125, v, (double)((48 - k) - 8) / 2.0);
        v = j_clamped_lerp(0.1171875, v, (double)(k - 0) / 3.0);
        return jclamp(v, -64.0, 64.0);
    }
    __device__ double noodle(int which, int x, int y, int z) const {
        // density.c:186-209
        const bool in = y >= -60 && y <= 320;
        const double sc = 2.6666666666666665;
        switch (which) {
        case 0: return in ? n3(T_NOODLE, (double)x, (double)y, (double)z) : -1.0;
        case 1: return in ? -0.07500000000000001 + -0.025 * n3(T_NOODLE_THICK, (double)x, (double)y, (double)z) : 0.0;
        case 2: return in ? n3(T_RIDGE_A, (double)x * sc, (double)y * sc, (double)z * sc) : 0.0;
        default: return in ? n3(T_RIDGE_B, (double)x * sc, (double)y * sc, (double)z * sc) : 0.0;
        }
    }

    // ---- preliminary surface level (column_top.c:28-122) ----
    // It depends on the block position through its quart position only (offset and factor are sampled there), as
    // the game's per-quart cache of it assumes; so does the deep-dark test of the aquifer. One evaluation gives
    // both: the surface, and the deep-dark offset (the offset where erosion is below -0.225, else -infinity, so
    // that y_gradient + it > 0.9 is the test of aquifer.c:218-228 at every y).
    __device__ void quart_surface(int qx, int qz, int &surface, double &dd_offset) const {
        double c, w, e, offset, factor;
        info(qx, qz, c, w, e);
        offset_factor(c, w, e, offset, factor);
        surface = surface_from(offset, factor);
        dd_offset = e < -0.22499999403953552 ? offset : -__builtin_huge_val();
    }
    __device__ static int surface_from(double offset, double factor) {
        const double inner = (1.0 / factor) * 0.2734375 + offset * -1.0;
        const double ub = jclamp(128.0 + -128.0 * inner, -40.0, 320.0);
        const int y0 = (int)(ub / 8.0) * 8;
        if (y0 <= -64) return -64;
        for (int y = y0; y >= -64; y -= 8) {
            const double yd = (double)y;
            const double depth = y_gradient(yd, -64.0, 320.0, 1.5, -1.5) + offset;
            const double dq = depth * factor;
            const double cheese = 4.0 * (dq > 0.0 ? dq : dq * 0.25);
            const double c = jclamp(-0.703125 + cheese, -64.0, 64.0);
            const double a1 = y_gradient(yd, 240.0, 256.0, 1.0, 0.0);
            const double a2 = y_gradient(yd, -64.0, -40.0, 0.0, 1.0);
            if (-0.390625 + (0.1171875 + a2 * (-0.1171875 + (-0.078125 + a1 * (0.078125 + c)))) > 0.0) return y;
        }
        return -64;
    }

    // ---- aquifer (aquifer.c) ----
    __device__ static void global_fluid(int y, int &level, int &type) {
        if (y < -54) { level = -54; type = BS_LAVA; }
        else { level = 63; type = BS_WATER; }
    }
    __device__ static int level_block(int level, int type, int y) { return y < level ? type : BS_AIR; }

    // computeFluid (aquifer.c:286-331) with computeSurfaceLevel (:248-273) and computeFluidType (:277-289); the
    // preliminary surfaces and the deep-dark test come from the chunk's quart table
    __device__ void fluid(const ChunkAq &A, int x, int y, int z, int &out_level, int &out_type) const {
        const int off[13][2] = {{0, 0}, {-2, -1}, {-1, -1}, {0, -1}, {1, -1}, {-3, 0}, {-2, 0},
                                {-1, 0}, {1, 0}, {-2, 1}, {-1, 1}, {0, 1}, {1, 1}};
        int gl, gt;
        global_fluid(y, gl, gt);
        int min_top = 0x7FFFFFFF;
        bool has_above = false;
        for (int i = 0; i < 13; i++) {
            const int top = A.prelim(x + off[i][0] * 16, z + off[i][1] * 16);
            const int adj = top + 8;
            const bool center = i == 0;
            if (center && y - 12 > adj) { out_level = gl; out_type = gt; return; }
            const bool above = y + 12 > adj;
            if (above || center) {
                int lv, ty;
                global_fluid(adj, lv, ty);
                if (level_block(lv, ty, adj) != BS_AIR) {
                    if (center) has_above = true;
                    if (above) { out_level = lv; out_type = ty; return; }
                }
            }
            if (top < min_top) min_top = top;
        }
        int level;
        if (A.deep_dark(x, y, z)) {
            level = kNoFluid;
        } else {
            const int d = (min_top + 8) - y;
            const double f = has_above ? j_clamped_lerp(1.0, 0.0, j_progress((double)d, 0.0, 64.0)) : 0.0;
            const double high = j_lerp_progress(f, 1.0, 0.0, -0.3, 0.8);
            const double low = j_lerp_progress(f, 1.0, 0.0, -0.8, 0.4);
            const double flood = jclamp(n3(T_AQ_FLOOD, (double)x, (double)y * 0.67, (double)z), -1.0, 1.0);
            if (flood - high > 0.0) {
                level = gl;
            } else if (flood - low > 0.0) {
                const int fx = floor_div(x, 16), fy = floor_div(y, 40), fz = floor_div(z, 16);
                const double n = n3(T_AQ_SPREAD, (double)fx, (double)fy * 0.7142857142857143, (double)fz) * 10.0;
                const int lvl = fy * 40 + 20 + (int)floor(n / 3.0) * 3;
                level = min_top < lvl ? min_top : lvl;
            } else {
                level = kNoFluid;
            }
        }
        int type = gt;
        if (level <= -10 && level != kNoFluid && gt != BS_LAVA) {
            const double n = n3(T_AQ_LAVA, (double)floor_div(x, 64), (double)floor_div(y, 40), (double)floor_div(z, 64));
            if (fabs(n) > 0.3) type = BS_LAVA;
        }
        out_level = level;
        out_type = type;
    }

    // the random anchor of an aquifer grid cell (aquifer.c:452-459, Mth.getSeed rng.c:170-175)
    __device__ void anchor(int gx, int gy, int gz, int &ax, int &ay, int &az) const {
        const int64_t xm = (int64_t)(int32_t)((uint32_t)gx * 3129871u);
        int64_t i = xm ^ ((int64_t)gz * 116129781ll) ^ (int64_t)gy;
        i = (int64_t)((uint64_t)i * (uint64_t)i * 42317861ull + (uint64_t)i * 11ull);
        Xr r = xr_make((uint64_t)(i >> 16) ^ aq_lo, aq_hi);
        ax = gx * 16 + (int)xr_next_bounded(r, 10u);
        ay = gy * 12 + (int)xr_next_bounded(r, 9u);
        az = gz * 16 + (int)xr_next_bounded(r, 10u);
    }

    __device__ double pressure(int x, int y, int z, double &barrier, bool &have_barrier, int la, int ta, int lb,
                               int tb) const {
        // aquifer.c:366-404
        const int sa = level_block(la, ta, y), sb = level_block(lb, tb, y);
        if ((sa == BS_LAVA && sb == BS_WATER) || (sa == BS_WATER && sb == BS_LAVA)) return 2.0;
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
        } else {
            if (!have_barrier) {
                barrier = n3(T_AQ_BARRIER, (double)x, (double)y * 0.5, (double)z);
                have_barrier = true;
            }
            inner = barrier;
        }
        return 2.0 * (inner + o);
    }

    // computeSubstance's block state (aquifer.c:416-563); skip_above is the column chunk's skipSamplingAboveY, the
    // grid cells' anchors and fluids (computed lazily by the reference, once per cell) come from the chunk's table
    __device__ int substance(const ChunkAq &A, int x, int y, int z, double density, int skip_above) const {
        if (density > 0.0) return BS_SOLID;
        int gl, gt;
        global_fluid(y, gl, gt);
        if (y > skip_above) return level_block(gl, gt, y);
        if (level_block(gl, gt, y) == BS_LAVA) return BS_LAVA;
        const int gx = floor_div(x - 5, 16), gy = floor_div(y + 1, 12), gz = floor_div(z - 5, 16);
        int d[4] = {0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF};
        int4 an[4] = {make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0), make_int4(0, 0, 0, 0)};
        for (int i = 0; i <= 1; i++)
            for (int j = -1; j <= 1; j++)
                for (int k = 0; k <= 1; k++) {
                    const int4 a = A.cell(gx + i, gy + j, gz + k);
                    const int ddx = a.x - x, ddy = a.y - y, ddz = a.z - z;
                    const int dist = ddx * ddx + ddy * ddy + ddz * ddz;
                    // the four nearest, a later cell taking a tie's place (aquifer.c:478-491)
                    int at = 4;
                    for (int s = 3; s >= 0; s--)
                        if (dist <= d[s]) at = s;
                    if (at == 4) continue;
                    for (int s = 3; s > at; s--) {
                        d[s] = d[s - 1];
                        an[s] = an[s - 1];
                    }
                    d[at] = dist;
                    an[at] = a;
                }
        const int l0 = an[0].w >> 2, t0 = an[0].w & 3;
        const double s01 = 1.0 - (double)(d[1] - d[0]) / 25.0;
        const int result = level_block(l0, t0, y);
        if (!(s01 > 0.0)) return result;
        {
            int bl, bt;
            global_fluid(y - 1, bl, bt);
            if (result == BS_WATER && level_block(bl, bt, y - 1) == BS_LAVA) return result;
        }
        double barrier = 0.0;
        bool have_barrier = false;
        const int l1 = an[1].w >> 2, t1 = an[1].w & 3;
        if (density + s01 * pressure(x, y, z, barrier, have_barrier, l0, t0, l1, t1) > 0.0) return BS_SOLID;
        const int l2 = an[2].w >> 2, t2 = an[2].w & 3;
        const double s02 = 1.0 - (double)(d[2] - d[0]) / 25.0;
        if (s02 > 0.0 && density + s01 * s02 * pressure(x, y, z, barrier, have_barrier, l0, t0, l2, t2) > 0.0)
            return BS_SOLID;
        const double s12 = 1.0 - (double)(d[2] - d[1]) / 25.0;
        if (s12 > 0.0 && density + s01 * s12 * pressure(x, y, z, barrier, have_barrier, l1, t1, l2, t2) > 0.0)
            return BS_SOLID;
        return result;
    }
};

// Mth.lerp3 over one cell's corner columns, x innermost and z outermost (phase2.c:55-65, 132-150).
__device__ __forceinline__ double cell_lerp(double dx, double dy, double dz, const double *c00, const double *c10,
                                            const double *c01, const double *c11, int cy) {
    const double a = j_lerp(dy, c00[cy], c00[cy + 1]);
    const double b = j_lerp(dy, c10[cy], c10[cy + 1]);
    const double c = j_lerp(dy, c01[cy], c01[cy + 1]);
    const double d = j_lerp(dy, c11[cy], c11[cy + 1]);
    return j_lerp(dz, j_lerp(dx, a, b), j_lerp(dx, c, d));
}

}  // namespace whf
