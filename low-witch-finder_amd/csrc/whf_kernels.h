// whf_kernels.h — kernel-side data layouts and the phase-1 gate functions shared by the scan kernels.
#pragma once

#include "whf_device.h"
#include "whf_tables.h"

namespace whf {

// One Perlin sampler to build: MD5 digests (as two big-endian longs) of the noise identifier and of the
// octave's "octave_<n>" name, and which of the two octave samplers of the double noise it belongs to.
struct BuildDesc {
    uint64_t na, nb;   // MD5(identifier)
    uint64_t oa, ob;   // MD5("octave_" + (firstOctave + slot))
    int sub;           // 0 = first sampler, 1 = second (second deriver of the same noise random)
    int pad;
};

// A double-Perlin noise as the gates sample it: per sub-sampler the set-relative table index of each
// non-empty octave in slot order, its lacunarity 2^(firstOctave+slot), amplitude and persistence.
struct NoiseDesc {
    int set;                 // SET_E / SET_C / SET_L
    int np;                  // non-empty octaves per sub-sampler
    int tab[2][16];
    double amplitude;        // (1/6) / (0.1 * (1 + 1/(span+1)))  (noise.c:116-130)
    double lac[16], amp[16], per[16];
};

// Tiered erosion: evaluation k uses set-E table k with coordinate scale sc[k] * lac[k] and weight w[k]
// (= amp * persistence of its slot); thr[t] is checked after evaluation t + 1 (noise.c:359-459).
struct EroDesc {
    double sc[8], lac[8], w[8];
    double w23[8];      // w * 2^-23: the weights of the FP32 gate, whose values come scaled by 2^23 (exact)
    double amplitude;
    double thr[7];
    double margin[8];   // FP32 gate: bound on |estimate - FP64 value| of sum * amplitude after evaluation k
    double coord_lim;   // |x|, |z| below this keep every scaled coordinate inside maintainPrecision's identity range
};

// Division by a divisor fixed per launch, by multiply-high (Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994, fig. 4.1): q = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(m, n).
struct FastDiv {
    uint32_t d, m, l;
    __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
        if (d == 1) return n;
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
        return (t + ((n - t) >> 1)) >> (l - 1);
    }
};
__host__ __device__ inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0u, 0u};
    if (d <= 1) { f.d = 1; return f; }
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    f.l = l;
    f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    return f;
}

// Everything the gate kernels read about a request's noises.
struct GateTables {
    NoiseDesc noise[14];
    EroDesc ero;
};

// Per-seed-slot table storage of one sampler set.
struct SetView {
    const uint8_t *tab;   // [slot][n][256]
    const SPar *par;      // [slot][n]
    int n;
};

struct GPerm {
    const uint8_t *p;
    __device__ __forceinline__ int operator()(int i) const { return p[i & 255]; }
};

// Byte pairs P(i) | P(i + 1) << 8 read from a 256-byte table in HBM.
struct GPair {
    const uint8_t *p;
    __device__ __forceinline__ uint32_t operator()(int i) const {
        return (uint32_t)p[i & 255] | (uint32_t)p[(i + 1) & 255] << 8;
    }
    template <int SEL>
    __device__ __forceinline__ uint32_t next(uint32_t pair, int a) const {
        return (*this)((int)((pair >> (8 * SEL)) & 255u) + a);
    }
};

// Double-Perlin 3D sample over tables in HBM (noise.c:263-289).
__device__ __forceinline__ double noise_3d(const NoiseDesc &d, const SetView &sv, uint32_t slot, double x,
                                           double y, double z) {
    double out[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double sc = s ? 1.0181268882175227 : 1.0;
        const double X = s ? x * sc : x, Y = s ? y * sc : y, Z = s ? z * sc : z;
        double v = 0.0;
        for (int p = 0; p < d.np; p++) {
            const int t = d.tab[s][p];
            const size_t ti = (size_t)slot * sv.n + t;
            const SPar sp = sv.par[ti];
            const double lac = d.lac[p];
            double pv = perlin_3d(GPerm{sv.tab + ti * 256}, sp, keep_precision(X * lac), keep_precision(Y * lac),
                                  keep_precision(Z * lac));
            v += (d.amp[p] * pv) * d.per[p];
        }
        out[s] = v;
    }
    return (out[0] + out[1]) * d.amplitude;
}

// Double-Perlin y = 0 sample over tables in HBM (noise.c:320-342).
__device__ __forceinline__ double noise_y0(const NoiseDesc &d, const SetView &sv, uint32_t slot, double x, double z) {
    double out[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double sc = s ? 1.0181268882175227 : 1.0;
        const double X = s ? x * sc : x, Z = s ? z * sc : z;
        double v = 0.0;
        for (int p = 0; p < d.np; p++) {
            const int t = d.tab[s][p];
            const size_t ti = (size_t)slot * sv.n + t;
            const SPar sp = sv.par[ti];
            const double lac = d.lac[p];
            double pv = perlin_y0(GPerm{sv.tab + ti * 256}, sp, keep_precision(X * lac), keep_precision(Z * lac));
            v += (d.amp[p] * pv) * d.per[p];
        }
        out[s] = v;
    }
    return (out[0] + out[1]) * d.amplitude;
}

// The three cave heuristics of the phase-1 ladder (caves.c:58-130) over sets C and L of one seed.
struct CaveCtx {
    const GateTables *gt;
    SetView C, L;
    uint32_t slotC, slotL;

    __device__ double n3(int g, double x, double y, double z) const {
        const NoiseDesc &d = gt->noise[g];
        return d.set == SET_C ? noise_3d(d, C, slotC, x, y, z) : noise_3d(d, L, slotL, x, y, z);
    }

    // Entrance's first term: cave_entrance + 0.37 + clampedLerp(0.3, 0, (10 + y) / 40)
    __device__ double entrance_c(int x, int y, int z) const {
        double ent = n3(G_CAVE_ENTRANCE, (double)x * 0.75, (double)y * 0.5, (double)z * 0.75);
        double t = (10.0 + (double)y) / 40.0;
        double ramp = t < 0.0 ? 0.3 : (t > 1.0 ? 0.0 : 0.3 + t * (0.0 - 0.3));
        return ent + 0.37 + ramp;
    }

    // The spaghetti pair shared by Entrance and Entrance2 (caves.c:85-104).
    __device__ double spaghetti(int x, int y, int z) const {
        const double X = (double)x, Y = (double)y, Z = (double)z;
        double rar = n3(G_SPAG_RARITY, (double)(int)((uint32_t)x * 2u), Y, (double)(int)((uint32_t)z * 2u));
        double e = rar < -0.5 ? 0.75 : (rar < 0.0 ? 1.0 : (rar < 0.5 ? 1.5 : 2.0));
        double tv = n3(G_SPAG_THICKNESS, X, Y, Z);
        double th = 0.065 + ((tv - -1.0) / (1.0 - -1.0)) * (0.088 - 0.065);
        double xe = X / e, ye = Y / e, ze = Z / e;
        double l = n3(G_SPAG_1, xe, ye, ze);
        double m = fabs(e * l) - th;
        double n = n3(G_SPAG_2, xe, ye, ze);
        double o = fabs(e * n) - th;
        double p = jclamp(jmax(m, o), -1.0, 1.0);
        double mod = n3(G_ROUGH_MOD, X, Y, Z);
        double rg = n3(G_ROUGH, X, Y, Z);
        double q = (-0.05 + (-0.05 * mod)) * (-0.4 + fabs(rg));
        return p + q;
    }

    __device__ double entrance(int x, int y, int z) const { return jmin(entrance_c(x, y, z), spaghetti(x, y, z)); }

    // Entrance(x, y, z) >= 0, evaluating the spaghetti pair only when the first term alone does not decide it:
    // min(c, s) >= 0 exactly when c >= 0 and s >= 0 (no NaN reaches here).
    __device__ bool entrance_nonneg(int x, int y, int z) const {
        double c = entrance_c(x, y, z);
        if (!(c >= 0.0)) return false;
        return jmin(c, spaghetti(x, y, z)) >= 0.0;
    }

    __device__ double cheese(int x, int y, int z) const {
        double layer = n3(G_CAVE_LAYER, (double)x, (double)(int)((uint32_t)y * 8u), (double)z);
        double a = 4.0 * layer * layer;
        double b = jclamp(0.27 + n3(G_CAVE_CHEESE, (double)x, (double)y * 0.6666666666666666, (double)z), -1.0, 1.0);
        return a + b;
    }

    __device__ double flooded(int hut_x, int hut_z, int y) const {
        return n3(G_FLOODEDNESS, (double)(hut_x + 3), (double)y * 0.67, (double)(hut_z + 3));
    }
};

}  // namespace whf
