// whf_device.h — device-side building blocks of the witch-hut Seed Check for gfx950 (CDNA4, wave64).
//
// Integer parts (SplitMix64, Xoroshiro128++, java.util.Random placement) are exact by construction. The
// noise parts are IEEE binary64 and must reproduce the reference CPU scanner bit for bit, so this code is
// compiled with -ffp-contract=off and keeps the reference's association order in every expression
// (lysh noise.c:36-58, 224-342; caves.c:58-130; mcmath.h), including x / s as a true division.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace whf {

// ---------------------------------------------------------------------------------------------- seeds

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

// ---------------------------------------------------------------------------------------------- Xoroshiro128++
// The game's XoroshiroRandomSource (lysh rng.c:75-181).

struct Xr {
    uint64_t lo, hi;
};

__device__ __forceinline__ uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ __forceinline__ Xr xr_make(uint64_t lo, uint64_t hi) {
    if ((lo | hi) == 0) { lo = kGolden; hi = 0x6A09E667F3BCC909ull; }
    return Xr{lo, hi};
}

__device__ __forceinline__ Xr xr_seed(uint64_t seed) {
    uint64_t l = seed ^ 0x6A09E667F3BCC909ull;
    return Xr{mix64(l), mix64(l - 0x61C8864680B583EBull)};
}

__device__ __forceinline__ uint64_t xr_next(Xr &s) {
    uint64_t lo = s.lo, hi = s.hi;
    uint64_t out = rotl(lo + hi, 17) + lo;
    hi ^= lo;
    s.lo = rotl(lo, 49) ^ hi ^ (hi << 21);
    s.hi = rotl(hi, 28);
    return out;
}

// Java 17 Random.nextInt(bound) over a 64-bit source: Lemire multiply-shift with rejection (rng.c:124-146).
__device__ __forceinline__ uint32_t xr_next_bounded(Xr &s, uint32_t bound) {
    uint64_t m = (xr_next(s) & 0xFFFFFFFFull) * (uint64_t)bound;
    if ((uint32_t)m < bound) {
        uint32_t lim = (0u - bound) % bound;
        while ((uint32_t)m < lim) m = (xr_next(s) & 0xFFFFFFFFull) * (uint64_t)bound;
    }
    return (uint32_t)(m >> 32);
}

// The same engine as four 32-bit words (lo = a1:a0, hi = b1:b0) for the permutation shuffle: each 64-bit rotation is
// two alignbits, and a draw yields only the low word of nextLong, which is what nextInt uses.
struct Xw {
    uint32_t a0, a1, b0, b1;
};
__device__ __forceinline__ Xw xw_of(const Xr &s) {
    return Xw{(uint32_t)s.lo, (uint32_t)(s.lo >> 32), (uint32_t)s.hi, (uint32_t)(s.hi >> 32)};
}
__device__ __forceinline__ Xr xr_of(const Xw &s) {
    return Xr{((uint64_t)s.a1 << 32) | s.a0, ((uint64_t)s.b1 << 32) | s.b0};
}
__device__ __forceinline__ uint32_t xw_next32(Xw &s) {
    const uint64_t sum = (((uint64_t)s.a1 << 32) | s.a0) + (((uint64_t)s.b1 << 32) | s.b0);
    const uint32_t out = __builtin_amdgcn_alignbit((uint32_t)sum, (uint32_t)(sum >> 32), 15) + s.a0;   // rotl 17
    const uint32_t h0 = s.b0 ^ s.a0, h1 = s.b1 ^ s.a1;                                                 // hi ^= lo
    const uint32_t r0 = __builtin_amdgcn_alignbit(s.a1, s.a0, 15), r1 = __builtin_amdgcn_alignbit(s.a0, s.a1, 15);
    // rotl(lo, 49) ^ hi ^ (hi << 21): three-input XOR (truth table 0x96) in one instruction
    s.a0 = __builtin_amdgcn_bitop3_b32(r0, h0, h0 << 21, 0x96);
    s.a1 = __builtin_amdgcn_bitop3_b32(r1, h1, __builtin_amdgcn_alignbit(h1, h0, 11), 0x96);
    s.b0 = __builtin_amdgcn_alignbit(h0, h1, 4);                        // rotl(hi, 28)
    s.b1 = __builtin_amdgcn_alignbit(h1, h0, 4);
    return out;
}

__device__ __forceinline__ double xr_next_double(Xr &s) {
    return (double)(xr_next(s) >> 11) * 0x1.0p-53;
}

// createRandomDeriver(): two longs of the parent (rng.c:157-161).
__device__ __forceinline__ Xr xr_fork(Xr &parent) {
    uint64_t a = xr_next(parent);
    uint64_t b = xr_next(parent);
    return xr_make(a, b);
}

// ---------------------------------------------------------------------------------------------- placement
// java.util.Random region placement of the swamp hut (structure.c:4-24, rng.c:187-227).

constexpr uint64_t kLcgMul = 0x5DEECE66Dull;
constexpr uint64_t kLcgM48 = 0xFFFFFFFFFFFFull;

__device__ __forceinline__ int lcg_next31(uint64_t &st) {
    st = (st * kLcgMul + 0xBull) & kLcgM48;
    return (int)(st >> 17);
}

__device__ __forceinline__ int lcg_int24(uint64_t &st) {
    for (;;) {
        int bits = lcg_next31(st);
        int val = bits % 24;
        if ((int32_t)((uint32_t)bits - (uint32_t)val + 23u) >= 0) return val;
    }
}

__device__ __forceinline__ void hut_chunk(uint64_t seed, int rx, int rz, int salt, int &cx, int &cz) {
    uint64_t s = (uint64_t)((int64_t)rx * 341873128712ll) + (uint64_t)((int64_t)rz * 132897987541ll) + seed +
                 (uint64_t)(int64_t)salt;
    uint64_t st = (s ^ kLcgMul) & kLcgM48;
    int j = lcg_int24(st);
    int k = lcg_int24(st);
    cx = rx * 32 + j;
    cz = rz * 32 + k;
}

// ---------------------------------------------------------------------------------------------- Perlin
// A sampler is a 256-byte permutation plus its origin; y-constants for the y = 0 climate path are
// precomputed at build time exactly as the reference does (noise.c:73-77).

struct SPar {
    double ox, oy, oz;
    double hy, hy1, fy;   // oy - floor(oy), hy - 1, fade(hy)
    int jy;               // (int)floor(oy)
    int pad;
};
static_assert(sizeof(SPar) == 56, "SPar layout");

// Gradient table of the game's improved noise (noise.c:15-20) as bit masks over the 4-bit hash: for each
// axis, which hashes give -1 and which give 0. Built at compile time from the table itself.
struct GradMasks {
    uint32_t neg[3], zero[3];
};
constexpr GradMasks make_grad_masks() {
    const int g[16][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
                          {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1},
                          {1, 1, 0}, {0, -1, 1}, {-1, 1, 0}, {0, -1, -1}};
    GradMasks m{{0, 0, 0}, {0, 0, 0}};
    for (int h = 0; h < 16; h++)
        for (int a = 0; a < 3; a++) {
            if (g[h][a] < 0) m.neg[a] |= 1u << h;
            if (g[h][a] == 0) m.zero[a] |= 1u << h;
        }
    return m;
}
constexpr GradMasks kGrad = make_grad_masks();

// The gradient component as a double (-1.0, 0.0 or 1.0) built from its bit pattern, so g * x below is the
// reference's (double)g[a] * x exactly, signed zeros included.
__device__ __forceinline__ double grad_comp(uint32_t h, uint32_t neg, uint32_t zero) {
    uint32_t hi = (((zero >> h) & 1u) ? 0u : 0x3FF00000u) | (((neg >> h) & 1u) << 31);
    return __hiloint2double((int)hi, 0);
}

__device__ __forceinline__ double grad(int hash, double x, double y, double z) {
    uint32_t h = (uint32_t)hash & 15u;
    return grad_comp(h, kGrad.neg[0], kGrad.zero[0]) * x + grad_comp(h, kGrad.neg[1], kGrad.zero[1]) * y +
           grad_comp(h, kGrad.neg[2], kGrad.zero[2]) * z;
}

__device__ __forceinline__ double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
__device__ __forceinline__ double lerp(double t, double a, double b) { return a + t * (b - a); }

// Fisher-Yates of PerlinNoiseSampler(random) (noise.c:61-78) over 0..255 with nextInt(256 - i), in the calling lane's
// column of a 64-lane LDS area laid out lane-interleaved (word w of lane l at words[w * 64 + l]), so the 64 lanes'
// byte accesses hit 64 distinct banks whatever the swap index. pr ends where the reference's random does.
__device__ __forceinline__ void shuffle_perm(Xr &pr, uint32_t *words, uint32_t lane) {
    uint8_t *bytes = reinterpret_cast<uint8_t *>(words);
    const uint32_t base = lane * 4;   // byte address of position 0 in this lane's column
    auto identity = [&] {
#pragma unroll 8
        for (int w = 0; w < 64; w++) words[w * 64 + lane] = 0x03020100u + (uint32_t)w * 0x04040404u;
    };
    // byte j of this lane's column: word j >> 2 (j < 256) at 256-byte stride, byte j & 3 within it
    auto addr = [&](uint32_t j) { return (__builtin_amdgcn_ubfe(j, 2u, 6u) << 8) + ((j & 3u) | base); };
    identity();
    // Fast path: nextInt(b) with no rejection test. A draw can only be rejected when the low word of its product
    // with b is below b <= 256 (rng.c:124-146), so the lane tracks the smallest such word; if it ends below 256 the
    // lane replays the shuffle below with the exact nextInt. The draw of step i + 1 is computed between the reads
    // and the writes of step i: it does not depend on the table, so it fills the LDS read latency. Four steps per
    // iteration keep the in-word byte position a constant offset.
    const Xr start = pr;
    Xw w = xw_of(pr);
    const uint32_t x0 = xw_next32(w);
    uint32_t lo_min = x0 << 8;        // the low word of x0 * 256
    uint32_t jj = x0 >> 24;           // nextInt(256)
    // step i (i < 255) draws nextInt(b) for step i + 1 with b = 255 - i
    auto step = [&](uint32_t ai, uint32_t b, bool draw) {
        const uint32_t aj = addr(jj);
        const uint8_t vi = bytes[ai];
        const uint8_t vj = bytes[aj];
        if (draw) {
            const uint64_t m = (uint64_t)xw_next32(w) * b;
            lo_min = min(lo_min, (uint32_t)m);
            jj = (256u - b) + (uint32_t)(m >> 32);
        }
        bytes[ai] = vj;
        bytes[aj] = vi;
    };
    uint32_t aq = base, bq = 255u;
    for (int q = 0; q < 63; q++, aq += 256u, bq -= 4u) {
        // an opaque bound: otherwise the compiler turns the products into 64-bit induction variables and full
        // 64-bit multiplies; opaque, each product is one 32 x 32 -> 64 multiply
        asm volatile("" : "+s"(bq));
#pragma unroll
        for (int r = 0; r < 4; r++) step(aq + (uint32_t)r, bq - (uint32_t)r, true);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) step(aq + (uint32_t)r, 3u - (uint32_t)r, r < 3);
    pr = xr_of(w);
    if (lo_min < 256u) {
        // exact replay (odds about 2^-16 per sampler): 256 draws nextInt(256 - i), as above but with the rejection
        pr = start;
        identity();
        uint32_t ai = base;
        uint32_t j = xr_next_bounded(pr, 256u);
        for (int i = 0; i < 256; i++) {
            const uint32_t aj = addr(j);
            const uint8_t vi = bytes[ai];
            const uint8_t vj = bytes[aj];
            if (i < 255) j = (uint32_t)(i + 1) + xr_next_bounded(pr, 255u - (uint32_t)i);
            bytes[ai] = vj;
            bytes[aj] = vi;
            ai += ((i & 3) == 3) ? 256 - 3 : 1;
        }
    }
}

// One Perlin sampler from random pr (PerlinNoiseSampler(random), noise.c:61-78): the origin, then the permutation
// (shuffle_perm); the table then goes to out_tab (256 bytes) and the y constants to out.
__device__ __forceinline__ void build_sampler(Xr &pr, uint32_t *words, uint32_t lane, uint8_t *out_tab, SPar &out) {
    const double ox = xr_next_double(pr) * 256.0;
    const double oy = xr_next_double(pr) * 256.0;
    const double oz = xr_next_double(pr) * 256.0;
    shuffle_perm(pr, words, lane);
    uint4 *dst = reinterpret_cast<uint4 *>(out_tab);
#pragma unroll
    for (int q = 0; q < 16; q++)
        dst[q] = make_uint4(words[(4 * q) * 64 + lane], words[(4 * q + 1) * 64 + lane], words[(4 * q + 2) * 64 + lane],
                            words[(4 * q + 3) * 64 + lane]);
    SPar sp;
    sp.ox = ox;
    sp.oy = oy;
    sp.oz = oz;
    const double fj = floor(oy);
    sp.jy = (int)fj;
    sp.hy = oy - fj;
    sp.hy1 = sp.hy - 1.0;
    sp.fy = fade(sp.hy);
    sp.pad = 0;
    out = sp;
}

// Corner blend shared by the y = 0 and 3D samplers (noise.c:224-244): P is the permutation, (i, j, k) the
// lattice cell, (x, y, z) the fractional position and ty the y value whose fade is used.
template <typename Perm>
__device__ __forceinline__ double perlin_cell(const Perm &P, int i, int j, int k, double x, double y, double z,
                                              double fy) {
    int a = P(i), b = P(i + 1);
    int c = P(a + j), d = P(a + j + 1), e = P(b + j), f = P(b + j + 1);
    double x1 = x - 1, y1 = y - 1, z1 = z - 1;
    double v000 = grad(P(c + k), x, y, z);
    double v100 = grad(P(e + k), x1, y, z);
    double v010 = grad(P(d + k), x, y1, z);
    double v110 = grad(P(f + k), x1, y1, z);
    double v001 = grad(P(c + k + 1), x, y, z1);
    double v101 = grad(P(e + k + 1), x1, y, z1);
    double v011 = grad(P(d + k + 1), x, y1, z1);
    double v111 = grad(P(f + k + 1), x1, y1, z1);
    double fx = fade(x), fz = fade(z);
    double lo = lerp(fy, lerp(fx, v000, v100), lerp(fx, v010, v110));
    double hi = lerp(fy, lerp(fx, v001, v101), lerp(fx, v011, v111));
    return lerp(fz, lo, hi);
}

// y = 0 sample with the precomputed y constants (noise.c:292-318).
template <typename Perm>
__device__ __forceinline__ double perlin_y0(const Perm &P, const SPar &s, double x, double z) {
    double d = x + s.ox, f = z + s.oz;
    double fi = floor(d), fk = floor(f);
    int i = (int)fi, k = (int)fk;
    double g = d - fi, l = f - fk;
    // The y-row uses hy and hy - 1 as its two corner offsets and fade(hy) as its weight.
    int a = P(i), b = P(i + 1);
    int j = s.jy;
    int c = P(a + j), dd = P(a + j + 1), e = P(b + j), ff = P(b + j + 1);
    double h = s.hy, h1 = s.hy1, g1 = g - 1, l1 = l - 1;
    double v000 = grad(P(c + k), g, h, l);
    double v100 = grad(P(e + k), g1, h, l);
    double v010 = grad(P(dd + k), g, h1, l);
    double v110 = grad(P(ff + k), g1, h1, l);
    double v001 = grad(P(c + k + 1), g, h, l1);
    double v101 = grad(P(e + k + 1), g1, h, l1);
    double v011 = grad(P(dd + k + 1), g, h1, l1);
    double v111 = grad(P(ff + k + 1), g1, h1, l1);
    double fx = fade(g), fz = fade(l), fy = s.fy;
    double lo = lerp(fy, lerp(fx, v000, v100), lerp(fx, v010, v110));
    double hi = lerp(fy, lerp(fx, v001, v101), lerp(fx, v011, v111));
    return lerp(fz, lo, hi);
}

// 3D sample with yScale = yMax = 0 (noise.c:246-261).
template <typename Perm>
__device__ __forceinline__ double perlin_3d(const Perm &P, const SPar &s, double x, double y, double z) {
    double d = x + s.ox, e = y + s.oy, f = z + s.oz;
    double fi = floor(d), fj = floor(e), fk = floor(f);
    double h = e - fj;
    return perlin_cell(P, (int)fi, (int)fj, (int)fk, d - fi, h, f - fk, fade(h));
}

// The same gradient dot product by selects: u = x or y, v = y, x or z, signs from the low two hash bits (Perlin's
// improved-noise formulation of the same 16-entry table). It equals the reference's three-term sum
// (g0*x + g1*y) + g2*z bit for bit unless both live terms are zeros: adding the zero component to a non-zero
// sum, or to the +0 of an exact cancellation, changes nothing, and reordering two terms is exact. Callers take
// the three-term form whenever a coordinate that can reach a live term is zero.
__device__ __forceinline__ double grad_sel(uint32_t h, double x, double y, double z) {
    h &= 15u;
    double u = h < 8u ? x : y;
    double v = h < 4u ? y : ((h == 12u || h == 14u) ? x : z);
    if (h & 1u) u = -u;
    if (h & 2u) v = -v;
    return u + v;
}

// ((byte SEL of v) + a) & 255 in one instruction: an SDWA add that reads byte SEL of its first source and writes
// byte 0 of the result with the other bits cleared. The permutation's lookup chain is all such index steps.
template <int SEL>
__device__ __forceinline__ uint32_t byte_add(uint32_t v, uint32_t a) {
    static_assert(SEL == 0 || SEL == 1, "byte 0 or 1 of a pair");
    uint32_t d;
    if constexpr (SEL == 0)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:BYTE_0 dst_unused:UNUSED_PAD src0_sel:BYTE_0 src1_sel:DWORD"
            : "=v"(d) : "v"(v), "v"(a));
    else
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:BYTE_0 dst_unused:UNUSED_PAD src0_sel:BYTE_1 src1_sel:DWORD"
            : "=v"(d) : "v"(v), "v"(a));
    return d;
}

// The pair tables the y = 0 evaluations read (LDS or HBM) share one interface: Q(i) is the byte pair
// P(i & 255) | P((i + 1) & 255) << 8 for a lattice coordinate i, and Q.next<SEL>(pair, a) is Q((byte SEL of pair) + a),
// the step from one level of the lattice hash to the next.

// y = 0 Perlin evaluation of two independent samples at once (two tables, two origins, two points), written as
// lock-step lanes of a 2-array so the compiler interleaves the two dependency chains. Q(i) returns the byte pair
// P(i & 255) | P((i + 1) & 255) << 8 of the permutation; every lookup and every arithmetic step is the one of
// perlin_y0 above, so the two values are bit-identical to two perlin_y0 calls.
template <typename Pairs>
__device__ __forceinline__ void perlin_y0_x2(const Pairs *Q, const SPar *const *s, const double *x, const double *z,
                                             double *out) {
    double g[2], l[2], fx[2], fz[2];
    int i[2], k[2];
    uint32_t ab[2], cd[2], ef[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        double d = x[t] + s[t]->ox, f = z[t] + s[t]->oz;
        double fi = floor(d), fk = floor(f);
        i[t] = (int)fi;
        k[t] = (int)fk;
        g[t] = d - fi;
        l[t] = f - fk;
    }
#pragma unroll
    for (int t = 0; t < 2; t++) ab[t] = Q[t](i[t]);
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int j = s[t]->jy;
        cd[t] = Q[t].template next<0>(ab[t], j);
        ef[t] = Q[t].template next<1>(ab[t], j);
    }
    uint32_t h0[2], h1[2], h2[2], h3[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        h0[t] = Q[t].template next<0>(cd[t], k[t]);   // P(c+k),  P(c+k+1)
        h1[t] = Q[t].template next<0>(ef[t], k[t]);   // P(e+k),  P(e+k+1)
        h2[t] = Q[t].template next<1>(cd[t], k[t]);   // P(d+k),  P(d+k+1)
        h3[t] = Q[t].template next<1>(ef[t], k[t]);   // P(f+k),  P(f+k+1)
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const double h = s[t]->hy, hm = s[t]->hy1, gm = g[t] - 1, lm = l[t] - 1;
        double c[8];
        c[0] = grad_sel(h0[t], g[t], h, l[t]);
        c[1] = grad_sel(h1[t], gm, h, l[t]);
        c[2] = grad_sel(h2[t], g[t], hm, l[t]);
        c[3] = grad_sel(h3[t], gm, hm, l[t]);
        c[4] = grad_sel(h0[t] >> 8, g[t], h, lm);
        c[5] = grad_sel(h1[t] >> 8, gm, h, lm);
        c[6] = grad_sel(h2[t] >> 8, g[t], hm, lm);
        c[7] = grad_sel(h3[t] >> 8, gm, hm, lm);
        // grad_sel differs from the three-term form only when both live terms of a corner are zero, which needs
        // a zero coordinate: g, h or l (g - 1, h - 1, l - 1 are never zero since g, h, l lie in [0, 1))
        if (g[t] == 0.0 || l[t] == 0.0 || h == 0.0) {
            c[0] = grad((int)h0[t], g[t], h, l[t]);
            c[1] = grad((int)h1[t], gm, h, l[t]);
            c[2] = grad((int)h2[t], g[t], hm, l[t]);
            c[3] = grad((int)h3[t], gm, hm, l[t]);
            c[4] = grad((int)(h0[t] >> 8), g[t], h, lm);
            c[5] = grad((int)(h1[t] >> 8), gm, h, lm);
            c[6] = grad((int)(h2[t] >> 8), g[t], hm, lm);
            c[7] = grad((int)(h3[t] >> 8), gm, hm, lm);
        }
        fx[t] = fade(g[t]);
        fz[t] = fade(l[t]);
        const double fy = s[t]->fy;
        const double lo = lerp(fy, lerp(fx[t], c[0], c[1]), lerp(fx[t], c[2], c[3]));
        const double hi = lerp(fy, lerp(fx[t], c[4], c[5]), lerp(fx[t], c[6], c[7]));
        out[t] = lerp(fz[t], lo, hi);
    }
}

// ---------------------------------------------------------------------------------------------- FP32 gate core
// The same two evaluations for the FP32 gate. The lattice part (origin add, floor, fraction) stays FP64, so the cell
// is the reference's. The fractional coordinates are then truncated to 23-bit fixed point, X = floor(x * 2^23), and
// each corner's gradient dot +-X +-Y (or Z) is an exact integer: the gradient's three components come from 2-bit
// fields of three 32-bit constants, indexed by the hash, and meet the coordinates in 24-bit multiply-adds. A dot is
// at most 2^24 in magnitude, so it converts to float exactly. Fades and lerps run in FP32 (with FMA) on the dots
// scaled by 2^23, which scales the result by 2^23 exactly; the caller's weights carry the 2^-23.
// Worst-case error against the FP64 value (absolute, unscaled):
//   fixed-point coordinates: each within 2^-23 below its double, so a dot is within 2^-22 = 2.4e-7;
//   fade(t) from the float of t (within 2^-25): |f'| <= 1.875 on the input, five roundings    <= 1.7e-6;
//   first lerp level: dots' error + rounding of b - a and of the FMA + fade error x 4           <= 7.3e-6;
//   second level (y fade, the float of the FP64 fade, within 2^-25 x 4):                        <= 7.6e-6;
//   third level: 7.6e-6 + roundings + fade error x 4                                            <= 1.5e-5.
// The gate uses four times that, kFp32PerlinBound = 2^-14, as the margin around every threshold; the measured
// maximum (whf_noise_error) is reported beside it.
constexpr double kFp32PerlinBound = 6.103515625e-05;   // 2^-14
constexpr int kFixBits = 23;
constexpr double kFixOne = 8388608.0;                   // 2^23
constexpr double kFixInv = 1.0 / 8388608.0;             // 2^-23, exact

// What the FP32 gate needs of a y = 0 sampler: origin x and z, the lattice y, the fixed-point y offsets of the two
// corner rows (hy and hy - 1) and the y fade as a float.
struct EPar {
    double ox, oz;
    int jy, hq, hmq;
    float fy;
};
static_assert(sizeof(EPar) == 32, "EPar layout");

__device__ __forceinline__ EPar make_epar(const SPar &s) {
    EPar e;
    e.ox = s.ox;
    e.oz = s.oz;
    e.jy = s.jy;
    e.hq = (int)(s.hy * kFixOne);   // hy in [0, 1): truncation = floor
    e.hmq = e.hq - (1 << kFixBits);
    e.fy = (float)s.fy;
    return e;
}

// The gradient table of kGrad as 2-bit two's-complement fields (bits 2h, 2h + 1 of c[axis]).
struct GradLut {
    uint32_t c[3];
};
constexpr GradLut make_grad_lut() {
    GradLut l{{0, 0, 0}};
    for (int h = 0; h < 16; h++)
        for (int a = 0; a < 3; a++) {
            const bool neg = (kGrad.neg[a] >> h) & 1u, zero = (kGrad.zero[a] >> h) & 1u;
            const uint32_t f = zero ? 0u : (neg ? 3u : 1u);
            l.c[a] |= f << (2 * h);
        }
    return l;
}
constexpr GradLut kGradLut = make_grad_lut();

// Gradient dot of hash h (off = 2 * (h & 15) in its low 5 bits) with fixed-point (X, Y, Z): exact.
__device__ __forceinline__ int grad_q(uint32_t off, int X, int Y, int Z) {
    const int gx = __builtin_amdgcn_sbfe((int)kGradLut.c[0], off, 2u);
    const int gy = __builtin_amdgcn_sbfe((int)kGradLut.c[1], off, 2u);
    const int gz = __builtin_amdgcn_sbfe((int)kGradLut.c[2], off, 2u);
    return __mul24(gx, X) + __mul24(gy, Y) + __mul24(gz, Z);
}
__device__ __forceinline__ float fade_q(float t) {
    const float p = __builtin_fmaf(t, __builtin_fmaf(t, 6.0f, -15.0f), 10.0f);
    return t * t * t * p;
}
__device__ __forceinline__ float lerp_q(float t, float a, float b) { return __builtin_fmaf(t, b - a, a); }

// N independent y = 0 evaluations (pair tables Q[t], samplers s[t], points (x[t], z[t])) scaled by 2^23, written as
// lock-step lanes of N-arrays so the compiler interleaves their dependency chains.
template <int N, typename Pairs>
__device__ __forceinline__ void perlin_y0_xn_q(const Pairs *Q, const EPar *const *s, const double *x, const double *z,
                                               double *out) {
    int X[N], Z[N], i[N], k[N];
    float gf[N], lf[N];
    uint32_t ab[N], cd[N], ef[N];
#pragma unroll
    for (int t = 0; t < N; t++) {
        const double d = x[t] + s[t]->ox, f = z[t] + s[t]->oz;
        const double fi = floor(d), fk = floor(f);
        i[t] = (int)fi;
        k[t] = (int)fk;
        const double gd = d - fi, ld = f - fk;
        X[t] = (int)(gd * kFixOne);
        Z[t] = (int)(ld * kFixOne);
        gf[t] = (float)gd;
        lf[t] = (float)ld;
    }
#pragma unroll
    for (int t = 0; t < N; t++) ab[t] = Q[t](i[t]);
    // the index steps stay in plain integer form here: the one-instruction steps (Q.next) take about 10
    // instructions per evaluation off this blend but raise the FP32 erosion kernel from 109 to 117-130 VGPRs,
    // which leaves no room beside its four waves per SIMD for the other stages' waves, and the pipelined step
    // was 2 % slower
#pragma unroll
    for (int t = 0; t < N; t++) {
        const int j = s[t]->jy;
        cd[t] = Q[t]((int)(ab[t] & 255u) + j);
        ef[t] = Q[t]((int)(ab[t] >> 8) + j);
    }
    uint32_t h0[N], h1[N], h2[N], h3[N];
#pragma unroll
    for (int t = 0; t < N; t++) {
        h0[t] = Q[t]((int)(cd[t] & 255u) + k[t]);   // P(c+k),  P(c+k+1)
        h1[t] = Q[t]((int)(ef[t] & 255u) + k[t]);   // P(e+k),  P(e+k+1)
        h2[t] = Q[t]((int)(cd[t] >> 8) + k[t]);     // P(d+k),  P(d+k+1)
        h3[t] = Q[t]((int)(ef[t] >> 8) + k[t]);     // P(f+k),  P(f+k+1)
    }
#pragma unroll
    for (int t = 0; t < N; t++) {
        const int Y = s[t]->hq, Ym = s[t]->hmq;
        const int Xm = X[t] - (1 << kFixBits), Zm = Z[t] - (1 << kFixBits);
        // low byte of a pair: hash bits at 0..3, field offset 2h; high byte: bits 8..11
        const float c0 = (float)grad_q(h0[t] << 1, X[t], Y, Z[t]);
        const float c1 = (float)grad_q(h1[t] << 1, Xm, Y, Z[t]);
        const float c2 = (float)grad_q(h2[t] << 1, X[t], Ym, Z[t]);
        const float c3 = (float)grad_q(h3[t] << 1, Xm, Ym, Z[t]);
        const float c4 = (float)grad_q((h0[t] >> 7) & 30u, X[t], Y, Zm);
        const float c5 = (float)grad_q((h1[t] >> 7) & 30u, Xm, Y, Zm);
        const float c6 = (float)grad_q((h2[t] >> 7) & 30u, X[t], Ym, Zm);
        const float c7 = (float)grad_q((h3[t] >> 7) & 30u, Xm, Ym, Zm);
        const float fx = fade_q(gf[t]), fz = fade_q(lf[t]), fy = s[t]->fy;
        const float lo = lerp_q(fy, lerp_q(fx, c0, c1), lerp_q(fx, c2, c3));
        const float hi = lerp_q(fy, lerp_q(fx, c4, c5), lerp_q(fx, c6, c7));
        out[t] = (double)lerp_q(fz, lo, hi);
    }
}

// OctavePerlinNoiseSampler.maintainPrecision (noise.c:28-34): identity inside +-1e7.
__device__ __forceinline__ double keep_precision(double v) {
    if (v > -1.0e7 && v < 1.0e7) return v;
    return v - floor(v * 2.9802322387695312E-08 + 0.5) * 3.3554432E7;
}

// ---------------------------------------------------------------------------------------------- Java math
// Math.min / Math.max / MathHelper.clamp with Java's NaN and signed-zero rules (mcmath.h).

__device__ __forceinline__ bool is_neg_zero(double v) {
    return __double_as_longlong(v) == (long long)0x8000000000000000ull;
}
__device__ __forceinline__ double jmax(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && is_neg_zero(a)) return b;
    return a >= b ? a : b;
}
__device__ __forceinline__ double jmin(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && is_neg_zero(b)) return b;
    return a <= b ? a : b;
}
__device__ __forceinline__ double jclamp(double v, double lo, double hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

}  // namespace whf
