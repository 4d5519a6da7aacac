/* whf_oracle.c — CPU oracle for the witch-hut phase-1 Seed Check (see whf_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: never linked into the HIP product path.
 *
 * Floating-point discipline (as the reference requires, build.sh:23 and caves.c:1-6): compiled with
 * -ffp-contract=off, every expression keeps the reference's association order, x/s stays a division.
 */
#include "whf_oracle.h"
#include "whf_oracle_int.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "whf_oracle_tables.h"

/* ======================================================================== seeds */

uint64_t wo_splitmix64_at(uint64_t stream, uint64_t index) {
    /* Vigna's SplitMix64: state advances by the golden gamma before each output. */
    uint64_t z = stream + (index + 1) * UINT64_C(0x9E3779B97F4A7C15);
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

uint64_t wo_scramble_index(uint64_t index) {
    /* hunt.py:76 SCRAMBLE and :121-124: seed = index * 6364136223846793005 mod 2^64 */
    return index * UINT64_C(6364136223846793005);
}

/* ======================================================================== MD5 (RFC 1321) */

static uint32_t md5_k[64];
static int md5_ready;
static pthread_once_t md5_once = PTHREAD_ONCE_INIT;

static void md5_init_table(void) {
    /* RFC 1321 section 3.4: T[i] = floor(2^32 * |sin(i)|), i = 1..64 */
    for (int i = 0; i < 64; i++) md5_k[i] = (uint32_t)floor(fabs(sin((double)(i + 1))) * 4294967296.0);
    md5_ready = 1;
}

static uint32_t rol32(uint32_t v, int s) { return (v << s) | (v >> (32 - s)); }

static void md5_compress(uint32_t st[4], const uint8_t blk[64]) {
    static const int shift[4][4] = { {7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21} };
    uint32_t w[16];
    for (int i = 0; i < 16; i++)
        w[i] = (uint32_t)blk[4 * i] | (uint32_t)blk[4 * i + 1] << 8 | (uint32_t)blk[4 * i + 2] << 16 |
               (uint32_t)blk[4 * i + 3] << 24;
    uint32_t a = st[0], b = st[1], c = st[2], d = st[3];
    for (int r = 0; r < 4; r++) {
        for (int q = 0; q < 16; q++) {
            int i = r * 16 + q, g;
            uint32_t f;
            switch (r) {
            case 0: f = (b & c) | (~b & d); g = q; break;
            case 1: f = (d & b) | (~d & c); g = (1 + 5 * q) & 15; break;
            case 2: f = b ^ c ^ d; g = (5 + 3 * q) & 15; break;
            default: f = c ^ (b | ~d); g = (7 * q) & 15; break;
            }
            uint32_t t = a + f + md5_k[i] + w[g];
            a = d; d = c; c = b;
            b = b + rol32(t, shift[r][q & 3]);
        }
    }
    st[0] += a; st[1] += b; st[2] += c; st[3] += d;
}

void wo_md5(const uint8_t *data, size_t len, uint8_t out[16]) {
    pthread_once(&md5_once, md5_init_table);
    uint32_t st[4] = { 0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u };
    size_t off = 0;
    for (; off + 64 <= len; off += 64) md5_compress(st, data + off);
    uint8_t last[128] = { 0 };
    size_t rem = len - off;
    memcpy(last, data + off, rem);
    last[rem] = 0x80;
    size_t total = rem + 9 <= 64 ? 64 : 128;
    uint64_t bits = (uint64_t)len << 3;
    for (int i = 0; i < 8; i++) last[total - 8 + i] = (uint8_t)(bits >> (8 * i));
    md5_compress(st, last);
    if (total == 128) md5_compress(st, last + 64);
    for (int i = 0; i < 16; i++) out[i] = (uint8_t)(st[i / 4] >> (8 * (i % 4)));
}

/* ======================================================================== Xoroshiro128++ (rng.c:75-181) */

static uint64_t stafford13(uint64_t z) {
    z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
    z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
    return z ^ (z >> 31);
}

void wo_xr_set(wo_xr *s, uint64_t lo, uint64_t hi) {
    /* two-long constructor; an all-zero state is replaced by the golden-ratio pair (rng.c:104-111) */
    if ((lo | hi) == 0) { lo = UINT64_C(0x9E3779B97F4A7C15); hi = UINT64_C(0x6A09E667F3BCC909); }
    s->lo = lo;
    s->hi = hi;
}

void wo_xr_seed(wo_xr *s, uint64_t seed) {
    /* RandomSupport.upgradeSeedTo128bit (rng.c:97-102) */
    uint64_t l = seed ^ UINT64_C(0x6A09E667F3BCC909);
    s->lo = stafford13(l);
    s->hi = stafford13(l - UINT64_C(0x61C8864680B583EB));
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t wo_xr_next(wo_xr *s) {
    uint64_t lo = s->lo, hi = s->hi;
    uint64_t out = rotl64(lo + hi, 17) + lo;
    hi ^= lo;
    s->lo = rotl64(lo, 49) ^ hi ^ (hi << 21);
    s->hi = rotl64(hi, 28);
    return out;
}

int wo_xr_next_int(wo_xr *s, int bound) {
    /* Java 17 bounded nextInt via Lemire's multiply-shift with rejection (rng.c:124-146) */
    uint64_t m = (wo_xr_next(s) & 0xFFFFFFFFu) * (uint64_t)(uint32_t)bound;
    if ((uint32_t)m < (uint32_t)bound) {
        uint32_t lim = (uint32_t)(0u - (uint32_t)bound) % (uint32_t)bound;
        while ((uint32_t)m < lim) m = (wo_xr_next(s) & 0xFFFFFFFFu) * (uint64_t)(uint32_t)bound;
    }
    return (int)(m >> 32);
}

double wo_xr_next_double(wo_xr *s) {
    return (double)(wo_xr_next(s) >> 11) * 0x1.0p-53;
}

static uint64_t be64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v = v << 8 | p[i];
    return v;
}

/* createRandom(name) of a deriver: MD5(name) split into two big-endian longs XORed into the state (rng.c:164-170) */
void wo_xr_fork_named(wo_xr *out, const wo_xr *deriver, const char *name) {
    uint8_t h[16];
    wo_md5((const uint8_t *)name, strlen(name), h);
    wo_xr_set(out, be64(h) ^ deriver->lo, be64(h + 8) ^ deriver->hi);
}

/* createRandomDeriver(): two longs of the parent become the new state (rng.c:157-161) */
void wo_xr_fork(wo_xr *out, wo_xr *parent) {
    uint64_t a = wo_xr_next(parent);
    uint64_t b = wo_xr_next(parent);
    wo_xr_set(out, a, b);
}

/* ======================================================================== java.util.Random (rng.c:187-227) */

#define LCG_MUL UINT64_C(0x5DEECE66D)
#define LCG_M48 UINT64_C(0xFFFFFFFFFFFF)

static inline int lcg_next(uint64_t *st, int bits) {
    *st = (*st * LCG_MUL + 0xB) & LCG_M48;
    return (int)(int64_t)(*st >> (48 - bits));
}

int wo_lcg_next_int_seeded(uint64_t *st, int bound) {
    if ((bound & -bound) == bound) return (int)(((int64_t)bound * (int64_t)lcg_next(st, 31)) >> 31);
    for (;;) {
        int bits = lcg_next(st, 31);
        int val = bits % bound;
        /* Java rejects when bits - val + (bound - 1) overflows int */
        if ((int32_t)((uint32_t)bits - (uint32_t)val + (uint32_t)(bound - 1)) >= 0) return val;
    }
}

uint64_t wo_lcg_next_long(uint64_t *st) {
    /* ((long)next(32) << 32) + next(32), the second word sign-extended (rng.c:197-211) */
    int64_t hi = lcg_next(st, 32);
    int64_t lo = lcg_next(st, 32);
    return ((uint64_t)hi << 32) + (uint64_t)lo;
}

void wo_hut_chunk(uint64_t world_seed, int rx, int rz, int salt, int *cx, int *cz) {
    /* region seed: rx*341873128712 + rz*132897987541 + seed + salt, then setSeed scrambles with the
     * multiplier; chunk = region*32 + nextInt(24) on each axis (structure.c:4-24) */
    uint64_t s = (uint64_t)((int64_t)rx * INT64_C(341873128712)) + (uint64_t)((int64_t)rz * INT64_C(132897987541)) +
                 world_seed + (uint64_t)(int64_t)salt;
    uint64_t st = (s ^ LCG_MUL) & LCG_M48;
    int j = wo_lcg_next_int_seeded(&st, 24);
    int k = wo_lcg_next_int_seeded(&st, 24);
    *cx = rx * 32 + j;
    *cz = rz * 32 + k;
}

/* ======================================================================== Perlin (noise.c:14-78, 223-342) */

static const int GRAD[16][3] = {
    {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
    {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}, {1, 1, 0}, {0, -1, 1}, {-1, 1, 0}, {0, -1, -1},
};

void wo_perlin_build(wo_perlin *p, wo_xr *r) {
    p->ox = wo_xr_next_double(r) * 256.0;
    p->oy = wo_xr_next_double(r) * 256.0;
    p->oz = wo_xr_next_double(r) * 256.0;
    for (int i = 0; i < 256; i++) p->perm[i] = (uint8_t)i;
    for (int i = 0; i < 256; i++) {
        int j = i + wo_xr_next_int(r, 256 - i);
        uint8_t t = p->perm[i];
        p->perm[i] = p->perm[j];
        p->perm[j] = t;
    }
}

static inline double fade5(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
static inline double lerp1(double t, double a, double b) { return a + t * (b - a); }
static inline double dotg(int h, double x, double y, double z) {
    const int *g = GRAD[h & 15];
    return (double)g[0] * x + (double)g[1] * y + (double)g[2] * z;
}

/* Improved-noise corner blend with the y fade supplied separately (noise.c:224-244). */
double wo_perlin_core(const wo_perlin *p, int i, int j, int k, double x, double y, double z, double ty) {
    const uint8_t *P = p->perm;
    int a = P[i & 255], b = P[(i + 1) & 255];
    int aa = P[(a + j) & 255], ab = P[(a + j + 1) & 255];
    int ba = P[(b + j) & 255], bb = P[(b + j + 1) & 255];
    double c000 = dotg(P[(aa + k) & 255], x, y, z);
    double c100 = dotg(P[(ba + k) & 255], x - 1, y, z);
    double c010 = dotg(P[(ab + k) & 255], x, y - 1, z);
    double c110 = dotg(P[(bb + k) & 255], x - 1, y - 1, z);
    double c001 = dotg(P[(aa + k + 1) & 255], x, y, z - 1);
    double c101 = dotg(P[(ba + k + 1) & 255], x - 1, y, z - 1);
    double c011 = dotg(P[(ab + k + 1) & 255], x, y - 1, z - 1);
    double c111 = dotg(P[(bb + k + 1) & 255], x - 1, y - 1, z - 1);
    double tx = fade5(x), fy = fade5(ty), tz = fade5(z);
    double lo = lerp1(fy, lerp1(tx, c000, c100), lerp1(tx, c010, c110));
    double hi = lerp1(fy, lerp1(tx, c001, c101), lerp1(tx, c011, c111));
    return lerp1(tz, lo, hi);
}

/* PerlinNoiseSampler.sample with yScale = yMax = 0 (noise.c:246-261) */
static double perlin_sample3(const wo_perlin *p, double x, double y, double z) {
    double d = x + p->ox, e = y + p->oy, f = z + p->oz;
    int i = (int)floor(d), j = (int)floor(e), k = (int)floor(f);
    double h = e - j;
    return wo_perlin_core(p, i, j, k, d - i, h, f - k, h);
}

/* OctavePerlinNoiseSampler.maintainPrecision (noise.c:28-34) */
double wo_keep_precision(double v) {
    return v - (double)(long)floor(v / 33554432.0 + 0.5) * 33554432.0;
}

/* ======================================================================== octave / double noise */

static void octave_build(wo_octave *o, wo_xr *noise_rng, const wo_noise_param *np) {
    /* non-legacy octave construction: a deriver forked from the noise's random, then each non-zero slot i
     * gets its own random createRandom("octave_" + (firstOctave + i)) (noise.c:80-109) */
    memset(o, 0, sizeof *o);
    o->n = np->n;
    o->first_octave = np->first_octave;
    o->lac0 = ldexp(1.0, np->first_octave);
    o->per0 = ldexp(1.0, np->n - 1) / (ldexp(1.0, np->n) - 1.0);
    wo_xr der;
    wo_xr_fork(&der, noise_rng);
    for (int i = 0; i < np->n; i++) {
        o->amp[i] = np->amps[i];
        if (np->amps[i] == 0.0) continue;
        char name[24];
        snprintf(name, sizeof name, "octave_%d", np->first_octave + i);
        wo_xr r;
        wo_xr_fork_named(&r, &der, name);
        wo_perlin_build(&o->oct[i], &r);
        o->present[i] = 1;
    }
}

void wo_noise_init(wo_noise *n, uint64_t world_seed, int noise_id) {
    /* worldSeed -> Xoroshiro -> deriver -> createRandom(identifier) -> two octave samplers sharing that random
     * (noise.c:111-143); the output amplitude is (1/6) / (0.1 * (1 + 1/(span+1))) with span the distance between
     * the first and last non-zero amplitude (noise.c:116-130) */
    const wo_noise_param *np = &WO_NOISE[noise_id];
    wo_xr r, der, nr;
    wo_xr_seed(&r, world_seed);
    wo_xr_fork(&der, &r);
    wo_xr_fork_named(&nr, &der, np->name);
    octave_build(&n->sub[0], &nr, np);
    octave_build(&n->sub[1], &nr, np);
    int first = -1, last = -1;
    for (int i = 0; i < np->n; i++)
        if (np->amps[i] != 0.0) { if (first < 0) first = i; last = i; }
    double span = (double)(last - first);
    n->amplitude = 0.16666666666666666 / (0.1 * (1.0 + 1.0 / (span + 1.0)));
}

#define WO_SECOND_SCALE 1.0181268882175227

static double octave_sample3(const wo_octave *o, double x, double y, double z) {
    /* noise.c:263-281: value += (amp * v) * persistence, lacunarity doubles and persistence halves per slot */
    double v = 0.0, lac = o->lac0, per = o->per0;
    for (int i = 0; i < o->n; i++) {
        if (o->present[i]) {
            double xs = x * lac, ys = y * lac, zs = z * lac;
            if (!(xs > -1.0e7 && xs < 1.0e7)) xs = wo_keep_precision(xs);
            if (!(ys > -1.0e7 && ys < 1.0e7)) ys = wo_keep_precision(ys);
            if (!(zs > -1.0e7 && zs < 1.0e7)) zs = wo_keep_precision(zs);
            v += (o->amp[i] * perlin_sample3(&o->oct[i], xs, ys, zs)) * per;
        }
        lac *= 2.0;
        per /= 2.0;
    }
    return v;
}

double wo_noise_sample(const wo_noise *n, double x, double y, double z) {
    /* noise.c:285-289 */
    return (octave_sample3(&n->sub[0], x, y, z) +
            octave_sample3(&n->sub[1], x * WO_SECOND_SCALE, y * WO_SECOND_SCALE, z * WO_SECOND_SCALE)) *
           n->amplitude;
}

double wo_noise_sample_y0(const wo_noise *n, double x, double z) {
    /* y = 0 is an exact special case of the 3D sample: 0 + oy = oy and the y fade uses the same fraction
     * (noise.c:291-342 restates this path separately; the oracle reuses the 3D path with y = 0). */
    return wo_noise_sample(n, x, 0.0, z);
}

/* Tiered erosion (noise.c:359-459): evaluations interleaved A0,B0,A1,B1,... over the non-empty slots, each
 * weighted by amp * persistence of its slot; after evaluation k (k = 1..7) the running sum times the output
 * amplitude is compared with threshold[k-1]; below it the region is rejected at that tier. */
static const double WO_TIER_T[7] = { 0.05, 0.22, 0.42, 0.45, 0.49, 0.51, 0.55 };

int wo_erosion_tiered(const wo_noise *n, double x, double z, double *value, int *tier) {
    const wo_octave *A = &n->sub[0], *B = &n->sub[1];
    double sum = 0.0;
    int k = 0, t = 0;
    double lac = A->lac0, per = A->per0;
    *tier = -1;
    for (int i = 0; i < A->n; i++, lac *= 2.0, per /= 2.0) {
        if (!A->present[i]) continue;
        for (int s = 0; s < 2; s++) {
            const wo_octave *o = s ? B : A;
            double sc = s ? WO_SECOND_SCALE : 1.0;
            double xs = x * sc * lac, zs = z * sc * lac;
            if (!(xs > -1.0e7 && xs < 1.0e7)) xs = wo_keep_precision(xs);
            if (!(zs > -1.0e7 && zs < 1.0e7)) zs = wo_keep_precision(zs);
            double w = o->amp[i] * per;
            sum += w * perlin_sample3(&o->oct[i], xs, 0.0, zs);
            if (k >= 1 && t < 7) {
                if (sum * n->amplitude < WO_TIER_T[t]) {
                    *tier = t;
                    *value = sum * n->amplitude;
                    return 1;
                }
                if (++t >= 7) { *value = sum * n->amplitude; return 0; }
            }
            k++;
        }
    }
    *value = sum * n->amplitude;
    return 0;
}

/* ======================================================================== Java math (mcmath.h) */

static inline int is_neg_zero(double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    return b == UINT64_C(0x8000000000000000);
}
static double jmax(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && is_neg_zero(a)) return b;
    return a >= b ? a : b;
}
static double jmin(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && is_neg_zero(b)) return b;
    return a <= b ? a : b;
}
static double jclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ======================================================================== caves (caves.c:58-130) */

static double tunnel_scale(double v) {
    if (v < -0.5) return 0.75;
    if (v < 0.0) return 1.0;
    if (v < 0.5) return 1.5;
    return 2.0;
}

static double spaghetti(const wo_phase1 *p, int x, int y, int z) {
    double X = (double)x, Y = (double)y, Z = (double)z;
    double e = tunnel_scale(wo_noise_sample(&p->spag_rarity, (double)(x * 2), Y, (double)(z * 2)));
    double tv = wo_noise_sample(&p->spag_thickness, X, Y, Z);
    double th = 0.065 + ((tv - -1.0) / (1.0 - -1.0)) * (0.088 - 0.065);
    double l = wo_noise_sample(&p->spag_1, X / e, Y / e, Z / e);
    double m = fabs(e * l) - th;
    double n = wo_noise_sample(&p->spag_2, X / e, Y / e, Z / e);
    double o = fabs(e * n) - th;
    double pp = jclamp(jmax(m, o), -1.0, 1.0);
    double mod = wo_noise_sample(&p->spag_rough_mod, X, Y, Z);
    double rg = wo_noise_sample(&p->spag_rough, X, Y, Z);
    double q = (-0.05 + (-0.05 * mod)) * (-0.4 + fabs(rg));
    return pp + q;
}

double wo_entrance(const wo_phase1 *p, int x, int y, int z) {
    double ent = wo_noise_sample(&p->cave_entrance, (double)x * 0.75, (double)y * 0.5, (double)z * 0.75);
    double t = (10.0 + (double)y) / 40.0;
    double ramp = t < 0.0 ? 0.3 : (t > 1.0 ? 0.0 : 0.3 + t * (0.0 - 0.3));
    double c = ent + 0.37 + ramp;
    return jmin(c, spaghetti(p, x, y, z));
}

double wo_entrance2(const wo_phase1 *p, int x, int y, int z) { return spaghetti(p, x, y, z); }

double wo_cheese(const wo_phase1 *p, int x, int y, int z) {
    double layer = wo_noise_sample(&p->cave_layer, (double)x, (double)(int32_t)((uint32_t)y * 8u), (double)z);
    double a = 4.0 * layer * layer;
    double b = jclamp(0.27 + wo_noise_sample(&p->cave_cheese, (double)x, (double)y * 0.6666666666666666, (double)z),
                      -1.0, 1.0);
    return a + b;
}

/* ======================================================================== phase 1 (phase1.c) */

void wo_phase1_init(wo_phase1 *p, uint64_t s, int vf) {
    int large = (vf & WO_V_LARGE) != 0;
    wo_noise_init(&p->erosion, s, large ? WO_N_EROSION_LARGE : WO_N_EROSION);
    wo_noise_init(&p->temperature, s, large ? WO_N_TEMPERATURE_LARGE : WO_N_TEMPERATURE);
    wo_noise_init(&p->continentalness, s, large ? WO_N_CONTINENTALNESS_LARGE : WO_N_CONTINENTALNESS);
    wo_noise_init(&p->ridge, s, WO_N_RIDGE);
    wo_noise_init(&p->floodedness, s, WO_N_FLOODEDNESS);
    wo_noise_init(&p->cave_entrance, s, WO_N_CAVE_ENTRANCE);
    wo_noise_init(&p->spag_rarity, s, WO_N_SPAG3D_RARITY);
    wo_noise_init(&p->spag_thickness, s, WO_N_SPAG3D_THICKNESS);
    wo_noise_init(&p->spag_1, s, WO_N_SPAG3D_1);
    wo_noise_init(&p->spag_2, s, WO_N_SPAG3D_2);
    wo_noise_init(&p->spag_rough_mod, s, WO_N_SPAG_ROUGH_MOD);
    wo_noise_init(&p->spag_rough, s, WO_N_SPAG_ROUGH);
    wo_noise_init(&p->cave_layer, s, WO_N_CAVE_LAYER);
    wo_noise_init(&p->cave_cheese, s, WO_N_CAVE_CHEESE);
}

static double climate_at(const wo_noise *n, int hut_x, int hut_z) {
    /* sampled at the quarter-resolution point of the hut's chunk-corner + 8 (phase1.c:21-29) */
    return wo_noise_sample_y0(n, (hut_x + 8) / 4.0, (hut_z + 8) / 4.0);
}

static double flooded_at(const wo_phase1 *p, int hut_x, int hut_z, int y) {
    return wo_noise_sample(&p->floodedness, (double)(hut_x + 3), (double)y * 0.67, (double)(hut_z + 3));
}

int wo_phase1_check(const wo_phase1 *p, int vf, int hut_x, int hut_z, int max_h, int *tier) {
    *tier = -1;
    if (!(vf & WO_V_SINGLE)) {
        double e;
        if (wo_erosion_tiered(&p->erosion, (hut_x + 8) / 4.0, (hut_z + 8) / 4.0, &e, tier)) return WO_REJ_EROSION;
        if (e < 0.55) return WO_REJ_EROSION;
        double t = climate_at(&p->temperature, hut_x, hut_z);
        if ((vf & WO_V_118) ? (t < -0.45) : (t > 0.2 || t < -0.45)) return WO_REJ_TEMPERATURE;
        double r = climate_at(&p->ridge, hut_x, hut_z);
        if ((r > 0.42 && r < 0.91) || (r < -0.42 && r > -0.91)) return WO_REJ_RIDGE;
        if ((vf & WO_V_262) && r <= -0.91) return WO_REJ_RIDGE;
    }
    int x = hut_x + 3, z = hut_z + 3;
    if (wo_entrance(p, x, 50, z) >= 0) return WO_REJ_ENTRANCE50;
    if (wo_entrance(p, x, 60, z) >= 0) return WO_REJ_ENTRANCE60;
    if (wo_entrance2(p, x, max_h, z) >= 0 && wo_cheese(p, x, max_h, z) >= 0) return WO_REJ_CAVE_MAXH;
    int floor_y = max_h < -40 ? max_h : -40;
    if (floor_y < -50) floor_y = -50;
    for (int y = 0; y >= floor_y; y -= 10)
        if (max_h < y && wo_entrance2(p, x, y, z) >= 0 && wo_cheese(p, x, y, z) >= 0) return WO_REJ_LADDER_NEG;
    for (int y = 10; y <= 40; y += 10)
        if (wo_entrance(p, x, y, z) >= 0 && wo_cheese(p, x, y, z) >= 0) return WO_REJ_LADDER_POS;
    if (!(vf & WO_V_SINGLE) && climate_at(&p->continentalness, hut_x, hut_z) < -0.11) return WO_REJ_CONTINENTALNESS;
    for (int y = max_h; y <= 60; y += 10)
        if (flooded_at(p, hut_x, hut_z, y) > 0.41) return WO_REJ_FLOODEDNESS;
    return WO_ACCEPT;
}

void wo_phase1_probe(const wo_phase1 *p, int hut_x, int hut_z, int max_h, double out[38]) {
    int x = hut_x + 3, z = hut_z + 3, k = 0;
    out[k++] = climate_at(&p->erosion, hut_x, hut_z);
    out[k++] = climate_at(&p->temperature, hut_x, hut_z);
    out[k++] = climate_at(&p->ridge, hut_x, hut_z);
    out[k++] = climate_at(&p->continentalness, hut_x, hut_z);
    out[k++] = wo_entrance(p, x, 50, z);
    out[k++] = wo_entrance(p, x, 60, z);
    out[k++] = wo_entrance2(p, x, max_h, z);
    out[k++] = wo_cheese(p, x, max_h, z);
    for (int y = 0; y >= -40; y -= 10) { out[k++] = wo_entrance2(p, x, y, z); out[k++] = wo_cheese(p, x, y, z); }
    for (int y = 10; y <= 40; y += 10) { out[k++] = wo_entrance(p, x, y, z); out[k++] = wo_cheese(p, x, y, z); }
    for (int y = -50; y <= 60; y += 10) out[k++] = flooded_at(p, hut_x, hut_z, y);
}

/* ======================================================================== versions */

/* Game version -> (1.18 rules, 26.2 rules), as lysh's version table (biome_trees.h:70499-70526) and its
 * aliases (main.c:104-125). The biome tree index is not needed by phase 1. */
static const struct { const char *name; int is118, is262; } WO_VERSIONS[] = {
    {"1.18", 1, 0}, {"1.18.1", 1, 0}, {"1.18.2", 1, 0}, {"1.18.x", 1, 0},
    {"1.19", 0, 0}, {"1.19.2", 0, 0}, {"1.19.3", 0, 0}, {"1.19.4", 0, 0}, {"1.19.x", 0, 0},
    {"1.20", 0, 0}, {"1.20.1", 0, 0}, {"1.20.2", 0, 0}, {"1.20.4", 0, 0}, {"1.20.5", 0, 0}, {"1.20.6", 0, 0},
    {"1.20.x", 0, 0}, {"1.21", 0, 0}, {"1.21.1", 0, 0}, {"1.21.2", 0, 0}, {"1.21.3", 0, 0}, {"1.21.4", 0, 0},
    {"1.21.5", 0, 0}, {"1.21.8", 0, 0}, {"1.21.11", 0, 0}, {"1.21.x~26.1", 0, 0}, {"1.21.5~26.1", 0, 0},
    {"1.19.x~26.1", 0, 0}, {"26.1", 0, 0}, {"26.2", 0, 1}, {"26.3", 0, 1}, {"26.2~26.3", 0, 1},
};

int wo_version_flags(const char *name, int *vf) {
    for (size_t i = 0; i < sizeof WO_VERSIONS / sizeof WO_VERSIONS[0]; i++)
        if (!strcmp(WO_VERSIONS[i].name, name)) {
            *vf = (WO_VERSIONS[i].is118 ? WO_V_118 : 0) | (WO_VERSIONS[i].is262 ? WO_V_262 : 0);
            return 0;
        }
    return 1;
}

/* ======================================================================== scan */

typedef struct {
    const int64_t *seeds;
    int64_t n_seeds;
    int rx0, rx1, rz0, rz1, max_h, vf;
    pthread_mutex_t *mu;
    int64_t *next;
    wo_funnel f;
    wo_hit *hits;
    int64_t cap, *n_hits;
} scan_job;

static void *scan_worker(void *arg) {
    scan_job *j = (scan_job *)arg;
    wo_phase1 *p = malloc(sizeof *p);
    for (;;) {
        pthread_mutex_lock(j->mu);
        int64_t i = (*j->next)++;
        pthread_mutex_unlock(j->mu);
        if (i >= j->n_seeds) break;
        uint64_t seed = (uint64_t)j->seeds[i];
        wo_phase1_init(p, seed, j->vf);
        for (int rz = j->rz0; rz < j->rz1; rz++)
            for (int rx = j->rx0; rx < j->rx1; rx++) {
                int cx, cz, tier;
                wo_hut_chunk(seed, rx, rz, 14357620, &cx, &cz);
                int kind = wo_phase1_check(p, j->vf, cx * 16, cz * 16, j->max_h, &tier);
                j->f.checks++;
                j->f.funnel[kind]++;
                if (tier >= 0) j->f.tiers[tier]++;
                if (kind > WO_REJ_RIDGE) j->f.reached_ladder++;
                if (kind != WO_ACCEPT) continue;
                j->f.accepted++;
                pthread_mutex_lock(j->mu);
                int64_t h = (*j->n_hits)++;
                if (h < j->cap) {
                    j->hits[h].seed = (int64_t)seed;
                    j->hits[h].rx = rx; j->hits[h].rz = rz;
                    j->hits[h].cx = cx; j->hits[h].cz = cz;
                }
                pthread_mutex_unlock(j->mu);
            }
    }
    free(p);
    return NULL;
}

int wo_scan(const int64_t *seeds, int64_t n_seeds, int rx0, int rx1, int rz0, int rz1, int max_h, int vf,
            int threads, wo_funnel *out, wo_hit *hits, int64_t cap, int64_t *n_hits) {
    if (rx1 <= rx0 || rz1 <= rz0 || n_seeds < 0) return 1;
    if (threads < 1) threads = 1;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    int64_t next = 0;
    *n_hits = 0;
    scan_job *jobs = calloc((size_t)threads, sizeof *jobs);
    pthread_t *th = calloc((size_t)threads, sizeof *th);
    for (int t = 0; t < threads; t++) {
        jobs[t] = (scan_job){ seeds, n_seeds, rx0, rx1, rz0, rz1, max_h, vf, &mu, &next, {0}, hits, cap, n_hits };
        pthread_create(&th[t], NULL, scan_worker, &jobs[t]);
    }
    memset(out, 0, sizeof *out);
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        out->checks += jobs[t].f.checks;
        out->accepted += jobs[t].f.accepted;
        out->reached_ladder += jobs[t].f.reached_ladder;
        for (int k = 0; k < WO_KIND_N; k++) out->funnel[k] += jobs[t].f.funnel[k];
        for (int k = 0; k < 8; k++) out->tiers[k] += jobs[t].f.tiers[k];
    }
    free(jobs);
    free(th);
    return 0;
}

int wo_scan_regions(int64_t seed, int rx0, int rx1, int rz0, int rz1, int max_h, int vf, int8_t *kind,
                    int8_t *tier) {
    if (rx1 <= rx0 || rz1 <= rz0) return 1;
    wo_phase1 *p = malloc(sizeof *p);
    wo_phase1_init(p, (uint64_t)seed, vf);
    int64_t idx = 0;
    for (int rz = rz0; rz < rz1; rz++)
        for (int rx = rx0; rx < rx1; rx++, idx++) {
            int cx, cz, t;
            wo_hut_chunk((uint64_t)seed, rx, rz, 14357620, &cx, &cz);
            kind[idx] = (int8_t)wo_phase1_check(p, vf, cx * 16, cz * 16, max_h, &t);
            tier[idx] = (int8_t)t;
        }
    free(p);
    return 0;
}

/* ======================================================================== fingerprints */

uint64_t wo_perlin_fingerprint(const wo_perlin *p) {
    uint64_t h = UINT64_C(0xCBF29CE484222325);
    uint8_t buf[256 + 24];
    memcpy(buf, p->perm, 256);
    memcpy(buf + 256, &p->ox, 8);
    memcpy(buf + 264, &p->oy, 8);
    memcpy(buf + 272, &p->oz, 8);
    for (size_t i = 0; i < sizeof buf; i++) { h ^= buf[i]; h *= UINT64_C(0x100000001B3); }
    return h;
}

uint64_t wo_sampler_fingerprint(uint64_t world_seed, int noise_id, int sub, int slot) {
    if (noise_id < 0 || noise_id >= WO_N_COUNT || sub < 0 || sub > 1 || slot < 0 || slot >= 16) return 0;
    wo_noise *n = malloc(sizeof *n);
    wo_noise_init(n, world_seed, noise_id);
    uint64_t h = n->sub[sub].present[slot] ? wo_perlin_fingerprint(&n->sub[sub].oct[slot]) : 0;
    free(n);
    return h;
}

/* ======================================================================== binding helpers */

void *wo_probe_alloc(uint64_t world_seed, int vflags) {
    wo_phase1 *p = malloc(sizeof *p);
    if (p) wo_phase1_init(p, world_seed, vflags);
    return p;
}

void *wo_noise_alloc(uint64_t world_seed, int noise_id) {
    if (noise_id < 0 || noise_id >= WO_N_COUNT) return NULL;
    wo_noise *n = malloc(sizeof *n);
    if (n) wo_noise_init(n, world_seed, noise_id);
    return n;
}

void wo_probe_free(void *p) { free(p); }

double wo_noise_amplitude(const void *noise) { return ((const wo_noise *)noise)->amplitude; }

void wo_noise_perm(const void *noise, int sub, int slot, uint8_t out[256]) {
    const wo_noise *n = (const wo_noise *)noise;
    memcpy(out, n->sub[sub & 1].oct[slot & 15].perm, 256);
}
