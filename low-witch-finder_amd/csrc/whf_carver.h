// whf_carver.h — device functions of the carvers: java.util.Random, the game's sine table, and the cave and canyon
// carvers of one origin chunk (engine/tools/lysh/src/carver.c), reduced to what a target chunk's surface needs:
// the set of its cells some ellipsoid claims.
//
// Why a set suffices: an ellipsoid claims each unclaimed cell it covers (and does not skip) before it tries to
// carve it, and claimed cells are never carved again (carver.c:233-241). Whether the first claimant carves depends
// only on the cell's own block state (solid or water can be replaced) and on the aquifer's substance at density 0
// there (carver.c:145-175), neither of which any carver changes. So a cell ends up carved exactly when some
// ellipsoid covers it and its state allows it, whatever the order of the 17 x 17 x 3 carvers: the claims of all
// of them can be gathered in parallel, and the states are read afterwards.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace whf {

// java.util.Random (rng.c:186-231).
struct JRand {
    uint64_t s;
    __host__ __device__ void seed(uint64_t v) { s = (v ^ 0x5DEECE66Dull) & 0xFFFFFFFFFFFFull; }
    __host__ __device__ int next(int bits) {
        s = (s * 0x5DEECE66Dull + 0xBull) & 0xFFFFFFFFFFFFull;
        return (int)(s >> (48 - bits));
    }
    __host__ __device__ int next_int(int bound) {
        if (bound <= 0) return 0;
        if ((bound & (bound - 1)) == 0) return (int)(((int64_t)bound * (int64_t)next(31)) >> 31);
        int u = next(31), v;
        for (;;) {
            v = u % bound;
            if ((int32_t)((uint32_t)u - (uint32_t)v + (uint32_t)(bound - 1)) >= 0) break;
            u = next(31);
        }
        return v;
    }
    __host__ __device__ uint64_t next_long() {
        const uint64_t hi = (uint64_t)(uint32_t)next(32);
        const int64_t lo = (int64_t)next(32);   // the low word is added sign-extended
        return (hi << 32) + (uint64_t)lo;
    }
    __host__ __device__ float next_float() { return (float)next(24) * 5.9604645e-8f; }
    __host__ __device__ float between(float lo, float hi) { return next_float() * (hi - lo) + lo; }
    // setLargeFeatureSeed (structure.c:26-42)
    __host__ __device__ void large_feature(uint64_t world, int cx, int cz) {
        seed(world);
        const uint64_t a = next_long();
        const uint64_t b = next_long();
        seed((uint64_t)((int64_t)cx * (int64_t)a) ^ (uint64_t)((int64_t)cz * (int64_t)b) ^ world);
    }
};

// A cave tunnel to walk: its random's seed, start, the system's radius multipliers and floor, its thickness and
// direction, and the step range [first, steps). Float fields hold values the reference draws as floats.
struct TunnelJob {
    uint64_t seed;
    double x, y, z;
    float hrm, vrm, thick, yaw, pitch, floor_level;
    int first, steps;
};

// An ellipsoid to claim: centre and radii; a cave's skips the cells at or below its floor level, a canyon's uses the
// width factors of its canyon (canyon >= 0: a shared width table; kOwnWidths: the table of the lane that made it).
struct Ell {
    double x, y, z, hr, vr;
    float floor_level;
    int canyon;
};
constexpr int kCaveEll = -1, kOwnWidths = -2;

// The claims of the carvers on one target chunk: a bit per cell (16 x 16 columns x 384 y, 6 words per column).
// The walks hand their work to callbacks: `spawn` takes a tunnel, `ell` an ellipsoid within reach of the chunk
// (with the canyon's width table), `widths` gives a canyon its width table (and its id) — so a caller can queue the
// pieces and spread them over lanes, or run each at once.
struct CarveClaims {
    const float *sin_tab;            // Mth.SIN, 65536 entries
    unsigned long long *bits;        // [256 columns][6]
    int min_x, min_z;                // the target chunk's first block
    double mid_x, mid_z;             // its middle (ChunkPos.getMiddleBlockX/Z)

    __device__ float msin(double a) const { return sin_tab[(int)((int64_t)(a * 10430.378350470453) & 65535)]; }
    __device__ float mcos(double a) const {
        return sin_tab[(int)((int64_t)(a * 10430.378350470453 + 16384.0) & 65535)];
    }

    // carveEllipsoid's early exit (carver.c:202-210): too far from the chunk to touch it
    __device__ bool in_reach(const Ell &e) const {
        const double reach = 16.0 + e.hr * 2.0;
        return !(fabs(e.x - mid_x) > reach || fabs(e.z - mid_z) > reach);
    }

    // carveEllipsoid's cells (carver.c:211-245), for the columns c of its box with c % stride == first; canyon: width
    // factors by y - minY, else the cave floor level
    __device__ void claim(const Ell &e, const float *width, int first, int stride) const {
        const double x = e.x, y = e.y, z = e.z, hr = e.hr, vr = e.vr, floor_level = (double)e.floor_level;
        const bool canyon = e.canyon != kCaveEll;
        const int i0 = max((int)floor(x - hr) - min_x - 1, 0);
        const int i1 = min((int)floor(x + hr) - min_x, 15);
        const int k0 = max((int)floor(y - vr) - 1, -64 + 1);
        const int k1 = min((int)floor(y + vr) + 1, -64 + 384 - 1 - 7);
        const int n0 = max((int)floor(z - hr) - min_z - 1, 0);
        const int n1 = min((int)floor(z + hr) - min_z, 15);
        const int nxc = i1 - i0 + 1, nzc = n1 - n0 + 1;
        if (nxc <= 0 || nzc <= 0) return;
        for (int c = first; c < nxc * nzc; c += stride) {
            const int p = i0 + c % nxc, s = n0 + c / nxc;
            const double nx = ((double)(min_x + p) + 0.5 - x) / hr;
            const double nz = ((double)(min_z + s) + 0.5 - z) / hr;
            if (nx * nx + nz * nz >= 1.0) continue;
            unsigned long long *col = bits + (size_t)(s * 16 + p) * 6;
            for (int v = k1; v > k0; v--) {
                const double ny = ((double)v - 0.5 - y) / vr;
                bool skip;
                if (!canyon) skip = ny <= floor_level || (nx * nx + ny * ny + nz * nz) >= 1.0;
                else skip = (nx * nx + nz * nz) * (double)width[v - -64 - 1] + ny * ny / 6.0 >= 1.0;
                if (skip) continue;
                const int b = v + 64;
                atomicOr(&col[b >> 6], 1ull << (b & 63));
            }
        }
    }

    // CaveWorldCarver.createTunnel (carver.c:268-329). A branch's two tunnels are thinner than 1 and do not branch
    // again, so the recursion is one level: kBranch tells which level this is. At its branch point a tunnel hands
    // the two branch tunnels to `spawn` (each draws from its own random, so the parent's draws do not depend on
    // them, and neither do the claims) and ends.
    template <bool kBranch, class Spawn, class EllSink>
    __device__ void tunnel(const TunnelJob &j, Spawn &&spawn, EllSink &&ell) const {
        JRand r;
        r.seed(j.seed);
        const int steps = j.steps;
        const float thickness = j.thick;
        const double hrm = (double)j.hrm, vrm = (double)j.vrm;
        double x = j.x, y = j.y, z = j.z;
        float yaw = j.yaw, pitch = j.pitch;
        const int branch = r.next_int(steps / 2) + steps / 4;
        const bool wide = r.next_int(6) == 0;
        float dyaw = 0.0f, dpitch = 0.0f;
        for (int i = j.first; i < steps; i++) {
            const float ang = (float)i * 3.1415927f / (float)steps;
            const double radius = 1.5 + (double)(msin((double)ang) * thickness);
            const double vert = radius * 1.0;
            const float cp = mcos((double)pitch);
            x += (double)(mcos((double)yaw) * cp);
            y += (double)msin((double)pitch);
            z += (double)(msin((double)yaw) * cp);
            pitch *= wide ? 0.92f : 0.7f;
            pitch += dpitch * 0.1f;
            yaw += dyaw * 0.1f;
            dpitch *= 0.9f;
            dyaw *= 0.75f;
            const float a1 = r.next_float(), b1 = r.next_float(), c1 = r.next_float();
            dpitch += (a1 - b1) * c1 * 2.0f;
            const float a2 = r.next_float(), b2 = r.next_float(), c2 = r.next_float();
            dyaw += (a2 - b2) * c2 * 4.0f;
            if (i == branch && thickness > 1.0f) {
                if constexpr (kBranch) {
                    const uint64_t s1 = r.next_long();
                    const float t1 = r.next_float() * 0.5f + 0.5f;
                    spawn(TunnelJob{s1, x, y, z, j.hrm, j.vrm, t1, yaw - 1.5707964f, pitch / 3.0f, j.floor_level, i, steps});
                    const uint64_t s2 = r.next_long();
                    const float t2 = r.next_float() * 0.5f + 0.5f;
                    spawn(TunnelJob{s2, x, y, z, j.hrm, j.vrm, t2, yaw + 1.5707964f, pitch / 3.0f, j.floor_level, i, steps});
                }
                return;
            }
            if (r.next_int(4) == 0) continue;
            const double dx = x - mid_x, dz = z - mid_z, dy = (double)(steps - i);
            const double rr = (double)(thickness + 2.0f + 16.0f);
            if (!(dx * dx + dz * dz - dy * dy <= rr * rr)) return;
            const Ell e{x, y, z, radius * hrm, vert * vrm, j.floor_level, kCaveEll};
            if (in_reach(e)) ell(e, (const float *)nullptr);
        }
    }

    // CaveWorldCarver.carve (carver.c:332-372): rooms go to `ell`, tunnels to `spawn` (each tunnel draws from its own
    // random seeded from this one, so the order they run in changes nothing)
    template <class Spawn, class EllSink>
    __device__ void caves(JRand &r, int ox, int oz, int y_hi, Spawn &&spawn, EllSink &&ell) const {
        const int n = r.next_int(r.next_int(r.next_int(15) + 1) + 1);
        for (int k = 0; k < n; k++) {
            const double x = (double)(ox * 16 + r.next_int(16));
            const int y = r.next_int(y_hi - -56 + 1) + -56;
            const double z = (double)(oz * 16 + r.next_int(16));
            const float hrm = r.between(0.7f, 1.4f);
            const float vrm = r.between(0.8f, 1.3f);
            const float floor_level = r.between(-1.0f, -0.4f);
            int tunnels = 1;
            if (r.next_int(4) == 0) {
                const double room_scale = (double)r.between(0.1f, 0.9f);
                const float f = 1.0f + r.next_float() * 6.0f;
                const double d17 = 1.5 + (double)(msin(1.5707963705062866) * f);
                const Ell e{x + 1.0, (double)y, z, d17, d17 * room_scale, floor_level, kCaveEll};
                if (in_reach(e)) ell(e, (const float *)nullptr);
                tunnels += r.next_int(4);
            }
            for (int t = 0; t < tunnels; t++) {
                const float yaw = r.next_float() * 6.2831855f;
                const float pitch = (r.next_float() - 0.5f) / 4.0f;
                const float t0 = r.next_float(), t1 = r.next_float();
                float thick = t0 * 2.0f + t1;
                if (r.next_int(10) == 0) {
                    const float a = r.next_float(), b = r.next_float();
                    thick = thick * (a * b * 3.0f + 1.0f);
                }
                const int steps = 112 - r.next_int(112 / 4);
                const uint64_t seed = r.next_long();
                spawn(TunnelJob{seed, x, (double)y, z, hrm, vrm, thick, yaw, pitch, floor_level, 0, steps});
            }
        }
    }

    // CanyonWorldCarver.carve and doCarve (carver.c:376-494); the width table comes from `widths(id)` (id >= 0), or
    // is the lane's own (id = kOwnWidths) when that gives none
    template <class EllSink, class Widths>
    __device__ void canyon(JRand &r, int ox, int oz, EllSink &&ell, Widths &&widths) const {
        double x = (double)(ox * 16 + r.next_int(16));
        double y = (double)(r.next_int(67 - 10 + 1) + 10);
        double z = (double)(oz * 16 + r.next_int(16));
        float yaw = r.next_float() * 6.2831855f;
        float pitch = r.between(-0.125f, 0.125f);
        const double y_scale = 3.0;
        const float d = 6.0f - 0.0f, e = (d - 2.0f) / 2.0f, f = d - e;   // TrapezoidFloat(0, 6, plateau 2)
        const float a = r.next_float() * f + 0.0f;
        const float thick = a + r.next_float() * e;
        const float dist = r.between(0.75f, 1.0f);
        const int steps = (int)((float)112 * dist);
        const uint64_t seed = r.next_long();
        JRand q;
        q.seed(seed);
        float own[384];
        int id = kOwnWidths;
        float *width = widths(id);
        if (!width) {
            width = own;
            id = kOwnWidths;
        }
        float fw = 1.0f;
        for (int i = 0; i < 384; i++) {
            if (i == 0 || q.next_int(3) == 0) {
                const float u = q.next_float(), v = q.next_float();
                fw = 1.0f + u * v;
            }
            width[i] = fw * fw;
        }
        float dyaw = 0.0f, dpitch = 0.0f;
        for (int i = 0; i < steps; i++) {
            const float ang = (float)i * 3.1415927f / (float)steps;
            double radius = 1.5 + (double)(msin((double)ang) * thick);
            double vert = radius * y_scale;
            radius *= (double)q.between(0.75f, 1.0f);
            const float f7 = 1.0f - fabsf(0.5f - (float)i / (float)steps) * 2.0f;
            const float f8 = 0.0f + 1.0f * f7;
            vert = (double)f8 * vert * (double)q.between(0.75f, 1.0f);
            const float cp = mcos((double)pitch), sp = msin((double)pitch);
            x += (double)(mcos((double)yaw) * cp);
            y += (double)sp;
            z += (double)(msin((double)yaw) * cp);
            pitch *= 0.7f;
            pitch += dpitch * 0.05f;
            yaw += dyaw * 0.05f;
            dpitch *= 0.8f;
            dyaw *= 0.5f;
            const float a1 = q.next_float(), b1 = q.next_float(), c1 = q.next_float();
            dpitch += (a1 - b1) * c1 * 2.0f;
            const float a2 = q.next_float(), b2 = q.next_float(), c2 = q.next_float();
            dyaw += (a2 - b2) * c2 * 4.0f;
            if (q.next_int(4) == 0) continue;
            const double dx = x - mid_x, dz = z - mid_z, dy = (double)(steps - i);
            const double rr = (double)(thick + 2.0f + 16.0f);
            if (!(dx * dx + dz * dz - dy * dy <= rr * rr)) return;
            const Ell el{x, y, z, radius, vert, 0.0f, id};
            if (in_reach(el)) ell(el, (const float *)width);
        }
    }

    // one carver of one origin chunk (carver.c:498-523): the start roll, then its caves or canyon
    template <class Spawn, class EllSink, class Widths>
    __device__ void carver(uint64_t world, int ox, int oz, int ci, Spawn &&spawn, EllSink &&ell, Widths &&widths) const {
        JRand r;
        r.large_feature(world + (uint64_t)ci, ox, oz);
        const float prob = ci == 0 ? 0.15f : (ci == 1 ? 0.07f : 0.01f);
        if (!(r.next_float() <= prob)) return;
        if (ci == 2) canyon(r, ox, oz, ell, widths);
        else caves(r, ox, oz, ci == 0 ? 180 : 47, spawn, ell);
    }
};

}  // namespace whf
