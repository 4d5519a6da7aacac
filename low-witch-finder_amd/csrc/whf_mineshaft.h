// whf_mineshaft.h — Deep Dive device code: the chunk carver seed, the mineshaft start test and the walk of the
// mineshaft layout that fixes the room floor. Integer only.
//
// The layout walk follows the game's MineshaftPieces as the reference restates it (engine/tools/mstree/
// mineshaft_tree.h:73-85 draws, :275-312 collision, :330-452 placement, :454-520 piece attempt, :529-565 move
// below sea level) and visits children in the order of its iterative form (mineshaft_tree_iter.h:18-160).
// A group of G lanes walks one layout: every lane of the group computes the same draws and decisions, and the
// collision test of a new box against the placed pieces is split over the G lanes (piece i on lane i mod G)
// and joined by a ballot. The group's pieces live in LDS.
#pragma once

#include "whf_device.h"

namespace whf {

// java.util.Random.nextDouble() < 0.004  <=>  (next(26) << 27) + next(27) < kStartLimit, since the double is that
// 53-bit integer times 2^-53 exactly and 0.004 (as a double) times 2^53 is 36028797018963.96875.
constexpr uint64_t kStartLimit = 36028797018964ull;
constexpr int kMsMaxPieces = 512;   // the reference's piece cap (mineshaft_tree.h:11-13)
constexpr int kMsMaxFrames = 16;    // depth <= 9 below the room keeps the walk at most 10 frames deep

__device__ __forceinline__ uint64_t lcg_step(uint64_t s) { return (s * kLcgMul + 0xBull) & kLcgM48; }

// Random.nextLong(): (next(32) << 32) + next(32), the low word sign-extended (rng.c:197-211).
__device__ __forceinline__ uint64_t lcg_next_long(uint64_t &st) {
    st = lcg_step(st);
    const int32_t hi = (int32_t)(uint32_t)(st >> 16);
    st = lcg_step(st);
    const int32_t lo = (int32_t)(uint32_t)(st >> 16);
    return ((uint64_t)(int64_t)hi << 32) + (uint64_t)(int64_t)lo;
}

// Chunk carver seed (room_core.h:55-58 with the setSeed mask): a, b are the first two nextLong of Random(S).
__device__ __forceinline__ uint64_t carver_seed(uint64_t s48, uint64_t a, uint64_t b, int cx, int cz) {
    return ((uint64_t)((int64_t)cx * (int64_t)a) ^ (uint64_t)((int64_t)cz * (int64_t)b) ^ s48) & kLcgM48;
}

// new Random(carver).nextDouble() < 0.004 (mstree.cpp:10-15); the second draw only when the first leaves it open.
__device__ __forceinline__ bool ms_is_start(uint64_t carver) {
    uint64_t s = lcg_step((carver ^ kLcgMul) & kLcgM48);
    const uint64_t hi = s >> 22;
    if (hi > ((kStartLimit - 1) >> 27)) return false;
    s = lcg_step(s);
    return (hi << 27) + (s >> 21) < kStartLimit;
}

// ---------------------------------------------------------------------------------------------- layout walk

enum { MS_ROOM = 0, MS_CORRIDOR = 1, MS_CROSSING = 2, MS_STAIRS = 3 };
enum { MS_N = 0, MS_S = 1, MS_W = 2, MS_E = 3 };
enum { MS_DONE = 0, MS_CUT = 1, MS_CAPPED = 2 };

struct MsOut {
    int status;        // MS_DONE, MS_CUT (a piece below the cut line), MS_CAPPED (piece cap or int8 box range)
    int floor;         // room floor after the move below sea level (MS_DONE only)
    int pieces;
    uint32_t draws;    // LCG steps, the start test's two included
};

// java.util.Random over the walk's 48-bit state.
struct MsRng {
    uint64_t rs;
    uint32_t draws;
    __device__ __forceinline__ int32_t next(int bits) {
        rs = lcg_step(rs);
        draws++;
        return (int32_t)(rs >> (48 - bits));
    }
    __device__ __forceinline__ int32_t next_int(int32_t bound) {
        if ((bound & -bound) == bound) return (int32_t)(((int64_t)bound * (int64_t)next(31)) >> 31);
        int32_t bits, val;
        do {
            bits = next(31);
            val = bits % bound;
        } while ((int32_t)((uint32_t)bits - (uint32_t)val + (uint32_t)(bound - 1)) < 0);
        return val;
    }
};

// Boxes are kept relative to the room's low corner (x, z) and to y = 50, as int8 offsets:
//   .x = (x0 + 128) | (z0 + 128) << 8 | (127 - x1) << 16 | (127 - z1) << 24
//   .y = (y0 + 128) | (127 - y1) << 16
// so a placed box P meets a query box Q (inclusive bounds) exactly when every one of the six byte fields of P
// is <= the matching field of Q's dual encoding (x1, z1, -x0, -z0 and y1, -y0). The test runs on 16-bit lanes:
// (256 + q) - p keeps bit 8 exactly when q >= p, for bytes p, q.
__device__ __forceinline__ bool fits8(int v) { return v >= -128 && v <= 127; }

__device__ __forceinline__ uint2 ms_encode(int x0, int y0, int z0, int x1, int y1, int z1) {
    return make_uint2((uint32_t)(x0 + 128) | (uint32_t)(z0 + 128) << 8 | (uint32_t)(127 - x1) << 16 |
                          (uint32_t)(127 - z1) << 24,
                      (uint32_t)(y0 + 128) | (uint32_t)(127 - y1) << 16);
}

struct MsQuery {
    uint32_t a, b, c;
};

__device__ __forceinline__ MsQuery ms_query(int x0, int y0, int z0, int x1, int y1, int z1) {
    const uint32_t q = (uint32_t)(x1 + 128) | (uint32_t)(z1 + 128) << 8 | (uint32_t)(127 - x0) << 16 |
                       (uint32_t)(127 - z0) << 24;
    return MsQuery{(q & 0x00FF00FFu) | 0x01000100u, ((q >> 8) & 0x00FF00FFu) | 0x01000100u,
                   ((uint32_t)(y1 + 128) | (uint32_t)(127 - y0) << 16) | 0x01000100u};
}

__device__ __forceinline__ bool ms_meets(uint2 p, const MsQuery &q) {
    const uint32_t t = (q.a - (p.x & 0x00FF00FFu)) & (q.b - ((p.x >> 8) & 0x00FF00FFu)) & (q.c - p.y);
    return (t & 0x01000100u) == 0x01000100u;
}

// Walk frame (one word): piece index, phase, side, k (offset along the piece), and the piece's kind.
__device__ __forceinline__ uint32_t ms_frame(int idx, int phase, int side, int k, int type, int facing, int depth) {
    return (uint32_t)idx | (uint32_t)phase << 9 | (uint32_t)side << 13 | (uint32_t)k << 15 | (uint32_t)type << 23 |
           (uint32_t)facing << 25 | (uint32_t)depth << 27;
}

// The room floor of the mineshaft started by `carver`. Pieces whose y0 falls below cut_rel (relative to y = 50)
// stop the walk with MS_CUT; INT_MIN never cuts. box/frm are the group's LDS arrays; lig is the lane's index in
// the group and gshift the group's first lane in the wave.
template <int G>
__device__ MsOut ms_room_floor(uint64_t carver, int cut_rel, uint2 *box, uint32_t *frm, int lig, int gshift) {
    constexpr uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    MsRng g{(carver ^ kLcgMul) & kLcgM48, 0};
    (void)g.next(26);   // the start test's nextDouble
    (void)g.next(27);
    const int rx1 = 7 + g.next_int(6);
    const int ry1 = 4 + g.next_int(6);
    const int rz1 = 7 + g.next_int(6);
    box[0] = ms_encode(0, 0, 0, rx1, ry1, rz1);
    frm[0] = ms_frame(0, 0, 0, 0, MS_ROOM, 0, 0);
    int n = 1, sp = 1, min_y = 0, max_y = ry1, stop = MS_DONE;

    while (sp > 0) {
        uint32_t f = frm[sp - 1];
        const int idx = (int)(f & 511u);
        int phase = (int)((f >> 9) & 15u), side = (int)((f >> 13) & 3u), k = (int)((f >> 15) & 255u);
        const int type = (int)((f >> 23) & 3u), facing = (int)((f >> 25) & 3u), d = (int)((f >> 27) & 15u);
        const uint2 pb = box[idx];
        const int x0 = (int)(pb.x & 255u) - 128, z0 = (int)((pb.x >> 8) & 255u) - 128;
        const int x1 = 127 - (int)((pb.x >> 16) & 255u), z1 = 127 - (int)(pb.x >> 24);
        const int y0 = (int)(pb.y & 0xFFFFu) - 128, y1 = 127 - (int)(pb.y >> 16);

        // ---- the next child request of this piece (mineshaft_tree_iter.h:18-152) ----
        bool have = false;
        int ri = 0, rj = 0, rk = 0, rdir = 0, rl = d;
        if (type == MS_STAIRS) {
            if (phase == 0) {
                phase = 1;
                have = true;
                rj = y0;
                if (facing == MS_N) { ri = x0; rk = z0 - 1; }
                else if (facing == MS_S) { ri = x0; rk = z1 + 1; }
                else if (facing == MS_W) { ri = x0 - 1; rk = z0; }
                else { ri = x1 + 1; rk = z0; }
                rdir = facing;
            }
        } else if (type == MS_CROSSING) {
            while (phase < 7) {
                const int ph = phase++;
                if (ph < 3) {
                    // straight on, then the two sides; the side the crossing was entered from is never tried
                    rj = y0;
                    if (facing == MS_N || facing == MS_S) {
                        if (ph == 0) { ri = x0 + 1; rk = facing == MS_N ? z0 - 1 : z1 + 1; rdir = facing; }
                        else if (ph == 1) { ri = x0 - 1; rk = z0 + 1; rdir = MS_W; }
                        else { ri = x1 + 1; rk = z0 + 1; rdir = MS_E; }
                    } else {
                        if (ph == 0) { ri = x0 + 1; rk = z0 - 1; rdir = MS_N; }
                        else if (ph == 1) { ri = x0 + 1; rk = z1 + 1; rdir = MS_S; }
                        else { ri = facing == MS_W ? x0 - 1 : x1 + 1; rk = z0 + 1; rdir = facing; }
                    }
                    have = true;
                    break;
                }
                if (y1 - y0 + 1 <= 3) { phase = 7; break; }
                if (!g.next(1)) continue;   // each exit of the upper storey is a coin flip
                const int q = ph - 3;
                rj = y0 + 4;
                if (q == 0) { ri = x0 + 1; rk = z0 - 1; rdir = MS_N; }
                else if (q == 1) { ri = x0 - 1; rk = z0 + 1; rdir = MS_W; }
                else if (q == 2) { ri = x1 + 1; rk = z0 + 1; rdir = MS_E; }
                else { ri = x0 + 1; rk = z1 + 1; rdir = MS_S; }
                have = true;
                break;
            }
        } else if (type == MS_CORRIDOR) {
            const bool ew = facing == MS_W || facing == MS_E;
            if (phase == 0) {
                // the continuation: on (j <= 1) or turning, one block up or down at random
                const int j = g.next_int(4);
                rj = y0 - 1 + g.next_int(3);
                if (facing == MS_N) {
                    if (j <= 1) { ri = x0; rk = z0 - 1; rdir = MS_N; }
                    else if (j == 2) { ri = x0 - 1; rk = z0; rdir = MS_W; }
                    else { ri = x1 + 1; rk = z0; rdir = MS_E; }
                } else if (facing == MS_S) {
                    if (j <= 1) { ri = x0; rk = z1 + 1; rdir = MS_S; }
                    else if (j == 2) { ri = x0 - 1; rk = z1 - 3; rdir = MS_W; }
                    else { ri = x1 + 1; rk = z1 - 3; rdir = MS_E; }
                } else if (facing == MS_W) {
                    if (j <= 1) { ri = x0 - 1; rk = z0; rdir = MS_W; }
                    else if (j == 2) { ri = x0; rk = z0 - 1; rdir = MS_N; }
                    else { ri = x0; rk = z1 + 1; rdir = MS_S; }
                } else {
                    if (j <= 1) { ri = x1 + 1; rk = z0; rdir = MS_E; }
                    else if (j == 2) { ri = x1 - 3; rk = z0 - 1; rdir = MS_N; }
                    else { ri = x1 - 3; rk = z1 + 1; rdir = MS_S; }
                }
                phase = 1;
                have = true;
            } else if (phase <= 2) {
                // side branches every 5 blocks, one depth further down
                if (phase == 1) {
                    if (d >= 8) phase = 9;
                    else k = 3;
                } else {
                    k += 5;
                }
                if (phase != 9) {
                    const int lim = ew ? x1 - x0 : z1 - z0;
                    while (k + 3 <= lim) {
                        const int l = g.next_int(5);
                        if (l <= 1) {
                            rj = y0;
                            rl = d + 1;
                            if (ew) { ri = x0 + k; rk = l == 0 ? z0 - 1 : z1 + 1; rdir = l == 0 ? MS_N : MS_S; }
                            else { ri = l == 0 ? x0 - 1 : x1 + 1; rk = z0 + k; rdir = l == 0 ? MS_W : MS_E; }
                            phase = 2;
                            have = true;
                            break;
                        }
                        k += 5;
                    }
                    if (!have) phase = 9;
                }
            }
        } else {
            // the room: exits along its sides in turn, N, S, W, E
            const int xs = x1 - x0 + 1, zs = z1 - z0 + 1;
            const int jj = max(1, (y1 - y0 + 1) - 4);
            while (side < 4) {
                const int span = side < 2 ? xs : zs;
                k = phase == 0 ? 0 : k + 4;
                if (k >= span) { side++; phase = 0; continue; }
                k += g.next_int(span);
                if (k + 3 > span) { side++; phase = 0; continue; }
                rj = y0 + g.next_int(jj) + 1;
                if (side == 0) { ri = x0 + k; rk = z0 - 1; rdir = MS_N; }
                else if (side == 1) { ri = x0 + k; rk = z1 + 1; rdir = MS_S; }
                else if (side == 2) { ri = x0 - 1; rk = z0 + k; rdir = MS_W; }
                else { ri = x1 + 1; rk = z0 + k; rdir = MS_E; }
                phase = 1;
                have = true;
                break;
            }
        }
        if (!have) {
            sp--;
            continue;
        }
        frm[sp - 1] = ms_frame(idx, phase, side, k, type, facing, d);

        // ---- the attempt (mineshaft_tree_iter.h:163-209) ----
        if (rl > 8 || abs(ri) > 80 || abs(rk) > 80) continue;
        int bx0 = ri, by0 = rj, bz0 = rk, bx1 = ri, by1 = rj + 2, bz1 = rk, nt;
        bool placed = false;
        const int m = g.next_int(100);
        if (m >= 70) {
            if (m >= 80) {
                nt = MS_CROSSING;
                if (g.next_int(4) == 0) by1 += 4;
                if (rdir == MS_N) { bx0 = ri - 1; bx1 = ri + 3; bz0 = rk - 4; }
                else if (rdir == MS_S) { bx0 = ri - 1; bx1 = ri + 3; bz1 = rk + 4; }
                else if (rdir == MS_W) { bx0 = ri - 4; bz0 = rk - 1; bz1 = rk + 3; }
                else { bx1 = ri + 4; bz0 = rk - 1; bz1 = rk + 3; }
            } else {
                nt = MS_STAIRS;
                by0 = rj - 5;
                if (rdir == MS_N) { bx1 = ri + 2; bz0 = rk - 8; }
                else if (rdir == MS_S) { bx1 = ri + 2; bz1 = rk + 8; }
                else if (rdir == MS_W) { bx0 = ri - 8; bz1 = rk + 2; }
                else { bx1 = ri + 8; bz1 = rk + 2; }
            }
            if (!(fits8(bx0) && fits8(bx1) && fits8(bz0) && fits8(bz1) && fits8(by0) && fits8(by1))) {
                stop = MS_CAPPED;
                break;
            }
            const MsQuery q = ms_query(bx0, by0, bz0, bx1, by1, bz1);
            bool hit = false;
            for (int i = lig; i < n; i += G) hit |= ms_meets(box[i], q);
            placed = ((__ballot(hit) >> gshift) & gmask) == 0;
        } else {
            nt = MS_CORRIDOR;
            // the longest of 4, 3 or 2 five-block sections that meets no placed piece
            int len = g.next_int(3) + 2;
            bool bad = false;
            for (; len > 0; --len) {
                const int L = len * 5;
                if (rdir == MS_N) { bx1 = ri + 2; bz0 = rk - (L - 1); }
                else if (rdir == MS_S) { bx1 = ri + 2; bz1 = rk + L - 1; }
                else if (rdir == MS_W) { bx0 = ri - (L - 1); bz1 = rk + 2; }
                else { bx1 = ri + L - 1; bz1 = rk + 2; }
                if (!(fits8(bx0) && fits8(bx1) && fits8(bz0) && fits8(bz1) && fits8(by0) && fits8(by1))) {
                    bad = true;
                    break;
                }
                const MsQuery q = ms_query(bx0, by0, bz0, bx1, by1, bz1);
                bool hit = false;
                for (int i = lig; i < n; i += G) hit |= ms_meets(box[i], q);
                if (((__ballot(hit) >> gshift) & gmask) == 0) break;
            }
            if (bad) {
                stop = MS_CAPPED;
                break;
            }
            if (len > 0) {
                placed = true;
                if (g.next_int(3) != 0) (void)g.next_int(23);   // no rails: the cobweb roll
            }
        }
        if (!placed) continue;
        if (n >= kMsMaxPieces) {
            stop = MS_CAPPED;
            break;
        }
        box[n] = ms_encode(bx0, by0, bz0, bx1, by1, bz1);
        n++;
        min_y = min(min_y, by0);
        max_y = max(max_y, by1);
        if (by0 < cut_rel) {
            stop = MS_CUT;
            break;
        }
        if (sp >= kMsMaxFrames) {   // unreachable: depth <= 9 bounds the walk
            stop = MS_CAPPED;
            break;
        }
        frm[sp++] = ms_frame(n - 1, 0, 0, 0, nt, rdir, rl + 1);
    }

    MsOut o{stop, 0, n, 0};
    if (stop == MS_DONE) {
        // moveBelowSeaLevel(63, -64, random, 10) (mineshaft_tree.h:556-564), in y relative to 50
        const int span = max_y - min_y + 1;
        int l = span - 63;
        if (l < 53) l += g.next_int(53 - l);
        o.floor = l - max_y;
    }
    o.draws = g.draws;
    return o;
}

}  // namespace whf
