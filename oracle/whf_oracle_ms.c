/* whf_oracle_ms.c — CPU oracle for the Deep Dive mineshaft stage.
 *
 * TEST INFRASTRUCTURE ONLY (like whf_oracle.c): a plain-C restatement of the reference's mineshaft start test
 * and room-layout generator, used by tests/, the smoke check and the benchmark's CPU-baseline leg as the
 * checker. The HIP product path never links, loads or calls it.
 *
 * What it restates:
 *   - carver seed of a chunk: Random(S48) -> a, b = two nextLong; X = (cx*a ^ cz*b ^ S48) mod 2^48
 *     (engine/gpu/room_core.h:55-58, engine/tools/mstree/mstree.cpp:43-46, swamphut.h:77-90);
 *   - mineshaft start: new Random(X).nextDouble() < 0.004 (mstree.cpp:10-15, room_core.h:64-72);
 *   - the mineshaft layout and its room floor after the move below sea level
 *     (engine/tools/mstree/mineshaft_tree.h:73-85 draws, :275-312 overlap, :330-452 placement,
 *     :454-520 piece attempt, :529-565 generate and move; the iterative walk mineshaft_tree_iter.h:18-296
 *     visits children in the same order as this recursion).
 * The layout is kept as full int32 boxes. The reference packs boxes as int8 offsets from the room corner and
 * reports status 2 when a box does not fit; the same test is made here so the status codes agree.
 */
#include <limits.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "whf_oracle.h"

#define MS_MUL UINT64_C(0x5DEECE66D)
#define MS_M48 UINT64_C(0xFFFFFFFFFFFF)

enum { DIR_N = 0, DIR_S = 1, DIR_W = 2, DIR_E = 3 };

/* ------------------------------------------------------------------ start test and carver seed */

uint64_t wo_ms_carver_seed(uint64_t s48, int32_t cx, int32_t cz) {
    uint64_t st = ((s48 & MS_M48) ^ MS_MUL) & MS_M48;
    uint64_t a = wo_lcg_next_long(&st);
    uint64_t b = wo_lcg_next_long(&st);
    return ((uint64_t)((int64_t)cx * (int64_t)a) ^ (uint64_t)((int64_t)cz * (int64_t)b) ^ s48) & MS_M48;
}

static inline int32_t ms_bits(uint64_t *st, int bits) {
    *st = (*st * MS_MUL + 0xB) & MS_M48;
    return (int32_t)(*st >> (48 - bits));
}

double wo_ms_start_double(uint64_t carver) {
    /* java.util.Random.nextDouble: ((long)next(26) << 27) + next(27), times 2^-53 */
    uint64_t st = (carver ^ MS_MUL) & MS_M48;
    int64_t hi = ms_bits(&st, 26);
    int64_t lo = ms_bits(&st, 27);
    return (double)((hi << 27) + lo) * 0x1.0p-53;
}

int wo_ms_is_start(uint64_t carver) { return wo_ms_start_double(carver) < 0.004; }

/* ------------------------------------------------------------------ layout generator */

typedef struct {
    uint64_t rs;
    wo_ms_piece *p;
    int n, cap;
    int sx, sz;
    int min_y, max_y;
    int abort_below;
    int stop;        /* 1: a piece went below the cut; 2: piece cap reached */
    int range_bad;   /* a box or a query box does not fit int8 offsets from (sx, 50, sz) */
    uint64_t draws;
    uint64_t overlap_tests;
} ms_gen;

static inline int32_t g_next(ms_gen *g, int bits) {
    /* draws made after the walk stopped (the recursion still rolls a crossing's coins) change no result and
     * are not counted */
    if (!g->stop) g->draws++;
    return ms_bits(&g->rs, bits);
}

/* java.util.Random.nextInt(bound) (mineshaft_tree.h:77-85) */
static int32_t g_int(ms_gen *g, int32_t bound) {
    if ((bound & -bound) == bound) return (int32_t)(((int64_t)bound * (int64_t)g_next(g, 31)) >> 31);
    int32_t bits, val;
    do {
        bits = g_next(g, 31);
        val = bits % bound;
    } while ((int32_t)((uint32_t)bits - (uint32_t)val + (uint32_t)(bound - 1)) < 0);
    return val;
}

static inline int fits8(int v) { return v >= -128 && v <= 127; }

static int box_fits(const ms_gen *g, int x0, int y0, int z0, int x1, int y1, int z1) {
    return fits8(x0 - g->sx) && fits8(x1 - g->sx) && fits8(z0 - g->sz) && fits8(z1 - g->sz) && fits8(y0 - 50) &&
           fits8(y1 - 50);
}

/* inclusive box intersection against every placed piece (StructurePiece.findCollisionPiece) */
static int g_overlaps(ms_gen *g, int x0, int y0, int z0, int x1, int y1, int z1) {
    g->overlap_tests++;
    if (!box_fits(g, x0, y0, z0, x1, y1, z1)) {
        g->range_bad = 1;   /* the reference cannot test such a box: it reports no overlap and flags the layout */
        return 0;
    }
    for (int i = 0; i < g->n; i++) {
        const wo_ms_piece *q = &g->p[i];
        if (q->x1 >= x0 && q->x0 <= x1 && q->z1 >= z0 && q->z0 <= z1 && q->y1 >= y0 && q->y0 <= y1) return 1;
    }
    return 0;
}

static void g_push(ms_gen *g, const wo_ms_piece *np) {
    if (!box_fits(g, np->x0, np->y0, np->z0, np->x1, np->y1, np->z1)) g->range_bad = 1;
    g->p[g->n++] = *np;
    if (np->y0 < g->min_y) g->min_y = np->y0;
    if (np->y1 > g->max_y) g->max_y = np->y1;
}

static void g_try(ms_gen *g, int i, int j, int k, int dir, int depth, int parent);

static void place_children(ms_gen *g, int idx) {
    const wo_ms_piece s = g->p[idx];
    const int d = s.depth;
    if (s.type == WO_MS_STAIRS) {
        switch (s.facing) {
        case DIR_N: g_try(g, s.x0, s.y0, s.z0 - 1, DIR_N, d, idx); break;
        case DIR_S: g_try(g, s.x0, s.y0, s.z1 + 1, DIR_S, d, idx); break;
        case DIR_W: g_try(g, s.x0 - 1, s.y0, s.z0, DIR_W, d, idx); break;
        default:    g_try(g, s.x1 + 1, s.y0, s.z0, DIR_E, d, idx); break;
        }
        return;
    }
    if (s.type == WO_MS_CROSSING) {
        /* straight on, then the two sides; the side opposite the entrance is never tried */
        switch (s.facing) {
        case DIR_N:
            g_try(g, s.x0 + 1, s.y0, s.z0 - 1, DIR_N, d, idx);
            g_try(g, s.x0 - 1, s.y0, s.z0 + 1, DIR_W, d, idx);
            g_try(g, s.x1 + 1, s.y0, s.z0 + 1, DIR_E, d, idx);
            break;
        case DIR_S:
            g_try(g, s.x0 + 1, s.y0, s.z1 + 1, DIR_S, d, idx);
            g_try(g, s.x0 - 1, s.y0, s.z0 + 1, DIR_W, d, idx);
            g_try(g, s.x1 + 1, s.y0, s.z0 + 1, DIR_E, d, idx);
            break;
        case DIR_W:
            g_try(g, s.x0 + 1, s.y0, s.z0 - 1, DIR_N, d, idx);
            g_try(g, s.x0 + 1, s.y0, s.z1 + 1, DIR_S, d, idx);
            g_try(g, s.x0 - 1, s.y0, s.z0 + 1, DIR_W, d, idx);
            break;
        default:
            g_try(g, s.x0 + 1, s.y0, s.z0 - 1, DIR_N, d, idx);
            g_try(g, s.x0 + 1, s.y0, s.z1 + 1, DIR_S, d, idx);
            g_try(g, s.x1 + 1, s.y0, s.z0 + 1, DIR_E, d, idx);
            break;
        }
        if (s.y1 - s.y0 + 1 > 3) {   /* two-storey crossing: each upper exit is a coin flip */
            const int uy = s.y0 + 4;
            if (g_next(g, 1)) g_try(g, s.x0 + 1, uy, s.z0 - 1, DIR_N, d, idx);
            if (g_next(g, 1)) g_try(g, s.x0 - 1, uy, s.z0 + 1, DIR_W, d, idx);
            if (g_next(g, 1)) g_try(g, s.x1 + 1, uy, s.z0 + 1, DIR_E, d, idx);
            if (g_next(g, 1)) g_try(g, s.x0 + 1, uy, s.z1 + 1, DIR_S, d, idx);
        }
        return;
    }
    if (s.type == WO_MS_CORRIDOR) {
        /* the continuation: straight on (j <= 1) or a turn, one block up or down */
        const int j = g_int(g, 4);
        const int y = s.y0 - 1 + g_int(g, 3);
        int ci, ck, cd;
        switch (s.facing) {
        case DIR_N:
            if (j <= 1) { ci = s.x0; ck = s.z0 - 1; cd = DIR_N; }
            else if (j == 2) { ci = s.x0 - 1; ck = s.z0; cd = DIR_W; }
            else { ci = s.x1 + 1; ck = s.z0; cd = DIR_E; }
            break;
        case DIR_S:
            if (j <= 1) { ci = s.x0; ck = s.z1 + 1; cd = DIR_S; }
            else if (j == 2) { ci = s.x0 - 1; ck = s.z1 - 3; cd = DIR_W; }
            else { ci = s.x1 + 1; ck = s.z1 - 3; cd = DIR_E; }
            break;
        case DIR_W:
            if (j <= 1) { ci = s.x0 - 1; ck = s.z0; cd = DIR_W; }
            else if (j == 2) { ci = s.x0; ck = s.z0 - 1; cd = DIR_N; }
            else { ci = s.x0; ck = s.z1 + 1; cd = DIR_S; }
            break;
        default:
            if (j <= 1) { ci = s.x1 + 1; ck = s.z0; cd = DIR_E; }
            else if (j == 2) { ci = s.x1 - 3; ck = s.z0 - 1; cd = DIR_N; }
            else { ci = s.x1 - 3; ck = s.z1 + 1; cd = DIR_S; }
            break;
        }
        g_try(g, ci, y, ck, cd, d, idx);
        if (g->stop || d >= 8) return;
        /* side branches every 5 blocks along the corridor, one depth further down the tree */
        if (s.facing == DIR_W || s.facing == DIR_E) {
            for (int kk = s.x0 + 3; kk + 3 <= s.x1; kk += 5) {
                const int l = g_int(g, 5);
                if (l == 0) g_try(g, kk, s.y0, s.z0 - 1, DIR_N, d + 1, idx);
                else if (l == 1) g_try(g, kk, s.y0, s.z1 + 1, DIR_S, d + 1, idx);
                if (g->stop) return;
            }
        } else {
            for (int kk = s.z0 + 3; kk + 3 <= s.z1; kk += 5) {
                const int l = g_int(g, 5);
                if (l == 0) g_try(g, s.x0 - 1, s.y0, kk, DIR_W, d + 1, idx);
                else if (l == 1) g_try(g, s.x1 + 1, s.y0, kk, DIR_E, d + 1, idx);
                if (g->stop) return;
            }
        }
        return;
    }
    /* the room: exits along each of its four sides in turn (N, S, W, E) */
    const int xs = s.x1 - s.x0 + 1, zs = s.z1 - s.z0 + 1;
    int jj = (s.y1 - s.y0 + 1) - 4;
    if (jj <= 0) jj = 1;
    for (int side = 0; side < 4; side++) {
        const int span = side < 2 ? xs : zs;
        for (int k = 0; k < span; k += 4) {
            k += g_int(g, span);
            if (k + 3 > span) break;
            const int y = s.y0 + g_int(g, jj) + 1;
            switch (side) {
            case 0: g_try(g, s.x0 + k, y, s.z0 - 1, DIR_N, d, idx); break;
            case 1: g_try(g, s.x0 + k, y, s.z1 + 1, DIR_S, d, idx); break;
            case 2: g_try(g, s.x0 - 1, y, s.z0 + k, DIR_W, d, idx); break;
            default: g_try(g, s.x1 + 1, y, s.z0 + k, DIR_E, d, idx); break;
            }
            if (g->stop) return;
        }
    }
}

static void g_try(ms_gen *g, int i, int j, int k, int dir, int depth, int parent) {
    if (g->stop) return;
    if (depth > 8 || abs(i - g->sx) > 80 || abs(k - g->sz) > 80) return;
    wo_ms_piece np;
    memset(&np, 0, sizeof np);
    np.facing = (int8_t)dir;
    np.depth = (int8_t)(depth + 1);
    np.parent = (int16_t)parent;
    int x0 = i, y0 = j, z0 = k, x1 = i, y1 = j + 2, z1 = k;
    const int m = g_int(g, 100);
    if (m >= 80) {
        np.type = WO_MS_CROSSING;
        if (g_int(g, 4) == 0) y1 += 4;
        switch (dir) {
        case DIR_N: x0 = i - 1; x1 = i + 3; z0 = k - 4; break;
        case DIR_S: x0 = i - 1; x1 = i + 3; z1 = k + 4; break;
        case DIR_W: x0 = i - 4; z0 = k - 1; z1 = k + 3; break;
        default:    x1 = i + 4; z0 = k - 1; z1 = k + 3; break;
        }
        if (g_overlaps(g, x0, y0, z0, x1, y1, z1)) return;
    } else if (m >= 70) {
        np.type = WO_MS_STAIRS;
        y0 = j - 5;
        switch (dir) {
        case DIR_N: x1 = i + 2; z0 = k - 8; break;
        case DIR_S: x1 = i + 2; z1 = k + 8; break;
        case DIR_W: x0 = i - 8; z1 = k + 2; break;
        default:    x1 = i + 8; z1 = k + 2; break;
        }
        if (g_overlaps(g, x0, y0, z0, x1, y1, z1)) return;
    } else {
        np.type = WO_MS_CORRIDOR;
        /* the longest of 4, 3 or 2 sections (5 blocks each) that does not collide */
        int len;
        for (len = g_int(g, 3) + 2; len > 0; --len) {
            const int L = len * 5;
            switch (dir) {
            case DIR_N: x1 = i + 2; z0 = k - (L - 1); break;
            case DIR_S: x1 = i + 2; z1 = k + L - 1; break;
            case DIR_W: x0 = i - (L - 1); z1 = k + 2; break;
            default:    x1 = i + L - 1; z1 = k + 2; break;
            }
            if (!g_overlaps(g, x0, y0, z0, x1, y1, z1)) break;
        }
        if (len <= 0) return;
        np.rails = (int8_t)(g_int(g, 3) == 0);
        if (!np.rails) (void)g_int(g, 23);   /* the cobweb roll of a corridor without rails */
    }
    if (g->n >= g->cap) { g->stop = 2; return; }
    np.x0 = x0; np.y0 = y0; np.z0 = z0; np.x1 = x1; np.y1 = y1; np.z1 = z1;
    g_push(g, &np);
    if (np.y0 < g->abort_below) { g->stop = 1; return; }
    place_children(g, g->n - 1);
}

int wo_ms_layout(uint64_t carver, int32_t cx, int32_t cz, int abort_below, wo_ms_piece *pieces, int cap,
                 wo_ms_result *out) {
    ms_gen g;
    memset(&g, 0, sizeof g);
    memset(out, 0, sizeof *out);
    g.rs = (carver ^ MS_MUL) & MS_M48;
    g.p = pieces;
    g.cap = cap;
    g.abort_below = abort_below;
    g.min_y = INT_MAX;
    g.max_y = INT_MIN;
    /* the start's first two draws were the nextDouble of the start test */
    (void)g_next(&g, 26);
    (void)g_next(&g, 27);
    wo_ms_piece room;
    memset(&room, 0, sizeof room);
    room.type = WO_MS_ROOM;
    room.facing = -1;
    room.depth = 0;
    room.parent = -1;
    room.x0 = cx * 16 + 2;
    room.y0 = 50;
    room.z0 = cz * 16 + 2;
    room.x1 = room.x0 + 7 + g_int(&g, 6);
    room.y1 = 54 + g_int(&g, 6);
    room.z1 = room.z0 + 7 + g_int(&g, 6);
    g.sx = room.x0;
    g.sz = room.z0;
    g_push(&g, &room);
    place_children(&g, 0);
    out->pieces = g.n;
    out->draws = g.draws;
    out->overlap_tests = g.overlap_tests;
    out->min_y = g.min_y;
    out->max_y = g.max_y;
    if (g.stop == 2 || g.range_bad) { out->status = 2; return 2; }
    if (g.stop == 1) { out->status = 1; return 1; }
    /* moveBelowSeaLevel(63, -64, random, 10): l = span - 63, plus nextInt(53 - l) when below 53 */
    const int span = g.max_y - g.min_y + 1;
    int l = span - 63;
    if (l < 53) l += g_int(&g, 53 - l);
    const int m = l - g.max_y;
    out->span_y = span;
    out->room_floor = 50 + m;
    out->draws = g.draws;
    out->status = 0;
    /* report the boxes where the world has them: moved by m */
    for (int i = 0; i < g.n; i++) {
        pieces[i].y0 += m;
        pieces[i].y1 += m;
    }
    return 0;
}

int wo_ms_room_floor(uint64_t carver, int abort_below, wo_ms_result *out) {
    wo_ms_piece buf[WO_MS_MAXP];
    return wo_ms_layout(carver, 0, 0, abort_below, buf, WO_MS_MAXP, out);
}

/* ------------------------------------------------------------------ forward sweep over structure seeds */

typedef struct {
    uint64_t s_first, count;
    int rx0, rx1, rz0, rz1, floor_max, abort_below;
    wo_dd_stats st;
    wo_dd_hit *hits;
    int64_t cap, n;
    pthread_mutex_t *mu;
    wo_dd_hit *shared_hits;
    int64_t *shared_n;
} dd_job;

static void dd_run(dd_job *J, uint64_t i0, uint64_t i1) {
    for (uint64_t i = i0; i < i1; i++) {
        const uint64_t s48 = (J->s_first + i) & MS_M48;
        J->st.structure_seeds++;
        for (int rz = J->rz0; rz < J->rz1; rz++) {
            for (int rx = J->rx0; rx < J->rx1; rx++) {
                int cx, cz;
                wo_hut_chunk(s48, rx, rz, 14357620, &cx, &cz);
                const uint64_t X = wo_ms_carver_seed(s48, cx, cz);
                J->st.tests++;
                if (!wo_ms_is_start(X)) continue;
                J->st.starts++;
                wo_ms_result r;
                wo_ms_room_floor(X, J->abort_below, &r);
                J->st.status[r.status]++;
                J->st.draws += r.draws;
                J->st.pieces += (uint64_t)r.pieces;
                if (r.status != 0 || r.room_floor > J->floor_max) continue;
                J->st.survivors++;
                wo_dd_hit h = {s48, rx, rz, cx, cz, X, r.room_floor, r.pieces};
                pthread_mutex_lock(J->mu);
                if (*J->shared_n < J->cap) J->shared_hits[*J->shared_n] = h;
                (*J->shared_n)++;
                pthread_mutex_unlock(J->mu);
            }
        }
    }
}

typedef struct { dd_job *job; uint64_t i0, i1; } dd_slice;
static void *dd_thread(void *arg) {
    dd_slice *s = (dd_slice *)arg;
    dd_run(s->job, s->i0, s->i1);
    return NULL;
}

int wo_deepdive_sweep(uint64_t s_first, uint64_t count, int rx0, int rx1, int rz0, int rz1, int floor_max,
                      int abort_below, int threads, wo_dd_stats *out, wo_dd_hit *hits, int64_t cap,
                      int64_t *n_hits) {
    if (rx0 >= rx1 || rz0 >= rz1 || threads < 1 || threads > 256) return 1;
    pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    int64_t n = 0;
    dd_job *jobs = calloc((size_t)threads, sizeof *jobs);
    dd_slice *sl = calloc((size_t)threads, sizeof *sl);
    pthread_t *th = calloc((size_t)threads, sizeof *th);
    if (!jobs || !sl || !th) { free(jobs); free(sl); free(th); return 1; }
    for (int t = 0; t < threads; t++) {
        dd_job *J = &jobs[t];
        J->s_first = s_first; J->count = count;
        J->rx0 = rx0; J->rx1 = rx1; J->rz0 = rz0; J->rz1 = rz1;
        J->floor_max = floor_max; J->abort_below = abort_below;
        J->mu = &mu; J->shared_hits = hits; J->shared_n = &n; J->cap = cap;
        sl[t].job = J;
        sl[t].i0 = count * (uint64_t)t / (uint64_t)threads;
        sl[t].i1 = count * (uint64_t)(t + 1) / (uint64_t)threads;
        if (threads == 1) dd_run(J, sl[t].i0, sl[t].i1);
        else pthread_create(&th[t], NULL, dd_thread, &sl[t]);
    }
    memset(out, 0, sizeof *out);
    for (int t = 0; t < threads; t++) {
        if (threads > 1) pthread_join(th[t], NULL);
        const wo_dd_stats *s = &jobs[t].st;
        out->structure_seeds += s->structure_seeds;
        out->tests += s->tests;
        out->starts += s->starts;
        out->survivors += s->survivors;
        out->draws += s->draws;
        out->pieces += s->pieces;
        for (int k = 0; k < 3; k++) out->status[k] += s->status[k];
    }
    free(jobs); free(sl); free(th);
    *n_hits = n;
    return 0;
}
