/* whf_oracle_int.h — oracle building blocks shared between its source files (RNG forks, Perlin construction and
 * the corner blend). TEST INFRASTRUCTURE ONLY. */
#ifndef WHF_ORACLE_INT_H
#define WHF_ORACLE_INT_H

#include "whf_oracle.h"

void   wo_xr_set(wo_xr *s, uint64_t lo, uint64_t hi);                           /* rng.c:88-95 */
void   wo_xr_fork(wo_xr *out, wo_xr *parent);                                   /* rng.c:143-147 */
void   wo_xr_fork_named(wo_xr *out, const wo_xr *deriver, const char *name);    /* rng.c:157-170 */
void   wo_perlin_build(wo_perlin *p, wo_xr *r);                                 /* noise.c:61-78 */
/* gradient corner blend at lattice cell (i, j, k) with offsets (x, y, z) and the y fade taken at ty (noise.c:224-244) */
double wo_perlin_core(const wo_perlin *p, int i, int j, int k, double x, double y, double z, double ty);
double wo_keep_precision(double v);                                             /* noise.c:28-34 */

/* Fluid statuses of aquifer anchors, memoized by grid cell, with the chunk's skipSamplingAboveY. */
#define WO_AQ_SLOTS 512
typedef struct {
    int valid, level, type;
    int gx, gy, gz;
} wo_aq_slot;
typedef struct {
    int skip_above;
    wo_aq_slot st[WO_AQ_SLOTS];
} wo_aq_memo;

/* The terrain of one chunk: its 5 x 5 corner columns on first use and its aquifer (phase2.c:108-130, 303-316). */
typedef struct {
    const wo_phase2 *p;
    int cx, cz;
    uint32_t have;
    double main[25][49];
    double noodle[25][3][49];
    wo_aq_memo aq;
} wo_p2_chunk;

void   wo_p2_chunk_init(wo_p2_chunk *ch, const wo_phase2 *p, int cx, int cz);
double wo_p2_chunk_density(wo_p2_chunk *ch, int x, int y, int z);
int    wo_p2_chunk_substance(wo_p2_chunk *ch, int x, int y, int z, double density);
int    wo_p2_chunk_state(wo_p2_chunk *ch, int x, int y, int z);
int    wo_p2_chunk_height(wo_p2_chunk *ch, int x, int z);

#endif
