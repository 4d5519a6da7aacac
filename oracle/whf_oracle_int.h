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

#endif
