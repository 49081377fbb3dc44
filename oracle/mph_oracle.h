/*
 * CPU oracle for the MPH-WC explicit step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's per-step algorithm
 * (source/main.cpp:836-2317), used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg as the checker.  The product library never links or calls it.
 *
 * Every arithmetic expression keeps the reference's operation order and is compiled
 * without FP contraction, so on the same inputs the oracle reproduces the reference
 * bit for bit (checked against the reference's own source built under oracle/_ref).
 */
#ifndef MPH_ORACLE_H
#define MPH_ORACLE_H

#include "../include/mph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state *orc_create(const mph_params *p, int n, const int *type, const double *x,
                      const double *v, int nthreads);
int orc_step(orc_state *s, int nsteps);          /* returns number of rebuilds */
int orc_last_rebuild(const orc_state *s);
int orc_count(const orc_state *s);
double orc_time(const orc_state *s);
long orc_neighbor_total(const orc_state *s);
void orc_scalars(const orc_state *s, double *out /* [16], same slots as mphref_scalars */);
void orc_get(const orc_state *s, int *index, int *type, double *x, double *v, double *force,
             double *dens_a, double *vol_strain, double *divergence, double *press_p,
             double *press_a, double *gcenter, int *ncount, double *mass);
void orc_neighbors(const orc_state *s, long *ptr, int *ind);
void orc_virial(orc_state *s, double *vp);
/* VolStrainP with long double weights and sums (not part of the restatement: the exact
 * value that the tests measure both codes' FP64 rounding against) */
void orc_vol_strain_exact(const orc_state *s, double *out);
void orc_destroy(orc_state *s);

#ifdef __cplusplus
}
#endif
#endif
