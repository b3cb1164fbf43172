/*
 * ptc_oracle.h -- CPU oracle for the particle-integration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference's per-particle algorithm (sources/interpolations.F90,
 * sources/accelerations.F90, sources/integrations.F90 of gkowal/particles),
 * kept in the reference's operation order so that, compiled without FP
 * contraction, it reproduces the reference bit for bit.  Only tests/, the
 * benchmark's CPU-baseline leg and the smoke check load it; the product
 * (particles_amd) never links or calls it.
 *
 * Conventions shared with the reference:
 *   - fields are Fortran column-major arrays f(dm1,dm2,dm3,3): x fastest,
 *     then y, z, component (sources/fields.F90:33-35,63-64);
 *   - params is the driver's 32-entry vector, entries 1..17 set
 *     (sources/particles.F90:317-333), here 0-based: params[0] = qom ...
 *     params[16] = zs;
 *   - state is state(8,m) per particle (sources/particles.F90:263).
 *
 * Field modes: PTC_MODE_VB is the reference's (velocity U and B stored,
 * force q/m (v-U) x B, sources/accelerations.F90:82-92).  PTC_MODE_EB stores
 * E in place of U and uses q/m (E + v x B); it has no reference counterpart.
 */
#ifndef PTC_ORACLE_H
#define PTC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* trilinear gather, sources/interpolations.F90:104-226 */
void orc_interpolate_lin(const int dm[3], const double xl[3], const double xs[3],
                         const double *uu, const double *bb, const double pos[3],
                         double u[3], double b[3]);

/* relativistic Lorentz-force RHS, sources/accelerations.F90:59-96 */
void orc_acceleration(int mode, double qom, double vun, const int dm[3],
                      const double xl[3], const double xs[3], const double *uu,
                      const double *bb, const double vec[6], double acc[6]);

/* initial step guess, sources/integrations.F90:524-589 */
double orc_dti_guess(int iord, int mode, const int dm[3], const double *params,
                     const double *uu, const double *bb, const double si[6]);

/* Cash-Karp RK5(4), sources/integrations.F90:57-200.
 * max_steps > 0 caps attempted steps (the reference has no cap); the return
 * value is a status code (0 = reached tmax, see PTC_STATUS_* below). */
int orc_integrate_rk5(int mode, int m, const int dm[3], const double *params,
                      const double *time, const double *uu, const double *bb,
                      double *state, int64_t cnt[3], int64_t max_steps);

/* DOP853, sources/integrations.F90:224-505 */
int orc_integrate_dp853(int mode, int m, const int dm[3], const double *params,
                        const double *time, const double *uu, const double *bb,
                        double *state, int64_t cnt[3], int64_t max_steps);

/* particle loop, sources/particles.F90:337-356 (OpenMP static schedule);
 * method 0 = rk5, 1 = dop853.  state is (8,m,np), cnt (3,np), status (np)
 * may be NULL.  nthreads <= 0 keeps the OpenMP default. */
void orc_integrate(int method, int mode, int64_t np, int m, const int dm[3],
                   const double *params, const double *time, const double *uu,
                   const double *bb, double *state, int64_t *cnt, int32_t *status,
                   int64_t max_steps, int nthreads);

/* ---- time-dependent fields (no counterpart in the reference, whose field
 * is static and whose RHS takes no time) ----
 * A series holds ns snapshots at strictly increasing times st; uu / bb are
 * ns consecutive (dm1, dm2, dm3, 3) arrays.  At time t the fields are
 * f_k + a (f_{k+1} - f_k) of the trilinear interpolants of snapshots k, k+1
 * with st[k] <= t < st[k+1], a = (t - st[k]) / (st[k+1] - st[k]); snapshot 0
 * alone at or before st[0], snapshot ns-1 alone at or after st[ns-1].
 * Stage i of a step from t of size dt is at t + c_i dt (orc_stage_nodes). */
void orc_interpolate_series(const int dm[3], const double xl[3], const double xs[3], int ns,
                            const double *st, const double *uu, const double *bb,
                            const double pos[3], double t, double u[3], double b[3]);
void orc_acceleration_series(int mode, double qom, double vun, const int dm[3],
                             const double xl[3], const double xs[3], int ns, const double *st,
                             const double *uu, const double *bb, const double vec[6], double t,
                             double acc[6]);
/* orc_integrate in a series (ns >= 1; the initial step guess sees t = 0) */
void orc_integrate_series(int method, int mode, int64_t np, int m, const int dm[3],
                          const double *params, const double *time, int ns, const double *st,
                          const double *uu, const double *bb, double *state, int64_t *cnt,
                          int32_t *status, int64_t max_steps, int nthreads);
/* the stage nodes c (6 for rk5, 12 for dop853) and the row sums of the
 * stage matrix the oracle holds; returns the number of stages */
int orc_stage_nodes(int method, double *c, double *rowsum);

int orc_max_threads(void);

/* status codes shared with the device integrator */
#define PTC_STATUS_DONE      0 /* reached tmax */
#define PTC_STATUS_MAXSTEPS  1 /* attempted-step cap reached */
#define PTC_STATUS_NONFINITE 2 /* error norm or state not finite */
#define PTC_STATUS_DTSMALL   3 /* step no longer advances t */

#define PTC_MODE_VB 0
#define PTC_MODE_EB 1

#ifdef __cplusplus
}
#endif
#endif
