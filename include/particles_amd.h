/*
 * particles_amd.h -- C-ABI of the MI355X test-particle integrator.
 *
 * Drop-in boundary for the integration path of gkowal/particles: the
 * reference's per-particle Fortran calls
 *
 *   call integrate_rk5  (m, dm, params, time, uu, bb, state(:,:,n), counter(:,n))
 *   call integrate_dp853(m, dm, params, time, uu, bb, state(:,:,n), counter(:,n))
 *
 * inside the particle loop at sources/particles.F90:337-356 (signatures at
 * sources/integrations.F90:57,66-72 and :224,233-239), and the field load of
 * sources/fields.F90:51-78, become batch calls on device-resident data.
 *
 * Plain C types only (pointers, sizes); every array argument is host memory
 * in the reference's Fortran layout:
 *   fields   uu/ee, bb : double (dm1, dm2, dm3, 3), x fastest
 *   params            : double params(nparams), nparams >= 17, 1-based
 *                       meaning of sources/particles.F90:317-333
 *   time              : double time(m)
 *   state             : double state(8, m, np); state(1:6,1,n) is the
 *                       initial position and momentum u = gamma v / c
 *   counter           : int64  counter(3, np) = (all, accepted, rejected)
 *
 * Every function returns PTC_OK (0) or a nonzero PTC_E* code and leaves a
 * message for ptc_last_error(); none calls exit().  One host thread per
 * handle; each particle batch owns one HIP stream.
 */
#ifndef PARTICLES_AMD_H
#define PARTICLES_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define PTC_OK        0
#define PTC_EARG      1  /* invalid argument (shape, pointer, parameter) */
#define PTC_EHIP      2  /* HIP runtime error */
#define PTC_ENOMEM    3  /* device allocation failed */
#define PTC_ESTATE    4  /* handle in the wrong state */

/* methods (sources/particles.F90:338 'rk5'; :347 'dop853','dp853','rk8') */
#define PTC_RK5       0
#define PTC_DOP853    1

/* field formulation */
#define PTC_MODE_VB   0  /* U and B stored; force q/m (v - U) x B (accelerations.F90:82-92) */
#define PTC_MODE_EB   1  /* E and B stored; force q/m (E + v x B) (no reference counterpart) */

/* field storage precision */
#define PTC_FP64      0
#define PTC_FP32      1

/* per-particle status (no counterpart in the reference, which loops
 * forever in these cases, sources/integrations.F90:129,361) */
#define PTC_STATUS_RUNNING   (-1) /* not yet at tmax (segmented runs) */
#define PTC_STATUS_DONE        0  /* reached tmax */
#define PTC_STATUS_MAXSTEPS    1  /* attempted-step cap reached */
#define PTC_STATUS_NONFINITE   2  /* error norm or state not finite */
#define PTC_STATUS_DTSMALL     3  /* step no longer advances t */

typedef struct ptc_fields ptc_fields;
typedef struct ptc_particles ptc_particles;

/* Replaces initialize_fields (sources/fields.F90:51-78): uploads one field
 * snapshot to HBM once; the handle owns the device copy.  `f1` is U (VB mode)
 * or E (EB mode), `bb` is B; unit scaling (fields.F90:73-74) is the caller's.
 * prec = PTC_FP32 stores the fields in float (mixed-precision mode). */
int ptc_fields_create(const int dm[3], const double *f1, const double *bb, int mode, int prec,
                      ptc_fields **out);
/* Replaces finalize_fields (sources/fields.F90:86-101). */
int ptc_fields_destroy(ptc_fields *f);

/* The whole particle loop of sources/particles.F90:337-356 in one call:
 * integrates every particle from t = 0 to tmax = params(3), writes the
 * snapshots into state and the step counts into counter, exactly like
 * integrate_rk5 / integrate_dp853 per particle.  status (np entries) may be
 * NULL.  Synchronous; host buffers stay the caller's. */
int ptc_integrate(const ptc_fields *f, int method, int64_t np, int m, const double *params,
                  int nparams, const double *time, double *state, int64_t *counter,
                  int32_t *status);

/* Device-resident particle batch: the same integration split into
 * segments, so callers can flush snapshots per interval and benchmark with
 * inputs already in HBM.  create uploads state(8,m,np) (only state(1:6,1,:)
 * is read, the rest is kept where no snapshot is written, as the reference's
 * intent(inout) state); advance integrates every particle until
 * t >= min(t_stop, tmax) (continuing a trajectory across segments gives the
 * same result as one uninterrupted call); download copies state, counter and
 * status (either may be NULL) back to the host. */
int ptc_particles_create(const ptc_fields *f, int method, int64_t np, int m, const double *params,
                         int nparams, const double *time, const double *state,
                         ptc_particles **out);
int ptc_particles_advance(ptc_particles *p, double t_stop);
int ptc_particles_download(ptc_particles *p, double *state, int64_t *counter, int32_t *status);
/* totals[0..4] = sum of all / accepted / rejected steps, number of RHS
 * evaluations, number of particles still running; reduced on the device. */
int ptc_particles_totals(ptc_particles *p, int64_t totals[5]);
/* Duration in ms of the integrator kernels of the last advance, measured
 * with HIP events on the batch's own stream. */
int ptc_particles_last_kernel_ms(const ptc_particles *p, double *ms);
int ptc_particles_destroy(ptc_particles *p);

/* Evaluates one building block of the device integrator for n inputs, so
 * each can be checked on its own against interpolate_lin
 * (sources/interpolations.F90:104), acceleration (sources/accelerations.F90:59)
 * and dti_guess (sources/integrations.F90:524):
 *   PTC_PROBE_GATHER : in pos(3,n)        -> out (6,n) = (U or E, B)
 *   PTC_PROBE_RHS    : in vec(6,n)        -> out (6,n) = d(vec)/dt
 *   PTC_PROBE_DTI5/8 : in vec(6,n)        -> out (n)   = initial step, order 5/8
 *   PTC_PROBE_POW    : in (x,y)(2,n)      -> out (n)   = x**y as the step
 *                                                       controller computes it
 * params as for ptc_integrate (geometry, units, tolerances); host arrays. */
#define PTC_PROBE_GATHER 0
#define PTC_PROBE_RHS    1
#define PTC_PROBE_DTI5   2
#define PTC_PROBE_DTI8   3
#define PTC_PROBE_POW    4
int ptc_probe(const ptc_fields *f, int op, int64_t n, const double *params, int nparams,
              const double *in, double *out);

/* Tuning and behaviour switches (process-wide), e.g.
 *   "block"       threads per block (64..1024, multiple of 64), default 256
 *   "grid_blocks" blocks of the persistent kernel (0 = fill the device)
 *   "layout"      field layout in HBM for the next ptc_fields_create:
 *                 "soa" (the reference's six planes), "aos" (six
 *                 components interleaved per node) or "brick" (aos plus
 *                 each cell's eight corners in one 384-byte block; 9x the
 *                 memory of "soa")
 *   "max_steps"   attempted-step cap per particle (0 = none)
 *   "cell_cache"  1 (default) keeps, per particle, the corner values of the
 *                 last cell it sampled (aos layout) in registers, so stage
 *                 points in the same cell skip the reload; 0 reloads every
 *                 time.  Results are identical either way.
 * Takes effect for the next create/advance. */
int ptc_set_option(const char *key, const char *value);
/* Reads an option back as text into buf (size len). */
int ptc_get_option(const char *key, char *buf, int len);

/* Message of the last failed call on this thread ("" if none). */
const char *ptc_last_error(void);

/* Library version string. */
const char *ptc_version(void);

/* x ** y computed on the host by the same code the device step controller
 * runs (PTC_PROBE_POW evaluates it on the device); it reproduces the host
 * libm's pow, which the reference's real ** real calls
 * (sources/integrations.F90:177,186,466,486,581). */
double ptc_pow_host(double x, double y);

#ifdef __cplusplus
}
#endif
#endif
