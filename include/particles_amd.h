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
/* A second device copy of a static field in another precision and layout
 * (those the "layout" / "f32_layout" options name for prec), built on the
 * device from the stored values, so no host round trip or plane copy is
 * needed: the FP32 field of mixed precision from an FP64 field (values
 * rounded to float), or other row forms.  The same values as creating the
 * new form from the same arrays.  The source stays valid; destroy both.
 * No counterpart in the reference (one precision, fields.F90:33-35). */
int ptc_fields_convert(const ptc_fields *src, int prec, ptc_fields **out);
/* The stored field values back on the host, as the arrays ptc_fields_create
 * takes: f1 (U or E) and bb, each double (dm1, dm2, dm3, 3), x fastest
 * (the reference's uu / bb of sources/fields.F90:63-64).  FP32 storage
 * comes back widened to double.  snapshot selects one snapshot of a series
 * (0 for a static field).  No counterpart in the reference (whose fields
 * stay on the host); for checks and CPU baselines on fields built on the
 * device. */
int ptc_fields_download(const ptc_fields *f, int snapshot, double *f1, double *bb);
/* Replaces finalize_fields (sources/fields.F90:86-101). */
int ptc_fields_destroy(ptc_fields *f);

/* Synthesises a benchmark field on the device instead of uploading one (no
 * counterpart in the reference, which reads FITS files; for grids too large
 * to build on the host).  Node i sits at i / dm on the unit periodic box.
 * Mode tables: n (3, nb) integer wave vectors, amp (nb), pol (3, nb),
 * phase (nb); a table adds sum_m amp_m pol_m cos(2 pi n_m . x + phase_m).
 *   PTC_RECIPE_TURBULENT_VB   : VB mode; B = scalars[0] z + B table,
 *                               U = V table
 *   PTC_RECIPE_RECONNECTION_EB: EB mode; B = s(y) x + scalars[1] z + B table,
 *                               s(y) = tanh((y-1/4)/d) - tanh((y-3/4)/d) - 1,
 *                               d = scalars[0]; E = -V x B + scalars[3] curl B
 *                               (periodic central differences) with
 *                               V = (0, -scalars[2] s(y), 0); no V table.
 * prec as for ptc_fields_create. */
#define PTC_RECIPE_TURBULENT_VB    0
#define PTC_RECIPE_RECONNECTION_EB 1
int ptc_fields_synthesize(const int dm[3], int recipe, const double *scalars, int nb,
                          const int32_t *b_n, const double *b_amp, const double *b_pol,
                          const double *b_phase, int nv, const int32_t *v_n, const double *v_amp,
                          const double *v_pol, const double *v_phase, int prec, ptc_fields **out);

/* Time-dependent fields (no counterpart in the reference, whose field is
 * static: sources/fields.F90:51-78 loads one snapshot and the RHS has no
 * time argument, sources/accelerations.F90:59).  Builds a series from n
 * snapshot handles of equal dimensions, mode, precision and layout, in node
 * rows (layout aos, f32_layout nodes) or column rows (ptc_fields_create or
 * ptc_fields_synthesize), at strictly increasing
 * finite times in the integrator's time unit (that of params(3)).  Their
 * device arrays are copied into one block, so the snapshot handles may be
 * destroyed afterwards; destroy the series with ptc_fields_destroy.
 * A stage at time t sees f = f_k + a (f_{k+1} - f_k), the trilinear
 * interpolants of snapshots k and k+1 blended with a = (t - times[k]) /
 * (times[k+1] - times[k]), times[k] <= t < times[k+1]; at or before
 * times[0] snapshot 0 alone, at or after times[n-1] snapshot n-1 alone.
 * Stage i of a step from t of size dt is at t + c_i dt, c_i the method's
 * published nodes (Cash-Karp 1990; Hairer-Norsett-Wanner dop853).  A
 * series integrates with the PI step-size controller; with the cell cache
 * (option cell_cache 1) each lane keeps the two snapshots' corner values of
 * its last cell, the second in LDS (48 values per thread), where the device
 * gives a block that much LDS.  A series in column rows runs the six-lane
 * group kernel, which keeps both snapshots' corner values in registers.
 * n = 1 is a static field taken through the series kernels. */
int ptc_fields_series(const ptc_fields *const *snaps, const double *times, int n,
                      ptc_fields **out);

/* The whole particle loop of sources/particles.F90:337-356 in one call:
 * integrates every particle from t = 0 to tmax = params(3), writes the
 * snapshots into state and the step counts into counter, exactly like
 * integrate_rk5 / integrate_dp853 per particle.  status (np entries) may be
 * NULL.  Synchronous; host buffers stay the caller's.  Internally the run
 * advances from snapshot time to snapshot time (at most 64 launches) so the
 * work queue can be ordered by cost; results are those of one launch. */
int ptc_integrate(const ptc_fields *f, int method, int64_t np, int m, const double *params,
                  int nparams, const double *time, double *state, int64_t *counter,
                  int32_t *status);

/* ptc_integrate without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  time, state
 * and counter (and status, which may be NULL) are DEVICE pointers with the
 * layouts of ptc_integrate; the batch state lives in a fixed workspace
 * compiled into the library (7.3 MiB per device, held until the process
 * ends), so np is limited to 0..65536 (PTC_EARG beyond; no hipMalloc,
 * hipFree or allocating library call is made).  The work is enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) and the call returns
 * without waiting: results are complete once the caller synchronises that
 * stream.  Results equal ptc_integrate's bit for bit (PI controller; the
 * work queue in index order, which changes no result).  Every call on one
 * device shares that device's workspace: calls that overlap on one device
 * (other streams or threads before the previous call's stream is
 * synchronised) race and give undefined results. */
int ptc_integrate_noalloc(const ptc_fields *f, int method, int64_t np, int m,
                          const double *params, int nparams, const double *time, double *state,
                          int64_t *counter, int32_t *status, void *stream);

/* Device-resident particle batch: the same integration split into
 * segments, so callers can flush snapshots per interval and benchmark with
 * inputs already in HBM.  create uploads state(8,m,np) (only state(1:6,1,:)
 * is read, the rest is kept where no snapshot is written, as the reference's
 * intent(inout) state; state = NULL starts from zeros, for
 * ptc_particles_generate); advance integrates every particle until
 * t >= min(t_stop, tmax) (continuing a trajectory across segments gives the
 * same result as one uninterrupted call); download copies state, counter and
 * status (either may be NULL) back to the host. */
int ptc_particles_create(const ptc_fields *f, int method, int64_t np, int m, const double *params,
                         int nparams, const double *time, const double *state,
                         ptc_particles **out);
int ptc_particles_advance(ptc_particles *p, double t_stop);
int ptc_particles_download(ptc_particles *p, double *state, int64_t *counter, int32_t *status);
/* Initial states generated on the device (replaces the host generation of
 * sources/particles.F90:291-301 for large batches; no bitwise counterpart:
 * the reference draws from the compiler's random_number).  Philox4x32-10
 * per particle (counter = particle index and draw, key = seed).  Positions
 * uniform in [lo, hi]; momenta u = gamma v / c:
 *   PTC_INIT_MAXWELL   : each component normal(0, a) (the reference's
 *                        normal(0, vth) start)
 *   PTC_INIT_GAMMA     : gamma = a, isotropic direction
 *   PTC_INIT_GAMMA_LOG : gamma log-uniform in [a, b], isotropic direction
 * Only before the batch's first advance (PTC_ESTATE otherwise); overwrites
 * state(1:6,1,:) on the device (download returns it). */
#define PTC_INIT_MAXWELL   0
#define PTC_INIT_GAMMA     1
#define PTC_INIT_GAMMA_LOG 2
int ptc_particles_generate(ptc_particles *p, int kind, uint64_t seed, const double lo[3],
                           const double hi[3], double a, double b);
/* Mean, standard deviation, skewness and excess kurtosis of each of the six
 * components of state(1:6,1,:) (the statistics sources/particles.F90:303-312
 * prints, random.F90:231-353), reduced on the device: moments[4 c + k]. */
int ptc_particles_moments(ptc_particles *p, double moments[24]);
/* n advances in one HIP graph: the launches for the stop times t_stop[0..n-1]
 * (typically the snapshot times) are captured and submitted together; the
 * same result as n calls of ptc_particles_advance.  The work-queue order is
 * computed once per call, before the graph, and kept for all n intervals.
 * The graph is instantiated once per batch and kept: a later call with the
 * same n and the same launch settings (options, step-size controller, queue
 * order on or off) replays it with its stop-time arguments replaced
 * (kernel-node parameter update); any other call captures it anew.
 * ptc_particles_last_kernel_ms then reports the whole graph.  (The reference
 * has no counterpart: it integrates every particle through all snapshots
 * in one loop, sources/particles.F90:337-356.) */
int ptc_particles_advance_graph(ptc_particles *p, int n, const double *t_stop);
/* stats[0]: graphs instantiated by ptc_particles_advance_graph for this
 * batch; stats[1]: calls that replayed the kept graph with new stop times. */
int ptc_particles_graph_stats(const ptc_particles *p, int64_t stats[2]);
/* Autotuner: integrates the batch up to t_end in 1 + 9 equal segments, the
 * first to warm up, then one per candidate launch setting (work-queue order
 * by cost at 256 threads per block with one lane per particle, with the
 * group kernel of six or three lanes per particle alone and with the last
 * 32768 particles handed to it ("tail" and "group_lanes" options); cost
 * order at 128 / 64 threads per block, spatial order at 256, index order at
 * 256), and keeps for this batch the one with the most RHS evaluations per
 * kernel millisecond.  The segments are real progress: every setting gives
 * the same results.  chosen (may be NULL) = block, sort cadence, order
 * (1 spatial, 2 cost), tail (0 lane per particle, -1 group kernel alone,
 * N > 0 hand-over count), lanes per particle of the group kernel (3 or 6).
 * No counterpart in the reference (its OpenMP loop has no launch settings,
 * sources/particles.F90:337-356). */
int ptc_particles_autotune(ptc_particles *p, double t_end, int chosen[5]);
/* totals[0..4] = sum of all / accepted / rejected steps, number of RHS
 * evaluations, number of particles still running; reduced on the device. */
int ptc_particles_totals(ptc_particles *p, int64_t totals[5]);
/* Cell loads since the batch was created: each is one fetch of a cell's
 * eight corner rows of one field snapshot from device memory (384 bytes of
 * field values in FP64, 192 in FP32), made where an RHS evaluation
 * (acceleration -> interpolate_lin, sources/accelerations.F90:82) leaves
 * the cell its particle last sampled; evaluations in the same cell reuse
 * the values (the cell cache).  Counted per wave on the device.  No
 * counterpart in the reference (diagnostics: the bytes the gathers must
 * move). */
int ptc_particles_cell_loads(ptc_particles *p, int64_t *loads);
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
/*   PTC_PROBE_FEATURES : in (vec(6), h, err, pi_ratio, rejected)(10,n)
 *                        -> out (16,n) = the step-size network's features
 *                        at vec after an accepted step h with error ratio
 *                        err (no counterpart in the reference; aos layout) */
#define PTC_PROBE_FEATURES 5
/*   PTC_PROBE_GATHER_T : in (pos(3), t)(4,n) -> out (6,n) = (U or E, B) at
 *                        time t in a snapshot series (ptc_fields_series)
 *   PTC_PROBE_RHS_T    : in (vec(6), t)(7,n) -> out (6,n) = d(vec)/dt at t
 *                        (aos layout; a static handle is its own snapshot 0;
 *                        the other ops read snapshot 0 of a series) */
#define PTC_PROBE_GATHER_T 6
#define PTC_PROBE_RHS_T    7
int ptc_probe(const ptc_fields *f, int op, int64_t n, const double *params, int nparams,
              const double *in, double *out);

/* ---- learned step-size proposal (no counterpart in the reference) ----
 *
 * The reference sizes steps with a PI controller (sources/integrations.F90:
 * 177,186 for RK5, :466-490 for DOP853).  A small network can propose the
 * step after each accepted step instead, from features of the local field
 * and orbit (PTC_PROBE_FEATURES lists them); the embedded error estimate
 * still accepts or rejects every step and the PI rule still sizes the step
 * after a rejection.  Results then differ from the reference's (other step
 * sequence, same tolerance).
 *
 * Network: 16 -> 64 -> 64 -> 1, GELU, BF16 weights and activations with FP32
 * accumulation on the matrix cores; output = log(h_next / h).  Parameters are
 * PTC_MLP_PARAMS floats, packed as w1[16][64], b1[64], w2[64][64], b2[64],
 * w3[64], b3, mu[16], sig[16] (weights [in][out]; inputs are normalised as
 * (f - mu) / sig; mu and sig are not trained).
 *
 * Training records (PTC_MLP_RECORD floats each): 16 features taken after an
 * accepted step of size h, log(h_next / h) of the next accepted step, log of
 * that step's error ratio, the rejections before it, and the order q of the
 * error estimate (err ~ h^q: 5 for RK5, 1 / (0.125 - 0.2 beta) for DOP853).
 * The training target is log(h_next / h) + alpha log(1 / err) / q, clamped
 * to [-4, 4]: the accepted step itself (alpha = 0) or extrapolated to an
 * error ratio of one (alpha = 1). */
#define PTC_MLP_PARAMS 5345
#define PTC_MLP_TRAINABLE 5313
#define PTC_MLP_RECORD 20
typedef struct ptc_mlp ptc_mlp;
/* params = NULL: random initialisation from seed (mu = 0, sig = 1). */
int ptc_mlp_create(const float *params, uint64_t seed, ptc_mlp **out);
int ptc_mlp_destroy(ptc_mlp *net);
int ptc_mlp_get_params(const ptc_mlp *net, float *params);
int ptc_mlp_set_params(ptc_mlp *net, const float *params);
/* out[i] = network(features[i][0:16]), host arrays */
int ptc_mlp_forward(const ptc_mlp *net, int64_t n, const float *features, float *out);
/* The same outputs, evaluated as the six-lanes-per-particle integrator
 * evaluates the learned controller (ten samples per wavefront in one
 * 32-sample matrix tile); identical to ptc_mlp_forward.  No counterpart in
 * the reference (its step size is the PI rule only). */
int ptc_mlp_forward_groups(const ptc_mlp *net, int64_t n, const float *features, float *out);
/* gradient (PTC_MLP_TRAINABLE floats, packed order) and value of the loss
 * mean((out - target)^2) over n host records */
int ptc_mlp_gradient(const ptc_mlp *net, int64_t n, const float *records, float alpha,
                     float *grad, double *loss);
/* sets mu, sig to the mean and standard deviation (floored at 1e-3) of the
 * records' features; records in host (on_device = 0) or device memory */
int ptc_mlp_normalize(ptc_mlp *net, int64_t n, const float *records, int on_device);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of `batch`
 * records drawn with replacement from n (stream `seed`); losses (steps
 * floats, may be NULL) receives each minibatch's loss before its update */
int ptc_mlp_train(ptc_mlp *net, int64_t n, const float *records, int on_device, int steps,
                  int64_t batch, float lr, float alpha, uint64_t seed, float *losses);
/* c[64][64] = a[64][64] bt[64][64]^T through the device's BF16 wave product
 * (inputs rounded to BF16, FP32 accumulation), host arrays */
int ptc_mlp_probe_gemm(const float *a, const float *bt, float *c);

/* Record mode: the batch keeps the PI controller and, for a fraction `rate`
 * of its accepted steps (hash-sampled per particle and step), writes a
 * training record into a device buffer of `capacity` records (later records
 * are dropped).  capacity = 0 leaves record mode.  Needs the aos layout with
 * cell_cache 1. */
int ptc_particles_record(ptc_particles *p, double rate, int64_t capacity);
/* Number of records written (at most capacity); host_out (count x
 * PTC_MLP_RECORD) and device_records (the device buffer, valid until the
 * batch is destroyed or leaves record mode) may be NULL. */
int ptc_particles_records(const ptc_particles *p, int64_t *count, float *host_out,
                          const float **device_records);
/* Learned mode: after each accepted step h the next step is h exp(network +
 * shift), kept within [facmin h_PI, facmax h] (h_PI: the PI controller's
 * proposal, which stays the floor) and capped at dtmax; net = NULL returns to
 * the PI controller.  The network must outlive the batch's use of it and
 * must have been trained or given parameters (PTC_ESTATE otherwise).  Needs
 * the aos layout with cell_cache 1. */
int ptc_particles_use_mlp(ptc_particles *p, const ptc_mlp *net, double shift);

/* Tuning and behaviour switches (process-wide), e.g.
 *   "block"       threads per block (64..1024, multiple of 64), default 256
 *   "grid_blocks" blocks of the persistent kernel (0 = fill the device)
 *   "layout"      field layout in HBM for the next ptc_fields_create:
 *                 "soa" (the reference's six planes), "aos" (six
 *                 components interleaved per node), "brick" (aos plus
 *                 each cell's eight corners in one 384-byte block; 9x the
 *                 memory of "soa"), "pairs" (128-byte rows holding a
 *                 node and its +x neighbour, so a gather reads four
 *                 aligned cache lines; 2.67x the memory of "aos") or
 *                 "columns" (192-byte rows holding a node and its +x, +y
 *                 and +x+y neighbours, stored z fastest, so a gather
 *                 reads one 384-byte span, three or four cache lines; 4x
 *                 the memory of "aos", 192 GiB at 1024^3; the fastest
 *                 FP64 form measured on 256^3 to 1024^3 fields); FP64
 *                 fields only
 *   "f32_layout"  FP32 fields: "nodes" (default; 32-byte node rows),
 *                 "pairs" (64-byte rows holding a node and its +x
 *                 neighbour, so a gather reads four aligned 64-byte rows;
 *                 twice the memory), "quads" (128-byte rows holding a
 *                 node and its +x, +y and +x+y neighbours, so a gather
 *                 reads two aligned rows; four times the memory) or
 *                 "columns" (those rows stored z fastest, so the two rows
 *                 are adjacent).  Pair, quad and column rows (FP64 and
 *                 FP32) are read by the group or trio kernel only: every
 *                 "tail" setting then runs it, and probes and record mode
 *                 refuse them; snapshot series take node rows or column
 *                 rows, pair and quad rows not.  Results are identical to
 *                 node rows with the group kernel.
 *   "max_steps"   attempted-step cap per particle (0 = none)
 *   "cell_cache"  1 (default) keeps, per particle, the corner values of the
 *                 last cell it sampled (aos layout) in registers, so stage
 *                 points in the same cell skip the reload; 0 reloads every
 *                 time.  Results are identical either way.  Applies to
 *                 the one-lane-per-particle kernel; the group kernel
 *                 ("tail") always keeps its cell.
 *   "sort"        N > 0: before every N-th advance (default 1), re-order
 *                 the work queue ("order"); 0: index order
 *   "order"       "cost": most RHS evaluations in the previous
 *                 segment first, so the cheap particles fill the tail (the
 *                 first advance has no history and keeps index order);
 *                 "spatial": Morton order of the particles' cells
 *                 (neighbouring lanes gather from neighbouring memory);
 *                 "auto" (default): cost for batches of at most 2,000,000
 *                 particles, index order (no re-ordering) above.
 *                 Results are identical for every order.
 *   "tail"        node rows with the PI controller: which kernel
 *                 integrates.  "0": one lane per particle (all 64 lanes of
 *                 a wave carry particles, each lane all six components);
 *                 "group": the group kernel, six lanes per particle (one
 *                 component per lane; fewer instructions per step and two
 *                 or three waves per SIMD, so single trajectories and gathers from
 *                 large fields finish sooner); N > 0: one lane per particle
 *                 until the queue is empty and at most N particles are in
 *                 flight, then the group kernel finishes them; "auto"
 *                 (default): the group kernel for RK5, for batches of at
 *                 most 32768 particles, for fields over 2 GB and for FP64
 *                 snapshot series, else one lane per particle (FP64
 *                 static fields: handing the last 32768 over; a series
 *                 never hands over).  Mixed precision (FP32 fields) takes
 *                 the same kernels in FP32.  Results are identical for
 *                 every setting (FP64) or within FP32 rounding of each
 *                 other.  The learned controller (ptc_particles_use_mlp)
 *                 on FP32 fields follows "group" and "auto" but never
 *                 hands over (N > 0 runs one lane per particle); on FP64
 *                 fields, and in record mode, it runs one lane per
 *                 particle.
 *   "group_lanes" static fields with the PI controller, wherever a group
 *                 kernel runs ("tail" group or a hand-over, and pair, quad
 *                 and column rows): "6" the group kernel (one component per
 *                 lane, ten particles per wave); "3" the trio kernel
 *                 (lane a holds position a and momentum a, 21 particles
 *                 per wave: less than half the instructions per particle,
 *                 more particles in flight, each trajectory slower);
 *                 "auto" (default): three on column rows and for FP32
 *                 DOP853 batches of at least 2,000,000 particles, six
 *                 elsewhere.  Results are
 *                 identical for both (FP64 and FP32); snapshot series
 *                 follow the option too ("auto": six in FP64, three in
 *                 FP32); the
 *                 learned controller (FP32 fields) follows the option as
 *                 the PI controller does.
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

/* Diagnostic, no counterpart in the reference: the rate at which the
 * device serves random, aligned 128-byte lines from a buffer of `bytes`
 * (allocated and freed here): groups of eight lanes each read one line per
 * load (16 bytes a lane), eight independent loads in flight per lane, for
 * `rounds` rounds of loads, timed with HIP events after one untimed launch.
 * This is the access pattern of the integrator's cell loads from pair and
 * quad rows of fields far larger than the caches (rows of one line each at
 * unrelated addresses), so its rate bounds a gather-bound launch on them;
 * column rows put a cell's lines side by side and are served faster.
 * result = GB/s, ms, bytes read. */
int ptc_diag_line_gather(int64_t bytes, int rounds, double result[3]);
/* Diagnostic, no counterpart in the reference: the rate of cell loads as the
 * integrator's group kernels make them (ten groups of six lanes per wave,
 * three waves per SIMD, 16-byte loads, one cell load at a time per group)
 * from a buffer of `bytes` (allocated and freed here), `cells` loads per
 * group, timed with HIP events after one untimed launch.  pattern 0: four
 * random aligned 128-byte lines per load (FP64 pair rows); pattern 1: one
 * 384-byte span at a random 192-byte-aligned offset (FP64 column rows:
 * three or four lines).  result = cell loads per second, ms, cell loads. */
int ptc_diag_cell_gather(int64_t bytes, int pattern, int cells, double result[3]);

#ifdef __cplusplus
}
#endif
#endif
