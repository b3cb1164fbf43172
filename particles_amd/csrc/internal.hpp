// Internal types shared by the C-ABI layer (api.hip) and the integrator
// kernels (integrator.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptc {

struct MlpPacked;

enum Layout : int {
    LAYOUT_SOA = 0, LAYOUT_AOS = 1, LAYOUT_BRICK = 2, LAYOUT_PAIRS = 3, LAYOUT_QUADS = 4,
    LAYOUT_COLUMNS = 5
};
// the group kernel's row forms (its ROWS template argument)
enum : int { ROWS_NODES = 0, ROWS_PAIRS = 1, ROWS_QUADS = 2, ROWS_COLUMNS = 3 };

// Device copy of one field snapshot.
//   SOA: six planes of dm1*dm2*dm3 values (U/E x,y,z then B x,y,z), each x
//        fastest -- the reference's uu(:,:,:,c) / bb(:,:,:,c) planes.
//   AOS: one array of nodes, six values per node (U/E x,y,z, B x,y,z), so
//        the eight corners of a cell are eight 48-byte rows.
//   PAIRS: per node a row of 16 values (64 bytes in FP32, 128 in FP64):
//        the node's six, then those of its +x neighbour (periodic), then 4
//        zeros; a cell's eight corners are four aligned rows, one cache
//        line each in FP64 (the group kernel).
//   QUADS (FP32 only): per node a 128-byte row of 32 floats: the six values
//        of the node, of its +x, +y and +x+y neighbours (periodic), then 8
//        zeros; a cell's eight corners are two aligned 128-byte rows (the
//        group kernel).
//   COLUMNS: quad rows stored z fastest (row of node (i, j, k) at
//        (i + dm1 j) dm3 + k), so the two rows of a cell, planes k and
//        k + 1, are adjacent.  FP64: a 192-byte row of 24 doubles, the six
//        values of the node, of its +x, +y and +x+y neighbours (periodic),
//        value c of node a at 4 c + a, no padding; a cell's rows are one
//        384-byte span, three 128-byte lines when k is even, four when it
//        is odd.  FP32: the 128-byte FP32 quad row (24 floats, 8 zeros); a
//        cell's rows are two adjacent lines.
//   BRICK: AOS nodes plus, per cell, the eight corner rows of that cell in
//        the reference's corner order (48 values, 384 bytes, three 128-byte
//        lines) -- eight copies of every node, so a gather reads one
//        contiguous block.  Cells whose right index is not left + 1 (a face
//        case of the reference's separate floor) read the AOS nodes.
struct DeviceFields {
    int dm[3];
    int mode;     // PTC_MODE_VB / PTC_MODE_EB
    int prec;     // PTC_FP64 / PTC_FP32
    int layout;   // Layout
    int64_t nodes;
    void *data;   // SOA: 6 * nodes values; AOS/BRICK: nodes * 6 values
    void *bricks; // BRICK: nodes * 48 values (cell-major), else nullptr
    // time series (AOS only): nsnap snapshots of node rows stored one after
    // another, sstride values apart, at the strictly increasing times
    // stime[nsnap] (device memory).  nsnap = 0: one static snapshot.
    int nsnap;
    int64_t sstride;
    const double *stime;
};

// Geometry and controller constants of one integration (params vector of
// sources/particles.F90:317-333, plus derived values).
struct StepParams {
    double qom, vun, tmx, dtm, atol, rtol, fcmn, fcmx, safe, beta, expo;
    double xl[3], xs[3];
    double inv_dm[3];  // 1.0 / dm, computed like sources/interpolations.F90:141
    double dmd[3];     // dm as double
    int dm[3];
    int m;             // number of snapshots
    int64_t max_steps; // 0 = no cap
    double rxs[3], rvun;  // 1 / xs, 1 / vun
    int pow2;             // every xs and vun a power of two (divisions are exact scalings)
};

// Structure-of-arrays particle state resident in HBM.
struct DeviceParticles {
    int64_t np;
    double *x[6];        // current state si (position, momentum)
    double *t, *dt, *errold;
    int32_t *nout;       // next snapshot index (1-based, as in the reference)
    int32_t *flags;      // bit 0: last attempt rejected; bit 1: started
    int32_t *status;     // PTC_STATUS_*
    int64_t *cnt;        // 3 planes of np: all, accepted, rejected
    int64_t *nrhs;       // RHS evaluations per particle (diagnostics)
    double *out;         // state(8, m, np), Fortran order
    const double *time;  // time(m)
    const int32_t *order;// work-queue permutation (nullptr = identity)
    int32_t *qcount;     // [0] work-queue head of the lane-per-particle kernel,
                         // [1] of the group kernel, [2] particles the former finished
                         // or found retired (not running) when it fetched them
    int32_t park;        // hand every particle in flight to the group kernel once the
                         // queue is empty and at most this many are left (0 = never)
    unsigned long long *loads;  // cell loads since creation (each: the eight corner
                                // rows of one snapshot), added per wave at launch end;
                                // may be null
    // step-size controller: the reference's PI rule (CTRL_PI), the same with
    // sampled training records (CTRL_RECORD), or network proposals after
    // accepted steps with the PI rule after rejections (CTRL_LEARNED)
    const struct MlpPacked *net;  // CTRL_LEARNED
    float ctrl_shift;             // CTRL_LEARNED: added to the network's log factor
    uint64_t rec_thresh;          // CTRL_RECORD: sample an accepted step when hash < thresh
    float *pend;                  // CTRL_RECORD: np x 16 features awaiting their target
    double *pend_h;               // CTRL_RECORD: np, step after which they were taken
    int64_t *pend_rej;            // CTRL_RECORD: np, rejections at that point
    float *rec;                   // CTRL_RECORD: rec_cap x kRecord
    unsigned long long *rec_count;// CTRL_RECORD: records claimed (may exceed rec_cap)
    int64_t rec_cap;
};

enum : int32_t { FLAG_REJECTED = 1, FLAG_STARTED = 2, FLAG_PENDING = 4 };
enum : int { CTRL_PI = 0, CTRL_RECORD = 1, CTRL_LEARNED = 2 };

struct LaunchConfig {
    int block;
    int grid_blocks;   // 0 = occupancy-derived
    int cell_cache;    // keep the last cell's corner values per lane (AOS only): 0 off, 1 on
    int ctrl;          // CTRL_*; RECORD and LEARNED need the AOS layout with the cell cache
    int tail;          // FP64 static PI runs: 0 lane per particle only; N > 0 hand the
                       // particles in flight to the group kernel once the queue is empty
                       // and at most N are left; kTailGroupOnly: the group kernel only
    int lanes;         // lanes per particle of the group kernel for static fields with
                       // the PI controller: 3 (integrate_trio_kernel) or 6
                       // (integrate_group_kernel, also every series and learned run)
};
constexpr int kTailGroupOnly = -1;
constexpr int kTailAuto = -2;  // option value: chosen per batch (ptc_particles_advance)

// Launches the integrator on `stream`; returns a hipError_t.  A series of
// snapshots (f.nsnap > 0) takes the time-interpolating kernels (PI
// controller).
hipError_t launch_integrate(int method, const DeviceFields &f, const StepParams &sp,
                            const DeviceParticles &p, double t_stop, const LaunchConfig &lc,
                            hipStream_t stream);
// The same for FP32 fields (integrator_mixed.hip).
hipError_t launch_integrate_mixed(int method, const DeviceFields &f, const StepParams &sp,
                                  const DeviceParticles &p, double t_stop, const LaunchConfig &lc,
                                  hipStream_t stream);
hipError_t launch_probe_mixed(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                              const double *in, double *out, hipStream_t stream);
// Evaluates one building block per input on the device (ptc_probe).
hipError_t launch_probe(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                        const double *in, double *out, hipStream_t stream);
// Sums counters on the device: out[0..4] = all, accepted, rejected, rhs, running.
hipError_t launch_totals(const DeviceParticles &p, int64_t *d_out, hipStream_t stream);
// Line-gather diagnostic (diag.hip): `blocks` blocks of 256 threads, each
// group of kDiagLineLanes lanes reading one random aligned 128-byte line of
// buf (nlines lines) per load, kDiagInFlight loads in flight, `rounds`
// rounds.
constexpr int kDiagLineLanes = 8;
constexpr int kDiagInFlight = 8;
hipError_t launch_line_gather(const void *buf, uint64_t nlines, uint64_t seed, int rounds,
                              uint32_t *sink, unsigned blocks, hipStream_t stream);
// Cell-load diagnostic: pattern 0 four random 128-byte lines per cell load
// (nunits lines), pattern 1 one 384-byte span at a random 192-byte-aligned
// offset (nunits 192-byte rows, at least 2).
hipError_t launch_cell_gather(const void *buf, uint64_t nunits, int pattern, uint64_t seed,
                              int cells, uint32_t *sink, unsigned blocks, hipStream_t stream);
// Zeroes n int32 values (graph-safe replacement for hipMemsetAsync).
hipError_t launch_zero_i32(int32_t *ptr, int64_t n, hipStream_t stream);
hipError_t launch_zero_i64(int64_t *ptr, int64_t n, hipStream_t stream);
// Clears bit in n int32 flags.
hipError_t launch_clear_flag(int32_t *flags, int64_t n, int32_t bit, hipStream_t stream);
// SOA planes -> AOS nodes (and FP64 -> FP32 when prec is PTC_FP32).
hipError_t launch_to_aos(const double *soa, int64_t nodes, void *aos, int prec, hipStream_t stream);
hipError_t launch_to_pairs_f32(const double *soa, const int dm[3], float *out, hipStream_t stream);
hipError_t launch_to_pairs_f64(const double *soa, const int dm[3], double *out, hipStream_t stream);
hipError_t launch_to_quads_f32(const double *soa, const int dm[3], float *out, hipStream_t stream);
hipError_t launch_to_columns_f64(const double *soa, const int dm[3], double *out, hipStream_t stream);
hipError_t launch_to_columns_f32(const double *soa, const int dm[3], float *out, hipStream_t stream);
// values per node row of a stored layout (node rows: 6 FP64 or 8 FP32)
constexpr int row_values(int layout, int prec)
{
    return layout == LAYOUT_QUADS     ? 32
           : layout == LAYOUT_COLUMNS ? (prec == 1 ? 32 : 24)
           : layout == LAYOUT_PAIRS   ? 16
           : prec == 1                ? 8
                                      : 6;
}
// nodes per row: a pair row holds node n and its +x neighbour, a quad row
// node n and its +x, +y and +x+y neighbours; value c of the row's node a
// (a < row_nodes) sits at position row_nodes * c + a, so the values of one
// component are adjacent (one vector load per row and component), and the
// row's own node's value c at row_nodes * c
constexpr int row_nodes(int layout)
{
    return layout == LAYOUT_QUADS || layout == LAYOUT_COLUMNS ? 4 : layout == LAYOUT_PAIRS ? 2 : 1;
}
// the row of node n (x fastest) in a stored layout: n itself, except in
// column rows, which are stored z fastest
__host__ __device__ inline int64_t row_of(int layout, const int dm[3], int64_t n)
{
    if (layout != LAYOUT_COLUMNS) return n;
    const int64_t plane = (int64_t)dm[0] * dm[1];
    const int64_t k = n / plane, ij = n - k * plane;
    return ij * dm[2] + k;
}
// layouts whose rows only the group and trio kernels read (static fields)
constexpr bool rows_layout(int layout)
{
    return layout == LAYOUT_PAIRS || layout == LAYOUT_QUADS || layout == LAYOUT_COLUMNS;
}
// AOS nodes -> per-cell corner bricks.
hipError_t launch_to_bricks(const DeviceFields &f, hipStream_t stream);
// Rows of a static field in another layout (node, pair, quad or column rows) and
// precision from a stored static field (not bricks); out holds
// row_values(layout, prec) values per node.
hipError_t launch_convert_rows(const DeviceFields &f, void *out, int layout, int prec,
                               hipStream_t stream);
// Component comp (0..5: U or E x, y, z, then B x, y, z) of snapshot snap
// (0 for a static field) as a plane of nodes doubles, x fastest.
hipError_t launch_extract_plane(const DeviceFields &f, int snap, int comp, double *out,
                                hipStream_t stream);

// field synthesis (fields.hip): fill / add a constant, add a mode sum to three
// planes, add the reconnection sheet and guide field to B, and E of the
// reconnection recipe from B
hipError_t synth_fill(double *p, int64_t n, double v, hipStream_t s);
hipError_t synth_add(double *p, int64_t n, double v, hipStream_t s);
hipError_t synth_modes(const int dm[3], int nm, const int32_t *n, const double *amp,
                       const double *pol, const double *phase, double *out, hipStream_t s);
hipError_t synth_sheet(const int dm[3], double delta, double guide, double *b, hipStream_t s);
hipError_t synth_reconnection_e(const int dm[3], double delta, double vin, double eta,
                                const double *b, double *e, hipStream_t s);

// spatial work-queue order (order.hip): Morton order of the running
// particles' cells into `order`; scratch of order_scratch_bytes(np)
size_t order_scratch_bytes(int64_t np);
hipError_t launch_spatial_order(const DeviceParticles &S, const StepParams &P, int32_t *order,
                                void *scratch, size_t scratch_bytes, hipStream_t stream);
// most RHS evaluations since the last call first; prev (np) holds the
// counts of the last call and is updated
hipError_t launch_cost_order(const DeviceParticles &S, int64_t *prev, int32_t *order, void *scratch,
                             size_t scratch_bytes, hipStream_t stream);

// initial states and their moments (init.hip)
hipError_t launch_init(const DeviceParticles &S, int m, int kind, uint64_t seed, const double lo[3],
                       const double hi[3], double a, double b, hipStream_t stream);
hipError_t launch_moments(const DeviceParticles &S, int m, const double *shift, double *partials,
                          int blocks, hipStream_t stream);

// step-size network (mlp.hip; layouts in mlp_device.hpp)
constexpr int kMlpMaxGradBlocks = 512;
hipError_t mlp_pack(const float *params, MlpPacked *packed, hipStream_t stream);
// groups = 1: evaluated as the group kernel evaluates it (mlp_forward_groups)
hipError_t mlp_forward(const MlpPacked *packed, int64_t n, const float *feat, int ld, float *out,
                       hipStream_t stream, int groups = 0);
int mlp_grad_blocks(int64_t batch);
hipError_t mlp_gradient(const MlpPacked *packed, const float *rec, int64_t n, int64_t batch,
                        uint64_t seed, uint64_t step, float alpha, float lo, float hi,
                        float *partials, hipStream_t stream);
hipError_t mlp_reduce(const float *partials, int nblocks, float *grad_out, int apply, float lr,
                      float bc1, float bc2, float *params, float *m1, float *m2, MlpPacked *packed,
                      float *loss_out, hipStream_t stream);
hipError_t mlp_stats(const float *rec, int64_t n, double *partials, int blocks, hipStream_t stream);
hipError_t mlp_gemm_probe(const float *a, const float *bt, float *c, hipStream_t stream);

}  // namespace ptc
