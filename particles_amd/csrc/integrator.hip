// Integrator kernels for gfx950 (MI355X).
//
// One particle per lane in a persistent kernel: every lane integrates a
// particle until it reaches the end of the segment, hands the state back to
// HBM and pulls the next particle index from a device work queue (one
// wave-aggregated atomic per refill), so particles whose step counts differ
// by decades share waves without idling lanes.  All stage arithmetic of an
// attempt is the same instruction stream for accepted and rejected lanes;
// only bookkeeping diverges.
//
// The arithmetic restates the reference per particle (gkowal/particles,
// sources/integrations.F90, accelerations.F90, interpolations.F90) in its
// operation order; this file is compiled with -ffp-contract=off so every
// FP64 add/mul/div/sqrt rounds like the reference's own build, and the
// step-size powers use ptc_libm_pow (libm_pow.hpp), which reproduces the
// host libm's pow bit for bit.  Trajectories and step counters then match
// the reference bit for bit.
//
// Redundant evaluations of the reference are not repeated, which leaves every
// result unchanged:
//   * RK5 recomputes k6 = dt f(s6) on accept (integrations.F90:156) -- the
//     stored k6 is identical;
//   * DOP853 recomputes f(si) (:445, equal to stage 1) and f(sf) (:446, equal
//     to the next step's stage 1) -- f(sf) is evaluated once and reused;
//   * dti_guess evaluates f at a trial point and never uses it (:576);
//   * a rejected attempt reuses f(si) (stage 1 does not depend on dt).
#include "internal.hpp"
#include "libm_pow.hpp"

#include "particles_amd.h"

#include <type_traits>

namespace ptc {
namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// trilinear gather, sources/interpolations.F90:104-226
// ---------------------------------------------------------------------------

// i == n wraps to 0 (see corners()); anything else outside [0, n) can only
// come from a non-finite position and is clamped to 0 so the gather stays in
// bounds (the non-finite result then retires the particle).
__device__ __forceinline__ int wrap_idx(int i, int n)
{
    i = i >= n ? i - n : i;
    return (unsigned)i < (unsigned)n ? i : 0;
}

template <class W>
struct CornersT {
    uint32_t o[8];  // node offsets of the 8 corners, reference order
    W w[8];         // trilinear weights, reference order
    bool adjacent;  // right index == left + 1 (periodic) in every direction
};
using Corners = CornersT<double>;

// Index and weight math of interpolations.F90:140-149 and :207-210.  The
// right index is its own floor of x + 1/dm, as in the reference, so at a
// cell face it can differ from left+1; an index equal to dm (x just below a
// periodic image, where the reference reads one past the array) wraps to 0.
// W = float (mixed precision): the in-cell fractions are formed in FP64 as
// above and rounded once; the weight products are FP32.
template <class W>
__device__ __forceinline__ void corners(const StepParams &P, const double *pos, CornersT<W> &c)
{
    int l[3], r[3];
    double xd[3], xu[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double x = (pos[i] - P.xl[i]) / P.xs[i];
        const double xq = x + P.inv_dm[i];
        double xp = P.dmd[i] * (xq - floor(xq));
        r[i] = wrap_idx((int)floor(xp), P.dm[i]);
        xp = P.dmd[i] * (x - floor(x));
        const double pl = floor(xp);
        xd[i] = xp - pl;
        xu[i] = 1.0 - xd[i];
        l[i] = wrap_idx((int)pl, P.dm[i]);
    }
    c.adjacent = true;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int lp = l[i] + 1 == P.dm[i] ? 0 : l[i] + 1;
        c.adjacent = c.adjacent && r[i] == lp;
    }
    const uint32_t sy = (uint32_t)P.dm[0], sz = (uint32_t)P.dm[0] * (uint32_t)P.dm[1];
    const uint32_t yl = (uint32_t)l[1] * sy, yr = (uint32_t)r[1] * sy;
    const uint32_t zl = (uint32_t)l[2] * sz, zr = (uint32_t)r[2] * sz;
    const uint32_t xl = (uint32_t)l[0], xr = (uint32_t)r[0];
    c.o[0] = xl + yl + zl; c.o[1] = xl + yl + zr; c.o[2] = xl + yr + zl; c.o[3] = xl + yr + zr;
    c.o[4] = xr + yl + zl; c.o[5] = xr + yl + zr; c.o[6] = xr + yr + zl; c.o[7] = xr + yr + zr;
    const W u0 = W(xu[0]), u1 = W(xu[1]), u2 = W(xu[2]);
    const W d0 = W(xd[0]), d1 = W(xd[1]), d2 = W(xd[2]);
    c.w[0] = u0 * u1 * u2;
    c.w[1] = u0 * u1 * d2;
    c.w[2] = u0 * d1 * u2;
    c.w[3] = u0 * d1 * d2;
    c.w[4] = d0 * u1 * u2;
    c.w[5] = d0 * u1 * d2;
    c.w[6] = d0 * d1 * u2;
    c.w[7] = d0 * d1 * d2;
}

// f = w0 v0 + w1 v1 + ... + w7 v7, summed left to right per component
template <class T>
__device__ __forceinline__ void accumulate_row(int j, T w, const T *v, T *f)
{
    if (j == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = w * v[k];
    } else {
#pragma unroll
        for (int k = 0; k < 6; k++) f[k] = f[k] + w * v[k];
    }
}

// eight 48-byte node rows of the AOS array
__device__ __forceinline__ void gather_rows(const double *__restrict__ d, const Corners &c, double *f)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const double2 *__restrict__ q = (const double2 *)(d + (size_t)c.o[j] * 6);
        const double2 a = q[0], b = q[1], e = q[2];
        const double v[6] = {a.x, a.y, b.x, b.y, e.x, e.y};
        accumulate_row(j, c.w[j], v, f);
    }
}

// f[0..2] = U (or E), f[3..5] = B; each a left-to-right weighted sum of the
// eight corners (interpolations.F90:212-220).
template <int LAYOUT>
__device__ __forceinline__ void gather(const DeviceFields &F, const StepParams &P, const double *pos,
                                       double *f)
{
    Corners c;
    corners(P, pos, c);
    const double *__restrict__ d = (const double *)F.data;
    if constexpr (LAYOUT == LAYOUT_SOA) {
        const int64_t N = F.nodes;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const double *__restrict__ p = d + k * N;
            double s = c.w[0] * p[c.o[0]];
#pragma unroll
            for (int j = 1; j < 8; j++) s = s + c.w[j] * p[c.o[j]];
            f[k] = s;
        }
    } else if constexpr (LAYOUT == LAYOUT_AOS) {
        gather_rows(d, c, f);
    } else {
        if (c.adjacent) {
            // the cell's own 384-byte block: row j holds corner j
            const double2 *__restrict__ q =
                (const double2 *)((const double *)F.bricks + (size_t)c.o[0] * 48);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const double2 a = q[3 * j], b = q[3 * j + 1], e = q[3 * j + 2];
                const double v[6] = {a.x, a.y, b.x, b.y, e.x, e.y};
                accumulate_row(j, c.w[j], v, f);
            }
        } else {
            gather_rows(d, c, f);
        }
    }
}

// ---------------------------------------------------------------------------
// relativistic Lorentz force, sources/accelerations.F90:59-96
// ---------------------------------------------------------------------------
template <int MODE>
__device__ __forceinline__ void force_from_fields(const StepParams &P, const double *s,
                                                  const double *f, double *a)
{
    const double g = sqrt(s[3] * s[3] + s[4] * s[4] + s[5] * s[5] + 1.0);
    const double v0 = s[3] / g, v1 = s[4] / g, v2 = s[5] / g;
    a[0] = v0 / P.vun;
    a[1] = v1 / P.vun;
    a[2] = v2 / P.vun;
    if constexpr (MODE == PTC_MODE_EB) {
        a[3] = P.qom * (f[0] + (v1 * f[5] - v2 * f[4]));
        a[4] = P.qom * (f[1] + (v2 * f[3] - v0 * f[5]));
        a[5] = P.qom * (f[2] + (v0 * f[4] - v1 * f[3]));
    } else {
        const double w0 = v0 - f[0], w1 = v1 - f[1], w2 = v2 - f[2];
        a[3] = P.qom * (w1 * f[5] - w2 * f[4]);
        a[4] = P.qom * (w2 * f[3] - w0 * f[5]);
        a[5] = P.qom * (w0 * f[4] - w1 * f[3]);
    }
}

// Mixed precision: the same force in FP32 from the FP64 stage point.
template <int MODE>
__device__ __forceinline__ void force_from_fields_f32(const StepParams &P, const double *s,
                                                      const float *f, float *a)
{
    const float ux = float(s[3]), uy = float(s[4]), uz = float(s[5]);
    const float g = sqrtf(ux * ux + uy * uy + uz * uz + 1.0f);
    const float v0 = ux / g, v1 = uy / g, v2 = uz / g;
    const float vun = float(P.vun), qom = float(P.qom);
    a[0] = v0 / vun;
    a[1] = v1 / vun;
    a[2] = v2 / vun;
    if constexpr (MODE == PTC_MODE_EB) {
        a[3] = qom * (f[0] + (v1 * f[5] - v2 * f[4]));
        a[4] = qom * (f[1] + (v2 * f[3] - v0 * f[5]));
        a[5] = qom * (f[2] + (v0 * f[4] - v1 * f[3]));
    } else {
        const float w0 = v0 - f[0], w1 = v1 - f[1], w2 = v2 - f[2];
        a[3] = qom * (w1 * f[5] - w2 * f[4]);
        a[4] = qom * (w2 * f[3] - w0 * f[5]);
        a[5] = qom * (w0 * f[4] - w1 * f[3]);
    }
}

// FP32 nodes: eight floats (U or E x,y,z, B x,y,z, |B|, 0), 32 bytes, so a
// corner is one 16-byte and one 8-byte load
__device__ __forceinline__ void load_node_f32(const float *__restrict__ d, uint32_t o, float *v)
{
    const float4 a = *(const float4 *)(d + (size_t)o * 8);
    const float2 b = *(const float2 *)(d + (size_t)o * 8 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
}

template <int MODE>
__device__ __forceinline__ void rhs_f32(const DeviceFields &F, const StepParams &P, const double *s,
                                        float *a)
{
    CornersT<float> c;
    corners(P, s, c);
    const float *__restrict__ d = (const float *)F.data;
    float f[6];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float v[6];
        load_node_f32(d, c.o[j], v);
        accumulate_row(j, c.w[j], v, f);
    }
    force_from_fields_f32<MODE>(P, s, f, a);
}

template <int MODE, int LAYOUT>
__device__ __forceinline__ void rhs(const DeviceFields &F, const StepParams &P, const double *s,
                                    double *a)
{
    double f[6];
    gather<LAYOUT>(F, P, s, f);
    force_from_fields<MODE>(P, s, f, a);
}

// The RHS as the integrator calls it.  With a cell cache, each lane keeps the
// 48 corner values of the last cell it gathered from (the node rows are
// identified by the offsets of corners 0 and 7, which fix both the left and
// the right index triples); a stage point in the same cell re-weights those
// values instead of loading them again.  The values and the summation order
// are the same either way, so the result is bit-identical.  CACHE 0 reloads
// every time (SOA and brick layouts).
// PREC = PTC_FP32 is the mixed-precision force: FP32 nodes, weights and
// arithmetic, slopes returned as float (K = float).
template <int MODE, int LAYOUT, int CACHE, int PREC>
struct Force {
    using K = std::conditional_t<PREC == PTC_FP32, float, double>;
    const DeviceFields &F;
    const StepParams &P;
    uint32_t key0 = 0xffffffffu, key7 = 0xffffffffu;
    K v[CACHE == 1 ? 48 : 1];

    __device__ __forceinline__ Force(const DeviceFields &F_, const StepParams &P_) : F(F_), P(P_) {}

    __device__ __forceinline__ void operator()(const double *s, K *a)
    {
        if constexpr (PREC == PTC_FP32) {
            static_assert(LAYOUT == LAYOUT_AOS, "FP32 fields are stored as AOS nodes");
            if constexpr (CACHE == 0) {
                rhs_f32<MODE>(F, P, s, a);
            } else {
                CornersT<float> c;
                corners(P, s, c);
                if (c.o[0] != key0 || c.o[7] != key7) {
                    const float *__restrict__ d = (const float *)F.data;
#pragma unroll
                    for (int j = 0; j < 8; j++) load_node_f32(d, c.o[j], &v[6 * j]);
                    key0 = c.o[0];
                    key7 = c.o[7];
                }
                float f[6];
#pragma unroll
                for (int j = 0; j < 8; j++) accumulate_row(j, c.w[j], &v[6 * j], f);
                force_from_fields_f32<MODE>(P, s, f, a);
            }
        } else if constexpr (CACHE == 0) {
            rhs<MODE, LAYOUT>(F, P, s, a);
        } else {
            static_assert(LAYOUT == LAYOUT_AOS, "the cell cache reads AOS node rows");
            Corners c;
            corners(P, s, c);
            if (c.o[0] != key0 || c.o[7] != key7) {
                const double *__restrict__ d = (const double *)F.data;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const double2 *__restrict__ q = (const double2 *)(d + (size_t)c.o[j] * 6);
                    const double2 x0 = q[0], x1 = q[1], x2 = q[2];
                    v[6 * j + 0] = x0.x; v[6 * j + 1] = x0.y; v[6 * j + 2] = x1.x;
                    v[6 * j + 3] = x1.y; v[6 * j + 4] = x2.x; v[6 * j + 5] = x2.y;
                }
                key0 = c.o[0];
                key7 = c.o[7];
            }
            double f[6];
#pragma unroll
            for (int j = 0; j < 8; j++) accumulate_row(j, c.w[j], &v[6 * j], f);
            force_from_fields<MODE>(P, s, f, a);
        }
    }
};

// ---------------------------------------------------------------------------
// initial step, sources/integrations.F90:524-589 (ki = f(si) is given)
// ---------------------------------------------------------------------------
template <class K>
__device__ __forceinline__ double dti_guess(const StepParams &P, int iord, const double *si,
                                            const K *ki, const PowTables &T)
{
    double sr[6], s, q, d0, d1, d2, dti, dtf;
#pragma unroll
    for (int i = 0; i < 6; i++) sr[i] = P.atol + P.rtol * fabs(si[i]);
    s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) { q = si[i] / sr[i]; s = s + q * q; }
    d0 = sqrt(s / 6.0);
    s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) { q = double(ki[i]) / sr[i]; s = s + q * q; }
    d1 = sqrt(s / 6.0);
    dti = (fmin(d0, d1) < 1.0e-10) ? 1.0e-06 : 1.0e-01 * (d0 / d1);
    s = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double sf = si[i] + dti * double(ki[i]);
        q = (sf - si[i]) / sr[i];
        s = s + q * q;
    }
    d2 = sqrt(s / 6.0) / dti;
    if (d1 > d2) d2 = d1;
    if (d2 <= 1.0e-15) {
        dtf = 1.0e-03 * dti;
        if (dtf < 1.0e-06) dtf = 1.0e-06;
    } else {
        dtf = ptc_pow_with(1.0e-02 / d2, 1.0 / (double)iord, T.log, T.exp);
    }
    dti = 1.0e+02 * dti;
    if (dtf < dti) dti = dtf;
    if (P.dtm < dti) dti = P.dtm;
    return dti;
}

// ---------------------------------------------------------------------------
// Cash-Karp RK5(4) attempt, sources/integrations.F90:131-148
// ---------------------------------------------------------------------------
// Published Cash-Karp (1990) rationals; the reference holds the same values
// over common denominators (integrations.F90:86-98).
#define CK_A21 (1.0 / 5.0)
#define CK_A31 (3.0 / 40.0)
#define CK_A32 (9.0 / 40.0)
#define CK_A41 (3.0 / 10.0)
#define CK_A42 (-9.0 / 10.0)
#define CK_A43 (6.0 / 5.0)
#define CK_A51 (-11.0 / 54.0)
#define CK_A52 (5.0 / 2.0)
#define CK_A53 (-70.0 / 27.0)
#define CK_A54 (35.0 / 27.0)
#define CK_A61 (1631.0 / 55296.0)
#define CK_A62 (175.0 / 512.0)
#define CK_A63 (575.0 / 13824.0)
#define CK_A64 (44275.0 / 110592.0)
#define CK_A65 (253.0 / 4096.0)
#define CK_C1 (37.0 / 378.0)
#define CK_C3 (250.0 / 621.0)
#define CK_C4 (125.0 / 594.0)
#define CK_C6 (512.0 / 1771.0)
#define CK_D1 (-277.0 / 64512.0)
#define CK_D3 (6925.0 / 370944.0)
#define CK_D4 (-6925.0 / 202752.0)
#define CK_D5 (-277.0 / 14336.0)
#define CK_D6 (277.0 / 7084.0)

// k1 = dt f(si) from the cached derivative f1; returns del (error / atol),
// sf, and the stage-6 slope k6 used by the dense output.  Stage slopes are
// kept in the RHS's type K (double, or float in mixed precision); stage
// points, sums and the error norm are always FP64.
template <class Rhs>
__device__ __forceinline__ double attempt_rk5(Rhs &rhsf, const StepParams &P, const double *si,
                                              const typename Rhs::K *f1, double dt, double *sf,
                                              double *k6d)
{
    using K = typename Rhs::K;
    K k1[6], k2[6], k3[6], k4[6], k5[6], k6[6], a[6];
    double s[6];
#pragma unroll
    for (int c = 0; c < 6; c++) k1[c] = K(dt * double(f1[c]));
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = si[c] + CK_A21 * double(k1[c]);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k2[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++) s[c] = si[c] + CK_A31 * double(k1[c]) + CK_A32 * double(k2[c]);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k3[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++)
        s[c] = si[c] + CK_A41 * double(k1[c]) + CK_A42 * double(k2[c]) + CK_A43 * double(k3[c]);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k4[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++)
        s[c] = si[c] + CK_A51 * double(k1[c]) + CK_A52 * double(k2[c]) + CK_A53 * double(k3[c]) +
               CK_A54 * double(k4[c]);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k5[c] = K(dt * double(a[c]));
#pragma unroll
    for (int c = 0; c < 6; c++)
        s[c] = si[c] + CK_A61 * double(k1[c]) + CK_A62 * double(k2[c]) + CK_A63 * double(k3[c]) +
               CK_A64 * double(k4[c]) + CK_A65 * double(k5[c]);
    rhsf(s, a);
#pragma unroll
    for (int c = 0; c < 6; c++) k6[c] = K(dt * double(a[c]));

    const double eps = 2.2250738585072014e-308;  // tiny(1d0), integrations.F90:99
    double del = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        // sum(c(1:6) k(1:6,i)) and sum(d(1:6) k(1:6,i)), zero weights kept
        double sc = 0.0;
        sc = sc + CK_C1 * double(k1[c]);
        sc = sc + 0.0 * double(k2[c]);
        sc = sc + CK_C3 * double(k3[c]);
        sc = sc + CK_C4 * double(k4[c]);
        sc = sc + 0.0 * double(k5[c]);
        sc = sc + CK_C6 * double(k6[c]);
        sf[c] = si[c] + sc;
        double se = 0.0;
        se = se + CK_D1 * double(k1[c]);
        se = se + 0.0 * double(k2[c]);
        se = se + CK_D3 * double(k3[c]);
        se = se + CK_D4 * double(k4[c]);
        se = se + CK_D5 * double(k5[c]);
        se = se + CK_D6 * double(k6[c]);
        const double q = fabs(se / (fabs(si[c]) + fabs(double(k1[c])) + eps));
        if (c == 0 || q > del) del = q;
        k6d[c] = double(k6[c]);
    }
    return del / P.atol;
}

// ---------------------------------------------------------------------------
// DOP853 attempt, sources/integrations.F90:365-437 (Hairer-Norsett-Wanner
// coefficients, public dop853.f; same published values as the reference)
// ---------------------------------------------------------------------------
#define A0201 5.26001519587677318785587544488e-02
#define A0301 1.97250569845378994544595329183e-02
#define A0302 5.91751709536136983633785987549e-02
#define A0401 2.95875854768068491816892993775e-02
#define A0403 8.87627564304205475450678981324e-02
#define A0501 2.41365134159266685502369798665e-01
#define A0503 -8.84549479328286085344864962717e-01
#define A0504 9.24834003261792003115737966543e-01
#define A0601 3.70370370370370370370370370370e-02
#define A0604 1.70828608729473871279604482173e-01
#define A0605 1.25467687566822425016691814123e-01
#define A0701 3.71093750000000000000000000000e-02
#define A0704 1.70252211019544039314978060272e-01
#define A0705 6.02165389804559606850219397283e-02
#define A0706 -1.75781250000000000000000000000e-02
#define A0801 3.70920001185047927108779319836e-02
#define A0804 1.70383925712239993810214054705e-01
#define A0805 1.07262030446373284651809199168e-01
#define A0806 -1.53194377486244017527936158236e-02
#define A0807 8.27378916381402288758473766002e-03
#define A0901 6.24110958716075717114429577812e-01
#define A0904 -3.36089262944694129406857109825e+00
#define A0905 -8.68219346841726006818189891453e-01
#define A0906 2.75920996994467083049415600797e+01
#define A0907 2.01540675504778934086186788979e+01
#define A0908 -4.34898841810699588477366255144e+01
#define A1001 4.77662536438264365890433908527e-01
#define A1004 -2.48811461997166764192642586468e+00
#define A1005 -5.90290826836842996371446475743e-01
#define A1006 2.12300514481811942347288949897e+01
#define A1007 1.52792336328824235832596922938e+01
#define A1008 -3.32882109689848629194453265587e+01
#define A1009 -2.03312017085086261358222928593e-02
#define A1101 -9.37142430085987325717040216580e-01
#define A1104 5.18637242884406370830023853209e+00
#define A1105 1.09143734899672957818500254654e+00
#define A1106 -8.14978701074692612513997267357e+00
#define A1107 -1.85200656599969598641566180701e+01
#define A1108 2.27394870993505042818970056734e+01
#define A1109 2.49360555267965238987089396762e+00
#define A1110 -3.04676447189821950038236690220e+00
#define A1201 2.27331014751653820792359768449e+00
#define A1204 -1.05344954667372501984066689879e+01
#define A1205 -2.00087205822486249909675718444e+00
#define A1206 -1.79589318631187989172765950534e+01
#define A1207 2.79488845294199600508499808837e+01
#define A1208 -2.85899827713502369474065508674e+00
#define A1209 -8.87285693353062954433549289258e+00
#define A1210 1.23605671757943030647266201528e+01
#define A1211 6.43392746015763530355970484046e-01
#define DB01 5.42937341165687622380535766363e-02
#define DB06 4.45031289275240888144113950566e+00
#define DB07 1.89151789931450038304281599044e+00
#define DB08 -5.80120396001058478146721142270e+00
#define DB09 3.11164366957819894408916062370e-01
#define DB10 -1.52160949662516078556178806805e-01
#define DB11 2.01365400804030348374776537501e-01
#define DB12 4.47106157277725905176885569043e-02
#define BH1 0.244094488188976377952755905512e+00
#define BH2 0.733846688281611857341361741547e+00
#define BH3 0.220588235294117647058823529412e-01
#define ER01 0.1312004499419488073250102996e-01
#define ER06 -0.1225156446376204440720569753e+01
#define ER07 -0.4957589496572501915214079952e+00
#define ER08 0.1664377182454986536961530415e+01
#define ER09 -0.3503288487499736816886487290e+00
#define ER10 0.3341791187130174790297318841e+00
#define ER11 0.8192320648511571246570742613e-01
#define ER12 -0.2235530786388629525884427845e-01

// k1 = f(si) is given (raw derivatives, no dt inside, as in the reference);
// returns err and sf.  Slopes are stored in the RHS's type K; stage points,
// sums and the error norm are FP64.
template <class Rhs>
__device__ __forceinline__ double attempt_dp853(Rhs &rhsf, const StepParams &P, const double *si,
                                                const typename Rhs::K *k1, double dt, double *sf)
{
    using K = typename Rhs::K;
    K k2[6], k3[6], k4[6], k5[6], k6[6], k7[6], k8[6], k9[6], k10[6], k11[6], k12[6];
    double ss[6];
#define RHS(out) rhsf(ss, out)
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * A0201 * double(k1[i]);
    RHS(k2);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0301 * double(k1[i]) + A0302 * double(k2[i]));
    RHS(k3);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0401 * double(k1[i]) + A0403 * double(k3[i]));
    RHS(k4);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0501 * double(k1[i]) + A0503 * double(k3[i]) + A0504 * double(k4[i]));
    RHS(k5);
#pragma unroll
    for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0601 * double(k1[i]) + A0604 * double(k4[i]) + A0605 * double(k5[i]));
    RHS(k6);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A0701 * double(k1[i]) + A0704 * double(k4[i]) + A0705 * double(k5[i]) + A0706 * double(k6[i]));
    RHS(k7);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A0801 * double(k1[i]) + A0804 * double(k4[i]) + A0805 * double(k5[i]) + A0806 * double(k6[i]) +
                              A0807 * double(k7[i]));
    RHS(k8);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A0901 * double(k1[i]) + A0904 * double(k4[i]) + A0905 * double(k5[i]) + A0906 * double(k6[i]) +
                              A0907 * double(k7[i]) + A0908 * double(k8[i]));
    RHS(k9);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A1001 * double(k1[i]) + A1004 * double(k4[i]) + A1005 * double(k5[i]) + A1006 * double(k6[i]) +
                              A1007 * double(k7[i]) + A1008 * double(k8[i]) + A1009 * double(k9[i]));
    RHS(k10);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A1101 * double(k1[i]) + A1104 * double(k4[i]) + A1105 * double(k5[i]) + A1106 * double(k6[i]) +
                              A1107 * double(k7[i]) + A1108 * double(k8[i]) + A1109 * double(k9[i]) + A1110 * double(k10[i]));
    RHS(k11);
#pragma unroll
    for (int i = 0; i < 6; i++)
        ss[i] = si[i] + dt * (A1201 * double(k1[i]) + A1204 * double(k4[i]) + A1205 * double(k5[i]) + A1206 * double(k6[i]) +
                              A1207 * double(k7[i]) + A1208 * double(k8[i]) + A1209 * double(k9[i]) + A1210 * double(k10[i]) +
                              A1211 * double(k11[i]));
    RHS(k12);
#undef RHS

    double err3 = 0.0, err5 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double kb = DB01 * double(k1[i]) + DB06 * double(k6[i]) + DB07 * double(k7[i]) + DB08 * double(k8[i]) + DB09 * double(k9[i]) +
                          DB10 * double(k10[i]) + DB11 * double(k11[i]) + DB12 * double(k12[i]);
        sf[i] = si[i] + dt * kb;
        const double a = fabs(si[i]), b = fabs(sf[i]);
        const double sr = P.atol + P.rtol * (a > b ? a : b);
        const double e3 = (kb - BH1 * double(k1[i]) - BH2 * double(k9[i]) - BH3 * double(k12[i])) / sr;
        err3 = err3 + e3 * e3;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double a = fabs(si[i]), b = fabs(sf[i]);
        const double sr = P.atol + P.rtol * (a > b ? a : b);
        const double e5 = (ER01 * double(k1[i]) + ER06 * double(k6[i]) + ER07 * double(k7[i]) + ER08 * double(k8[i]) + ER09 * double(k9[i]) +
                           ER10 * double(k10[i]) + ER11 * double(k11[i]) + ER12 * double(k12[i])) / sr;
        err5 = err5 + e5 * e5;
    }
    const double deno = err5 + 1.0e-02 * err3;
    return deno <= 0.0 ? fabs(dt) * err5 / sqrt(6.0) : fabs(dt) * err5 / sqrt(6.0 * deno);
}

// cubic Hermite snapshot (integrations.F90:158-168 / :447-457); ka, kb are
// the end slopes times dt
__device__ __forceinline__ void hermite_out(double *st, double t, double tn, double dt,
                                            const double *si, const double *sf, const double *ka,
                                            const double *kb)
{
    const double fr = (t - tn) / dt, fl = 1.0 - fr;
    double o[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double ds = sf[i] - si[i];
        o[i] = fr * si[i] + fl * sf[i] + fl * fr * ((ka[i] - ds) * fr + (ds - kb[i]) * fl);
    }
#pragma unroll
    for (int i = 0; i < 6; i++) st[i] = o[i];
    st[6] = sqrt(o[3] * o[3] + o[4] * o[4] + o[5] * o[5] + 1.0);
    st[7] = dt;
}

__device__ __forceinline__ bool finite6(const double *a)
{
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 6; i++) ok = ok && isfinite(a[i]);
    return ok;
}

// ---------------------------------------------------------------------------
// persistent work-queue integrator
// ---------------------------------------------------------------------------
template <int METHOD, int MODE, int LAYOUT, int CACHE, int PREC>
__global__ void __launch_bounds__(kBlock)
integrate_kernel(const DeviceFields F, const StepParams P, const DeviceParticles S, double t_stop)
{
    // the power's lookup tables (5 KB) in LDS: per-lane indices make every
    // lookup a divergent gather, which LDS serves far faster than memory
    __shared__ double s_logt[128][3];
    __shared__ unsigned long long s_expt[256];
    for (int i = threadIdx.x; i < 128 * 3; i += blockDim.x)
        (&s_logt[0][0])[i] = (&ptc_pow_log_tab[0][0])[i];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_expt[i] = ptc_pow_exp_tab[i];
    __syncthreads();
    const PowTables T{s_logt, s_expt};

    Force<MODE, LAYOUT, CACHE, PREC> force(F, P);
    using K = typename decltype(force)::K;
    const int lane = (int)__lane_id();
    const double t_end = t_stop < P.tmx ? t_stop : P.tmx;
    const int64_t N = S.np;

    bool active = false, exhausted = false, need_rhs = false, fresh = false;
    int64_t idx = 0;
    double si[6];
    K f1[6];
    double t = 0.0, dt = 0.0, errold = 1.0e-04, tnext = 0.0;
    int32_t n = 2, flags = 0, status = PTC_STATUS_RUNNING;
    int64_t c_all = 0, c_acc = 0, c_rej = 0, c_rhs = 0;

    for (;;) {
        // ---- refill lanes without a particle (one atomic per wave) ----
        const bool want = !active && !exhausted;
        const uint64_t wmask = __ballot(want);
        if (wmask) {
            const int leader = __ffsll((unsigned long long)wmask) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(S.qcount, __popcll(wmask));
            base = __shfl(base, leader);
            if (want) {
                const int64_t my = (int64_t)base + __popcll(wmask & ((1ull << lane) - 1ull));
                if (my >= N) {
                    exhausted = true;
                } else {
                    idx = S.order ? (int64_t)S.order[my] : my;
                    status = S.status[idx];
                    if (status == PTC_STATUS_RUNNING) {
                        active = true;
                        fresh = true;
                        need_rhs = true;
#pragma unroll
                        for (int c = 0; c < 6; c++) si[c] = S.x[c][idx];
                        t = S.t[idx];
                        dt = S.dt[idx];
                        errold = S.errold[idx];
                        n = S.nout[idx];
                        flags = S.flags[idx];
                        c_all = S.cnt[idx];
                        c_acc = S.cnt[N + idx];
                        c_rej = S.cnt[2 * N + idx];
                        c_rhs = S.nrhs[idx];
                    }
                }
            }
        }
        if (!__any(active)) {
            if (__all(exhausted)) break;
            continue;
        }

        // ---- one shared RHS slot: f(si) for refilled lanes and for lanes
        //      that accepted last iteration (the next step's stage 1) ----
        if (need_rhs) {
            force(si, f1);
            c_rhs++;
            need_rhs = false;
        }

        bool finish = false;
        if (fresh) {
            fresh = false;
            if (!(flags & FLAG_STARTED)) {
                // integrations.F90:120-127 / :352-359
                dt = dti_guess(P, METHOD == PTC_RK5 ? 5 : 8, si, f1, T);
                double *o = S.out + (size_t)8 * P.m * idx;
                o[6] = sqrt(si[3] * si[3] + si[4] * si[4] + si[5] * si[5] + 1.0);
                o[7] = dt;
                t = 0.0;
                n = 2;
                errold = 1.0e-04;
                flags = FLAG_STARTED;
                c_all = c_acc = c_rej = 0;
            }
            tnext = n <= P.m ? S.time[n - 1] : 0.0;
            if (!(t < t_end)) {
                if (!(t < P.tmx)) status = PTC_STATUS_DONE;
                finish = true;
            }
        }

        if (active && !finish) {
            if (P.max_steps > 0 && c_all >= P.max_steps) {
                status = PTC_STATUS_MAXSTEPS;
                finish = true;
            } else if (!isfinite(dt)) {
                status = PTC_STATUS_NONFINITE;
                finish = true;
            } else if (!(dt > 0.0) || t + dt == t) {
                status = PTC_STATUS_DTSMALL;
                finish = true;
            } else {
                double sf[6];
                double dtn;
                if constexpr (METHOD == PTC_RK5) {
                    double k6[6];
                    const double del = attempt_rk5(force, P, si, f1, dt, sf, k6);
                    c_rhs += 5;
                    if (!isfinite(del) || !finite6(sf)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                      // one power per lane serves both outcomes:
                      // del**(-0.2) on accept, del**(-0.25) on reject
                      // (integrations.F90:177,186)
                      const bool accept = del <= 1.0;
                      const double q = P.safe * ptc_pow_with(del, accept ? -0.20 : -0.25, T.log, T.exp);
                      if (accept) {
                        t = t + dt;
                        if (n <= P.m && t >= tnext) {
                            double k1[6];
#pragma unroll
                            for (int c = 0; c < 6; c++) k1[c] = dt * double(f1[c]);
                            do {
                                hermite_out(S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), t,
                                            tnext, dt, si, sf, k1, k6);
                                n++;
                                tnext = n <= P.m ? S.time[n - 1] : 0.0;
                            } while (n <= P.m && t >= tnext);
                        }
#pragma unroll
                        for (int c = 0; c < 6; c++) si[c] = sf[c];
                        c_acc++;
                        dtn = dt * (P.fcmx < q ? P.fcmx : q);
                        need_rhs = true;
                      } else {
                        c_rej++;
                        dtn = dt * (P.fcmn > q ? P.fcmn : q);
                      }
                    }
                } else {
                    const double err = attempt_dp853(force, P, si, f1, dt, sf);
                    c_rhs += 11;
                    if (!isfinite(err) || !finite6(sf)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                      // safe * err**expo serves both outcomes (integrations.F90:466,486)
                      const double qe = P.safe * ptc_pow_with(err, P.expo, T.log, T.exp);
                      if (err <= 1.0) {
                        t = t + dt;
                        if (n <= P.m && t >= tnext) {
                            // the end slope f(sf) is needed now; it is also
                            // the next step's stage 1
                            K fe[6];
                            double ka[6], kb[6];
                            force(sf, fe);
                            c_rhs++;
#pragma unroll
                            for (int c = 0; c < 6; c++) {
                                ka[c] = dt * double(f1[c]);
                                kb[c] = dt * double(fe[c]);
                            }
                            do {
                                hermite_out(S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), t,
                                            tnext, dt, si, sf, ka, kb);
                                n++;
                                tnext = n <= P.m ? S.time[n - 1] : 0.0;
                            } while (n <= P.m && t >= tnext);
#pragma unroll
                            for (int c = 0; c < 6; c++) f1[c] = fe[c];
                        } else {
                            need_rhs = true;
                        }
#pragma unroll
                        for (int c = 0; c < 6; c++) si[c] = sf[c];
                        c_acc++;
                        double q = qe;
                        // errold**beta; x**0 is exactly 1 (and q * 1 == q),
                        // so the beta = 0 default skips the power
                        if (P.beta != 0.0) q = q * ptc_pow_with(errold, P.beta, T.log, T.exp);
                        if (q < P.fcmn) q = P.fcmn;
                        if (flags & FLAG_REJECTED) dtn = dt * (q < 1.0 ? q : 1.0);
                        else dtn = dt * (q < P.fcmx ? q : P.fcmx);
                        errold = err > 1.0e-04 ? err : 1.0e-04;
                        flags &= ~FLAG_REJECTED;
                      } else {
                        c_rej++;
                        dtn = dt * (P.fcmn > qe ? P.fcmn : qe);
                        flags |= FLAG_REJECTED;
                      }
                    }
                }
                if (!finish) {
                    c_all++;
                    dt = dtn < P.dtm ? dtn : P.dtm;
                    if (!(t < t_end)) {
                        if (!(t < P.tmx)) status = PTC_STATUS_DONE;
                        finish = true;
                    }
                }
            }
        }

        if (active && finish) {
#pragma unroll
            for (int c = 0; c < 6; c++) S.x[c][idx] = si[c];
            S.t[idx] = t;
            S.dt[idx] = dt;
            S.errold[idx] = errold;
            S.nout[idx] = n;
            S.flags[idx] = flags;
            S.status[idx] = status;
            S.cnt[idx] = c_all;
            S.cnt[N + idx] = c_acc;
            S.cnt[2 * N + idx] = c_rej;
            S.nrhs[idx] = c_rhs;
            active = false;
            need_rhs = false;
        }
    }
}

// Single-call probes of the device building blocks, one input per thread:
//   op 0: gather at pos(3)          -> f(6)  (U or E, then B)
//   op 1: RHS at vec(6)             -> acc(6)
//   op 2: dti_guess, order 5, vec(6) -> dt(1) (op 3: order 8)
//   op 4: step-size power x^y, (x, y) -> (1)
template <int MODE, int LAYOUT, int PREC>
__global__ void probe_kernel(const DeviceFields F, const StepParams P, int op, int64_t n,
                             const double *__restrict__ in, double *__restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[6];
    Force<MODE, LAYOUT, 0, PREC> force(F, P);
    typename decltype(force)::K r[6];
    if (op == 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) v[k] = in[3 * i + k];
        if constexpr (PREC == PTC_FP32) {
            CornersT<float> c;
            corners(P, v, c);
            float f[6];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                float w[6];
                load_node_f32((const float *)F.data, c.o[j], w);
                accumulate_row(j, c.w[j], w, f);
            }
#pragma unroll
            for (int k = 0; k < 6; k++) r[k] = f[k];
        } else {
            gather<LAYOUT>(F, P, v, r);
        }
#pragma unroll
        for (int k = 0; k < 6; k++) out[6 * i + k] = double(r[k]);
    } else if (op == 1) {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] = in[6 * i + k];
        force(v, r);
#pragma unroll
        for (int k = 0; k < 6; k++) out[6 * i + k] = double(r[k]);
    } else if (op == 2 || op == 3) {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] = in[6 * i + k];
        force(v, r);
        out[i] = dti_guess(P, op == 2 ? 5 : 8, v, r, PowTables{ptc_pow_log_tab, ptc_pow_exp_tab});
    } else if (op == 4) {
        out[i] = ptc_libm_pow(in[2 * i], in[2 * i + 1]);
    }
}

template <int MODE, int LAYOUT, int PREC = PTC_FP64>
hipError_t launch_probe_t(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                          const double *in, double *out, hipStream_t stream)
{
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL((probe_kernel<MODE, LAYOUT, PREC>), dim3((unsigned)blocks), dim3(256), 0,
                       stream, f, sp, op, n, in, out);
    return hipGetLastError();
}

__global__ void totals_kernel(const DeviceParticles S, unsigned long long *out)
{
    const int64_t N = S.np;
    unsigned long long a = 0, b = 0, c = 0, d = 0, e = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        a += (unsigned long long)S.cnt[i];
        b += (unsigned long long)S.cnt[N + i];
        c += (unsigned long long)S.cnt[2 * N + i];
        d += (unsigned long long)S.nrhs[i];
        e += S.status[i] == PTC_STATUS_RUNNING ? 1ull : 0ull;
    }
    // wave reduction, then one atomic per wave and counter
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        b += __shfl_down(b, off);
        c += __shfl_down(c, off);
        d += __shfl_down(d, off);
        e += __shfl_down(e, off);
    }
    if (__lane_id() == 0) {
        atomicAdd(out + 0, a);
        atomicAdd(out + 1, b);
        atomicAdd(out + 2, c);
        atomicAdd(out + 3, d);
        atomicAdd(out + 4, e);
    }
}

__global__ void zero_i32_kernel(int32_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void zero_i64_kernel(int64_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ void to_aos_kernel(const double *__restrict__ soa, int64_t nodes, double *__restrict__ aos)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 6; c++) aos[i * 6 + c] = soa[c * nodes + i];
    }
}

// one thread per value of the brick array: cell-major, then corner, then
// component, so the writes are contiguous
__global__ void to_bricks_kernel(const double *__restrict__ aos, int dm0, int dm1, int dm2,
                                 int64_t nodes, double *__restrict__ bricks)
{
    const int64_t total = nodes * 48;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cell = g / 48;
        const int rem = (int)(g - cell * 48), j = rem / 6, comp = rem - j * 6;
        const int i = (int)(cell % dm0);
        const int jy = (int)((cell / dm0) % dm1);
        const int kz = (int)(cell / ((int64_t)dm0 * dm1));
        // corner j = (x, y, z) offsets (j>>2, (j>>1)&1, j&1): x slowest
        const int ii = (j >> 2) ? (i + 1 == dm0 ? 0 : i + 1) : i;
        const int jj = ((j >> 1) & 1) ? (jy + 1 == dm1 ? 0 : jy + 1) : jy;
        const int kk = (j & 1) ? (kz + 1 == dm2 ? 0 : kz + 1) : kz;
        const int64_t node = ii + (int64_t)dm0 * (jj + (int64_t)dm1 * kk);
        bricks[g] = aos[node * 6 + comp];
    }
}

// SOA planes (FP64) -> FP32 nodes (U or E, B, |B| from the FP64 values, 0)
__global__ void to_nodes_f32_kernel(const double *__restrict__ soa, int64_t nodes,
                                    float *__restrict__ out)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double bx = soa[3 * nodes + i], by = soa[4 * nodes + i], bz = soa[5 * nodes + i];
        float4 a, b;
        a.x = float(soa[i]);
        a.y = float(soa[nodes + i]);
        a.z = float(soa[2 * nodes + i]);
        a.w = float(bx);
        b.x = float(by);
        b.y = float(bz);
        b.z = float(sqrt(bx * bx + by * by + bz * bz));
        b.w = 0.0f;
        ((float4 *)out)[2 * i] = a;
        ((float4 *)out)[2 * i + 1] = b;
    }
}

template <int METHOD, int MODE, int LAYOUT, int CACHE, int PREC>
hipError_t launch_t(const DeviceFields &f, const StepParams &sp, const DeviceParticles &p,
                    double t_stop, LaunchConfig lc, hipStream_t stream)
{
    auto kern = integrate_kernel<METHOD, MODE, LAYOUT, CACHE, PREC>;
    int grid = lc.grid_blocks;
    if (grid <= 0) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, lc.block, 0);
        if (e != hipSuccess) return e;
        grid = cus * (per_cu > 0 ? per_cu : 1);
    }
    const int64_t need = (p.np + lc.block - 1) / lc.block;
    if ((int64_t)grid > need) grid = (int)need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(lc.block), 0, stream, f, sp, p, t_stop);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_integrate(int method, const DeviceFields &f, const StepParams &sp,
                            const DeviceParticles &p, double t_stop, const LaunchConfig &lc,
                            hipStream_t stream)
{
#define DISPATCH(M, MO, L, C, PR)                                              \
    if (method == M && f.mode == MO && f.layout == L && lc.cell_cache == C &&  \
        f.prec == PR)                                                          \
        return launch_t<M, MO, L, C, PR>(f, sp, p, t_stop, lc, stream);
#define DISPATCH_MODES(L, C, PR)                                               \
    DISPATCH(PTC_RK5, PTC_MODE_VB, L, C, PR)                                   \
    DISPATCH(PTC_RK5, PTC_MODE_EB, L, C, PR)                                   \
    DISPATCH(PTC_DOP853, PTC_MODE_VB, L, C, PR)                                \
    DISPATCH(PTC_DOP853, PTC_MODE_EB, L, C, PR)
    DISPATCH_MODES(LAYOUT_AOS, 1, PTC_FP64)
    DISPATCH_MODES(LAYOUT_AOS, 0, PTC_FP64)
    DISPATCH_MODES(LAYOUT_SOA, 0, PTC_FP64)
    DISPATCH_MODES(LAYOUT_BRICK, 0, PTC_FP64)
    DISPATCH_MODES(LAYOUT_AOS, 1, PTC_FP32)
    DISPATCH_MODES(LAYOUT_AOS, 0, PTC_FP32)
#undef DISPATCH_MODES
#undef DISPATCH
    return hipErrorInvalidValue;
}

hipError_t launch_probe(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                        const double *in, double *out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    if (op < 0 || op > 4) return hipErrorInvalidValue;
    if (f.prec == PTC_FP32) {
        if (f.mode == PTC_MODE_VB)
            return launch_probe_t<PTC_MODE_VB, LAYOUT_AOS, PTC_FP32>(f, sp, op, n, in, out, stream);
        return launch_probe_t<PTC_MODE_EB, LAYOUT_AOS, PTC_FP32>(f, sp, op, n, in, out, stream);
    }
    if (f.mode == PTC_MODE_VB && f.layout == LAYOUT_SOA)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_SOA>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_VB && f.layout == LAYOUT_AOS)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_AOS>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_EB && f.layout == LAYOUT_SOA)
        return launch_probe_t<PTC_MODE_EB, LAYOUT_SOA>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_EB && f.layout == LAYOUT_AOS)
        return launch_probe_t<PTC_MODE_EB, LAYOUT_AOS>(f, sp, op, n, in, out, stream);
    if (f.mode == PTC_MODE_VB)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_BRICK>(f, sp, op, n, in, out, stream);
    return launch_probe_t<PTC_MODE_EB, LAYOUT_BRICK>(f, sp, op, n, in, out, stream);
}

hipError_t launch_to_bricks(const DeviceFields &f, hipStream_t stream)
{
    if (f.prec != PTC_FP64 || !f.bricks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(to_bricks_kernel, dim3(65536), dim3(256), 0, stream, (const double *)f.data,
                       f.dm[0], f.dm[1], f.dm[2], f.nodes, (double *)f.bricks);
    return hipGetLastError();
}

hipError_t launch_totals(const DeviceParticles &p, int64_t *d_out, hipStream_t stream)
{
    hipError_t e = launch_zero_i64(d_out, 5, stream);
    if (e != hipSuccess) return e;
    int64_t blocks = (p.np + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(totals_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p,
                       (unsigned long long *)d_out);
    return hipGetLastError();
}

hipError_t launch_zero_i32(int32_t *ptr, int64_t n, hipStream_t stream)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(zero_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ptr, n);
    return hipGetLastError();
}

hipError_t launch_zero_i64(int64_t *ptr, int64_t n, hipStream_t stream)
{
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(zero_i64_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ptr, n);
    return hipGetLastError();
}

hipError_t launch_to_aos(const double *soa, int64_t nodes, void *aos, int prec, hipStream_t stream)
{
    int64_t blocks = (nodes + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (prec == PTC_FP32) {
        hipLaunchKernelGGL(to_nodes_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa,
                           nodes, (float *)aos);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(to_aos_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, soa, nodes,
                       (double *)aos);
    return hipGetLastError();
}

}  // namespace ptc
