// Trio kernel: three lanes integrate one particle, lane a (0..2) of the
// trio holding two components of the state and of its slopes, position
// x_a (component a) and momentum u_a (component a + 3); 21 trios per wave,
// lane 63 idle.  The arithmetic of every component is the group kernel's
// (group_impl.hpp) operation for operation, and every sum over the six
// components is formed by each lane from the values shuffled across the
// trio in component order, so the results are bit-identical to the group
// kernel, the lane-per-particle kernel and the reference
// (integrations.F90:57-200, 224-505, 524-589; accelerations.F90:59-96;
// interpolations.F90:104-226).
//
// Why: a group of six lanes repeats the per-particle work of every RHS
// evaluation (cell lookup, trilinear weights, gamma's square root, the
// shared division, the corner addresses) on six lanes for ten particles per
// wave.  Three lanes per particle do it for 21 particles per wave, and each
// lane's per-component work (stage sums, the gather and weighted sum of its
// two field components, its two force components) doubles: about 40 % fewer
// instructions per particle and evaluation, and more particles in flight
// per SIMD to hide the latency of gathers from fields far larger than the
// caches.  The cell lookup fits the trio exactly: lane a owns axis a's
// position, so the stage point needs no shuffle for it.
//
// Static fields, PI controller, FP64 (bit-identical) or FP32 (the mixed
// arithmetic of the group kernel); node, pair or quad rows.
#pragma once
#include "group_impl.hpp"

namespace ptc {
namespace {

// lanes per particle, and particles per wave (lane 63 stays idle)
constexpr int kTrio = 3;
constexpr int kTriosPerWave = 64 / kTrio;

// The RHS as a trio evaluates it: lane a keeps components a and a + 3 of the
// eight corner rows of the cell it last gathered from (the cell cache, split
// over the trio).  The corner lookup is group_corners with the trio's three
// lanes as the three axes.
// FEAT = 1 also keeps, from the last evaluation, the lane's two interpolated
// field components and the in-cell fraction (step-size features).  TIME = 1
// (a snapshot series in node or column rows): the lane's two components of
// the cell's corner rows in snapshot k + 1 as well, the stage time and the
// cached interval key, as GroupForce.
template <int MODE, int GEOM, int PREC, int ROWS, int FEAT = 0, int TIME = 0>
struct TrioForce {
    static_assert(ROWS != ROWS_QUADS || PREC == PTC_FP32, "quad rows hold FP32 fields");
    static_assert(TIME == 0 || (FEAT == 0 && (ROWS == ROWS_NODES || ROWS == ROWS_COLUMNS)),
                  "a series is read from node or column rows, with the PI controller");
    using K = std::conditional_t<PREC == PTC_FP32, float, double>;
    const DeviceFields &F;
    const StepParams &P;
    const int base, a;
    const int r1, r2;  // lanes of the trio holding axes a + 1 and a + 2 (mod 3)
    uint32_t key0 = 0xffffffffu, key7 = 0xffffffffu;
    uint32_t loads = 0;  // cell loads of the trio's particle (the same on each lane)
    K v[2][8];           // components a and a + 3 of the cached cell's corner rows
    K v1[TIME ? 2 : 1][8];  // TIME: the same in snapshot k + 1
    double ts = 0.0;        // TIME: the stage time
    int kh = 0, kc = -1;    // TIME: the last interval, the cached interval key
    K fint0 = K(0), fint1 = K(0);          // FEAT: U or E component a, B component a
    float xdl[3] = {0.0f, 0.0f, 0.0f};     // FEAT: the in-cell fraction

    __device__ __forceinline__ TrioForce(const DeviceFields &F_, const StepParams &P_, int b,
                                         int a_)
        : F(F_), P(P_), base(b), a(a_), r1(b + (a_ + 1) % 3), r2(b + (a_ + 2) % 3)
    {
    }

    __device__ __forceinline__ void at(double t)
    {
        if constexpr (TIME == 1) ts = t;
    }

    // node rows (six FP64 values per node, or eight FP32: six and |B|, pad)
    // of snapshot k into dst
    __device__ __forceinline__ void load_nodes(const uint32_t *o, K (*dst)[8], int k = 0) const
    {
        constexpr int kRow = PREC == PTC_FP32 ? 8 : 6;
        const K *__restrict__ d = (const K *)F.data + (size_t)k * F.sstride;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int j = 0; j < 8; j++) dst[h][j] = d[(size_t)o[j] * kRow + a + 3 * h];
    }

    // pair rows: as GroupForce::load_pairs, for both components
    __device__ __forceinline__ void load_pairs(const uint32_t *o)
    {
        using K2 = std::conditional_t<PREC == PTC_FP32, float2, double2>;
        const K2 *__restrict__ d2 = (const K2 *)F.data;
        const uint32_t dx = o[4] - o[0];
        const bool succ = dx == 1u || dx == (uint32_t)(1 - P.dm[0]);
        K2 p[2][4];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int j = 0; j < 4; j++) p[h][j] = d2[(size_t)o[j] * 8 + a + 3 * h];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                v[h][j] = p[h][j].x;
                v[h][4 + j] = p[h][j].y;
            }
        if (!succ) {
            const K *__restrict__ d = (const K *)F.data;
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    v[h][4 + j] = d[(size_t)o[4 + j] * 16 + 2 * (a + 3 * h)];
        }
    }

    // FP32 quad rows (node order, or column storage; snapshot k of a
    // series) into dst: as GroupForce::load_quads, for both components
    __device__ __forceinline__ void load_quads(const uint32_t *o, K (*dst)[8], int k = 0) const
    {
        const uint32_t sx = ROWS == ROWS_COLUMNS ? (uint32_t)P.dm[2] : 1u;
        const uint32_t sy = (uint32_t)P.dm[0] * sx;
        const uint32_t dx = o[4] - o[0], dy = o[2] - o[0];
        const bool xs = dx == sx || dx == (uint32_t)((int64_t)(1 - P.dm[0]) * sx);
        const bool ys = dy == sy || dy == (uint32_t)((int64_t)(1 - P.dm[1]) * sy);
        const float *__restrict__ d = (const float *)F.data + (size_t)k * F.sstride;
        const float4 *__restrict__ d4 = (const float4 *)d;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = a + 3 * h;
            const float4 q0 = d4[(size_t)o[0] * 8 + c], q1 = d4[(size_t)o[1] * 8 + c];
            dst[h][0] = q0.x; dst[h][1] = q1.x; dst[h][2] = q0.z; dst[h][3] = q1.z;
            dst[h][4] = q0.y; dst[h][5] = q1.y; dst[h][6] = q0.w; dst[h][7] = q1.w;
        }
        if (!(xs && ys)) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int j = 2; j < 8; j++)
                    dst[h][j] = d[(size_t)o[j] * 32 + 4 * (a + 3 * h)];
        }
    }

    // column rows (snapshot k of a series) into dst: as
    // GroupForce::load_columns, for both components
    __device__ __forceinline__ void load_columns(const uint32_t *o, K (*dst)[8], int k = 0) const
    {
        const uint32_t sx = (uint32_t)P.dm[2], sy = (uint32_t)P.dm[0] * (uint32_t)P.dm[2];
        const uint32_t dx = o[4] - o[0], dy = o[2] - o[0];
        const bool xs = dx == sx || dx == (uint32_t)((int64_t)(1 - P.dm[0]) * P.dm[2]);
        const bool ys = dy == sy || dy == (uint32_t)((int64_t)(1 - P.dm[1]) * sy);
        const double *__restrict__ d = (const double *)F.data + (size_t)k * F.sstride;
        const double2 *__restrict__ d2 = (const double2 *)d;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = a + 3 * h;
            const double2 a0 = d2[(size_t)o[0] * 12 + 2 * c], a1 = d2[(size_t)o[0] * 12 + 2 * c + 1];
            const double2 b0 = d2[(size_t)o[1] * 12 + 2 * c], b1 = d2[(size_t)o[1] * 12 + 2 * c + 1];
            dst[h][0] = a0.x; dst[h][1] = b0.x; dst[h][2] = a1.x; dst[h][3] = b1.x;
            dst[h][4] = a0.y; dst[h][5] = b0.y; dst[h][6] = a1.y; dst[h][7] = b1.y;
        }
        if (!(xs && ys)) {
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int j = 2; j < 8; j++)
                    dst[h][j] = d[(size_t)o[j] * 24 + 4 * (a + 3 * h)];
        }
    }

    // snapshot k's rows into dst, in the kernel's row form (a series)
    __device__ __forceinline__ void load_snapshot(const uint32_t *o, K (*dst)[8], int k) const
    {
        if constexpr (ROWS == ROWS_COLUMNS && PREC == PTC_FP32) load_quads(o, dst, k);
        else if constexpr (ROWS == ROWS_COLUMNS) load_columns(o, dst, k);
        else load_nodes(o, dst, k);
    }

    // components a and a + 3 of f(s), s given by its components a and a + 3
    // on each lane
    __device__ __forceinline__ void operator()(const double *s, K *out)
    {
        CornersT<K> c;
        group_corners<K, GEOM, ROWS>(P, s[0], base, a, c, FEAT ? xdl : nullptr);
        K f0, f1;
        if constexpr (TIME == 1) {
            // the two bracketing snapshots' interpolants blended in time, as
            // GroupForce
            double al;
            bool two;
            const int k = series_interval(F, ts, kh, al, two);
            const int key = 2 * k + (two ? 1 : 0);
            if (!(c.o[0] == key0 && c.o[7] == key7 && (key == kc || (!two && key + 1 == kc)))) {
                loads += two ? 2u : 1u;
                load_snapshot(c.o, v, k);
                if (two) load_snapshot(c.o, v1, k + 1);
                key0 = c.o[0];
                key7 = c.o[7];
                kc = key;
            }
            f0 = c.w[0] * v[0][0];
            f1 = c.w[0] * v[1][0];
#pragma unroll
            for (int j = 1; j < 8; j++) {
                f0 = f0 + c.w[j] * v[0][j];
                f1 = f1 + c.w[j] * v[1][j];
            }
            if (two) {
                K g0 = c.w[0] * v1[0][0], g1 = c.w[0] * v1[1][0];
#pragma unroll
                for (int j = 1; j < 8; j++) {
                    g0 = g0 + c.w[j] * v1[0][j];
                    g1 = g1 + c.w[j] * v1[1][j];
                }
                const K ak = K(al);
                f0 = f0 + ak * (g0 - f0);
                f1 = f1 + ak * (g1 - f1);
            }
        } else {
            if (c.o[0] != key0 || c.o[7] != key7) {
                loads++;
                if constexpr (ROWS == ROWS_PAIRS) load_pairs(c.o);
                else if constexpr (ROWS == ROWS_QUADS) load_quads(c.o, v);
                else if constexpr (ROWS == ROWS_COLUMNS && PREC == PTC_FP32) load_quads(c.o, v);
                else if constexpr (ROWS == ROWS_COLUMNS) load_columns(c.o, v);
                else load_nodes(c.o, v);
                key0 = c.o[0];
                key7 = c.o[7];
            }
            // interpolations.F90:212-220, corners summed left to right: U or E
            // component a in f0, B component a in f1
            f0 = c.w[0] * v[0][0];
            f1 = c.w[0] * v[1][0];
#pragma unroll
            for (int j = 1; j < 8; j++) {
                f0 = f0 + c.w[j] * v[0][j];
                f1 = f1 + c.w[j] * v[1][j];
            }
        }
        if constexpr (FEAT == 1) {
            fint0 = f0;
            fint1 = f1;
        }
        const double u0 = __shfl(s[1], base), u1 = __shfl(s[1], base + 1),
                     u2 = __shfl(s[1], base + 2);
        // accelerations.F90:84-92: component a is v_a / vun, component a + 3
        // q/m (E_a + (v_{a+1} B_{a+2} - v_{a+2} B_{a+1})) (EB) or
        // q/m (w_{a+1} B_{a+2} - w_{a+2} B_{a+1}), w = v - U (VB)
        if constexpr (PREC == PTC_FP32) {
            const float ux = float(u0), uy = float(u1), uz = float(u2);
            const float rg = __builtin_amdgcn_rsqf(ux * ux + uy * uy + uz * uz + 1.0f);
            const float v0 = ux * rg, v1 = uy * rg, v2 = uz * rg;
            const float vs = a == 0 ? v0 : a == 1 ? v1 : v2;
            const float va = a == 0 ? v1 : a == 1 ? v2 : v0;  // v_{a+1}
            const float vd = a == 0 ? v2 : a == 1 ? v0 : v1;  // v_{a+2}
            const float bd = __shfl(f1, r2), ba = __shfl(f1, r1);
            const float qom = float(P.qom);
            out[0] = vs * float(P.rvun);
            if constexpr (MODE == PTC_MODE_EB) {
                out[1] = qom * (f0 + (va * bd - vd * ba));
            } else {
                const float ua = __shfl(f0, r1), ud = __shfl(f0, r2);
                out[1] = qom * ((va - ua) * bd - (vd - ud) * ba);
            }
        } else {
            const double g = sqrt(u0 * u0 + u1 * u1 + u2 * u2 + 1.0);
            const double q = s[1] / g;  // v_a, the trio's shared divisions
            out[0] = GEOM ? q * P.rvun : q / P.vun;
            const double bd = __shfl(f1, r2), ba = __shfl(f1, r1);
            if constexpr (MODE == PTC_MODE_EB) {
                const double va = __shfl(q, r1), vd = __shfl(q, r2);
                out[1] = P.qom * (f0 + (va * bd - vd * ba));
            } else {
                const double w = q - f0;
                const double wa = __shfl(w, r1), wd = __shfl(w, r2);
                out[1] = P.qom * (wa * bd - wd * ba);
            }
        }
    }
};

// The step-size features (features_bc) of the trio's accepted state si, on
// every lane of the trio, from the force's last evaluation, which was at si:
// the interpolated fields and the in-cell fraction it kept, and |B| at the
// eight corners from the B components the three lanes hold (the same values
// and operations as group_features).
template <int MODE, class TF>
__device__ __forceinline__ void trio_features(const TF &force, const StepParams &P,
                                              const double *si, int base, double h, float err,
                                              float pi_ratio, bool rej, float *f)
{
    const double s[6] = {0.0, 0.0, 0.0, __shfl(si[1], base), __shfl(si[1], base + 1),
                         __shfl(si[1], base + 2)};
    float ub[6], bc[8];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        ub[j] = float(__shfl(force.fint0, base + j));
        ub[3 + j] = float(__shfl(force.fint1, base + j));
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const float b = float(force.v[1][j]);
        bc[j] = corner_bmag(__shfl(b, base), __shfl(b, base + 1), __shfl(b, base + 2));
    }
    features_bc<MODE>(P, s, ub, bc, force.xdl, h, err, pi_ratio, rej, f);
}

// component j (0..5) of a value held as [0] (j < 3) or [1] (j >= 3) on lane
// j % 3 of the trio
__device__ __forceinline__ double trio_comp(const double *x, int base, int j)
{
    return j < 3 ? __shfl(x[0], base + j) : __shfl(x[1], base + j - 3);
}

// true if every component of x in the trio is finite
__device__ __forceinline__ bool trio_finite(const double *x, int base)
{
    const uint64_t bad = __ballot(!isfinite(x[0]) || !isfinite(x[1]));
    return ((bad >> base) & 0x7ull) == 0;
}

// dti_guess (integrations.F90:524-589), as group_dti_guess; ki = f(si)
template <class K>
__device__ __forceinline__ double trio_dti_guess(const StepParams &P, int iord, int base,
                                                 const double *si, const K *kik,
                                                 const PowTables &T)
{
    double ki[2], sr[2], q0[2], q1[2], q2[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        ki[h] = double(kik[h]);
        sr[h] = P.atol + P.rtol * fabs(si[h]);
        q0[h] = si[h] / sr[h];
        q1[h] = ki[h] / sr[h];
    }
    double s = 0.0, q, d0, d1, d2, dti, dtf;
#pragma unroll
    for (int j = 0; j < 6; j++) { q = trio_comp(q0, base, j); s = s + q * q; }
    d0 = sqrt(s / 6.0);
    s = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) { q = trio_comp(q1, base, j); s = s + q * q; }
    d1 = sqrt(s / 6.0);
    dti = (fmin(d0, d1) < 1.0e-10) ? 1.0e-06 : 1.0e-01 * (d0 / d1);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const double sf = si[h] + dti * ki[h];
        q2[h] = (sf - si[h]) / sr[h];
    }
    s = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) { q = trio_comp(q2, base, j); s = s + q * q; }
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

// Cash-Karp attempt (integrations.F90:131-148) on two components, as
// group_attempt_rk5; returns del
template <class TF>
__device__ __forceinline__ double trio_attempt_rk5(TF &rhsf, const StepParams &P, int base,
                                                   const double *si, const typename TF::K *f1,
                                                   double t, double dt, double *sf, double *k6d)
{
    using K = typename TF::K;
    double k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], s[2];
    K r[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k1[h] = double(K(dt * double(f1[h])));
        s[h] = si[h] + CK_A21 * k1[h];
    }
    rhsf.at(t + CK_N2 * dt);
    rhsf(s, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k2[h] = double(K(dt * double(r[h])));
        s[h] = si[h] + CK_A31 * k1[h] + CK_A32 * k2[h];
    }
    rhsf.at(t + CK_N3 * dt);
    rhsf(s, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k3[h] = double(K(dt * double(r[h])));
        s[h] = si[h] + CK_A41 * k1[h] + CK_A42 * k2[h] + CK_A43 * k3[h];
    }
    rhsf.at(t + CK_N4 * dt);
    rhsf(s, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k4[h] = double(K(dt * double(r[h])));
        s[h] = si[h] + CK_A51 * k1[h] + CK_A52 * k2[h] + CK_A53 * k3[h] + CK_A54 * k4[h];
    }
    rhsf.at(t + dt);
    rhsf(s, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k5[h] = double(K(dt * double(r[h])));
        s[h] = si[h] + CK_A61 * k1[h] + CK_A62 * k2[h] + CK_A63 * k3[h] + CK_A64 * k4[h] +
               CK_A65 * k5[h];
    }
    rhsf.at(t + CK_N6 * dt);
    rhsf(s, r);
    const double eps = 2.2250738585072014e-308;  // tiny(1d0), integrations.F90:99
    double qc[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k6[h] = double(K(dt * double(r[h])));
        double sc = 0.0;
        sc = sc + CK_C1 * k1[h];
        sc = sc + 0.0 * k2[h];
        sc = sc + CK_C3 * k3[h];
        sc = sc + CK_C4 * k4[h];
        sc = sc + 0.0 * k5[h];
        sc = sc + CK_C6 * k6[h];
        sf[h] = si[h] + sc;
        double se = 0.0;
        se = se + CK_D1 * k1[h];
        se = se + 0.0 * k2[h];
        se = se + CK_D3 * k3[h];
        se = se + CK_D4 * k4[h];
        se = se + CK_D5 * k5[h];
        se = se + CK_D6 * k6[h];
        qc[h] = fabs(se / (fabs(si[h]) + fabs(k1[h]) + eps));
        k6d[h] = k6[h];
    }
    double del = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double q = trio_comp(qc, base, j);
        if (j == 0 || q > del) del = q;
    }
    return del / P.atol;
}

// DOP853 attempt (integrations.F90:365-437) on two components, as
// group_attempt_dp853; k1 = f(si); returns err
template <class TF>
__device__ __forceinline__ double trio_attempt_dp853(TF &rhsf, const StepParams &P, int base,
                                                     const double *si,
                                                     const typename TF::K *k1k, double t,
                                                     double dt, double *sf)
{
    using K = typename TF::K;
    double k1[2], k2[2], k3[2], k4[2], k5[2], k6[2], k7[2], k8[2], k9[2], k10[2], k11[2],
        k12[2], ss[2];
    K r[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k1[h] = double(k1k[h]);
        ss[h] = si[h] + dt * A0201 * k1[h];
    }
    rhsf.at(t + DN02 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k2[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0301 * k1[h] + A0302 * k2[h]);
    }
    rhsf.at(t + DN03 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k3[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0401 * k1[h] + A0403 * k3[h]);
    }
    rhsf.at(t + DN04 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k4[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0501 * k1[h] + A0503 * k3[h] + A0504 * k4[h]);
    }
    rhsf.at(t + DN05 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k5[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0601 * k1[h] + A0604 * k4[h] + A0605 * k5[h]);
    }
    rhsf.at(t + DN06 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k6[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0701 * k1[h] + A0704 * k4[h] + A0705 * k5[h] + A0706 * k6[h]);
    }
    rhsf.at(t + DN07 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k7[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0801 * k1[h] + A0804 * k4[h] + A0805 * k5[h] + A0806 * k6[h] +
                              A0807 * k7[h]);
    }
    rhsf.at(t + DN08 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k8[h] = double(r[h]);
        ss[h] = si[h] + dt * (A0901 * k1[h] + A0904 * k4[h] + A0905 * k5[h] + A0906 * k6[h] +
                              A0907 * k7[h] + A0908 * k8[h]);
    }
    rhsf.at(t + DN09 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k9[h] = double(r[h]);
        ss[h] = si[h] + dt * (A1001 * k1[h] + A1004 * k4[h] + A1005 * k5[h] + A1006 * k6[h] +
                              A1007 * k7[h] + A1008 * k8[h] + A1009 * k9[h]);
    }
    rhsf.at(t + DN10 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k10[h] = double(r[h]);
        ss[h] = si[h] + dt * (A1101 * k1[h] + A1104 * k4[h] + A1105 * k5[h] + A1106 * k6[h] +
                              A1107 * k7[h] + A1108 * k8[h] + A1109 * k9[h] + A1110 * k10[h]);
    }
    rhsf.at(t + DN11 * dt);
    rhsf(ss, r);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k11[h] = double(r[h]);
        ss[h] = si[h] + dt * (A1201 * k1[h] + A1204 * k4[h] + A1205 * k5[h] + A1206 * k6[h] +
                              A1207 * k7[h] + A1208 * k8[h] + A1209 * k9[h] + A1210 * k10[h] +
                              A1211 * k11[h]);
    }
    rhsf.at(t + dt);
    rhsf(ss, r);
    double e3[2], e5[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        k12[h] = double(r[h]);
        const double kb = DB01 * k1[h] + DB06 * k6[h] + DB07 * k7[h] + DB08 * k8[h] +
                          DB09 * k9[h] + DB10 * k10[h] + DB11 * k11[h] + DB12 * k12[h];
        sf[h] = si[h] + dt * kb;
        const double aa = fabs(si[h]), bb = fabs(sf[h]);
        const double sr = P.atol + P.rtol * (aa > bb ? aa : bb);
        e3[h] = (kb - BH1 * k1[h] - BH2 * k9[h] - BH3 * k12[h]) / sr;
        e5[h] = (ER01 * k1[h] + ER06 * k6[h] + ER07 * k7[h] + ER08 * k8[h] + ER09 * k9[h] +
                 ER10 * k10[h] + ER11 * k11[h] + ER12 * k12[h]) / sr;
    }
    double err3 = 0.0, err5 = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double q = trio_comp(e3, base, j);
        err3 = err3 + q * q;
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double q = trio_comp(e5, base, j);
        err5 = err5 + q * q;
    }
    const double deno = err5 + 1.0e-02 * err3;
    return deno <= 0.0 ? fabs(dt) * err5 / sqrt(6.0) : fabs(dt) * err5 / sqrt(6.0 * deno);
}

// cubic Hermite snapshot (integrations.F90:158-168 / :447-457) of two
// components; lane a of the trio writes components a + 1 and a + 4 of
// state(:, n), lane 0 also gamma and dt
__device__ __forceinline__ void trio_hermite_out(double *st, int a, int base, double t,
                                                 double tn, double dt, const double *si,
                                                 const double *sf, const double *ka,
                                                 const double *kb)
{
    const double fr = (t - tn) / dt, fl = 1.0 - fr;
    double o[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const double ds = sf[h] - si[h];
        o[h] = fr * si[h] + fl * sf[h] + fl * fr * ((ka[h] - ds) * fr + (ds - kb[h]) * fl);
    }
    const double o3 = __shfl(o[1], base), o4 = __shfl(o[1], base + 1),
                 o5 = __shfl(o[1], base + 2);
    st[a] = o[0];
    st[a + 3] = o[1];
    if (a == 0) {
        st[6] = sqrt(o3 * o3 + o4 * o4 + o5 * o5 + 1.0);
        st[7] = dt;
    }
}

// Waves per SIMD the kernel is compiled for: two.  With two components per
// lane the FP32 forms need 231-238 registers and no scratch, the FP64 forms
// 256 registers with 36-68 B of spills; held to 168 registers for a third
// wave, the FP32 forms would spill 212-236 B per lane.
constexpr int kTrioWaves = 2;

// The particle loop of integrate_group_kernel (static field), one particle
// per trio, particles taken from the second work-queue head S.qcount[1]: it
// runs a whole launch or takes the particles the lane-per-particle kernel
// hands over.  Particles whose status is not running, or that were started
// and already reached the end of the segment, are skipped.  CTRL_PI: the
// reference's controller; CTRL_LEARNED (FP32 fields): after an accepted step
// the network proposes the next step, as in the group kernel, evaluated once
// per wave for its 21 particles (mlp_forward_groups<3>; the network and a
// 64x72 BF16 buffer per wave in dynamic LDS).  TIME = 1: a snapshot series
// in node or column rows (PI controller), every RHS at its stage time as in
// the group kernel.
template <int METHOD, int MODE, int GEOM, int PREC, int ROWS, int CTRL = CTRL_PI, int TIME = 0>
__global__ void __launch_bounds__(kBlock, kTrioWaves)
integrate_trio_kernel(const DeviceFields F, const StepParams P, const DeviceParticles S,
                      double t_stop)
{
    static_assert(CTRL == CTRL_PI || (CTRL == CTRL_LEARNED && PREC == PTC_FP32 && TIME == 0),
                  "the trio kernel's learned controller takes static FP32 fields");
    __shared__ double s_logt[128][3];
    __shared__ unsigned long long s_expt[256];
    for (int i = threadIdx.x; i < 128 * 3; i += blockDim.x)
        (&s_logt[0][0])[i] = (&ptc_pow_log_tab[0][0])[i];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s_expt[i] = ptc_pow_exp_tab[i];
    extern __shared__ __attribute__((aligned(16))) char s_dyn[];
    if constexpr (CTRL == CTRL_LEARNED) {
        for (int i = threadIdx.x; i < (int)(sizeof(MlpInfer) / 4); i += blockDim.x)
            ((uint32_t *)s_dyn)[i] = ((const uint32_t *)&S.net->f)[i];
    }
    __syncthreads();
    const PowTables T{s_logt, s_expt};

    const int lane = (int)__lane_id();
    const int base = (lane / kTrio) * kTrio, a = lane - base;
    const double t_end = t_stop < P.tmx ? t_stop : P.tmx;
    const int64_t N = S.np;
    TrioForce<MODE, GEOM, PREC, ROWS, CTRL == CTRL_LEARNED ? 1 : 0, TIME> force(F, P, base, a);
    using K = typename decltype(force)::K;
    // this lane's two state planes, chosen once (indexing the kernel
    // argument's pointer array by a per-lane value would copy it to scratch)
    double *xp = a == 0 ? S.x[0] : a == 1 ? S.x[1] : S.x[2];
    double *up = a == 0 ? S.x[3] : a == 1 ? S.x[4] : S.x[5];

    // the lane past the last whole trio never takes a particle
    bool active = false, exhausted = lane >= kTrio * kTriosPerWave, need_rhs = false,
         fresh = false;
    int64_t idx = 0;
    double si[2] = {0.0, 0.0};
    K f1[2] = {K(0), K(0)};
    double t = 0.0, dt = 0.0, errold = 1.0e-04, tnext = 0.0;
    int32_t n = 2, flags = 0, status = PTC_STATUS_RUNNING;
    int64_t c_all = 0, c_acc = 0, c_rej = 0, c_rhs = 0;
    // learned controller: features are taken at the top of the iteration
    // after an accepted step of size h_acc with error ratio err_acc
    bool want_feat = false, rej_since = false, rej_before = false;
    double h_acc = 0.0;
    float err_acc = 0.0f, pi_ratio = 0.0f;

    for (;;) {
        // ---- refill trios without a particle (one atomic per wave) ----
        const bool want = !active && !exhausted;
        const uint64_t wmask = __ballot(want && a == 0);
        if (wmask) {
            const int first = __ffsll((unsigned long long)wmask) - 1;
            int b0 = 0;
            if (lane == first) b0 = atomicAdd(S.qcount + 1, __popcll(wmask));
            b0 = __shfl(b0, first);
            if (want) {
                const int64_t my = (int64_t)b0 + __popcll(wmask & ((1ull << base) - 1ull));
                if (my >= N) {
                    exhausted = true;
                } else {
                    idx = S.order ? (int64_t)S.order[my] : my;
                    status = S.status[idx];
                    t = S.t[idx];
                    flags = S.flags[idx];
                    if (status == PTC_STATUS_RUNNING && (t < t_end || !(flags & FLAG_STARTED))) {
                        active = true;
                        fresh = true;
                        need_rhs = true;
                        want_feat = rej_since = false;
                        si[0] = xp[idx];
                        si[1] = up[idx];
                        dt = S.dt[idx];
                        errold = S.errold[idx];
                        n = S.nout[idx];
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

        // ---- f(si) for new particles and after accepted steps ----
        if (need_rhs) {
            force.at(t);
            force(si, f1);
            c_rhs++;
            need_rhs = false;
        }

        if constexpr (CTRL == CTRL_LEARNED) {
            // (the force's last evaluation was at si: the one above, or the
            // end slope of the accepted step, which became si)
            if (__any(want_feat)) {
                float feat[kFeatures];
                if (want_feat) {
                    trio_features<MODE>(force, P, si, base, h_acc, err_acc, pi_ratio, rej_before,
                                        feat);
                } else {
#pragma unroll
                    for (int i = 0; i < kFeatures; i++) feat[i] = 0.0f;
                }
                const MlpInfer &W = *(const MlpInfer *)s_dyn;
                MlpWaveBuf &wb = ((MlpWaveBuf *)(s_dyn + kNetLds))[threadIdx.x >> 6];
                const float y = mlp_forward_groups<kTrio>(W, wb, feat);
                if (want_feat && isfinite(y)) {
                    // as the group kernel: growth at most facmax, never below
                    // facmin times the PI proposal
                    float lg = y + S.ctrl_shift;
                    const float lfmn = logf(float(P.fcmn)), hi = logf(float(P.fcmx));
                    const float lo = pi_ratio + lfmn > lfmn ? pi_ratio + lfmn : lfmn;
                    lg = lg < lo ? lo : lg;
                    lg = lg > hi ? hi : lg;
                    const double dtl = h_acc * double(__expf(lg));
                    dt = dtl < P.dtm ? dtl : P.dtm;
                }
            }
            want_feat = false;
        }

        bool finish = false;
        if (fresh) {
            fresh = false;
            if (!(flags & FLAG_STARTED)) {
                // integrations.F90:120-127 / :352-359
                dt = trio_dti_guess(P, METHOD == PTC_RK5 ? 5 : 8, base, si, f1, T);
                const double u3 = __shfl(si[1], base), u4 = __shfl(si[1], base + 1),
                             u5 = __shfl(si[1], base + 2);
                if (a == 0) {
                    double *o = S.out + (size_t)8 * P.m * idx;
                    o[6] = sqrt(u3 * u3 + u4 * u4 + u5 * u5 + 1.0);
                    o[7] = dt;
                }
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
                double sf[2], dtn = 0.0;
                if constexpr (METHOD == PTC_RK5) {
                    double k6[2];
                    const double del = trio_attempt_rk5(force, P, base, si, f1, t, dt, sf, k6);
                    c_rhs += 5;
                    if (!isfinite(del) || !trio_finite(sf, base)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                        // integrations.F90:177,186
                        const bool accept = del <= 1.0;
                        const double q =
                            P.safe * ptc_pow_with(del, accept ? -0.20 : -0.25, T.log, T.exp);
                        if (accept) {
                            t = t + dt;
                            if (n <= P.m && t >= tnext) {
                                double k1[2];
#pragma unroll
                                for (int h = 0; h < 2; h++) k1[h] = dt * double(f1[h]);
                                do {
                                    trio_hermite_out(
                                        S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), a,
                                        base, t, tnext, dt, si, sf, k1, k6);
                                    n++;
                                    tnext = n <= P.m ? S.time[n - 1] : 0.0;
                                } while (n <= P.m && t >= tnext);
                            }
                            si[0] = sf[0];
                            si[1] = sf[1];
                            if constexpr (CTRL != CTRL_PI)
                                on_accept<CTRL>(S, idx, flags, c_rej, dt, del, 5.0f, rej_since,
                                                want_feat, h_acc, err_acc, rej_before);
                            c_acc++;
                            dtn = dt * (P.fcmx < q ? P.fcmx : q);
                            need_rhs = true;
                        } else {
                            c_rej++;
                            rej_since = true;
                            dtn = dt * (P.fcmn > q ? P.fcmn : q);
                        }
                    }
                } else {
                    const double err = trio_attempt_dp853(force, P, base, si, f1, t, dt, sf);
                    c_rhs += 11;
                    if (!isfinite(err) || !trio_finite(sf, base)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                        // integrations.F90:466,486
                        const double qe = P.safe * ptc_pow_with(err, P.expo, T.log, T.exp);
                        if (err <= 1.0) {
                            t = t + dt;
                            if (n <= P.m && t >= tnext) {
                                K fe[2];
                                force.at(t);
                                force(sf, fe);
                                c_rhs++;
                                double ka[2], kb[2];
#pragma unroll
                                for (int h = 0; h < 2; h++) {
                                    ka[h] = dt * double(f1[h]);
                                    kb[h] = dt * double(fe[h]);
                                }
                                do {
                                    trio_hermite_out(
                                        S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), a,
                                        base, t, tnext, dt, si, sf, ka, kb);
                                    n++;
                                    tnext = n <= P.m ? S.time[n - 1] : 0.0;
                                } while (n <= P.m && t >= tnext);
                                f1[0] = fe[0];
                                f1[1] = fe[1];
                            } else {
                                need_rhs = true;
                            }
                            si[0] = sf[0];
                            si[1] = sf[1];
                            if constexpr (CTRL != CTRL_PI)
                                on_accept<CTRL>(S, idx, flags, c_rej, dt, err,
                                                float(-1.0 / P.expo), rej_since, want_feat, h_acc,
                                                err_acc, rej_before);
                            c_acc++;
                            double q = qe;
                            if (P.beta != 0.0) q = q * ptc_pow_with(errold, P.beta, T.log, T.exp);
                            if (q < P.fcmn) q = P.fcmn;
                            if (flags & FLAG_REJECTED) dtn = dt * (q < 1.0 ? q : 1.0);
                            else dtn = dt * (q < P.fcmx ? q : P.fcmx);
                            errold = err > 1.0e-04 ? err : 1.0e-04;
                            flags &= ~FLAG_REJECTED;
                        } else {
                            c_rej++;
                            rej_since = true;
                            dtn = dt * (P.fcmn > qe ? P.fcmn : qe);
                            flags |= FLAG_REJECTED;
                        }
                    }
                }
                if (!finish) {
                    c_all++;
                    dt = dtn < P.dtm ? dtn : P.dtm;
                    if constexpr (CTRL != CTRL_PI) {
                        if (want_feat) pi_ratio = float(log(dt / h_acc));
                    }
                    if (!(t < t_end)) {
                        if (!(t < P.tmx)) status = PTC_STATUS_DONE;
                        finish = true;
                    }
                }
            }
        }

        if (active && finish) {
            xp[idx] = si[0];
            up[idx] = si[1];
            if (a == 0) {
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
            }
            active = false;
            need_rhs = false;
            want_feat = false;
        }
    }
    // one count per trio (its lane 0); lane 63 never loads
    add_cell_loads(S, a == 0 && lane < kTrio * kTriosPerWave ? force.loads : 0u);
}

template <int METHOD, int MODE, int GEOM, int PREC, int ROWS, int CTRL = CTRL_PI, int TIME = 0>
hipError_t launch_trio_t(const DeviceFields &f, const StepParams &sp, const DeviceParticles &p,
                         double t_stop, int block, hipStream_t stream)
{
    auto kern = integrate_trio_kernel<METHOD, MODE, GEOM, PREC, ROWS, CTRL, TIME>;
    const size_t lds =
        CTRL == CTRL_LEARNED ? kNetLds + (size_t)(block / 64) * sizeof(MlpWaveBuf) : 0;
    if (CTRL == CTRL_LEARNED && !p.net) return hipErrorInvalidValue;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds);
    if (e != hipSuccess) return e;
    int64_t grid = (int64_t)cus * (per_cu > 0 ? per_cu : 1);
    const int64_t per_block = (int64_t)kTriosPerWave * (block / 64);
    const int64_t need = (p.np + per_block - 1) / per_block;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), lds, stream, f, sp, p, t_stop);
    return hipGetLastError();
}

}  // namespace
}  // namespace ptc
