// Group kernel: six lanes integrate one particle, lane c of the group
// holding component c (0..5) of the state and its slopes; ten groups per
// wave, lanes 60..63 idle.  Same arithmetic as integrate_kernel, per
// component, and every sum over components (error norms, the momentum's
// norm) is formed by every lane of the group from the six values shuffled
// across it, in the same order, so every lane of a group holds the same
// step-control state and the results are bit-identical to the lane-per-
// particle kernel and to the reference (integrations.F90:57-200, 224-505,
// 524-589; accelerations.F90:59-96; interpolations.F90:104-226).
//
// Why: one lane carrying a whole particle issues every component's
// arithmetic in sequence, so the slowest trajectory of a launch runs at one
// wave's issue rate and sets the launch time once the work queue is empty.
// A group issues about a quarter of the instructions per attempt on each of
// its lanes (per-particle work such as the cell lookup and the divisions by
// gamma is split over the group, not repeated), so those last trajectories
// finish sooner, and with under 256 registers per lane two waves share a
// SIMD (three for FP32 static fields and FP64 pair rows, group_waves), which
// hides the latency
// of gathers from fields too large for the caches.  The lane-per-particle kernel hands its particles over once the
// queue is exhausted and few are left in flight (DeviceParticles::park);
// this kernel then finishes them.  It can also run a whole launch.
//
// Static field or snapshot series in node rows, PI controller; FP64 fields
// (bit-identical) or FP32 nodes (mixed precision: FP32 weights, force and
// slopes, FP64 stage sums, as the mixed lane-per-particle kernel).  A static
// field of FP32 nodes can also take the learned step-size controller.
#pragma once
#include "integrator_impl.hpp"

namespace ptc {
namespace {

// lanes per particle, and particles per wave (lanes 60..63 stay idle)
constexpr int kGroup = 6;
constexpr int kGroupsPerWave = 64 / kGroup;

// component j of x across the group starting at lane base
template <class T>
__device__ __forceinline__ T gsh(T x, int base, int j) { return __shfl(x, base + j); }

// The group lanes that lane c (0..5) of a group reads for its component of
// the Lorentz force (accelerations.F90:84-92).  Component c + 1 of the
// force is, for c >= 3, q/m (E_k + (v_a B_d - v_d B_a)) in EB mode and
// q/m (w_a B_d - w_d B_a), w = v - U, in VB mode, with k = c - 3,
// a = (k + 1) % 3, d = (k + 2) % 3; for c < 3 it is v_c / vun.  Lanes k
// and k + 3 both hold v_k (the shared divisions) and w_k, and lane k + 3
// holds B_k, so lane c reads
//   own: lane k (its axis, for the stage point's position and E_k or U_k),
//   ra:  lane 3 + a (v_a, w_a, B_a), itself for c < 3 (v_c),
//   rd:  lane 3 + d (v_d, w_d, B_d),
// each lane exactly the values its own term uses, from three lanes fixed
// for the whole kernel.
struct ForceLanes {
    int own, ra, rd;
    __device__ __forceinline__ ForceLanes(int base, int c)
    {
        const int k = c < 3 ? c : c - 3;
        own = base + k;
        ra = c < 3 ? base + c : base + 3 + (k + 1) % 3;
        rd = base + 3 + (k + 2) % 3;
    }
};

// Component c of the Lorentz force (accelerations.F90:59-96, the operations
// of force_from_fields) from the momentum (u0, u1, u2) of the stage point,
// known on every lane, and the lane's component fc of the interpolated
// fields (U or E for c < 3, B for c >= 3).  The three divisions v = u /
// gamma are shared: lane c (and c + 3) of the group divides component
// c % 3, and each lane takes the quotients and field components its own
// term uses from the lanes ForceLanes names.
template <int MODE, int GEOM>
__device__ __forceinline__ double force_component(const StepParams &P, double u0, double u1,
                                                  double u2, double fc, int c,
                                                  const ForceLanes &L)
{
    const double g = sqrt(u0 * u0 + u1 * u1 + u2 * u2 + 1.0);
    const int c3 = c < 3 ? c : c - 3;
    const double q = (c3 == 0 ? u0 : c3 == 1 ? u1 : u2) / g;
    const double va = __shfl(q, L.ra), vd = __shfl(q, L.rd);
    const double lin = GEOM ? va * P.rvun : va / P.vun;
    const double bd = __shfl(fc, L.rd), ba = __shfl(fc, L.ra);
    double rot;
    if constexpr (MODE == PTC_MODE_EB) {
        // E + v x B
        const double ek = __shfl(fc, L.own);
        rot = P.qom * (ek + (va * bd - vd * ba));
    } else {
        // (v - U) x B: w_k on lanes k and k + 3
        const double w = q - __shfl(fc, L.own);
        const double wa = __shfl(w, L.ra), wd = __shfl(w, L.rd);
        rot = P.qom * (wa * bd - wd * ba);
    }
    return c < 3 ? lin : rot;
}

// Mixed precision: component c of the force in FP32 (the operations of
// force_from_fields_f32: FP32 momentum, hardware reciprocal square root of
// gamma squared); every lane forms the three velocity components itself.
template <int MODE>
__device__ __forceinline__ float force_component_f32(const StepParams &P, double u0, double u1,
                                                     double u2, float fc, int c,
                                                     const ForceLanes &L)
{
    const float ux = float(u0), uy = float(u1), uz = float(u2);
    const float rg = __builtin_amdgcn_rsqf(ux * ux + uy * uy + uz * uz + 1.0f);
    const float v0 = ux * rg, v1 = uy * rg, v2 = uz * rg;
    const float rvun = float(P.rvun), qom = float(P.qom);
    // v_a and v_d of component c (for c < 3, v_c in va)
    const bool a0 = c == 0 || c == 5, a1 = c == 1 || c == 3;
    const float va = a0 ? v0 : a1 ? v1 : v2;
    const float vd = c == 4 ? v0 : c == 5 ? v1 : v2;
    const float lin = va * rvun;
    const float bd = __shfl(fc, L.rd), ba = __shfl(fc, L.ra);
    float rot;
    if constexpr (MODE == PTC_MODE_EB) {
        const float ek = __shfl(fc, L.own);
        rot = qom * (ek + (va * bd - vd * ba));
    } else {
        // U_a and U_d from lanes a and d (3 below ra and rd), which hold
        // them as fc; w = v - U formed here
        const float ua = __shfl(fc, L.ra - 3), ud = __shfl(fc, L.rd - 3);
        rot = qom * ((va - ua) * bd - (vd - ud) * ba);
    }
    return c < 3 ? lin : rot;
}

// true if every component of x in the group is finite
__device__ __forceinline__ bool group_finite(double x, int base)
{
    const uint64_t bad = __ballot(!isfinite(x));
    return ((bad >> base) & 0x3full) == 0;
}

// corners() (interpolations.F90:140-149, :207-210) with the per-axis work
// shared: lane c (and c + 3) of the group computes axis c % 3 from pos, the
// stage point's position on that axis, the group exchanges the left and right indices
// and the in-cell fraction, and every lane forms the eight offsets and
// weights from them in corners()'s order (same operations, same results).
// xdo (if given): the in-cell fraction of the three axes.  ROWS =
// ROWS_COLUMNS: the offsets are rows of column storage (z fastest: strides
// dm3, dm1 dm3 and 1 along x, y and z) instead of nodes (x fastest).
template <class W, int GEOM, int ROWS = ROWS_NODES>
__device__ __forceinline__ void group_corners(const StepParams &P, double pos, int base,
                                              int c, CornersT<W> &cr, float *xdo = nullptr)
{
    const int a = c < 3 ? c : c - 3;
    const double xl = a == 0 ? P.xl[0] : a == 1 ? P.xl[1] : P.xl[2];
    const double inv_dm = a == 0 ? P.inv_dm[0] : a == 1 ? P.inv_dm[1] : P.inv_dm[2];
    const double dmd = a == 0 ? P.dmd[0] : a == 1 ? P.dmd[1] : P.dmd[2];
    const int dm = a == 0 ? P.dm[0] : a == 1 ? P.dm[1] : P.dm[2];
    double xs;
    if constexpr (GEOM) xs = a == 0 ? P.rxs[0] : a == 1 ? P.rxs[1] : P.rxs[2];
    else xs = a == 0 ? P.xs[0] : a == 1 ? P.xs[1] : P.xs[2];
    const double xa = GEOM ? (pos - xl) * xs : (pos - xl) / xs;
    const double xq = xa + inv_dm;
    double xp = dmd * (xq - floor(xq));
    const int r = wrap_idx((int)floor(xp), dm);
    xp = dmd * (xa - floor(xa));
    const double pl = floor(xp);
    const double xdm = xp - pl;
    const int l = wrap_idx((int)pl, dm);
    int li[3], ri[3];
    double xd[3], xu[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        li[j] = gsh(l, base, j);
        ri[j] = gsh(r, base, j);
        xd[j] = gsh(xdm, base, j);
        xu[j] = 1.0 - xd[j];
        if (xdo) xdo[j] = float(xd[j]);
    }
    uint32_t x0, x1, yl, yr, zl, zr;
    if constexpr (ROWS == ROWS_COLUMNS) {
        const uint32_t sx = (uint32_t)P.dm[2], sy = (uint32_t)P.dm[0] * (uint32_t)P.dm[2];
        x0 = (uint32_t)li[0] * sx; x1 = (uint32_t)ri[0] * sx;
        yl = (uint32_t)li[1] * sy; yr = (uint32_t)ri[1] * sy;
        zl = (uint32_t)li[2]; zr = (uint32_t)ri[2];
    } else {
        const uint32_t sy = (uint32_t)P.dm[0], sz = (uint32_t)P.dm[0] * (uint32_t)P.dm[1];
        yl = (uint32_t)li[1] * sy; yr = (uint32_t)ri[1] * sy;
        zl = (uint32_t)li[2] * sz; zr = (uint32_t)ri[2] * sz;
        x0 = (uint32_t)li[0]; x1 = (uint32_t)ri[0];
    }
    cr.o[0] = x0 + yl + zl; cr.o[1] = x0 + yl + zr; cr.o[2] = x0 + yr + zl; cr.o[3] = x0 + yr + zr;
    cr.o[4] = x1 + yl + zl; cr.o[5] = x1 + yl + zr; cr.o[6] = x1 + yr + zl; cr.o[7] = x1 + yr + zr;
    const W u0 = W(xu[0]), u1 = W(xu[1]), u2 = W(xu[2]);
    const W d0 = W(xd[0]), d1 = W(xd[1]), d2 = W(xd[2]);
    cr.w[0] = u0 * u1 * u2;
    cr.w[1] = u0 * u1 * d2;
    cr.w[2] = u0 * d1 * u2;
    cr.w[3] = u0 * d1 * d2;
    cr.w[4] = d0 * u1 * u2;
    cr.w[5] = d0 * u1 * d2;
    cr.w[6] = d0 * d1 * u2;
    cr.w[7] = d0 * d1 * d2;
}

// The snapshot interval of stage time ts in a series, as Force::interval:
// k with stime[k] <= ts < stime[k + 1], a = (ts - stime[k]) / (stime[k + 1] -
// stime[k]) and two = true; before or at the first snapshot time k = 0, at
// or after the last k = nsnap - 1, a = 0 and two = false.  kh holds the
// lane's last interval: the search walks from it (0 or 1 steps nearly
// always) and bisects otherwise.
__device__ __forceinline__ int series_interval(const DeviceFields &F, double ts, int &kh,
                                               double &a, bool &two)
{
    const int ns = F.nsnap;
    const double *__restrict__ st = F.stime;
    a = 0.0;
    two = false;
    if (ns <= 1 || !(ts > st[0])) return 0;
    if (ts >= st[ns - 1]) return ns - 1;
    int k = kh;
    if (!(st[k] <= ts && ts < st[k + 1])) {
        if (k + 2 < ns && st[k + 1] <= ts && ts < st[k + 2]) {
            k = k + 1;
        } else {
            int lo = 0, hi = ns - 1;  // st[lo] <= ts < st[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (st[mid] <= ts) lo = mid;
                else hi = mid;
            }
            k = lo;
        }
    }
    kh = k;
    a = (ts - st[k]) / (st[k + 1] - st[k]);
    two = true;
    return k;
}

// The RHS as the group evaluates it: each lane keeps component comp of the
// eight corner rows of the cell it last gathered from (the lane-per-particle
// kernel's cell cache, split over the group).  PREC = PTC_FP32: FP32 node
// rows of eight floats, FP32 weights, sums and force; the slope is a float.
// FEAT = 1 also keeps, from the last evaluation, component comp of the
// interpolated fields and the in-cell fraction (step-size features).
// ROWS = ROWS_PAIRS (static field in LAYOUT_PAIRS, FP32 or FP64): the four
// x-left corner rows also hold the x-right corners, read from them when the
// cell's right x index is the left one's periodic successor (else from their
// own rows).  ROWS = ROWS_QUADS (static FP32 field in LAYOUT_QUADS): the two
// (x-left, y-left) corner rows hold all four corners of their z plane, read
// from them where the right x and y indices are the left ones' successors.
// ROWS = ROWS_COLUMNS (a static field or a snapshot series in
// LAYOUT_COLUMNS): as quad rows, the offsets rows of the column storage
// (group_corners), so the two rows read are adjacent whenever the right z
// index is the left one's successor.
template <int MODE, int GEOM, int PREC, int TIME = 0, int FEAT = 0, int ROWS = ROWS_NODES>
struct GroupForce {
    static_assert(ROWS == ROWS_NODES || ROWS == ROWS_COLUMNS ||
                      (TIME == 0 && (ROWS == ROWS_PAIRS || (ROWS == ROWS_QUADS && PREC == PTC_FP32))),
                  "pair rows hold a static field, quad rows a static FP32 field");
    using K = std::conditional_t<PREC == PTC_FP32, float, double>;
    const DeviceFields &F;
    const StepParams &P;
    const int base, comp;
    uint32_t key0 = 0xffffffffu, key7 = 0xffffffffu;
    uint32_t loads = 0;  // cell loads of the group's particle (the same on each lane)
    K v[8];
    // TIME = 1 (a snapshot series, Force<..., TIME>): component comp of the
    // cell's corner rows in snapshot k + 1 as well, the stage time, the
    // lane's last interval and the cached interval key 2 k + (k + 1 loaded)
    K v1[TIME ? 8 : 1];
    double ts = 0.0;
    int kh = 0, kc = -1;
    K fint = K(0);
    float xdl[3] = {0.0f, 0.0f, 0.0f};

    const ForceLanes L;

    __device__ __forceinline__ GroupForce(const DeviceFields &F_, const StepParams &P_, int b, int c)
        : F(F_), P(P_), base(b), comp(c), L(b, c)
    {
    }

    __device__ __forceinline__ void at(double t)
    {
        if constexpr (TIME == 1) ts = t;
    }

    // the snapshot interval of the stage time (series_interval)
    __device__ __forceinline__ int interval(double &a, bool &two)
    {
        return series_interval(F, ts, kh, a, two);
    }

    // component comp of the eight corner rows of snapshot k at offsets o
    __device__ __forceinline__ void load(const uint32_t *o, int k, K *dst) const
    {
        if constexpr (PREC == PTC_FP32) {
            const float *__restrict__ d = (const float *)F.data + (size_t)k * F.sstride + comp;
#pragma unroll
            for (int j = 0; j < 8; j++) dst[j] = d[(size_t)o[j] * 8];
        } else {
            const double *__restrict__ d = (const double *)F.data + (size_t)k * F.sstride + comp;
#pragma unroll
            for (int j = 0; j < 8; j++) dst[j] = d[(size_t)o[j] * 6];
        }
    }

    // the same from pair rows (ROWS_PAIRS): the x-left corner rows hold the
    // lane's component of the corner and of its +x neighbour side by side
    // (one 8- or 16-byte load per row); those are the x-right corners when
    // the cell's right x index is the left one's periodic successor, else
    // the x-right corners come from their own rows
    __device__ __forceinline__ void load_pairs(const uint32_t *o, K *dst) const
    {
        using K2 = std::conditional_t<PREC == PTC_FP32, float2, double2>;
        const K2 *__restrict__ d2 = (const K2 *)F.data + comp;
        // x1 - x0 (mod 2^32) from the x-right and x-left offsets
        const uint32_t dx = o[4] - o[0];
        const bool succ = dx == 1u || dx == (uint32_t)(1 - P.dm[0]);
        K2 p[4];
#pragma unroll
        for (int j = 0; j < 4; j++) p[j] = d2[(size_t)o[j] * 8];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            dst[j] = p[j].x;
            dst[4 + j] = p[j].y;
        }
        if (!succ) {
            const K *__restrict__ d = (const K *)F.data + 2 * comp;
#pragma unroll
            for (int j = 0; j < 4; j++) dst[4 + j] = d[(size_t)o[4 + j] * 16];
        }
    }

    // the same from FP32 quad rows (ROWS_QUADS, or ROWS_COLUMNS in FP32: the
    // rows in column storage, of snapshot k): the (x-left, y-left) corner row
    // of each z plane holds the lane's component of its four corners side by
    // side (one 16-byte load per plane), corner (x, y) at x + 2 y, when the
    // right x and y indices are the left ones' periodic successors; else
    // every corner comes from its own row
    __device__ __forceinline__ void load_quads(const uint32_t *o, K *dst, int k = 0) const
    {
        // row strides along x and y (node order, or column storage)
        const uint32_t sx = ROWS == ROWS_COLUMNS ? (uint32_t)P.dm[2] : 1u;
        const uint32_t sy = (uint32_t)P.dm[0] * sx;
        const uint32_t dx = o[4] - o[0], dy = o[2] - o[0];
        const bool xs = dx == sx || dx == (uint32_t)((int64_t)(1 - P.dm[0]) * sx);
        const bool ys = dy == sy || dy == (uint32_t)((int64_t)(1 - P.dm[1]) * sy);
        const float *__restrict__ snap = (const float *)F.data + (size_t)k * F.sstride;
        const float4 *__restrict__ d4 = (const float4 *)snap + comp;
        // corner j: x = j >> 2, y = (j >> 1) & 1, z = j & 1
        const float4 q0 = d4[(size_t)o[0] * 8], q1 = d4[(size_t)o[1] * 8];
        dst[0] = q0.x; dst[1] = q1.x; dst[2] = q0.z; dst[3] = q1.z;
        dst[4] = q0.y; dst[5] = q1.y; dst[6] = q0.w; dst[7] = q1.w;
        if (!(xs && ys)) {
            const float *__restrict__ d = snap + 4 * comp;
#pragma unroll
            for (int j = 2; j < 8; j++) dst[j] = d[(size_t)o[j] * 32];
        }
    }

    // the same from column rows (ROWS_COLUMNS): the (x-left, y-left) corner
    // row of each z plane holds the lane's component of its four corners
    // side by side (two 16-byte loads per plane), corner (x, y) at x + 2 y,
    // when the right x and y indices are the left ones' periodic successors;
    // else every corner comes from its own row (snapshot k of a series)
    __device__ __forceinline__ void load_columns(const uint32_t *o, K *dst, int k = 0) const
    {
        const uint32_t sx = (uint32_t)P.dm[2], sy = (uint32_t)P.dm[0] * (uint32_t)P.dm[2];
        const uint32_t dx = o[4] - o[0], dy = o[2] - o[0];
        const bool xs = dx == sx || dx == (uint32_t)((int64_t)(1 - P.dm[0]) * P.dm[2]);
        const bool ys = dy == sy || dy == (uint32_t)((int64_t)(1 - P.dm[1]) * sy);
        const double *__restrict__ snap = (const double *)F.data + (size_t)k * F.sstride;
        const double2 *__restrict__ d2 = (const double2 *)snap + 2 * comp;
        // corner j: x = j >> 2, y = (j >> 1) & 1, z = j & 1
        const double2 a0 = d2[(size_t)o[0] * 12], a1 = d2[(size_t)o[0] * 12 + 1];
        const double2 b0 = d2[(size_t)o[1] * 12], b1 = d2[(size_t)o[1] * 12 + 1];
        dst[0] = a0.x; dst[1] = b0.x; dst[2] = a1.x; dst[3] = b1.x;
        dst[4] = a0.y; dst[5] = b0.y; dst[6] = a1.y; dst[7] = b1.y;
        if (!(xs && ys)) {
            const double *__restrict__ d = snap + 4 * comp;
#pragma unroll
            for (int j = 2; j < 8; j++) dst[j] = d[(size_t)o[j] * 24];
        }
    }

    // component comp of the cell's eight corner rows of snapshot k, in the
    // kernel's row form
    __device__ __forceinline__ void load_snapshot(const uint32_t *o, int k, K *dst) const
    {
        if constexpr (ROWS == ROWS_COLUMNS && PREC == PTC_FP32) load_quads(o, dst, k);
        else if constexpr (ROWS == ROWS_COLUMNS) load_columns(o, dst, k);
        else load(o, k, dst);
    }

    // component comp of f(s), s given by its component comp on each lane
    __device__ __forceinline__ K operator()(double s)
    {
        // the position on this lane's axis (comp % 3) and the momentum
        const double pos = __shfl(s, L.own);
        const double u0 = gsh(s, base, 3), u1 = gsh(s, base, 4), u2 = gsh(s, base, 5);
        CornersT<K> c;
        group_corners<K, GEOM, ROWS>(P, pos, base, comp, c, FEAT ? xdl : nullptr);
        K f;
        if constexpr (TIME == 1) {
            // the two bracketing snapshots' interpolants blended in time,
            // as Force::fields_at
            double a;
            bool two;
            const int k = interval(a, two);
            const int key = 2 * k + (two ? 1 : 0);
            if (!(c.o[0] == key0 && c.o[7] == key7 && (key == kc || (!two && key + 1 == kc)))) {
                loads += two ? 2u : 1u;
                load_snapshot(c.o, k, v);
                if (two) load_snapshot(c.o, k + 1, v1);
                key0 = c.o[0];
                key7 = c.o[7];
                kc = key;
            }
            f = c.w[0] * v[0];
#pragma unroll
            for (int j = 1; j < 8; j++) f = f + c.w[j] * v[j];
            if (two) {
                K g = c.w[0] * v1[0];
#pragma unroll
                for (int j = 1; j < 8; j++) g = g + c.w[j] * v1[j];
                const K ak = K(a);
                f = f + ak * (g - f);
            }
        } else {
            if (c.o[0] != key0 || c.o[7] != key7) {
                loads++;
                if constexpr (ROWS == ROWS_PAIRS) load_pairs(c.o, v);
                else if constexpr (ROWS == ROWS_QUADS) load_quads(c.o, v);
                else if constexpr (ROWS == ROWS_COLUMNS && PREC == PTC_FP32) load_quads(c.o, v);
                else if constexpr (ROWS == ROWS_COLUMNS) load_columns(c.o, v);
                else load(c.o, 0, v);
                key0 = c.o[0];
                key7 = c.o[7];
            }
            // interpolations.F90:212-220, corners summed left to right
            f = c.w[0] * v[0];
#pragma unroll
            for (int j = 1; j < 8; j++) f = f + c.w[j] * v[j];
        }
        if constexpr (FEAT == 1) fint = f;
        if constexpr (PREC == PTC_FP32)
            return force_component_f32<MODE>(P, u0, u1, u2, f, comp, L);
        else
            return force_component<MODE, GEOM>(P, u0, u1, u2, f, comp, L);
    }
};

// dti_guess (integrations.F90:524-589) with the component sums formed over
// the group; ki = f(si)
template <class K>
__device__ __forceinline__ double group_dti_guess(const StepParams &P, int iord, int base, double si,
                                                  K kik, const PowTables &T)
{
    const double ki = double(kik);
    const double sr = P.atol + P.rtol * fabs(si);
    const double q0 = si / sr, q1 = ki / sr;
    double s = 0.0, q, d0, d1, d2, dti, dtf;
#pragma unroll
    for (int j = 0; j < 6; j++) { q = gsh(q0, base, j); s = s + q * q; }
    d0 = sqrt(s / 6.0);
    s = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) { q = gsh(q1, base, j); s = s + q * q; }
    d1 = sqrt(s / 6.0);
    dti = (fmin(d0, d1) < 1.0e-10) ? 1.0e-06 : 1.0e-01 * (d0 / d1);
    const double sf = si + dti * ki;
    const double q2 = (sf - si) / sr;
    s = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) { q = gsh(q2, base, j); s = s + q * q; }
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

// Cash-Karp attempt (integrations.F90:131-148) on one component; returns del
template <class GF>
__device__ __forceinline__ double group_attempt_rk5(GF &rhsf, const StepParams &P, int base,
                                                    double si, typename GF::K f1, double t,
                                                    double dt, double &sf, double &k6d)
{
    using K = typename GF::K;
    const double k1 = double(K(dt * double(f1)));
    double s = si + CK_A21 * k1;
    rhsf.at(t + CK_N2 * dt);
    const double k2 = double(K(dt * double(rhsf(s))));
    s = si + CK_A31 * k1 + CK_A32 * k2;
    rhsf.at(t + CK_N3 * dt);
    const double k3 = double(K(dt * double(rhsf(s))));
    s = si + CK_A41 * k1 + CK_A42 * k2 + CK_A43 * k3;
    rhsf.at(t + CK_N4 * dt);
    const double k4 = double(K(dt * double(rhsf(s))));
    s = si + CK_A51 * k1 + CK_A52 * k2 + CK_A53 * k3 + CK_A54 * k4;
    rhsf.at(t + dt);
    const double k5 = double(K(dt * double(rhsf(s))));
    s = si + CK_A61 * k1 + CK_A62 * k2 + CK_A63 * k3 + CK_A64 * k4 + CK_A65 * k5;
    rhsf.at(t + CK_N6 * dt);
    const double k6 = double(K(dt * double(rhsf(s))));

    const double eps = 2.2250738585072014e-308;  // tiny(1d0), integrations.F90:99
    double sc = 0.0;
    sc = sc + CK_C1 * k1;
    sc = sc + 0.0 * k2;
    sc = sc + CK_C3 * k3;
    sc = sc + CK_C4 * k4;
    sc = sc + 0.0 * k5;
    sc = sc + CK_C6 * k6;
    sf = si + sc;
    double se = 0.0;
    se = se + CK_D1 * k1;
    se = se + 0.0 * k2;
    se = se + CK_D3 * k3;
    se = se + CK_D4 * k4;
    se = se + CK_D5 * k5;
    se = se + CK_D6 * k6;
    const double qc = fabs(se / (fabs(si) + fabs(k1) + eps));
    double del = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double q = gsh(qc, base, j);
        if (j == 0 || q > del) del = q;
    }
    k6d = k6;
    return del / P.atol;
}

// DOP853 attempt (integrations.F90:365-437) on one component; k1 = f(si);
// returns err
template <class GF>
__device__ __forceinline__ double group_attempt_dp853(GF &rhsf, const StepParams &P, int base,
                                                      double si, typename GF::K k1k, double t,
                                                      double dt, double &sf)
{
    const double k1 = double(k1k);
    double ss = si + dt * A0201 * k1;
    rhsf.at(t + DN02 * dt);
    const double k2 = double(rhsf(ss));
    ss = si + dt * (A0301 * k1 + A0302 * k2);
    rhsf.at(t + DN03 * dt);
    const double k3 = double(rhsf(ss));
    ss = si + dt * (A0401 * k1 + A0403 * k3);
    rhsf.at(t + DN04 * dt);
    const double k4 = double(rhsf(ss));
    ss = si + dt * (A0501 * k1 + A0503 * k3 + A0504 * k4);
    rhsf.at(t + DN05 * dt);
    const double k5 = double(rhsf(ss));
    ss = si + dt * (A0601 * k1 + A0604 * k4 + A0605 * k5);
    rhsf.at(t + DN06 * dt);
    const double k6 = double(rhsf(ss));
    ss = si + dt * (A0701 * k1 + A0704 * k4 + A0705 * k5 + A0706 * k6);
    rhsf.at(t + DN07 * dt);
    const double k7 = double(rhsf(ss));
    ss = si + dt * (A0801 * k1 + A0804 * k4 + A0805 * k5 + A0806 * k6 + A0807 * k7);
    rhsf.at(t + DN08 * dt);
    const double k8 = double(rhsf(ss));
    ss = si + dt * (A0901 * k1 + A0904 * k4 + A0905 * k5 + A0906 * k6 + A0907 * k7 + A0908 * k8);
    rhsf.at(t + DN09 * dt);
    const double k9 = double(rhsf(ss));
    ss = si + dt * (A1001 * k1 + A1004 * k4 + A1005 * k5 + A1006 * k6 + A1007 * k7 + A1008 * k8 +
                    A1009 * k9);
    rhsf.at(t + DN10 * dt);
    const double k10 = double(rhsf(ss));
    ss = si + dt * (A1101 * k1 + A1104 * k4 + A1105 * k5 + A1106 * k6 + A1107 * k7 + A1108 * k8 +
                    A1109 * k9 + A1110 * k10);
    rhsf.at(t + DN11 * dt);
    const double k11 = double(rhsf(ss));
    ss = si + dt * (A1201 * k1 + A1204 * k4 + A1205 * k5 + A1206 * k6 + A1207 * k7 + A1208 * k8 +
                    A1209 * k9 + A1210 * k10 + A1211 * k11);
    rhsf.at(t + dt);
    const double k12 = double(rhsf(ss));

    const double kb = DB01 * k1 + DB06 * k6 + DB07 * k7 + DB08 * k8 + DB09 * k9 + DB10 * k10 +
                      DB11 * k11 + DB12 * k12;
    sf = si + dt * kb;
    const double a = fabs(si), b = fabs(sf);
    const double sr = P.atol + P.rtol * (a > b ? a : b);
    const double e3 = (kb - BH1 * k1 - BH2 * k9 - BH3 * k12) / sr;
    const double e5 = (ER01 * k1 + ER06 * k6 + ER07 * k7 + ER08 * k8 + ER09 * k9 + ER10 * k10 +
                       ER11 * k11 + ER12 * k12) / sr;
    double err3 = 0.0, err5 = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double q = gsh(e3, base, j);
        err3 = err3 + q * q;
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
        const double q = gsh(e5, base, j);
        err5 = err5 + q * q;
    }
    const double deno = err5 + 1.0e-02 * err3;
    return deno <= 0.0 ? fabs(dt) * err5 / sqrt(6.0) : fabs(dt) * err5 / sqrt(6.0 * deno);
}

// cubic Hermite snapshot (integrations.F90:158-168 / :447-457) of one
// component; the group's lanes 0..5 write components 1..6 of state(:, n),
// lane 0 also gamma and dt
__device__ __forceinline__ void group_hermite_out(double *st, int comp, int cl, int base, double t,
                                                  double tn, double dt, double si, double sf,
                                                  double ka, double kb)
{
    const double fr = (t - tn) / dt, fl = 1.0 - fr;
    const double ds = sf - si;
    const double o = fr * si + fl * sf + fl * fr * ((ka - ds) * fr + (ds - kb) * fl);
    const double o3 = gsh(o, base, 3), o4 = gsh(o, base, 4), o5 = gsh(o, base, 5);
    st[comp] = o;
    if (cl == 0) {
        st[6] = sqrt(o3 * o3 + o4 * o4 + o5 * o5 + 1.0);
        st[7] = dt;
    }
}

// The step-size features (features_bc) of the group's accepted state si,
// on every lane of the group, from the force's last evaluation, which was
// at si: the interpolated fields and the in-cell fraction it kept, and |B|
// at the eight corners from the B components held by lanes 3..5.
template <int MODE, class GF>
__device__ __forceinline__ void group_features(const GF &force, const StepParams &P, double si,
                                               int base, double h, float err, float pi_ratio,
                                               bool rej, float *f)
{
    const double s[6] = {0.0, 0.0, 0.0, gsh(si, base, 3), gsh(si, base, 4), gsh(si, base, 5)};
    float ub[6], bc[8];
#pragma unroll
    for (int j = 0; j < 6; j++) ub[j] = float(gsh(force.fint, base, j));
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const float b = float(force.v[j]);
        bc[j] = corner_bmag(gsh(b, base, 3), gsh(b, base, 4), gsh(b, base, 5));
    }
    features_bc<MODE>(P, s, ub, bc, force.xdl, h, err, pi_ratio, rej, f);
}

// Waves per SIMD the kernel is compiled for.  FP32 static fields with the
// PI controller: three (168 registers, 12-44 B of spills on quad rows),
// 1.04-1.24x the accepted steps per second of two (1024^3 DOP853 +15 %,
// 512^3 RK5 +22 %).  FP64 static fields in pair rows with the PI
// controller: three as well (168 registers; RK5 spills 76 B, DOP853 124 B,
// mostly loop-invariant constants reloaded once per attempt), 6.5 % faster
// than two at 1024^3 (DOP853, 4,000,000 particles); column rows alike.  The other forms need 180-230 registers
// and run two (FP64 node rows forced to three measured within 1.3 % of two;
// keeping the FP64 cell cache in LDS instead of registers, to cut the
// spills, was 5 % slower).
constexpr int group_waves(int prec, int time, int ctrl, int rows)
{
    return time == 0 && ctrl == CTRL_PI &&
                   (prec == PTC_FP32 || rows == ROWS_PAIRS || rows == ROWS_COLUMNS)
               ? 3
               : 1;
}

// The same particle loop as integrate_kernel (static field or snapshot
// series in node rows), one particle per group of six lanes, particles
// taken from the second work-queue head S.qcount[1].  Particles whose
// status is not running, or that were started and already reached the end
// of the segment, are skipped.  CTRL_PI: the reference's controller;
// CTRL_LEARNED (static FP32 nodes): after an accepted step the network
// proposes the next step, as in integrate_kernel, evaluated once per wave
// for its ten particles (mlp_forward_groups; the network and a 64x72 BF16
// buffer per wave in dynamic LDS).
template <int METHOD, int MODE, int GEOM, int PREC, int TIME, int CTRL = CTRL_PI,
          int ROWS = ROWS_NODES>
__global__ void __launch_bounds__(kBlock, group_waves(PREC, TIME, CTRL, ROWS))
integrate_group_kernel(const DeviceFields F, const StepParams P, const DeviceParticles S,
                       double t_stop)
{
    static_assert(CTRL == CTRL_PI || (CTRL == CTRL_LEARNED && TIME == 0 && PREC == PTC_FP32),
                  "the group kernel's learned controller takes static FP32 node rows");
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
    const int base = (lane / kGroup) * kGroup, cl = lane - base;
    const int comp = cl;
    const double t_end = t_stop < P.tmx ? t_stop : P.tmx;
    const int64_t N = S.np;
    GroupForce<MODE, GEOM, PREC, TIME, CTRL == CTRL_LEARNED ? 1 : 0, ROWS> force(F, P, base,
                                                                                comp);
    using K = typename decltype(force)::K;
    // this lane's state plane, chosen once (indexing the kernel argument's
    // pointer array by a per-lane value would copy it to scratch)
    double *xc = S.x[0];
    xc = comp == 1 ? S.x[1] : xc;
    xc = comp == 2 ? S.x[2] : xc;
    xc = comp == 3 ? S.x[3] : xc;
    xc = comp == 4 ? S.x[4] : xc;
    xc = comp == 5 ? S.x[5] : xc;

    // the lanes past the last whole group never take a particle
    bool active = false, exhausted = lane >= kGroup * kGroupsPerWave, need_rhs = false,
         fresh = false;
    int64_t idx = 0;
    double si = 0.0;
    K f1 = K(0);
    double t = 0.0, dt = 0.0, errold = 1.0e-04, tnext = 0.0;
    int32_t n = 2, flags = 0, status = PTC_STATUS_RUNNING;
    int64_t c_all = 0, c_acc = 0, c_rej = 0, c_rhs = 0;
    // learned controller: features are taken at the top of the iteration
    // after an accepted step of size h_acc with error ratio err_acc
    bool want_feat = false, rej_since = false, rej_before = false;
    double h_acc = 0.0;
    float err_acc = 0.0f, pi_ratio = 0.0f;

    for (;;) {
        // ---- refill groups without a particle (one atomic per wave) ----
        const bool want = !active && !exhausted;
        const uint64_t wmask = __ballot(want && cl == 0);
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
                        si = xc[idx];
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
            f1 = force(si);
            c_rhs++;
            need_rhs = false;
        }

        if constexpr (CTRL == CTRL_LEARNED) {
            // (the force's last evaluation was at si: the one above, or the
            // end slope of the accepted step, which became si)
            if (__any(want_feat)) {
                float feat[kFeatures];
                if (want_feat) {
                    group_features<MODE>(force, P, si, base, h_acc, err_acc, pi_ratio, rej_before,
                                         feat);
                } else {
#pragma unroll
                    for (int i = 0; i < kFeatures; i++) feat[i] = 0.0f;
                }
                const MlpInfer &W = *(const MlpInfer *)s_dyn;
                MlpWaveBuf &wb = ((MlpWaveBuf *)(s_dyn + kNetLds))[threadIdx.x >> 6];
                const float y = mlp_forward_groups(W, wb, feat);
                if (want_feat && isfinite(y)) {
                    // as integrate_kernel: growth at most facmax, never below
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
                dt = group_dti_guess(P, METHOD == PTC_RK5 ? 5 : 8, base, si, f1, T);
                const double u3 = gsh(si, base, 3), u4 = gsh(si, base, 4), u5 = gsh(si, base, 5);
                if (cl == 0) {
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
                double sf = 0.0, dtn = 0.0;
                if constexpr (METHOD == PTC_RK5) {
                    double k6 = 0.0;
                    const double del = group_attempt_rk5(force, P, base, si, f1, t, dt, sf, k6);
                    c_rhs += 5;
                    if (!isfinite(del) || !group_finite(sf, base)) {
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
                                const double k1 = dt * double(f1);
                                do {
                                    group_hermite_out(
                                        S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), comp,
                                        cl, base, t, tnext, dt, si, sf, k1, k6);
                                    n++;
                                    tnext = n <= P.m ? S.time[n - 1] : 0.0;
                                } while (n <= P.m && t >= tnext);
                            }
                            si = sf;
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
                    const double err = group_attempt_dp853(force, P, base, si, f1, t, dt, sf);
                    c_rhs += 11;
                    if (!isfinite(err) || !group_finite(sf, base)) {
                        status = PTC_STATUS_NONFINITE;
                        finish = true;
                    } else {
                        // integrations.F90:466,486
                        const double qe = P.safe * ptc_pow_with(err, P.expo, T.log, T.exp);
                        if (err <= 1.0) {
                            t = t + dt;
                            if (n <= P.m && t >= tnext) {
                                force.at(t);
                                const K fe = force(sf);
                                c_rhs++;
                                const double ka = dt * double(f1), kb = dt * double(fe);
                                do {
                                    group_hermite_out(
                                        S.out + (size_t)8 * ((n - 1) + (size_t)P.m * idx), comp,
                                        cl, base, t, tnext, dt, si, sf, ka, kb);
                                    n++;
                                    tnext = n <= P.m ? S.time[n - 1] : 0.0;
                                } while (n <= P.m && t >= tnext);
                                f1 = fe;
                            } else {
                                need_rhs = true;
                            }
                            si = sf;
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
            xc[idx] = si;
            if (cl == 0) {
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
    // one count per group (its lane 0); lanes 60..63 never load
    add_cell_loads(S, cl == 0 && lane < kGroup * kGroupsPerWave ? force.loads : 0u);
}

template <int METHOD, int MODE, int GEOM, int PREC, int TIME = 0, int CTRL = CTRL_PI,
          int ROWS = ROWS_NODES>
hipError_t launch_group_t(const DeviceFields &f, const StepParams &sp, const DeviceParticles &p,
                          double t_stop, int block, hipStream_t stream)
{
    auto kern = integrate_group_kernel<METHOD, MODE, GEOM, PREC, TIME, CTRL, ROWS>;
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
    const int64_t per_block = (int64_t)kGroupsPerWave * (block / 64);
    const int64_t need = (p.np + per_block - 1) / per_block;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), lds, stream, f, sp, p, t_stop);
    return hipGetLastError();
}

}  // namespace
}  // namespace ptc
