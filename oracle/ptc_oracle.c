/*
 * ptc_oracle.c -- CPU oracle (FP64) for the particle-integration hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, by the smoke check and by the
 * benchmark's CPU-baseline leg, never by the product.  Each function restates
 * the reference algorithm it cites, in the reference's operation order.
 * Build with -O2 -ffp-contract=off (no FMA contraction, no -ffast-math) so
 * every rounding step matches the reference's own FP64 build; the only libm
 * call on the path is pow(), as in the reference (Fortran real ** real).
 *
 * The guards (max_steps, non-finite error norm, dt that no longer advances t)
 * do not exist in the reference; they only fire where the reference would
 * loop forever (NaN fields, dtmax <= 0, ...), so finite runs are unaffected.
 */
#include "ptc_oracle.h"

/* ORC_PLANT != 0 builds the oracle with a known mistake in the EB-mode
 * force (tests/test_eb_closed_form.py compiles such copies into a temporary
 * directory to show that its closed-form checks catch each one); the
 * library built by the Makefile leaves it 0. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

#include <math.h>
#include <float.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Fortran FLOOR(real) -> default integer (sources/interpolations.F90:142-147) */
static inline int ffloor(double x) { return (int)floor(x); }

/* offset of f(i,j,k,c), 1-based i,j,k (Fortran), 0-based component c */
static inline int64_t fidx(const int dm[3], int i, int j, int k, int c)
{
    return (int64_t)(i - 1) +
           (int64_t)dm[0] * ((int64_t)(j - 1) +
                             (int64_t)dm[1] * ((int64_t)(k - 1) + (int64_t)dm[2] * c));
}

/* The reference indexes l = dm+1 / r = dm+1 when x - floor(x) rounds up to
 * 1.0 (x = -tiny), reading past the array; here such an index wraps to 1.
 * Indices outside [1, n] otherwise only come from non-finite positions and
 * are clamped to 1 (the reference reads out of bounds there too). */
static inline int wrap1(int i, int n)
{
    i = i > n ? i - n : i;
    return (i >= 1 && i <= n) ? i : 1;
}

/* sources/interpolations.F90:104-226 */
void orc_interpolate_lin(const int dm[3], const double xl[3], const double xs[3],
                         const double *uu, const double *bb, const double pos[3],
                         double u[3], double b[3])
{
    double x[3], xd[3], xp[3], xu[3], c[8];
    int l[3], r[3], pl[3];

    for (int i = 0; i < 3; i++) x[i] = (pos[i] - xl[i]) / xs[i];          /* :140 */
    for (int i = 0; i < 3; i++) {                                          /* :141-143 */
        xd[i] = x[i] + 1.0 / (double)dm[i];
        xp[i] = (double)dm[i] * (xd[i] - (double)ffloor(xd[i]));
        r[i] = wrap1(ffloor(xp[i]) + 1, dm[i]);
    }
    for (int i = 0; i < 3; i++) {                                          /* :145-149 */
        xp[i] = (double)dm[i] * (x[i] - (double)ffloor(x[i]));
        pl[i] = ffloor(xp[i]);
        xd[i] = xp[i] - (double)pl[i];
        xu[i] = 1.0 - xd[i];
        l[i] = wrap1(pl[i] + 1, dm[i]);
    }

    /* corner weights, x slowest ... z fastest (:207-210) */
    c[0] = xu[0] * xu[1] * xu[2];
    c[1] = xu[0] * xu[1] * xd[2];
    c[2] = xu[0] * xd[1] * xu[2];
    c[3] = xu[0] * xd[1] * xd[2];
    c[4] = xd[0] * xu[1] * xu[2];
    c[5] = xd[0] * xu[1] * xd[2];
    c[6] = xd[0] * xd[1] * xu[2];
    c[7] = xd[0] * xd[1] * xd[2];

    /* weighted corner sums, left to right (:212-220) */
    for (int m = 0; m < 3; m++) {
        u[m] = c[0] * uu[fidx(dm, l[0], l[1], l[2], m)] + c[1] * uu[fidx(dm, l[0], l[1], r[2], m)]
             + c[2] * uu[fidx(dm, l[0], r[1], l[2], m)] + c[3] * uu[fidx(dm, l[0], r[1], r[2], m)]
             + c[4] * uu[fidx(dm, r[0], l[1], l[2], m)] + c[5] * uu[fidx(dm, r[0], l[1], r[2], m)]
             + c[6] * uu[fidx(dm, r[0], r[1], l[2], m)] + c[7] * uu[fidx(dm, r[0], r[1], r[2], m)];
    }
    for (int m = 0; m < 3; m++) {
        b[m] = c[0] * bb[fidx(dm, l[0], l[1], l[2], m)] + c[1] * bb[fidx(dm, l[0], l[1], r[2], m)]
             + c[2] * bb[fidx(dm, l[0], r[1], l[2], m)] + c[3] * bb[fidx(dm, l[0], r[1], r[2], m)]
             + c[4] * bb[fidx(dm, r[0], l[1], l[2], m)] + c[5] * bb[fidx(dm, r[0], l[1], r[2], m)]
             + c[6] * bb[fidx(dm, r[0], r[1], l[2], m)] + c[7] * bb[fidx(dm, r[0], r[1], r[2], m)];
    }
}

/* the force of sources/accelerations.F90:84-92 from interpolated fields u
 * (U, or E in EB mode) and b at vec */
static void lorentz(int mode, double qom, double vun, const double u[3], const double b[3],
                    const double vec[6], double acc[6])
{
    double v[3], w[3], g;

    g = sqrt(vec[3] * vec[3] + vec[4] * vec[4] + vec[5] * vec[5] + 1.0);  /* :84 */
    v[0] = vec[3] / g;                                                     /* :85 */
    v[1] = vec[4] / g;
    v[2] = vec[5] / g;

    acc[0] = v[0] / vun;                                                   /* :89 */
    acc[1] = v[1] / vun;
    acc[2] = v[2] / vun;
    if (mode == PTC_MODE_EB) {
#if ORC_PLANT == 1
        /* planted mistake (tests only): v x B with the wrong sign */
        acc[3] = qom * (u[0] - (v[1] * b[2] - v[2] * b[1]));
        acc[4] = qom * (u[1] - (v[2] * b[0] - v[0] * b[2]));
        acc[5] = qom * (u[2] - (v[0] * b[1] - v[1] * b[0]));
#elif ORC_PLANT == 2
        /* planted mistake (tests only): two components of v x B swapped */
        acc[3] = qom * (u[0] + (v[2] * b[0] - v[0] * b[2]));
        acc[4] = qom * (u[1] + (v[1] * b[2] - v[2] * b[1]));
        acc[5] = qom * (u[2] + (v[0] * b[1] - v[1] * b[0]));
#elif ORC_PLANT == 3
        /* planted mistake (tests only): E taken with the wrong sign */
        acc[3] = qom * (-u[0] + (v[1] * b[2] - v[2] * b[1]));
        acc[4] = qom * (-u[1] + (v[2] * b[0] - v[0] * b[2]));
        acc[5] = qom * (-u[2] + (v[0] * b[1] - v[1] * b[0]));
#else
        acc[3] = qom * (u[0] + (v[1] * b[2] - v[2] * b[1]));
        acc[4] = qom * (u[1] + (v[2] * b[0] - v[0] * b[2]));
        acc[5] = qom * (u[2] + (v[0] * b[1] - v[1] * b[0]));
#endif
    } else {
        w[0] = v[0] - u[0];                                                /* :87 */
        w[1] = v[1] - u[1];
        w[2] = v[2] - u[2];
        acc[3] = qom * (w[1] * b[2] - w[2] * b[1]);                        /* :90-92 */
        acc[4] = qom * (w[2] * b[0] - w[0] * b[2]);
        acc[5] = qom * (w[0] * b[1] - w[1] * b[0]);
    }
}

/* sources/accelerations.F90:59-96 (VB mode); EB mode replaces (v-U) x B by
 * E + v x B with E taken from the first field array. */
void orc_acceleration(int mode, double qom, double vun, const int dm[3],
                      const double xl[3], const double xs[3], const double *uu,
                      const double *bb, const double vec[6], double acc[6])
{
    double u[3], b[3];

    orc_interpolate_lin(dm, xl, xs, uu, bb, vec, u, b);                    /* :82 */
    lorentz(mode, qom, vun, u, b, vec, acc);
}

/* Snapshot interval of time t in a series of ns snapshots at strictly
 * increasing times st (no counterpart in the reference, whose field is
 * static): *k with st[k] <= t < st[k+1] and *a = (t - st[k]) / (st[k+1] -
 * st[k]), *two = 1; at or before st[0] (or t NaN) snapshot 0 alone, at or
 * after st[ns-1] snapshot ns-1 alone (*two = 0).  Bisection. */
static void series_interval(int ns, const double *st, double t, int *k, double *a, int *two)
{
    int lo = 0, hi = ns - 1;
    *a = 0.0;
    *two = 0;
    if (ns <= 1 || !(t > st[0])) { *k = 0; return; }
    if (t >= st[ns - 1]) { *k = ns - 1; return; }
    while (hi - lo > 1) {           /* st[lo] <= t < st[hi] */
        const int mid = lo + (hi - lo) / 2;
        if (st[mid] <= t) lo = mid;
        else hi = mid;
    }
    *k = lo;
    *a = (t - st[lo]) / (st[lo + 1] - st[lo]);
    *two = 1;
}

void orc_interpolate_series(const int dm[3], const double xl[3], const double xs[3], int ns,
                            const double *st, const double *uu, const double *bb,
                            const double pos[3], double t, double u[3], double b[3])
{
    const int64_t n3 = (int64_t)3 * dm[0] * dm[1] * dm[2];
    int k, two;
    double a, u2[3], b2[3];

    series_interval(ns, st, t, &k, &a, &two);
    orc_interpolate_lin(dm, xl, xs, uu + k * n3, bb + k * n3, pos, u, b);
    if (two) {
        orc_interpolate_lin(dm, xl, xs, uu + (k + 1) * n3, bb + (k + 1) * n3, pos, u2, b2);
        for (int i = 0; i < 3; i++) {
            u[i] = u[i] + a * (u2[i] - u[i]);
            b[i] = b[i] + a * (b2[i] - b[i]);
        }
    }
}

void orc_acceleration_series(int mode, double qom, double vun, const int dm[3],
                             const double xl[3], const double xs[3], int ns, const double *st,
                             const double *uu, const double *bb, const double vec[6], double t,
                             double acc[6])
{
    double u[3], b[3];

    orc_interpolate_series(dm, xl, xs, ns, st, uu, bb, vec, t, u, b);
    lorentz(mode, qom, vun, u, b, vec, acc);
}

/* The field one integration runs in: a static snapshot (ns = 0, the
 * reference's case) or a series of ns snapshots at times st, uu / bb then
 * holding ns consecutive (dm1, dm2, dm3, 3) arrays. */
typedef struct {
    int mode, ns;
    const int *dm;
    const double *xl, *xs, *st, *uu, *bb;
    double qom, vun;
} fsrc;

static fsrc fsrc_make(int mode, const int dm[3], const double *params, int ns, const double *st,
                      const double *uu, const double *bb)
{
    fsrc F;
    F.mode = mode;
    F.ns = ns;
    F.dm = dm;
    F.xl = params + 11;
    F.xs = params + 14;
    F.st = st;
    F.uu = uu;
    F.bb = bb;
    F.qom = params[0];
    F.vun = params[1];
    return F;
}

/* the RHS at vec and time t (t is unused for a static field) */
static void fsrc_rhs(const fsrc *F, double t, const double vec[6], double acc[6])
{
    if (F->ns == 0)
        orc_acceleration(F->mode, F->qom, F->vun, F->dm, F->xl, F->xs, F->uu, F->bb, vec, acc);
    else
        orc_acceleration_series(F->mode, F->qom, F->vun, F->dm, F->xl, F->xs, F->ns, F->st,
                                F->uu, F->bb, vec, t, acc);
}

/* sources/integrations.F90:524-589.  The reference also evaluates the RHS at
 * sf (:576) but never uses it; that evaluation is omitted (no side effects). */
static double dti_guess_f(const fsrc *F, int iord, const double *params, const double si[6])
{
    const double dtm = params[4];
    const double atol = params[5], rtol = params[6];
    double ki[6], sr[6], sf[6], d0, d1, d2, dti, dtf, s, q;

    fsrc_rhs(F, 0.0, si, ki);                                              /* :563 */
    for (int i = 0; i < 6; i++) sr[i] = atol + rtol * fabs(si[i]);         /* :565 */

    s = 0.0;
    for (int i = 0; i < 6; i++) { q = si[i] / sr[i]; s += q * q; }
    d0 = sqrt(s / 6.0);                                                    /* :567 */
    s = 0.0;
    for (int i = 0; i < 6; i++) { q = ki[i] / sr[i]; s += q * q; }
    d1 = sqrt(s / 6.0);                                                    /* :568 */
    if (fmin(d0, d1) < 1.0e-10) dti = 1.0e-06;                             /* :569-573 */
    else dti = 1.0e-01 * (d0 / d1);

    for (int i = 0; i < 6; i++) sf[i] = si[i] + dti * ki[i];               /* :575 */
    s = 0.0;
    for (int i = 0; i < 6; i++) { q = (sf[i] - si[i]) / sr[i]; s += q * q; }
    d2 = sqrt(s / 6.0) / dti;                                              /* :577 */
    if (d1 > d2) d2 = d1;
    if (d2 <= 1.0e-15) {                                                   /* :578-582 */
        dtf = 1.0e-03 * dti;
        if (dtf < 1.0e-06) dtf = 1.0e-06;
    } else {
        dtf = pow(1.0e-02 / d2, 1.0 / (double)iord);
    }
    dti = 1.0e+02 * dti;                                                   /* :583 */
    if (dtf < dti) dti = dtf;
    if (dtm < dti) dti = dtm;
    return dti;
}

double orc_dti_guess(int iord, int mode, const int dm[3], const double *params,
                     const double *uu, const double *bb, const double si[6])
{
    const fsrc F = fsrc_make(mode, dm, params, 0, NULL, uu, bb);
    return dti_guess_f(&F, iord, params, si);
}

/* Cash-Karp 1990 tableau (Numerical Recipes 2nd ed. sec. 16.2): stage
 * matrix, 5th-order weights and 5th-minus-4th weights.  The reference holds
 * the same rationals over common denominators (integrations.F90:86-98);
 * single divisions of exact integers round to the same doubles. */
static const double CK_A[6][5] = {
    {0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
    {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0},
    {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0},
    {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
static const double CK_C[6] = {37.0 / 378.0, 0.0, 250.0 / 621.0, 125.0 / 594.0, 0.0,
                               512.0 / 1771.0};
static const double CK_D[6] = {-277.0 / 64512.0, 0.0, 6925.0 / 370944.0, -6925.0 / 202752.0,
                               -277.0 / 14336.0, 277.0 / 7084.0};
/* stage nodes (row sums of CK_A): stage j at t + CK_N[j] dt, for a series */
static const double CK_N[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 3.0 / 5.0, 1.0, 7.0 / 8.0};

static inline void hermite_out(double *st, double t, double tn, double dt, const double si[6],
                               const double sf[6], const double ka[6], const double kb[6])
{
    /* cubic Hermite between si (fr = 1) and sf (fr = 0) with end slopes
     * ka, kb already multiplied by dt (integrations.F90:159-166,448-455) */
    double fr = (t - tn) / dt, fl = 1.0 - fr, ds, s2 = 0.0;
    for (int i = 0; i < 6; i++) {
        ds = sf[i] - si[i];
        st[i] = fr * si[i] + fl * sf[i] + fl * fr * ((ka[i] - ds) * fr + (ds - kb[i]) * fl);
    }
    s2 = st[3] * st[3] + st[4] * st[4] + st[5] * st[5];
    st[6] = sqrt(s2 + 1.0);
    st[7] = dt;
}

static inline int all_finite6(const double *a)
{
    for (int i = 0; i < 6; i++)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* sources/integrations.F90:57-200; in a series stage j is evaluated at
 * t + CK_N[j] dt */
static int rk5_run(const fsrc *F, int m, const double *params, const double *time,
                   double *state, int64_t cnt[3], int64_t max_steps)
{
    const double tmx = params[2], dtm = params[4];
    const double atol = params[5], fcmn = params[7], fcmx = params[8], safe = params[9];
    const double eps = DBL_MIN;                                            /* tiny(), :99 */
    double t, dt, dtn, del, q, si[6], sf[6], er[6], ac[6], s[6][6], k[6][6], sum;
    int n, status = PTC_STATUS_DONE;

    cnt[0] = cnt[1] = cnt[2] = 0;
    t = 0.0;
    dt = dti_guess_f(F, 5, params, state);                                 /* :122 */
    for (int i = 0; i < 6; i++) si[i] = state[i];
    state[6] = sqrt(si[3] * si[3] + si[4] * si[4] + si[5] * si[5] + 1.0);
    state[7] = dt;
    n = 2;                                                                 /* 1-based */

    while (t < tmx) {                                                      /* :129 */
        if (max_steps > 0 && cnt[0] >= max_steps) { status = PTC_STATUS_MAXSTEPS; break; }
        if (!isfinite(dt)) { status = PTC_STATUS_NONFINITE; break; }
        if (!(dt > 0.0) || t + dt == t) { status = PTC_STATUS_DTSMALL; break; }

        for (int j = 0; j < 6; j++) {                                      /* :131-138 */
            for (int c = 0; c < 6; c++) s[j][c] = si[c];
            for (int i = 0; i < j; i++)
                for (int c = 0; c < 6; c++) s[j][c] = s[j][c] + CK_A[j][i] * k[i][c];
            fsrc_rhs(F, t + CK_N[j] * dt, s[j], ac);
            for (int c = 0; c < 6; c++) k[j][c] = dt * ac[c];
        }
        for (int c = 0; c < 6; c++) {                                      /* :140-143 */
            sum = 0.0;
            for (int j = 0; j < 6; j++) sum += CK_C[j] * k[j][c];
            sf[c] = si[c] + sum;
            sum = 0.0;
            for (int j = 0; j < 6; j++) sum += CK_D[j] * k[j][c];
            er[c] = sum;
        }
        del = 0.0;                                                         /* :147-148 */
        for (int c = 0; c < 6; c++) {
            q = fabs(er[c] / (fabs(si[c]) + fabs(k[0][c]) + eps));
            if (c == 0 || q > del) del = q;
        }
        del = del / atol;
        if (!isfinite(del) || !all_finite6(sf)) { status = PTC_STATUS_NONFINITE; break; }

        if (del <= 1.0) {                                                  /* :150 */
            t = t + dt;
            /* the reference recomputes k6 = dt f(s6) here (:156-157): same
             * input, same value, so the stored k6 is used */
            while (n <= m && t >= time[n - 1]) {                           /* :158-168 */
                hermite_out(state + 8 * (int64_t)(n - 1), t, time[n - 1], dt, si, sf, k[0], k[5]);
                n++;
            }
            for (int c = 0; c < 6; c++) si[c] = sf[c];
            cnt[1]++;
            q = safe * pow(del, -0.20);                                    /* :177 */
            dtn = dt * (fcmx < q ? fcmx : q);
        } else {
            cnt[2]++;
            q = safe * pow(del, -0.25);                                    /* :186 */
            dtn = dt * (fcmn > q ? fcmn : q);
        }
        cnt[0]++;
        dt = dtn < dtm ? dtn : dtm;                                        /* :195 */
    }
    return status;
}

int orc_integrate_rk5(int mode, int m, const int dm[3], const double *params,
                      const double *time, const double *uu, const double *bb,
                      double *state, int64_t cnt[3], int64_t max_steps)
{
    const fsrc F = fsrc_make(mode, dm, params, 0, NULL, uu, bb);
    return rk5_run(&F, m, params, time, state, cnt, max_steps);
}

/* Hairer-Norsett-Wanner DOP853 coefficients (public dop853.f; the reference
 * lists the same published values at integrations.F90:255-326). */
#define A0201  5.26001519587677318785587544488e-02
#define A0301  1.97250569845378994544595329183e-02
#define A0302  5.91751709536136983633785987549e-02
#define A0401  2.95875854768068491816892993775e-02
#define A0403  8.87627564304205475450678981324e-02
#define A0501  2.41365134159266685502369798665e-01
#define A0503 -8.84549479328286085344864962717e-01
#define A0504  9.24834003261792003115737966543e-01
#define A0601  3.70370370370370370370370370370e-02
#define A0604  1.70828608729473871279604482173e-01
#define A0605  1.25467687566822425016691814123e-01
#define A0701  3.71093750000000000000000000000e-02
#define A0704  1.70252211019544039314978060272e-01
#define A0705  6.02165389804559606850219397283e-02
#define A0706 -1.75781250000000000000000000000e-02
#define A0801  3.70920001185047927108779319836e-02
#define A0804  1.70383925712239993810214054705e-01
#define A0805  1.07262030446373284651809199168e-01
#define A0806 -1.53194377486244017527936158236e-02
#define A0807  8.27378916381402288758473766002e-03
#define A0901  6.24110958716075717114429577812e-01
#define A0904 -3.36089262944694129406857109825e+00
#define A0905 -8.68219346841726006818189891453e-01
#define A0906  2.75920996994467083049415600797e+01
#define A0907  2.01540675504778934086186788979e+01
#define A0908 -4.34898841810699588477366255144e+01
#define A1001  4.77662536438264365890433908527e-01
#define A1004 -2.48811461997166764192642586468e+00
#define A1005 -5.90290826836842996371446475743e-01
#define A1006  2.12300514481811942347288949897e+01
#define A1007  1.52792336328824235832596922938e+01
#define A1008 -3.32882109689848629194453265587e+01
#define A1009 -2.03312017085086261358222928593e-02
#define A1101 -9.37142430085987325717040216580e-01
#define A1104  5.18637242884406370830023853209e+00
#define A1105  1.09143734899672957818500254654e+00
#define A1106 -8.14978701074692612513997267357e+00
#define A1107 -1.85200656599969598641566180701e+01
#define A1108  2.27394870993505042818970056734e+01
#define A1109  2.49360555267965238987089396762e+00
#define A1110 -3.04676447189821950038236690220e+00
#define A1201  2.27331014751653820792359768449e+00
#define A1204 -1.05344954667372501984066689879e+01
#define A1205 -2.00087205822486249909675718444e+00
#define A1206 -1.79589318631187989172765950534e+01
#define A1207  2.79488845294199600508499808837e+01
#define A1208 -2.85899827713502369474065508674e+00
#define A1209 -8.87285693353062954433549289258e+00
#define A1210  1.23605671757943030647266201528e+01
#define A1211  6.43392746015763530355970484046e-01
#define B01  5.42937341165687622380535766363e-02
#define B06  4.45031289275240888144113950566e+00
#define B07  1.89151789931450038304281599044e+00
#define B08 -5.80120396001058478146721142270e+00
#define B09  3.11164366957819894408916062370e-01
#define B10 -1.52160949662516078556178806805e-01
#define B11  2.01365400804030348374776537501e-01
#define B12  4.47106157277725905176885569043e-02
#define BH1  0.244094488188976377952755905512e+00
#define BH2  0.733846688281611857341361741547e+00
#define BH3  0.220588235294117647058823529412e-01
#define ER01  0.1312004499419488073250102996e-01
#define ER06 -0.1225156446376204440720569753e+01
#define ER07 -0.4957589496572501915214079952e+00
#define ER08  0.1664377182454986536961530415e+01
#define ER09 -0.3503288487499736816886487290e+00
#define ER10  0.3341791187130174790297318841e+00
#define ER11  0.8192320648511571246570742613e-01
#define ER12 -0.2235530786388629525884427845e-01
/* stage nodes c_i (public dop853.f; the row sums of the A rows), for a series */
#define DN02 0.526001519587677318785587544488e-01
#define DN03 0.789002279381515978178381316732e-01
#define DN04 0.118350341907227396726757197510e+00
#define DN05 0.281649658092772603273242802490e+00
#define DN06 0.333333333333333333333333333333e+00
#define DN07 0.25e+00
#define DN08 0.307692307692307692307692307692e+00
#define DN09 0.651282051282051282051282051282e+00
#define DN10 0.6e+00
#define DN11 0.857142857142857142857142857142e+00

/* sources/integrations.F90:224-505.  Slot names follow the stage numbers:
 * k11 and k12 are what the reference stores in k(2) and k(3). */
static int dp853_run(const fsrc *F, int m, const double *params, const double *time,
                     double *state, int64_t cnt[3], int64_t max_steps)
{
    const double tmx = params[2], dtm = params[4];
    const double atol = params[5], rtol = params[6], fcmn = params[7], fcmx = params[8];
    const double safe = params[9], beta = params[10];
    const double expo = 2.0e-01 * beta - 1.25e-01;                         /* :348 */
    double errold = 1.0e-04, t, dt, dtn, err3, err5, deno, err, q, a, b;
    double si[6], ss[6], sf[6], sr[6], er[6], kb[6], ka[6], ke[6];
    double k1[6], k2[6], k3[6], k4[6], k5[6], k6[6], k7[6], k8[6], k9[6], k10[6], k11[6], k12[6];
    int rejected = 0, n, status = PTC_STATUS_DONE;

#define RHS(in, out, tt) fsrc_rhs(F, tt, in, out)

    cnt[0] = cnt[1] = cnt[2] = 0;
    t = 0.0;
    dt = dti_guess_f(F, 8, params, state);                                 /* :354 */
    for (int i = 0; i < 6; i++) si[i] = state[i];
    state[6] = sqrt(si[3] * si[3] + si[4] * si[4] + si[5] * si[5] + 1.0);
    state[7] = dt;
    n = 2;

    while (t < tmx) {                                                      /* :361 */
        if (max_steps > 0 && cnt[0] >= max_steps) { status = PTC_STATUS_MAXSTEPS; break; }
        if (!isfinite(dt)) { status = PTC_STATUS_NONFINITE; break; }
        if (!(dt > 0.0) || t + dt == t) { status = PTC_STATUS_DTSMALL; break; }

        RHS(si, k1, t);                                                       /* :365-416 */
        for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * A0201 * k1[i];
        RHS(ss, k2, t + DN02 * dt);
        for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0301 * k1[i] + A0302 * k2[i]);
        RHS(ss, k3, t + DN03 * dt);
        for (int i = 0; i < 6; i++) ss[i] = si[i] + dt * (A0401 * k1[i] + A0403 * k3[i]);
        RHS(ss, k4, t + DN04 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A0501 * k1[i] + A0503 * k3[i] + A0504 * k4[i]);
        RHS(ss, k5, t + DN05 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A0601 * k1[i] + A0604 * k4[i] + A0605 * k5[i]);
        RHS(ss, k6, t + DN06 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A0701 * k1[i] + A0704 * k4[i] + A0705 * k5[i] + A0706 * k6[i]);
        RHS(ss, k7, t + DN07 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A0801 * k1[i] + A0804 * k4[i] + A0805 * k5[i] + A0806 * k6[i] +
                                  A0807 * k7[i]);
        RHS(ss, k8, t + DN08 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A0901 * k1[i] + A0904 * k4[i] + A0905 * k5[i] + A0906 * k6[i] +
                                  A0907 * k7[i] + A0908 * k8[i]);
        RHS(ss, k9, t + DN09 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A1001 * k1[i] + A1004 * k4[i] + A1005 * k5[i] + A1006 * k6[i] +
                                  A1007 * k7[i] + A1008 * k8[i] + A1009 * k9[i]);
        RHS(ss, k10, t + DN10 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A1101 * k1[i] + A1104 * k4[i] + A1105 * k5[i] + A1106 * k6[i] +
                                  A1107 * k7[i] + A1108 * k8[i] + A1109 * k9[i] + A1110 * k10[i]);
        RHS(ss, k11, t + DN11 * dt);
        for (int i = 0; i < 6; i++)
            ss[i] = si[i] + dt * (A1201 * k1[i] + A1204 * k4[i] + A1205 * k5[i] + A1206 * k6[i] +
                                  A1207 * k7[i] + A1208 * k8[i] + A1209 * k9[i] + A1210 * k10[i] +
                                  A1211 * k11[i]);
        RHS(ss, k12, t + dt);

        for (int i = 0; i < 6; i++) {                                      /* :418-421 */
            kb[i] = B01 * k1[i] + B06 * k6[i] + B07 * k7[i] + B08 * k8[i] + B09 * k9[i] +
                    B10 * k10[i] + B11 * k11[i] + B12 * k12[i];
            sf[i] = si[i] + dt * kb[i];
        }

        for (int i = 0; i < 6; i++) {                                      /* :425 */
            a = fabs(si[i]);
            b = fabs(sf[i]);
            sr[i] = atol + rtol * (a > b ? a : b);
        }
        err3 = 0.0;                                                        /* :426-427 */
        for (int i = 0; i < 6; i++) {
            er[i] = kb[i] - BH1 * k1[i] - BH2 * k9[i] - BH3 * k12[i];
            q = er[i] / sr[i];
            err3 += q * q;
        }
        err5 = 0.0;                                                        /* :428-431 */
        for (int i = 0; i < 6; i++) {
            er[i] = ER01 * k1[i] + ER06 * k6[i] + ER07 * k7[i] + ER08 * k8[i] + ER09 * k9[i] +
                    ER10 * k10[i] + ER11 * k11[i] + ER12 * k12[i];
            q = er[i] / sr[i];
            err5 += q * q;
        }
        deno = err5 + 1.0e-02 * err3;                                      /* :432-437 */
        if (deno <= 0.0) err = fabs(dt) * err5 / sqrt(6.0);
        else err = fabs(dt) * err5 / sqrt(6.0 * deno);
        if (!isfinite(err) || !all_finite6(sf)) { status = PTC_STATUS_NONFINITE; break; }

        if (err <= 1.0) {                                                  /* :439 */
            t = t + dt;
            /* the reference re-evaluates f(si) (= k1, :445) and f(sf) (:446);
             * only f(sf) is new */
            RHS(sf, ke, t);
            for (int i = 0; i < 6; i++) { ka[i] = dt * k1[i]; kb[i] = dt * ke[i]; }
            while (n <= m && t >= time[n - 1]) {                           /* :447-457 */
                hermite_out(state + 8 * (int64_t)(n - 1), t, time[n - 1], dt, si, sf, ka, kb);
                n++;
            }
            for (int i = 0; i < 6; i++) si[i] = sf[i];
            cnt[1]++;
            q = safe * pow(err, expo) * pow(errold, beta);                 /* :466-471 */
            if (q < fcmn) q = fcmn;
            if (rejected) dtn = dt * (q < 1.0 ? q : 1.0);
            else dtn = dt * (q < fcmx ? q : fcmx);
            errold = err > 1.0e-04 ? err : 1.0e-04;                        /* :472 */
            rejected = 0;
        } else {
            cnt[2]++;                                                      /* :482-490 */
            q = safe * pow(err, expo);
            dtn = dt * (fcmn > q ? fcmn : q);
            rejected = 1;
        }
        cnt[0]++;
        dt = dtn < dtm ? dtn : dtm;                                        /* :500 */
    }
#undef RHS
    return status;
}

int orc_integrate_dp853(int mode, int m, const int dm[3], const double *params,
                        const double *time, const double *uu, const double *bb,
                        double *state, int64_t cnt[3], int64_t max_steps)
{
    const fsrc F = fsrc_make(mode, dm, params, 0, NULL, uu, bb);
    return dp853_run(&F, m, params, time, state, cnt, max_steps);
}

int orc_stage_nodes(int method, double *c, double *rowsum)
{
    if (method == 0) {
        for (int j = 0; j < 6; j++) {
            c[j] = CK_N[j];
            rowsum[j] = 0.0;
            for (int i = 0; i < j; i++) rowsum[j] += CK_A[j][i];
        }
        return 6;
    }
    static const double cn[12] = {0.0, DN02, DN03, DN04, DN05, DN06, DN07, DN08, DN09, DN10,
                                  DN11, 1.0};
    const double rs[12] = {
        0.0, A0201, A0301 + A0302, A0401 + A0403, A0501 + A0503 + A0504,
        A0601 + A0604 + A0605, A0701 + A0704 + A0705 + A0706,
        A0801 + A0804 + A0805 + A0806 + A0807,
        A0901 + A0904 + A0905 + A0906 + A0907 + A0908,
        A1001 + A1004 + A1005 + A1006 + A1007 + A1008 + A1009,
        A1101 + A1104 + A1105 + A1106 + A1107 + A1108 + A1109 + A1110,
        A1201 + A1204 + A1205 + A1206 + A1207 + A1208 + A1209 + A1210 + A1211};
    for (int j = 0; j < 12; j++) {
        c[j] = cn[j];
        rowsum[j] = rs[j];
    }
    return 12;
}

/* sources/particles.F90:337-356 */
void orc_integrate(int method, int mode, int64_t np, int m, const int dm[3],
                   const double *params, const double *time, const double *uu,
                   const double *bb, double *state, int64_t *cnt, int32_t *status,
                   int64_t max_steps, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t n;
#pragma omp parallel for schedule(static)
    for (n = 0; n < np; n++) {
        double *st = state + (int64_t)8 * m * n;
        int s = method == 0
                    ? orc_integrate_rk5(mode, m, dm, params, time, uu, bb, st, cnt + 3 * n, max_steps)
                    : orc_integrate_dp853(mode, m, dm, params, time, uu, bb, st, cnt + 3 * n, max_steps);
        if (status) status[n] = s;
    }
}

void orc_integrate_series(int method, int mode, int64_t np, int m, const int dm[3],
                          const double *params, const double *time, int ns, const double *st,
                          const double *uu, const double *bb, double *state, int64_t *cnt,
                          int32_t *status, int64_t max_steps, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const fsrc F = fsrc_make(mode, dm, params, ns, st, uu, bb);
    int64_t n;
#pragma omp parallel for schedule(static)
    for (n = 0; n < np; n++) {
        double *sp = state + (int64_t)8 * m * n;
        int s = method == 0 ? rk5_run(&F, m, params, time, sp, cnt + 3 * n, max_steps)
                            : dp853_run(&F, m, params, time, sp, cnt + 3 * n, max_steps);
        if (status) status[n] = s;
    }
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
