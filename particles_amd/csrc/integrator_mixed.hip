// Mixed-precision kernels (FP32 fields): the integrator of integrator.hip
// instantiated for FP32 node storage, FP32 trilinear weights and RHS, with
// FP64 positions, time, stage accumulation and error norm.  This unit is
// compiled with FMA contraction within expressions (-ffp-contract=on): the
// mixed path is not bitwise comparable to the reference anyway, fused
// multiply-adds in the FP64 stage sums and the FP32 gathers both cut
// instructions and round less, and contracting only what one expression
// holds makes every instantiation (node, pair, quad or column rows; trio or
// group kernel) round the same way.
#include "trio_impl.hpp"

namespace ptc {

hipError_t launch_integrate_mixed(int method, const DeviceFields &f, const StepParams &sp,
                                  const DeviceParticles &p, double t_stop, const LaunchConfig &lc,
                                  hipStream_t stream)
{
    if (f.prec != PTC_FP32 || (f.layout != LAYOUT_AOS && !rows_layout(f.layout)))
        return hipErrorInvalidValue;
    const int geom = sp.pow2 ? 1 : 0;
    if (rows_layout(f.layout)) {
        // pair, quad or column rows: the trio or group kernel only (a
        // snapshot series in column rows with the PI controller as well)
        if (lc.tail != kTailGroupOnly) return hipErrorInvalidValue;
        if (f.nsnap > 0) {
            if (f.layout != LAYOUT_COLUMNS || lc.ctrl != CTRL_PI) return hipErrorInvalidValue;
#define GROUP_TC(M, MO, G)                                                     \
    if (method == M && f.mode == MO && geom == G)                              \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP32, ROWS_COLUMNS, CTRL_PI, 1>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP32, 1, CTRL_PI, ROWS_COLUMNS>(f, sp, p, t_stop, kBlock, stream);
            GROUP_TC(PTC_RK5, PTC_MODE_VB, 0) GROUP_TC(PTC_RK5, PTC_MODE_EB, 0)
            GROUP_TC(PTC_DOP853, PTC_MODE_VB, 0) GROUP_TC(PTC_DOP853, PTC_MODE_EB, 0)
            GROUP_TC(PTC_RK5, PTC_MODE_VB, 1) GROUP_TC(PTC_RK5, PTC_MODE_EB, 1)
            GROUP_TC(PTC_DOP853, PTC_MODE_VB, 1) GROUP_TC(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP_TC
            return hipErrorInvalidValue;
        }
        const int rows = f.layout == LAYOUT_QUADS     ? ROWS_QUADS
                         : f.layout == LAYOUT_COLUMNS ? ROWS_COLUMNS
                                                      : ROWS_PAIRS;
#define GROUP_R(M, MO, G, CT, R)                                               \
    if (method == M && f.mode == MO && geom == G && lc.ctrl == CT && rows == R) \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP32, R, CT>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP32, 0, CT, R>(f, sp, p, t_stop, kBlock, stream);
#define GROUP_P(M, MO, G, CT) \
    GROUP_R(M, MO, G, CT, ROWS_PAIRS) GROUP_R(M, MO, G, CT, ROWS_QUADS) GROUP_R(M, MO, G, CT, ROWS_COLUMNS)
#define GROUP_P_ALL(CT)                                                        \
        GROUP_P(PTC_RK5, PTC_MODE_VB, 0, CT) GROUP_P(PTC_RK5, PTC_MODE_EB, 0, CT) \
        GROUP_P(PTC_DOP853, PTC_MODE_VB, 0, CT) GROUP_P(PTC_DOP853, PTC_MODE_EB, 0, CT) \
        GROUP_P(PTC_RK5, PTC_MODE_VB, 1, CT) GROUP_P(PTC_RK5, PTC_MODE_EB, 1, CT) \
        GROUP_P(PTC_DOP853, PTC_MODE_VB, 1, CT) GROUP_P(PTC_DOP853, PTC_MODE_EB, 1, CT)
        GROUP_P_ALL(CTRL_PI)
        GROUP_P_ALL(CTRL_LEARNED)
#undef GROUP_P_ALL
#undef GROUP_P
#undef GROUP_R
        return hipErrorInvalidValue;
    }
    if (f.nsnap > 0) {
        // snapshot series: PI controller; the cell cache keeps both
        // snapshots' corner values in registers
        if (lc.ctrl != CTRL_PI) return hipErrorInvalidValue;
        if (lc.tail == kTailGroupOnly) {
#define GROUP_T(M, MO, G)                                                      \
    if (method == M && f.mode == MO && geom == G)                              \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP32, ROWS_NODES, CTRL_PI, 1>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP32, 1>(f, sp, p, t_stop, kBlock, stream);
            GROUP_T(PTC_RK5, PTC_MODE_VB, 0) GROUP_T(PTC_RK5, PTC_MODE_EB, 0)
            GROUP_T(PTC_DOP853, PTC_MODE_VB, 0) GROUP_T(PTC_DOP853, PTC_MODE_EB, 0)
            GROUP_T(PTC_RK5, PTC_MODE_VB, 1) GROUP_T(PTC_RK5, PTC_MODE_EB, 1)
            GROUP_T(PTC_DOP853, PTC_MODE_VB, 1) GROUP_T(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP_T
            return hipErrorInvalidValue;
        }
        const int c = lc.cell_cache == 1 ? 1 : 0;
#define SERIES(M, MO, C, G)                                                    \
    if (method == M && f.mode == MO && c == C && geom == G)                    \
        return launch_t<M, MO, LAYOUT_AOS, C, PTC_FP32, CTRL_PI, G, 1>(f, sp, p, t_stop, lc, stream);
#define SERIES_MODES(C, G)                                                     \
    SERIES(PTC_RK5, PTC_MODE_VB, C, G) SERIES(PTC_RK5, PTC_MODE_EB, C, G)      \
    SERIES(PTC_DOP853, PTC_MODE_VB, C, G) SERIES(PTC_DOP853, PTC_MODE_EB, C, G)
        SERIES_MODES(0, 0) SERIES_MODES(0, 1) SERIES_MODES(1, 0) SERIES_MODES(1, 1)
#undef SERIES_MODES
#undef SERIES
        return hipErrorInvalidValue;
    }
    // learned controller: the group or trio kernel ("group_lanes") runs the
    // whole launch ("tail" group) or one lane per particle does (no hand-over)
    if (lc.ctrl == CTRL_LEARNED && lc.tail == kTailGroupOnly) {
#define GROUP_L(M, MO, G)                                                      \
    if (method == M && f.mode == MO && geom == G)                              \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP32, ROWS_NODES, CTRL_LEARNED>(f, sp, p, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP32, 0, CTRL_LEARNED>(f, sp, p, t_stop, kBlock, stream);
        GROUP_L(PTC_RK5, PTC_MODE_VB, 0) GROUP_L(PTC_RK5, PTC_MODE_EB, 0)
        GROUP_L(PTC_DOP853, PTC_MODE_VB, 0) GROUP_L(PTC_DOP853, PTC_MODE_EB, 0)
        GROUP_L(PTC_RK5, PTC_MODE_VB, 1) GROUP_L(PTC_RK5, PTC_MODE_EB, 1)
        GROUP_L(PTC_DOP853, PTC_MODE_VB, 1) GROUP_L(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP_L
        return hipErrorInvalidValue;
    }
    // PI controller, static field: the group kernel (six lanes per
    // particle) takes the tail of the launch, or all of it ("tail" option)
    if (lc.tail != 0 && lc.ctrl == CTRL_PI) {
        DeviceParticles q = p;
        q.park = lc.tail > 0 ? lc.tail : 0;
        if (lc.tail != kTailGroupOnly) {
            hipError_t e = hipErrorInvalidValue;
#define PARKED(M, MO, C, G)                                                    \
    if (method == M && f.mode == MO && lc.cell_cache == C && geom == G)        \
        e = launch_t<M, MO, LAYOUT_AOS, C, PTC_FP32, CTRL_PI, G, 0, 1>(f, sp, q, t_stop, lc, stream); \
    else
// without the cell cache the kernels take the general geometry
#define PARKED0(M, MO)                                                         \
    if (method == M && f.mode == MO && lc.cell_cache == 0)                     \
        e = launch_t<M, MO, LAYOUT_AOS, 0, PTC_FP32, CTRL_PI, 0, 0, 1>(f, sp, q, t_stop, lc, stream); \
    else
            PARKED(PTC_RK5, PTC_MODE_VB, 1, 0) PARKED(PTC_RK5, PTC_MODE_EB, 1, 0)
            PARKED(PTC_DOP853, PTC_MODE_VB, 1, 0) PARKED(PTC_DOP853, PTC_MODE_EB, 1, 0)
            PARKED(PTC_RK5, PTC_MODE_VB, 1, 1) PARKED(PTC_RK5, PTC_MODE_EB, 1, 1)
            PARKED(PTC_DOP853, PTC_MODE_VB, 1, 1) PARKED(PTC_DOP853, PTC_MODE_EB, 1, 1)
            PARKED0(PTC_RK5, PTC_MODE_VB) PARKED0(PTC_RK5, PTC_MODE_EB)
            PARKED0(PTC_DOP853, PTC_MODE_VB) PARKED0(PTC_DOP853, PTC_MODE_EB)
            e = hipErrorInvalidValue;
#undef PARKED0
#undef PARKED
            if (e != hipSuccess) return e;
        }
#define GROUP(M, MO, G)                                                        \
    if (method == M && f.mode == MO && geom == G)                              \
        return lc.lanes == 3                                                   \
                   ? launch_trio_t<M, MO, G, PTC_FP32, ROWS_NODES>(f, sp, q, t_stop, kBlock, stream) \
                   : launch_group_t<M, MO, G, PTC_FP32>(f, sp, q, t_stop, kBlock, stream);
        GROUP(PTC_RK5, PTC_MODE_VB, 0) GROUP(PTC_RK5, PTC_MODE_EB, 0)
        GROUP(PTC_DOP853, PTC_MODE_VB, 0) GROUP(PTC_DOP853, PTC_MODE_EB, 0)
        GROUP(PTC_RK5, PTC_MODE_VB, 1) GROUP(PTC_RK5, PTC_MODE_EB, 1)
        GROUP(PTC_DOP853, PTC_MODE_VB, 1) GROUP(PTC_DOP853, PTC_MODE_EB, 1)
#undef GROUP
        return hipErrorInvalidValue;
    }
#define DISPATCH(M, MO, C, CT, G)                                              \
    if (method == M && f.mode == MO && lc.cell_cache == C && lc.ctrl == CT && geom == G) \
        return launch_t<M, MO, LAYOUT_AOS, C, PTC_FP32, CT, G>(f, sp, p, t_stop, lc, stream);
#define DISPATCH_MODES(C, CT, G)                                               \
    DISPATCH(PTC_RK5, PTC_MODE_VB, C, CT, G)                                   \
    DISPATCH(PTC_RK5, PTC_MODE_EB, C, CT, G)                                   \
    DISPATCH(PTC_DOP853, PTC_MODE_VB, C, CT, G)                                \
    DISPATCH(PTC_DOP853, PTC_MODE_EB, C, CT, G)
    DISPATCH_MODES(1, CTRL_PI, 1)
    DISPATCH_MODES(1, CTRL_RECORD, 1)
    DISPATCH_MODES(1, CTRL_LEARNED, 1)
    DISPATCH_MODES(1, CTRL_PI, 0)
    DISPATCH_MODES(1, CTRL_RECORD, 0)
    DISPATCH_MODES(1, CTRL_LEARNED, 0)
    if (lc.ctrl == CTRL_PI && lc.cell_cache == 0) {
        if (method == PTC_RK5 && f.mode == PTC_MODE_VB)
            return launch_t<PTC_RK5, PTC_MODE_VB, LAYOUT_AOS, 0, PTC_FP32, CTRL_PI, 0>(f, sp, p, t_stop, lc, stream);
        if (method == PTC_RK5)
            return launch_t<PTC_RK5, PTC_MODE_EB, LAYOUT_AOS, 0, PTC_FP32, CTRL_PI, 0>(f, sp, p, t_stop, lc, stream);
        if (f.mode == PTC_MODE_VB)
            return launch_t<PTC_DOP853, PTC_MODE_VB, LAYOUT_AOS, 0, PTC_FP32, CTRL_PI, 0>(f, sp, p, t_stop, lc, stream);
        return launch_t<PTC_DOP853, PTC_MODE_EB, LAYOUT_AOS, 0, PTC_FP32, CTRL_PI, 0>(f, sp, p, t_stop, lc, stream);
    }
#undef DISPATCH_MODES
#undef DISPATCH
    return hipErrorInvalidValue;
}

hipError_t launch_probe_mixed(const DeviceFields &f, const StepParams &sp, int op, int64_t n,
                              const double *in, double *out, hipStream_t stream)
{
    if (f.mode == PTC_MODE_VB)
        return launch_probe_t<PTC_MODE_VB, LAYOUT_AOS, PTC_FP32>(f, sp, op, n, in, out, stream);
    return launch_probe_t<PTC_MODE_EB, LAYOUT_AOS, PTC_FP32>(f, sp, op, n, in, out, stream);
}

}  // namespace ptc
