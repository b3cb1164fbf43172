// C-ABI of the integrator (include/particles_amd.h): handles, validation,
// host<->device transposition and launch bookkeeping.  The compute lives in
// integrator.hip.
#include "internal.hpp"
#include "libm_pow.hpp"
#include "mlp_device.hpp"

#include "particles_amd.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace ptc;

struct ptc_fields {
    DeviceFields d;
};

struct ptc_particles {
    const ptc_fields *fields;
    int method;
    StepParams sp;
    DeviceParticles d;
    void *block;          // one allocation holding every per-particle array
    double *d_time;
    int64_t *d_totals;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    double last_ms;
    int ctrl;             // CTRL_*
    void *rec_block;      // record mode: pending features, records, counter
    void *order_block;    // spatial order: order (np int32) + sort scratch
    size_t order_bytes;
    int64_t advances;     // advance calls so far (sort cadence)
    int tuned;            // launch settings below chosen by ptc_particles_autotune
    int tuned_block, tuned_sort, tuned_order, tuned_tail, tuned_lanes;
    // the graph of ptc_particles_advance_graph, kept for the batch's life and
    // replayed with new stop times while everything else its launches depend
    // on (the key below) is unchanged
    struct AdvanceGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int n = 0;                          // intervals
        LaunchConfig lc{};
        int64_t max_steps = 0;
        const int32_t *order = nullptr;
        int ctrl = 0;
        const void *net = nullptr;
        float shift = 0.f;
        uint64_t rec_thresh = 0;
        std::vector<hipGraphNode_t> nodes;  // the integrator kernel nodes
        std::vector<int> interval;          // the stop time each one takes
        bool replayable = false;            // the nodes' arguments checked at capture
        int64_t instantiated = 0, replayed = 0;
    } graph;
};

struct ptc_mlp {
    float *params;        // PTC_MLP_PARAMS FP32 master copy
    float *m1, *m2;       // Adam moments (PTC_MLP_TRAINABLE each)
    float *partials;      // kMlpMaxGradBlocks x (PTC_MLP_TRAINABLE + 1)
    float *grad;          // PTC_MLP_TRAINABLE + 1
    double *stats;        // feature statistics partials
    MlpPacked *packed;    // BF16 layouts read by the kernels
    int64_t t;            // Adam steps taken
    bool trained;         // parameters given or trained (not the random start)
    hipStream_t stream;
};

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(e == hipErrorOutOfMemory ? PTC_ENOMEM : PTC_EHIP, "%s: %s", what,
                hipGetErrorString(e));
}

#define HIP_TRY(expr)                                           \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);       \
    } while (0)

struct Options {
    int block = 256;
    int grid_blocks = 0;
    int layout = LAYOUT_AOS;
    int64_t max_steps = 0;
    int cell_cache = 1;
    int sort = 1;   // re-order the work queue every `sort` advances (0 = never)
    int order = 3;  // the order then: 1 spatial (Morton), 2 cost (most work first), 3 auto
    int tail = kTailAuto;  // LaunchConfig::tail, or chosen per batch
    int f32_layout = LAYOUT_AOS;  // FP32 fields: node rows, LAYOUT_PAIRS, _QUADS or _COLUMNS
    int lanes = 0;                // LaunchConfig::lanes, or 0: chosen per batch (choose_lanes)
};

std::mutex g_opt_mu;
Options g_opt;

Options options()
{
    std::lock_guard<std::mutex> lk(g_opt_mu);
    return g_opt;
}

void ok() { g_err.clear(); }

}  // namespace

extern "C" {

const char *ptc_last_error(void) { return g_err.c_str(); }

const char *ptc_version(void) { return "particles_amd 0.1 (gfx950)"; }

double ptc_pow_host(double x, double y) { return ptc_libm_pow(x, y); }

int ptc_diag_line_gather(int64_t bytes, int rounds, double result[3])
{
    if (!result || bytes < 128 * 1024 || rounds < 1)
        return fail(PTC_EARG, "ptc_diag_line_gather: needs bytes >= 128 KiB and rounds >= 1");
    const uint64_t nlines = (uint64_t)bytes / 128;
    void *buf = nullptr;
    uint32_t *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // the buffer's contents do not matter: it is read, never used
    if (e == hipSuccess) e = hipMalloc(&buf, nlines * 128);
    if (e == hipSuccess) e = hipMalloc((void **)&sink, sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // eight 256-thread blocks per CU, 32 waves per CU
    const unsigned blocks = (unsigned)(cus > 0 ? cus : 1) * 8u;
    const double lines = (double)blocks * 256.0 / kDiagLineLanes * rounds * kDiagInFlight;
    float ms = 0.f;
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        // pass 0 warms the code and the address translation; pass 1 is timed
        e = hipEventRecord(e0, nullptr);
        if (e == hipSuccess)
            e = launch_line_gather(buf, nlines, (uint64_t)(pass + 1), rounds, sink, blocks, nullptr);
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (buf) (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(e, "ptc_diag_line_gather");
    result[0] = lines * 128.0 / (ms * 1e-3) / 1e9;
    result[1] = ms;
    result[2] = lines * 128.0;
    ok();
    return PTC_OK;
}

int ptc_diag_cell_gather(int64_t bytes, int pattern, int cells, double result[3])
{
    if (!result || bytes < 1024 * 1024 || cells < 1 || (pattern != 0 && pattern != 1))
        return fail(PTC_EARG, "ptc_diag_cell_gather: needs bytes >= 1 MiB, pattern 0 or 1 and "
                              "cells >= 1");
    const uint64_t unit = pattern == 0 ? 128 : 192;
    const uint64_t nunits = (uint64_t)bytes / unit;
    void *buf = nullptr;
    uint32_t *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // the buffer's contents do not matter: it is read, never used
    if (e == hipSuccess) e = hipMalloc(&buf, nunits * unit);
    if (e == hipSuccess) e = hipMalloc((void **)&sink, sizeof(uint32_t));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // three 256-thread blocks per CU: three waves of ten groups per SIMD, as
    // the six-lane group kernel runs
    const unsigned blocks = (unsigned)(cus > 0 ? cus : 1) * 3u;
    const double loads = (double)blocks * 4.0 * 10.0 * cells;
    float ms = 0.f;
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        // pass 0 warms the code and the address translation; pass 1 is timed
        e = hipEventRecord(e0, nullptr);
        if (e == hipSuccess)
            e = launch_cell_gather(buf, nunits, pattern, (uint64_t)(pass + 1), cells, sink, blocks,
                                   nullptr);
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (buf) (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(e, "ptc_diag_cell_gather");
    result[0] = loads / (ms * 1e-3);
    result[1] = ms;
    result[2] = loads;
    ok();
    return PTC_OK;
}

int ptc_set_option(const char *key, const char *value)
{
    if (!key || !value) return fail(PTC_EARG, "ptc_set_option: null key or value");
    std::lock_guard<std::mutex> lk(g_opt_mu);
    char *end = nullptr;
    if (!strcmp(key, "block")) {
        long v = strtol(value, &end, 10);
        if (*end || v < 64 || v > 256 || v % 64)
            return fail(PTC_EARG, "block must be 64, 128, 192 or 256 (got %s)", value);
        g_opt.block = (int)v;
    } else if (!strcmp(key, "grid_blocks")) {
        long v = strtol(value, &end, 10);
        if (*end || v < 0 || v > (1 << 20)) return fail(PTC_EARG, "bad grid_blocks %s", value);
        g_opt.grid_blocks = (int)v;
    } else if (!strcmp(key, "layout")) {
        if (!strcmp(value, "soa")) g_opt.layout = LAYOUT_SOA;
        else if (!strcmp(value, "aos")) g_opt.layout = LAYOUT_AOS;
        else if (!strcmp(value, "brick")) g_opt.layout = LAYOUT_BRICK;
        else if (!strcmp(value, "pairs")) g_opt.layout = LAYOUT_PAIRS;
        else if (!strcmp(value, "columns")) g_opt.layout = LAYOUT_COLUMNS;
        else return fail(PTC_EARG, "layout must be soa, aos, brick, pairs or columns (got %s)",
                         value);
    } else if (!strcmp(key, "f32_layout")) {
        if (!strcmp(value, "nodes")) g_opt.f32_layout = LAYOUT_AOS;
        else if (!strcmp(value, "pairs")) g_opt.f32_layout = LAYOUT_PAIRS;
        else if (!strcmp(value, "quads")) g_opt.f32_layout = LAYOUT_QUADS;
        else if (!strcmp(value, "columns")) g_opt.f32_layout = LAYOUT_COLUMNS;
        else return fail(PTC_EARG, "f32_layout must be nodes, pairs, quads or columns (got %s)",
                         value);
    } else if (!strcmp(key, "max_steps")) {
        long long v = strtoll(value, &end, 10);
        if (*end || v < 0) return fail(PTC_EARG, "bad max_steps %s", value);
        g_opt.max_steps = v;
    } else if (!strcmp(key, "cell_cache")) {
        if (strcmp(value, "0") && strcmp(value, "1"))
            return fail(PTC_EARG, "cell_cache must be 0 or 1 (got %s)", value);
        g_opt.cell_cache = value[0] - '0';
    } else if (!strcmp(key, "sort")) {
        long v = strtol(value, &end, 10);
        if (*end || v < 0 || v > 1000000) return fail(PTC_EARG, "bad sort cadence %s", value);
        g_opt.sort = (int)v;
    } else if (!strcmp(key, "order")) {
        if (!strcmp(value, "spatial")) g_opt.order = 1;
        else if (!strcmp(value, "cost")) g_opt.order = 2;
        else if (!strcmp(value, "auto")) g_opt.order = 3;
        else return fail(PTC_EARG, "order must be spatial, cost or auto (got %s)", value);
    } else if (!strcmp(key, "group_lanes")) {
        if (!strcmp(value, "3")) g_opt.lanes = 3;
        else if (!strcmp(value, "6")) g_opt.lanes = 6;
        else if (!strcmp(value, "auto")) g_opt.lanes = 0;
        else return fail(PTC_EARG, "group_lanes must be 3, 6 or auto (got %s)", value);
    } else if (!strcmp(key, "tail")) {
        if (!strcmp(value, "group")) {
            g_opt.tail = kTailGroupOnly;
        } else if (!strcmp(value, "auto")) {
            g_opt.tail = kTailAuto;
        } else {
            long v = strtol(value, &end, 10);
            if (*end || v < 0 || v > (1l << 30))
                return fail(PTC_EARG, "tail must be a particle count >= 0 or group (got %s)",
                            value);
            g_opt.tail = (int)v;
        }
    } else {
        return fail(PTC_EARG, "unknown option %s", key);
    }
    ok();
    return PTC_OK;
}

int ptc_get_option(const char *key, char *buf, int len)
{
    if (!key || !buf || len <= 0) return fail(PTC_EARG, "ptc_get_option: bad arguments");
    Options o = options();
    if (!strcmp(key, "block")) snprintf(buf, len, "%d", o.block);
    else if (!strcmp(key, "grid_blocks")) snprintf(buf, len, "%d", o.grid_blocks);
    else if (!strcmp(key, "layout"))
        snprintf(buf, len, "%s", o.layout == LAYOUT_SOA     ? "soa"
                                 : o.layout == LAYOUT_AOS   ? "aos"
                                 : o.layout == LAYOUT_PAIRS ? "pairs"
                                 : o.layout == LAYOUT_COLUMNS ? "columns"
                                                            : "brick");
    else if (!strcmp(key, "f32_layout"))
        snprintf(buf, len, "%s", o.f32_layout == LAYOUT_PAIRS     ? "pairs"
                                 : o.f32_layout == LAYOUT_QUADS   ? "quads"
                                 : o.f32_layout == LAYOUT_COLUMNS ? "columns"
                                                                  : "nodes");
    else if (!strcmp(key, "max_steps")) snprintf(buf, len, "%lld", (long long)o.max_steps);
    else if (!strcmp(key, "cell_cache")) snprintf(buf, len, "%d", o.cell_cache);
    else if (!strcmp(key, "sort")) snprintf(buf, len, "%d", o.sort);
    else if (!strcmp(key, "order"))
        snprintf(buf, len, "%s", o.order == 2 ? "cost" : o.order == 3 ? "auto" : "spatial");
    else if (!strcmp(key, "group_lanes")) {
        if (o.lanes == 0) snprintf(buf, len, "auto");
        else snprintf(buf, len, "%d", o.lanes);
    }
    else if (!strcmp(key, "tail")) {
        if (o.tail == kTailGroupOnly) snprintf(buf, len, "group");
        else if (o.tail == kTailAuto) snprintf(buf, len, "auto");
        else snprintf(buf, len, "%d", o.tail);
    }
    else return fail(PTC_EARG, "unknown option %s", key);
    ok();
    return PTC_OK;
}

}  // extern "C"

// Builds the handle from six SOA planes already on the device (`soa` is
// taken over: freed here, or kept as the field array for the SOA layout).
static int fields_from_device_soa(const int dm[3], double *soa, int mode, int prec,
                                  ptc_fields **out)
{
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    Options o = options();
    if (prec == PTC_FP32) o.layout = o.f32_layout;  // FP32: node, pair, quad or column rows
    void *data = soa, *bricks = nullptr;
    hipError_t e = hipSuccess;
    if (rows_layout(o.layout)) {
        void *rows = nullptr;
        const size_t esz = prec == PTC_FP32 ? sizeof(float) : sizeof(double);
        e = hipMalloc(&rows, esz * (size_t)row_values(o.layout, prec) * (size_t)nodes);
        if (e == hipSuccess) {
            if (o.layout == LAYOUT_QUADS) e = launch_to_quads_f32(soa, dm, (float *)rows, nullptr);
            else if (o.layout == LAYOUT_COLUMNS && prec == PTC_FP32)
                e = launch_to_columns_f32(soa, dm, (float *)rows, nullptr);
            else if (o.layout == LAYOUT_COLUMNS) e = launch_to_columns_f64(soa, dm, (double *)rows, nullptr);
            else if (prec == PTC_FP32) e = launch_to_pairs_f32(soa, dm, (float *)rows, nullptr);
            else e = launch_to_pairs_f64(soa, dm, (double *)rows, nullptr);
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(soa);
        if (e != hipSuccess) {
            if (rows) (void)hipFree(rows);
            return hip_fail(e, "ptc_fields: pair rows");
        }
        data = rows;
    }
    if (o.layout == LAYOUT_AOS || o.layout == LAYOUT_BRICK) {
        void *aos = nullptr;
        e = hipMalloc(&aos, prec == PTC_FP32 ? sizeof(float) * 8 * nodes : sizeof(double) * 6 * nodes);
        if (e == hipSuccess) e = launch_to_aos(soa, nodes, aos, prec, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(soa);
        if (e != hipSuccess) {
            if (aos) (void)hipFree(aos);
            return hip_fail(e, "ptc_fields: interleave");
        }
        data = aos;
    }
    if (o.layout == LAYOUT_BRICK) {
        e = hipMalloc(&bricks, sizeof(double) * 48 * nodes);
        if (e == hipSuccess) {
            DeviceFields tmp{};
            tmp.dm[0] = dm[0]; tmp.dm[1] = dm[1]; tmp.dm[2] = dm[2];
            tmp.prec = prec; tmp.nodes = nodes; tmp.data = data; tmp.bricks = bricks;
            e = launch_to_bricks(tmp, nullptr);
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            if (bricks) (void)hipFree(bricks);
            (void)hipFree(data);
            return hip_fail(e, "ptc_fields: bricks");
        }
    }
    ptc_fields *f = new ptc_fields();
    f->d.dm[0] = dm[0];
    f->d.dm[1] = dm[1];
    f->d.dm[2] = dm[2];
    f->d.mode = mode;
    f->d.prec = prec;
    f->d.layout = o.layout;
    f->d.nodes = nodes;
    f->d.data = data;
    f->d.bricks = bricks;
    *out = f;
    ok();
    return PTC_OK;
}

static int check_field_args(const int dm[3], int mode, int prec)
{
    if (dm[0] < 2 || dm[1] < 2 || dm[2] < 2)
        return fail(PTC_EARG, "ptc_fields: dimensions must be >= 2 (got %d %d %d)", dm[0],
                    dm[1], dm[2]);
    if (mode != PTC_MODE_VB && mode != PTC_MODE_EB) return fail(PTC_EARG, "bad field mode %d", mode);
    if (prec != PTC_FP64 && prec != PTC_FP32)
        return fail(PTC_EARG, "field precision %d not supported", prec);
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    if (nodes > (int64_t)0xFFFFFFFFll)
        return fail(PTC_EARG, "ptc_fields: %lld nodes exceed 32-bit node offsets", (long long)nodes);
    return PTC_OK;
}

extern "C" {

int ptc_fields_create(const int dm[3], const double *f1, const double *bb, int mode, int prec,
                      ptc_fields **out)
{
    if (!out) return fail(PTC_EARG, "ptc_fields_create: out is null");
    *out = nullptr;
    if (!dm || !f1 || !bb) return fail(PTC_EARG, "ptc_fields_create: null dm/field pointer");
    int rc = check_field_args(dm, mode, prec);
    if (rc) return rc;
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    double *soa = nullptr;
    HIP_TRY(hipMalloc(&soa, sizeof(double) * 6 * nodes));
    hipError_t e = hipMemcpy(soa, f1, sizeof(double) * 3 * nodes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(soa + 3 * nodes, bb, sizeof(double) * 3 * nodes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(soa);
        return hip_fail(e, "ptc_fields_create: upload");
    }
    return fields_from_device_soa(dm, soa, mode, prec, out);
}

int ptc_fields_synthesize(const int dm[3], int recipe, const double *scalars, int nb,
                          const int32_t *b_n, const double *b_amp, const double *b_pol,
                          const double *b_phase, int nv, const int32_t *v_n, const double *v_amp,
                          const double *v_pol, const double *v_phase, int prec, ptc_fields **out)
{
    if (!out) return fail(PTC_EARG, "ptc_fields_synthesize: out is null");
    *out = nullptr;
    if (!dm || !scalars) return fail(PTC_EARG, "ptc_fields_synthesize: null dm/scalars");
    if (recipe != PTC_RECIPE_TURBULENT_VB && recipe != PTC_RECIPE_RECONNECTION_EB)
        return fail(PTC_EARG, "ptc_fields_synthesize: unknown recipe %d", recipe);
    const int mode = recipe == PTC_RECIPE_TURBULENT_VB ? PTC_MODE_VB : PTC_MODE_EB;
    int rc = check_field_args(dm, mode, prec);
    if (rc) return rc;
    if (nb < 0 || nv < 0 || (nb > 0 && (!b_n || !b_amp || !b_pol || !b_phase)) ||
        (nv > 0 && (!v_n || !v_amp || !v_pol || !v_phase)))
        return fail(PTC_EARG, "ptc_fields_synthesize: bad mode tables");
    if (recipe == PTC_RECIPE_RECONNECTION_EB && !(scalars[0] > 0.0))
        return fail(PTC_EARG, "ptc_fields_synthesize: sheet width must be > 0");
    const int64_t nodes = (int64_t)dm[0] * dm[1] * dm[2];
    double *soa = nullptr;
    HIP_TRY(hipMalloc(&soa, sizeof(double) * 6 * nodes));
    // mode tables: n (int32), then amp, pol, phase (double), for B then V
    const int nt = nb + nv;
    void *tab = nullptr;
    hipError_t e = hipMalloc(&tab, (sizeof(int32_t) * 4 + sizeof(double) * 5) * (nt > 0 ? nt : 1));
    auto fin = [&](hipError_t err, const char *what) {
        (void)hipFree(soa);
        if (tab) (void)hipFree(tab);
        return hip_fail(err, what);
    };
    if (e != hipSuccess) return fin(e, "ptc_fields_synthesize: tables");
    double *dd = (double *)tab;          // amp (nt), pol (3 nt), phase (nt)
    int32_t *dn = (int32_t *)(dd + 5 * (size_t)(nt > 0 ? nt : 1));
    std::vector<double> h(5 * (size_t)(nt > 0 ? nt : 1), 0.0);
    std::vector<int32_t> hn(3 * (size_t)(nt > 0 ? nt : 1), 0);
    for (int m = 0; m < nt; m++) {
        const bool isb = m < nb;
        const int q = isb ? m : m - nb;
        h[m] = isb ? b_amp[q] : v_amp[q];
        for (int c = 0; c < 3; c++) {
            h[nt + 3 * m + c] = isb ? b_pol[3 * q + c] : v_pol[3 * q + c];
            hn[3 * m + c] = isb ? b_n[3 * q + c] : v_n[3 * q + c];
        }
        h[4 * nt + m] = isb ? b_phase[q] : v_phase[q];
    }
    if (nt > 0) {
        e = hipMemcpy(dd, h.data(), sizeof(double) * 5 * nt, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dn, hn.data(), sizeof(int32_t) * 3 * nt, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fin(e, "ptc_fields_synthesize: tables");
    }
    double *f1 = soa, *b = soa + 3 * nodes;
    e = synth_fill(soa, 6 * nodes, 0.0, nullptr);
    if (e == hipSuccess)
        e = synth_modes(dm, nb, dn, dd, dd + nt, dd + 4 * nt, b, nullptr);
    if (recipe == PTC_RECIPE_TURBULENT_VB) {
        // U = V modes, B = b0 z + B modes
        if (e == hipSuccess)
            e = synth_modes(dm, nv, dn + 3 * nb, dd + nb, dd + nt + 3 * nb, dd + 4 * nt + nb, f1,
                            nullptr);
        if (e == hipSuccess) e = synth_add(b + 2 * nodes, nodes, scalars[0], nullptr);
    } else {
        // B += sheet(y) x + guide z; E = -V x B + eta curl B
        if (e == hipSuccess) e = synth_sheet(dm, scalars[0], scalars[1], b, nullptr);
        if (e == hipSuccess)
            e = synth_reconnection_e(dm, scalars[0], scalars[2], scalars[3], b, f1, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(tab);
    tab = nullptr;
    if (e != hipSuccess) {
        (void)hipFree(soa);
        return hip_fail(e, "ptc_fields_synthesize");
    }
    return fields_from_device_soa(dm, soa, mode, prec, out);
}

int ptc_fields_series(const ptc_fields *const *snaps, const double *times, int n,
                      ptc_fields **out)
{
    if (!out) return fail(PTC_EARG, "ptc_fields_series: out is null");
    *out = nullptr;
    if (!snaps || !times || n < 1) return fail(PTC_EARG, "ptc_fields_series: need n >= 1 snapshots and times");
    if (!snaps[0]) return fail(PTC_EARG, "ptc_fields_series: snapshot 0 is null");
    const DeviceFields &f0 = snaps[0]->d;
    for (int i = 0; i < n; i++) {
        if (!snaps[i]) return fail(PTC_EARG, "ptc_fields_series: snapshot %d is null", i);
        const DeviceFields &fi = snaps[i]->d;
        if (fi.nsnap > 0) return fail(PTC_EARG, "ptc_fields_series: snapshot %d is a series", i);
        if (fi.layout != LAYOUT_AOS && fi.layout != LAYOUT_COLUMNS)
            return fail(PTC_EARG, "ptc_fields_series: snapshot %d is not in node rows (layout "
                                  "aos, f32_layout nodes) or column rows", i);
        if (fi.dm[0] != f0.dm[0] || fi.dm[1] != f0.dm[1] || fi.dm[2] != f0.dm[2] ||
            fi.mode != f0.mode || fi.prec != f0.prec || fi.layout != f0.layout)
            return fail(PTC_EARG, "ptc_fields_series: snapshot %d differs from snapshot 0 in "
                                  "dimensions, mode, precision or layout", i);
        if (!std::isfinite(times[i]))
            return fail(PTC_EARG, "ptc_fields_series: time %d is not finite", i);
        if (i > 0 && !(times[i] > times[i - 1]))
            return fail(PTC_EARG, "ptc_fields_series: times must increase strictly (%d)", i);
    }
    // node rows (6 doubles or 8 floats) or column rows (24 doubles or 32
    // floats) per node
    const int64_t stride = (int64_t)row_values(f0.layout, f0.prec) * f0.nodes;
    const size_t bytes = (size_t)stride * (f0.prec == PTC_FP32 ? sizeof(float) : sizeof(double));
    void *data = nullptr;
    double *st = nullptr;
    HIP_TRY(hipMalloc(&data, bytes * (size_t)n));
    hipError_t e = hipMalloc(&st, sizeof(double) * (size_t)n);
    for (int i = 0; e == hipSuccess && i < n; i++)
        e = hipMemcpy((char *)data + bytes * (size_t)i, snaps[i]->d.data, bytes,
                      hipMemcpyDeviceToDevice);
    if (e == hipSuccess) e = hipMemcpy(st, times, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(data);
        if (st) (void)hipFree(st);
        return hip_fail(e, "ptc_fields_series");
    }
    ptc_fields *f = new ptc_fields();
    f->d = f0;
    f->d.data = data;
    f->d.bricks = nullptr;
    f->d.nsnap = n;
    f->d.sstride = stride;
    f->d.stime = st;
    *out = f;
    ok();
    return PTC_OK;
}

int ptc_fields_convert(const ptc_fields *src, int prec, ptc_fields **out)
{
    if (!out) return fail(PTC_EARG, "ptc_fields_convert: out is null");
    *out = nullptr;
    if (!src) return fail(PTC_EARG, "ptc_fields_convert: source handle is null");
    if (prec != PTC_FP64 && prec != PTC_FP32)
        return fail(PTC_EARG, "ptc_fields_convert: precision %d not supported", prec);
    const DeviceFields &s = src->d;
    if (s.nsnap > 0 || s.layout == LAYOUT_BRICK)
        return fail(PTC_EARG, "ptc_fields_convert: the source must be a static field in node, "
                              "pair, quad, column or plane (soa) storage");
    Options o = options();
    const int layout = prec == PTC_FP32 ? o.f32_layout : o.layout;
    if (layout == LAYOUT_SOA || layout == LAYOUT_BRICK)
        return fail(PTC_EARG, "ptc_fields_convert: the target layout must be node, pair, quad "
                              "or column rows (layout aos, pairs or columns; f32_layout nodes, "
                              "pairs, quads or columns)");
    const size_t esz = prec == PTC_FP32 ? sizeof(float) : sizeof(double);
    void *rows = nullptr;
    HIP_TRY(hipMalloc(&rows, esz * (size_t)row_values(layout, prec) * (size_t)s.nodes));
    hipError_t e = launch_convert_rows(s, rows, layout, prec, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(rows);
        return hip_fail(e, "ptc_fields_convert");
    }
    ptc_fields *f = new ptc_fields();
    f->d = s;
    f->d.prec = prec;
    f->d.layout = layout;
    f->d.data = rows;
    f->d.bricks = nullptr;
    f->d.nsnap = 0;
    f->d.sstride = 0;
    f->d.stime = nullptr;
    *out = f;
    ok();
    return PTC_OK;
}

int ptc_fields_download(const ptc_fields *f, int snapshot, double *f1, double *bb)
{
    if (!f || !f1 || !bb) return fail(PTC_EARG, "ptc_fields_download: null argument");
    const DeviceFields &d = f->d;
    const int ns = d.nsnap > 0 ? d.nsnap : 1;
    if (snapshot < 0 || snapshot >= ns)
        return fail(PTC_EARG, "ptc_fields_download: snapshot %d outside [0, %d)", snapshot, ns);
    double *plane = nullptr;
    HIP_TRY(hipMalloc(&plane, sizeof(double) * (size_t)d.nodes));
    hipError_t e = hipSuccess;
    for (int c = 0; c < 6 && e == hipSuccess; c++) {
        double *dst = (c < 3 ? f1 : bb) + (size_t)(c % 3) * (size_t)d.nodes;
        e = launch_extract_plane(d, snapshot, c, plane, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess)
            e = hipMemcpy(dst, plane, sizeof(double) * (size_t)d.nodes, hipMemcpyDeviceToHost);
    }
    (void)hipFree(plane);
    if (e != hipSuccess) return hip_fail(e, "ptc_fields_download");
    ok();
    return PTC_OK;
}

int ptc_fields_destroy(ptc_fields *f)
{
    if (!f) return PTC_OK;
    hipError_t e = hipFree(f->d.data);
    if (f->d.bricks) {
        hipError_t e2 = hipFree(f->d.bricks);
        if (e == hipSuccess) e = e2;
    }
    if (f->d.stime) {
        hipError_t e2 = hipFree((void *)f->d.stime);
        if (e == hipSuccess) e = e2;
    }
    delete f;
    if (e != hipSuccess) return hip_fail(e, "ptc_fields_destroy");
    ok();
    return PTC_OK;
}

// The "auto" tail setting for a batch (measured on the MI355X, every choice
// giving the same results): the group kernel alone for Cash-Karp (its
// trajectories differ most in length, and the lane-per-particle launch is set
// by the slowest), for batches that leave most lanes of the lane-per-particle
// kernel idle (64^3 RK5 with 1,000 particles: 12.4 ms per T0 against 29.0
// for one lane per particle, and 14.4 for the trio kernel, whose single
// trajectories run slower), and for fields beyond 2 GB (the gathers then miss the caches
// and the group kernel's two waves per SIMD hide their latency better);
// otherwise one lane per particle, in FP64 handing the last 32768 over (4 %
// faster on the 256^3 DOP853 workload; in FP32 the hand-over costs 4 %, and
// 10 % with the trio kernel taking the tail: 34.0 against 30.7 ms per T0).
// A snapshot series (two snapshots per gather; no hand-over, the lane
// kernel's series form keeps its particles to the end) takes the group
// kernel in FP64 as well (measured on 8 snapshots of 256^3 with 100,000
// particles: DOP853 1.15x, RK5 2.3x the lane kernel; FP32 DOP853 0.62x).
static int choose_tail(int tail, const DeviceFields &f, int method, int64_t np, int ctrl)
{
    // the learned controller runs in the group kernel on FP32 node rows
    // only, and never hands over; training records come from one lane per
    // particle
    const bool learned = ctrl == CTRL_LEARNED;
    if (rows_layout(f.layout)) return kTailGroupOnly;  // read by the group kernel only
    if (learned && f.prec != PTC_FP32) return 0;
    if (tail != kTailAuto) return learned && tail > 0 ? 0 : tail;
    if (f.layout != LAYOUT_AOS || (ctrl != CTRL_PI && !learned)) return 0;
    const double snap = (f.prec == PTC_FP32 ? 32.0 : 48.0) * (double)f.nodes;
    const double bytes = f.nsnap > 0 ? 2.0 * snap : snap;
    if (method == PTC_RK5 || np <= 32768 || bytes > 2.0 * 1024 * 1024 * 1024 ||
        (f.nsnap > 0 && f.prec == PTC_FP64))
        return kTailGroupOnly;
    if (f.nsnap > 0 || learned) return 0;
    return f.prec == PTC_FP32 ? 0 : 32768;
}

// The "auto" work-queue order (measured on the MI355X, every order giving
// the same results): the cost order (most RHS evaluations in the previous
// segment first) for batches whose launch is set by its slowest
// trajectories (100,000 particles at 256^3: +30 %; 1,000,000 at 512^3:
// +8 %); index order beyond 2,000,000 particles, where the tail is a small
// part of the launch and grouping the costly particles into the same waves
// costs gathers their latency hiding (4,000,000 at 1024^3: index order
// 3-8 % faster in FP32, equal in FP64).
static void resolve_order(Options &o, int64_t np)
{
    if (o.order != 3) return;
    if (np > 2000000) o.sort = 0;
    else o.order = 2;
}

static int resolve_tail(int tail, const ptc_particles *p)
{
    return choose_tail(tail, p->fields->d, p->method, p->d.np, p->ctrl);
}

// Lanes per particle of the kernel that runs a static field with the PI
// controller where particles share a wave's lanes ("group_lanes" auto,
// measured on the MI355X, both giving the same results): three (the trio
// kernel) for column rows (FP64: 1.25-1.44x the six-lane kernel on 256^3
// DOP853 with 100,000 particles, 512^3 RK5 with 1,000,000, 1024^3 DOP853
// and RK5 with 1,000,000 and 4,000,000; FP32: 1.32x on 256^3 DOP853, 1.25x
// on 512^3 RK5, 1.36x on 1024^3 DOP853) and for FP32 DOP853 batches of at
// least 2,000,000 particles on other rows (1024^3, 4,000,000 particles on
// quad rows: 408 against 472 ms per gyro-period); six elsewhere (the trio
// kernel measured 6-13 % slower on FP32 quad rows with 512^3 RK5 and with
// DOP853 batches of 1,000,000 particles, and 4-7 % slower on FP64 pair rows
// at 1024^3).  Snapshot series (8 snapshots of 256^3, DOP853, 100,000
// particles): six in FP64, where the trio's series form spills 268-292 B
// per lane and measured 1.17-1.46x slower (node and column rows); three in
// FP32 (36-156 B of spills), 1.11-1.38x faster than six lanes.
static int choose_lanes(int lanes, const DeviceFields &f, int method, int64_t np)
{
    if (lanes != 0) return lanes;
    if (f.nsnap > 0) return f.prec == PTC_FP32 ? 3 : 6;
    if (f.layout == LAYOUT_COLUMNS) return 3;
    return f.prec == PTC_FP32 && method == PTC_DOP853 && np >= 2000000 ? 3 : 6;
}

static int build_step_params(const ptc_fields *f, int m, const double *params, int nparams,
                             StepParams &sp)
{
    if (!params || nparams < 17) return fail(PTC_EARG, "params must have at least 17 entries");
    for (int i = 0; i < 17; i++)
        if (!std::isfinite(params[i])) return fail(PTC_EARG, "params(%d) is not finite", i + 1);
    sp.qom = params[0];
    sp.vun = params[1];
    sp.tmx = params[2];
    sp.dtm = params[4];
    sp.atol = params[5];
    sp.rtol = params[6];
    sp.fcmn = params[7];
    sp.fcmx = params[8];
    sp.safe = params[9];
    sp.beta = params[10];
    sp.expo = 2.0e-01 * sp.beta - 1.25e-01;  // integrations.F90:348
    // a power of two scales exactly: division by it equals multiplication by
    // its reciprocal (both round the same real number), when that reciprocal
    // is itself a normal double
    auto pow2 = [](double x) {
        int e = 0;
        return std::isnormal(x) && std::isnormal(1.0 / x) && std::fabs(std::frexp(x, &e)) == 0.5;
    };
    sp.pow2 = pow2(params[1]);
    sp.rvun = 1.0 / params[1];
    for (int i = 0; i < 3; i++) {
        sp.xl[i] = params[11 + i];
        sp.xs[i] = params[14 + i];
        sp.rxs[i] = 1.0 / sp.xs[i];
        sp.pow2 = sp.pow2 && pow2(sp.xs[i]);
        sp.dm[i] = f->d.dm[i];
        sp.dmd[i] = (double)f->d.dm[i];
        sp.inv_dm[i] = 1.0 / (double)f->d.dm[i];
    }
    // cases where the reference would never return (no step-size floor)
    if (!(sp.dtm > 0.0)) return fail(PTC_EARG, "dtmax = params(5) must be > 0 (got %g)", sp.dtm);
    if (sp.xs[0] == 0.0 || sp.xs[1] == 0.0 || sp.xs[2] == 0.0)
        return fail(PTC_EARG, "domain size params(15:17) must be nonzero");
    if (sp.vun == 0.0) return fail(PTC_EARG, "velocity unit params(2) must be nonzero");
    sp.m = m;
    sp.max_steps = options().max_steps;
    return PTC_OK;
}

int ptc_particles_create(const ptc_fields *f, int method, int64_t np, int m, const double *params,
                         int nparams, const double *time, const double *state, ptc_particles **out)
{
    if (!out) return fail(PTC_EARG, "ptc_particles_create: out is null");
    *out = nullptr;
    if (!f) return fail(PTC_EARG, "ptc_particles_create: fields handle is null");
    if (method != PTC_RK5 && method != PTC_DOP853) return fail(PTC_EARG, "bad method %d", method);
    if (np < 0 || np > (int64_t)0x7F000000) return fail(PTC_EARG, "bad particle count %lld", (long long)np);
    if (m < 1) return fail(PTC_EARG, "need at least one snapshot (m = %d)", m);
    if (np > 0 && !time) return fail(PTC_EARG, "null time");
    StepParams sp;
    int rc = build_step_params(f, m, params, nparams, sp);
    if (rc) return rc;

    ptc_particles *p = new ptc_particles();
    p->fields = f;
    p->method = method;
    p->sp = sp;
    p->d.np = np;
    p->last_ms = 0.0;

    const int64_t n = np > 0 ? np : 1;
    const size_t bytes = sizeof(double) * (9 * n + (size_t)8 * m * n) + sizeof(int32_t) * 3 * n +
                         sizeof(int64_t) * 4 * n;
    hipError_t e = hipMalloc(&p->block, bytes);
    if (e != hipSuccess) { delete p; return hip_fail(e, "ptc_particles_create: state"); }
    char *c = (char *)p->block;
    double *dx = (double *)c;
    for (int k = 0; k < 6; k++) p->d.x[k] = dx + k * n;
    p->d.t = dx + 6 * n;
    p->d.dt = dx + 7 * n;
    p->d.errold = dx + 8 * n;
    p->d.out = dx + 9 * n;
    c += sizeof(double) * (9 * n + (size_t)8 * m * n);
    p->d.cnt = (int64_t *)c;
    p->d.nrhs = p->d.cnt + 3 * n;
    c += sizeof(int64_t) * 4 * n;
    p->d.nout = (int32_t *)c;
    p->d.flags = p->d.nout + n;
    p->d.status = p->d.nout + 2 * n;
    p->d.order = nullptr;

    e = hipMalloc(&p->d_time, sizeof(double) * m + 64 * sizeof(int64_t));
    if (e != hipSuccess) { (void)hipFree(p->block); delete p; return hip_fail(e, "ptc_particles_create: time"); }
    p->d.time = p->d_time;
    p->d_totals = (int64_t *)(p->d_time + m);
    p->d.qcount = (int32_t *)(p->d_totals + 8);
    p->d.loads = (unsigned long long *)(p->d_totals + 16);

    auto cleanup = [&](hipError_t err, const char *what) {
        (void)hipFree(p->d_time);
        (void)hipFree(p->block);
        delete p;
        return hip_fail(err, what);
    };
    // totals, queue heads and the cell-load counter start at zero
    e = hipMemset(p->d_totals, 0, 64 * sizeof(int64_t));
    if (e != hipSuccess) return cleanup(e, "ptc_particles_create: counters");
    e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(e, "hipStreamCreate");
    if ((e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess)
        return cleanup(e, "hipEventCreate");

    // initial state: SoA copy of state(1:6,1,:), full state(8,m,np) image
    if (np > 0) {
        std::vector<double> h(9 * (size_t)np, 0.0);
        for (int64_t i = 0; i < np && state; i++) {
            const double *s = state + (size_t)8 * m * i;
            for (int k = 0; k < 6; k++) h[k * np + i] = s[k];
        }
        for (int64_t i = 0; i < np; i++) {
            h[6 * np + i] = 0.0;        // t
            h[7 * np + i] = 0.0;        // dt
            h[8 * np + i] = 1.0e-04;    // errold (integrations.F90:349)
        }
        e = hipMemcpy(dx, h.data(), sizeof(double) * 9 * np, hipMemcpyHostToDevice);
        if (e == hipSuccess && state)
            e = hipMemcpy(p->d.out, state, sizeof(double) * 8 * m * np, hipMemcpyHostToDevice);
        if (e == hipSuccess && !state)
            e = hipMemset(p->d.out, 0, sizeof(double) * 8 * m * np);
        std::vector<int32_t> hi(3 * (size_t)np);
        for (int64_t i = 0; i < np; i++) {
            hi[i] = 2;                            // next snapshot (1-based)
            hi[np + i] = 0;                       // flags
            hi[2 * np + i] = PTC_STATUS_RUNNING;  // status
        }
        if (e == hipSuccess)
            e = hipMemcpy(p->d.nout, hi.data(), sizeof(int32_t) * 3 * np, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(p->d.cnt, 0, sizeof(int64_t) * 4 * np);
        if (e == hipSuccess)
            e = hipMemcpy(p->d_time, time, sizeof(double) * m, hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(e, "ptc_particles_create: upload");
    }
    *out = p;
    ok();
    return PTC_OK;
}

int ptc_particles_advance(ptc_particles *p, double t_stop)
{
    if (!p) return fail(PTC_EARG, "ptc_particles_advance: null handle");
    if (std::isnan(t_stop)) return fail(PTC_EARG, "ptc_particles_advance: t_stop is NaN");
    p->last_ms = 0.0;
    if (p->d.np == 0) { ok(); return PTC_OK; }
    Options o = options();
    if (p->tuned) {
        o.block = p->tuned_block;
        o.sort = p->tuned_sort;
        o.order = p->tuned_order;
        o.tail = p->tuned_tail;
        o.lanes = p->tuned_lanes;
    }
    resolve_order(o, p->d.np);
    p->sp.max_steps = o.max_steps;
    LaunchConfig lc{o.block, o.grid_blocks,
                    p->fields->d.layout != LAYOUT_SOA && p->fields->d.layout != LAYOUT_BRICK
                        ? o.cell_cache
                        : 0,
                    p->ctrl, resolve_tail(o.tail, p),
                    choose_lanes(o.lanes, p->fields->d, p->method, p->d.np)};
    if (p->ctrl != CTRL_PI && lc.cell_cache != 1)
        return fail(PTC_ESTATE, "record and learned step-size modes need the aos layout with "
                                "cell_cache 1");
    if (o.sort > 0) {
        // order (np int32), previous RHS counts (np int64), sort scratch
        const size_t head = sizeof(int32_t) * (size_t)p->d.np + sizeof(int64_t) * (size_t)p->d.np;
        if (!p->order_block) {
            p->order_bytes = head + 256 + order_scratch_bytes(p->d.np);
            HIP_TRY(hipMalloc(&p->order_block, p->order_bytes));
            HIP_TRY(launch_zero_i64((int64_t *)((int32_t *)p->order_block + p->d.np), p->d.np,
                                    p->stream));
        }
        int32_t *order = (int32_t *)p->order_block;
        int64_t *prev = (int64_t *)(order + p->d.np);
        char *scratch = (char *)p->order_block + ((head + 255) & ~(size_t)255);
        const size_t sbytes = p->order_bytes - (size_t)(scratch - (char *)p->order_block);
        if (p->advances % o.sort == 0 || !p->d.order) {
            p->d.order = nullptr;  // the sort reads particles by their own index
            if (o.order == 2)
                HIP_TRY(launch_cost_order(p->d, prev, order, scratch, sbytes, p->stream));
            else
                HIP_TRY(launch_spatial_order(p->d, p->sp, order, scratch, sbytes, p->stream));
        }
        p->d.order = order;
    } else {
        p->d.order = nullptr;
    }
    p->advances++;
    HIP_TRY(launch_zero_i32(p->d.qcount, 3, p->stream));
    HIP_TRY(hipEventRecord(p->ev0, p->stream));
    HIP_TRY(launch_integrate(p->method, p->fields->d, p->sp, p->d, t_stop, lc, p->stream));
    HIP_TRY(hipEventRecord(p->ev1, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p->ev0, p->ev1));
    p->last_ms = ms;
    ok();
    return PTC_OK;
}

int ptc_particles_download(ptc_particles *p, double *state, int64_t *counter, int32_t *status)
{
    if (!p) return fail(PTC_EARG, "ptc_particles_download: null handle");
    const int64_t np = p->d.np;
    if (np == 0) { ok(); return PTC_OK; }
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (state)
        HIP_TRY(hipMemcpy(state, p->d.out, sizeof(double) * 8 * (size_t)p->sp.m * np,
                          hipMemcpyDeviceToHost));
    if (counter) {
        std::vector<int64_t> h(3 * (size_t)np);
        HIP_TRY(hipMemcpy(h.data(), p->d.cnt, sizeof(int64_t) * 3 * np, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < np; i++)
            for (int k = 0; k < 3; k++) counter[3 * i + k] = h[k * np + i];
    }
    if (status)
        HIP_TRY(hipMemcpy(status, p->d.status, sizeof(int32_t) * np, hipMemcpyDeviceToHost));
    ok();
    return PTC_OK;
}

int ptc_particles_generate(ptc_particles *p, int kind, uint64_t seed, const double lo[3],
                           const double hi[3], double a, double b)
{
    if (!p || !lo || !hi) return fail(PTC_EARG, "ptc_particles_generate: null argument");
    if (p->advances > 0)
        return fail(PTC_ESTATE, "ptc_particles_generate: the batch has already been advanced");
    if (kind != PTC_INIT_MAXWELL && kind != PTC_INIT_GAMMA && kind != PTC_INIT_GAMMA_LOG)
        return fail(PTC_EARG, "ptc_particles_generate: unknown kind %d", kind);
    for (int i = 0; i < 3; i++)
        if (!std::isfinite(lo[i]) || !std::isfinite(hi[i]))
            return fail(PTC_EARG, "ptc_particles_generate: region not finite");
    if (!std::isfinite(a) || (kind == PTC_INIT_MAXWELL && a < 0.0) ||
        (kind != PTC_INIT_MAXWELL && a < 1.0) ||
        (kind == PTC_INIT_GAMMA_LOG && !(b >= a && std::isfinite(b))))
        return fail(PTC_EARG, "ptc_particles_generate: bad distribution parameters %g %g", a, b);
    if (p->d.np == 0) { ok(); return PTC_OK; }
    HIP_TRY(launch_init(p->d, p->sp.m, kind, seed, lo, hi, a, b, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    ok();
    return PTC_OK;
}

int ptc_particles_moments(ptc_particles *p, double moments[24])
{
    if (!p || !moments) return fail(PTC_EARG, "ptc_particles_moments: null argument");
    for (int j = 0; j < 24; j++) moments[j] = 0.0;
    const int64_t n = p->d.np;
    if (n == 0) { ok(); return PTC_OK; }
    int blocks = (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    double *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(double) * (6 + 24 * (size_t)blocks)));
    std::vector<double> h(24 * (size_t)blocks);
    double shift[6] = {0, 0, 0, 0, 0, 0}, mean[6];
    hipError_t e = hipSuccess;
    // pass 1: means; pass 2: central moments about them
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        e = hipMemcpy(d, shift, sizeof shift, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = launch_moments(p->d, p->sp.m, d, d + 6, blocks, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e == hipSuccess)
            e = hipMemcpy(h.data(), d + 6, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) break;
        double sum[24] = {0};
        for (int b = 0; b < blocks; b++)
            for (int j = 0; j < 24; j++) sum[j] += h[(size_t)b * 24 + j];
        for (int c = 0; c < 6; c++) {
            if (pass == 0) {
                mean[c] = sum[4 * c] / (double)n;
                shift[c] = mean[c];
            } else {
                const double m2 = sum[4 * c + 1] / (double)n;
                const double sd = std::sqrt(m2);
                moments[4 * c] = mean[c];
                moments[4 * c + 1] = sd;
                moments[4 * c + 2] = sd > 0.0 ? sum[4 * c + 2] / (double)n / (sd * sd * sd) : 0.0;
                moments[4 * c + 3] = sd > 0.0 ? sum[4 * c + 3] / (double)n / (m2 * m2) - 3.0 : -3.0;
            }
        }
    }
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "ptc_particles_moments");
    ok();
    return PTC_OK;
}

namespace {

void drop_graph(ptc_particles::AdvanceGraph &g)
{
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
    g.n = 0;
    g.nodes.clear();
    g.interval.clear();
    g.replayable = false;
}

// the node a capture on stream s currently ends in (the captured work is one
// chain, so there is exactly one)
hipError_t capture_end(hipStream_t s, hipGraphNode_t *node)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    unsigned long long id = 0;
    hipGraph_t graph = nullptr;
    const hipGraphNode_t *deps = nullptr;
    size_t nd = 0;
    const hipError_t e = hipStreamGetCaptureInfo_v2(s, &st, &id, &graph, &deps, &nd);
    if (e != hipSuccess) return e;
    if (st != hipStreamCaptureStatusActive || nd != 1 || !deps) return hipErrorInvalidValue;
    *node = deps[0];
    return hipSuccess;
}

// Captures the n intervals (per interval: reset the queue heads with a fill
// kernel, then the integrator kernels to that interval's stop time) into
// p->graph, instantiates it, and records each integrator kernel node with its
// interval: the nodes between the interval's fill node and the capture's end
// after its launches.
hipError_t capture_graph(ptc_particles *p, int n, const double *t_stop, const LaunchConfig &lc)
{
    ptc_particles::AdvanceGraph &g = p->graph;
    std::vector<hipGraphNode_t> fill(n), last(n);
    hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    for (int k = 0; k < n && e == hipSuccess; k++) {
        e = launch_zero_i32(p->d.qcount, 3, p->stream);
        if (e == hipSuccess) e = capture_end(p->stream, &fill[k]);
        if (e == hipSuccess)
            e = launch_integrate(p->method, p->fields->d, p->sp, p->d, t_stop[k], lc, p->stream);
        if (e == hipSuccess) e = capture_end(p->stream, &last[k]);
    }
    const hipError_t e2 = hipStreamEndCapture(p->stream, &g.graph);
    if (e == hipSuccess) e = e2;
    for (int k = 0; k < n && e == hipSuccess; k++) {
        hipGraphNode_t node = last[k];
        while (e == hipSuccess && node != fill[k]) {
            hipGraphNodeType type;
            e = hipGraphNodeGetType(node, &type);
            if (e == hipSuccess && type != hipGraphNodeTypeKernel) e = hipErrorInvalidValue;
            if (e != hipSuccess) break;
            g.nodes.push_back(node);
            g.interval.push_back(k);
            size_t nd = 0;
            e = hipGraphNodeGetDependencies(node, nullptr, &nd);
            if (e == hipSuccess && nd != 1) e = hipErrorInvalidValue;
            if (e == hipSuccess) e = hipGraphNodeGetDependencies(node, &node, &nd);
        }
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
    // replays rewrite argument 3 of these nodes: only where each node holds
    // one pointer per argument and they are the values just launched
    g.replayable = e == hipSuccess && !g.nodes.empty();
    for (size_t j = 0; j < g.nodes.size() && g.replayable; j++) {
        hipKernelNodeParams kp{};
        if (hipGraphKernelNodeGetParams(g.nodes[j], &kp) != hipSuccess || !kp.kernelParams ||
            kp.extra) {
            g.replayable = false;
            break;
        }
        const DeviceFields *fa = (const DeviceFields *)kp.kernelParams[0];
        const DeviceParticles *pa = (const DeviceParticles *)kp.kernelParams[2];
        g.replayable = fa && pa && kp.kernelParams[3] && fa->data == p->fields->d.data &&
                       pa->np == p->d.np && pa->out == p->d.out && pa->qcount == p->d.qcount &&
                       std::memcmp(kp.kernelParams[3], &t_stop[g.interval[j]], sizeof(double)) == 0;
    }
    return e;
}

// New stop times into the instantiated graph: every integrator kernel takes
// (fields, step parameters, particles, stop time); the first three keep the
// values the node was captured with.
hipError_t set_graph_stops(ptc_particles::AdvanceGraph &g, const double *t_stop)
{
    for (size_t j = 0; j < g.nodes.size(); j++) {
        hipKernelNodeParams kp{};
        hipError_t e = hipGraphKernelNodeGetParams(g.nodes[j], &kp);
        if (e != hipSuccess) return e;
        if (!kp.kernelParams || kp.extra) return hipErrorNotSupported;
        double stop = t_stop[g.interval[j]];
        void *args[4] = {kp.kernelParams[0], kp.kernelParams[1], kp.kernelParams[2], &stop};
        kp.kernelParams = args;
        e = hipGraphExecKernelNodeSetParams(g.exec, g.nodes[j], &kp);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

int ptc_particles_advance_graph(ptc_particles *p, int n, const double *t_stop)
{
    if (!p || (n > 0 && !t_stop)) return fail(PTC_EARG, "ptc_particles_advance_graph: null argument");
    if (n < 0 || n > 100000) return fail(PTC_EARG, "ptc_particles_advance_graph: bad count %d", n);
    for (int k = 0; k < n; k++)
        if (std::isnan(t_stop[k])) return fail(PTC_EARG, "ptc_particles_advance_graph: NaN t_stop");
    p->last_ms = 0.0;
    if (p->d.np == 0 || n == 0) { ok(); return PTC_OK; }
    Options o = options();
    if (p->tuned) {
        o.block = p->tuned_block;
        o.sort = p->tuned_sort;
        o.order = p->tuned_order;
        o.tail = p->tuned_tail;
        o.lanes = p->tuned_lanes;
    }
    resolve_order(o, p->d.np);
    p->sp.max_steps = o.max_steps;
    LaunchConfig lc{o.block, o.grid_blocks,
                    p->fields->d.layout != LAYOUT_SOA && p->fields->d.layout != LAYOUT_BRICK
                        ? o.cell_cache
                        : 0,
                    p->ctrl, resolve_tail(o.tail, p),
                    choose_lanes(o.lanes, p->fields->d, p->method, p->d.np)};
    if (p->ctrl != CTRL_PI && lc.cell_cache != 1)
        return fail(PTC_ESTATE, "record and learned step-size modes need the aos layout with "
                                "cell_cache 1");
    // the queue order is set once, before capture (the library sort stays
    // out of the graph); every interval of the graph uses it
    if (o.sort > 0) {
        const size_t head = sizeof(int32_t) * (size_t)p->d.np + sizeof(int64_t) * (size_t)p->d.np;
        if (!p->order_block) {
            p->order_bytes = head + 256 + order_scratch_bytes(p->d.np);
            HIP_TRY(hipMalloc(&p->order_block, p->order_bytes));
            HIP_TRY(launch_zero_i64((int64_t *)((int32_t *)p->order_block + p->d.np), p->d.np,
                                    p->stream));
        }
        int32_t *order = (int32_t *)p->order_block;
        int64_t *prev = (int64_t *)(order + p->d.np);
        char *scratch = (char *)p->order_block + ((head + 255) & ~(size_t)255);
        const size_t sbytes = p->order_bytes - (size_t)(scratch - (char *)p->order_block);
        p->d.order = nullptr;
        if (o.order == 2)
            HIP_TRY(launch_cost_order(p->d, prev, order, scratch, sbytes, p->stream));
        else
            HIP_TRY(launch_spatial_order(p->d, p->sp, order, scratch, sbytes, p->stream));
        p->d.order = order;
    } else {
        p->d.order = nullptr;
    }
    HIP_TRY(hipStreamSynchronize(p->stream));
    ptc_particles::AdvanceGraph &g = p->graph;
    const bool same = g.exec && g.replayable && g.n == n && g.lc.block == lc.block &&
                      g.lc.grid_blocks == lc.grid_blocks && g.lc.cell_cache == lc.cell_cache &&
                      g.lc.ctrl == lc.ctrl && g.lc.tail == lc.tail && g.lc.lanes == lc.lanes &&
                      g.max_steps == p->sp.max_steps && g.order == p->d.order &&
                      g.ctrl == p->ctrl && g.net == (const void *)p->d.net &&
                      g.shift == p->d.ctrl_shift && g.rec_thresh == p->d.rec_thresh;
    hipError_t e = hipSuccess;
    if (same) {
        // replay: only the stop-time argument of each integrator node changes
        e = set_graph_stops(g, t_stop);
        if (e == hipSuccess) g.replayed++;
    }
    if (!same || e != hipSuccess) {
        drop_graph(g);
        e = capture_graph(p, n, t_stop, lc);
        if (e == hipSuccess) {
            g.n = n;
            g.lc = lc;
            g.max_steps = p->sp.max_steps;
            g.order = p->d.order;
            g.ctrl = p->ctrl;
            g.net = (const void *)p->d.net;
            g.shift = p->d.ctrl_shift;
            g.rec_thresh = p->d.rec_thresh;
            g.instantiated++;
        } else {
            drop_graph(g);
        }
    }
    if (e == hipSuccess) e = hipEventRecord(p->ev0, p->stream);
    if (e == hipSuccess) e = hipGraphLaunch(g.exec, p->stream);
    if (e == hipSuccess) e = hipEventRecord(p->ev1, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p->ev0, p->ev1);
    if (e != hipSuccess) return hip_fail(e, "ptc_particles_advance_graph");
    p->advances += n;
    p->last_ms = ms;
    ok();
    return PTC_OK;
}

int ptc_particles_graph_stats(const ptc_particles *p, int64_t stats[2])
{
    if (!p || !stats) return fail(PTC_EARG, "ptc_particles_graph_stats: null argument");
    stats[0] = p->graph.instantiated;
    stats[1] = p->graph.replayed;
    ok();
    return PTC_OK;
}

int ptc_particles_autotune(ptc_particles *p, double t_end, int chosen[5])
{
    if (!p) return fail(PTC_EARG, "ptc_particles_autotune: null handle");
    if (!std::isfinite(t_end)) return fail(PTC_EARG, "ptc_particles_autotune: t_end not finite");
    if (p->d.np == 0) { ok(); return PTC_OK; }
    // the batch's current time: the earliest particle time (all start at 0
    // unless the batch was advanced)
    std::vector<double> ht((size_t)p->d.np);
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(ht.data(), p->d.t, sizeof(double) * ht.size(), hipMemcpyDeviceToHost));
    double t0 = ht[0];
    for (double v : ht) t0 = v < t0 ? v : t0;
    struct Cand { int block, sort, order, tail, lanes; };
    const Cand cand[] = {{256, 1, 2, 0, 6},     {256, 1, 2, kTailGroupOnly, 6},
                         {256, 1, 2, 32768, 6}, {256, 1, 2, kTailGroupOnly, 3},
                         {256, 1, 2, 32768, 3}, {128, 1, 2, 0, 6},
                         {64, 1, 2, 0, 6},      {256, 1, 1, 0, 6},
                         {256, 0, 2, 0, 6}};
    const int nc = (int)(sizeof cand / sizeof cand[0]);
    if (!(t_end > t0)) return fail(PTC_EARG, "ptc_particles_autotune: t_end must exceed the "
                                             "batch's time %g", t0);
    const int saved = p->tuned;
    const int sb = p->tuned_block, ss = p->tuned_sort, so = p->tuned_order, st = p->tuned_tail,
              sl = p->tuned_lanes;
    double best = -1.0;
    int pick = 0;
    int64_t tot[5];
    for (int k = 0; k <= nc; k++) {
        // segment 0 warms up (and gives the cost order its history)
        const Cand c = cand[k == 0 ? 0 : k - 1];
        p->tuned = 1;
        p->tuned_block = c.block;
        p->tuned_sort = c.sort;
        p->tuned_order = c.order;
        p->tuned_tail = c.tail;
        p->tuned_lanes = c.lanes;
        int rc = ptc_particles_totals(p, tot);
        const int64_t rhs0 = tot[3];
        if (!rc) rc = ptc_particles_advance(p, t0 + (t_end - t0) * (k + 1) / (nc + 1));
        if (!rc) rc = ptc_particles_totals(p, tot);
        if (rc) {
            p->tuned = saved;
            p->tuned_block = sb; p->tuned_sort = ss; p->tuned_order = so; p->tuned_tail = st;
            p->tuned_lanes = sl;
            return rc;
        }
        const double rate = p->last_ms > 0.0 ? (double)(tot[3] - rhs0) / p->last_ms : 0.0;
        if (k > 0 && rate > best) {
            best = rate;
            pick = k - 1;
        }
    }
    p->tuned = 1;
    p->tuned_block = cand[pick].block;
    p->tuned_sort = cand[pick].sort;
    p->tuned_order = cand[pick].order;
    p->tuned_tail = cand[pick].tail;
    p->tuned_lanes = cand[pick].lanes;
    if (chosen) {
        chosen[0] = p->tuned_block;
        chosen[1] = p->tuned_sort;
        chosen[2] = p->tuned_order;
        chosen[3] = p->tuned_tail;
        chosen[4] = p->tuned_lanes;
    }
    ok();
    return PTC_OK;
}

int ptc_particles_totals(ptc_particles *p, int64_t totals[5])
{
    if (!p || !totals) return fail(PTC_EARG, "ptc_particles_totals: null argument");
    for (int i = 0; i < 5; i++) totals[i] = 0;
    if (p->d.np == 0) { ok(); return PTC_OK; }
    HIP_TRY(launch_totals(p->d, p->d_totals, p->stream));
    HIP_TRY(hipMemcpyAsync(totals, p->d_totals, sizeof(int64_t) * 5, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    ok();
    return PTC_OK;
}

int ptc_particles_cell_loads(ptc_particles *p, int64_t *loads)
{
    if (!p || !loads) return fail(PTC_EARG, "ptc_particles_cell_loads: null argument");
    *loads = 0;
    if (p->d.np == 0) { ok(); return PTC_OK; }
    unsigned long long v = 0;
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(&v, p->d.loads, sizeof v, hipMemcpyDeviceToHost));
    *loads = (int64_t)v;
    ok();
    return PTC_OK;
}

int ptc_particles_last_kernel_ms(const ptc_particles *p, double *ms)
{
    if (!p || !ms) return fail(PTC_EARG, "ptc_particles_last_kernel_ms: null argument");
    *ms = p->last_ms;
    ok();
    return PTC_OK;
}

int ptc_particles_destroy(ptc_particles *p)
{
    if (!p) return PTC_OK;
    (void)hipStreamSynchronize(p->stream);
    (void)hipEventDestroy(p->ev0);
    (void)hipEventDestroy(p->ev1);
    (void)hipStreamDestroy(p->stream);
    hipError_t e1 = hipFree(p->d_time);
    hipError_t e2 = hipFree(p->block);
    if (p->rec_block) (void)hipFree(p->rec_block);
    if (p->order_block) (void)hipFree(p->order_block);
    drop_graph(p->graph);
    delete p;
    if (e1 != hipSuccess) return hip_fail(e1, "ptc_particles_destroy");
    if (e2 != hipSuccess) return hip_fail(e2, "ptc_particles_destroy");
    ok();
    return PTC_OK;
}

int ptc_probe(const ptc_fields *f, int op, int64_t n, const double *params, int nparams,
              const double *in, double *out)
{
    if (!f) return fail(PTC_EARG, "ptc_probe: fields handle is null");
    if (op < PTC_PROBE_GATHER || op > PTC_PROBE_RHS_T)
        return fail(PTC_EARG, "ptc_probe: bad op %d", op);
    if (n < 0) return fail(PTC_EARG, "ptc_probe: negative count");
    if (n == 0) { ok(); return PTC_OK; }
    if (!in || !out) return fail(PTC_EARG, "ptc_probe: null in/out");
    StepParams sp;
    int rc = build_step_params(f, 1, params, nparams, sp);
    if (rc) return rc;
    if (op >= PTC_PROBE_FEATURES && f->d.layout != LAYOUT_AOS)
        return fail(PTC_EARG, "ptc_probe: op %d needs the aos layout", op);
    if (rows_layout(f->d.layout))
        return fail(PTC_EARG, "ptc_probe: the probes read node rows (layout aos, f32_layout "
                              "nodes)");
    const int nin = op == PTC_PROBE_GATHER     ? 3
                    : op == PTC_PROBE_POW      ? 2
                    : op == PTC_PROBE_FEATURES ? 10
                    : op == PTC_PROBE_GATHER_T ? 4
                    : op == PTC_PROBE_RHS_T    ? 7
                                               : 6;
    const int nout = (op == PTC_PROBE_GATHER || op == PTC_PROBE_RHS || op == PTC_PROBE_GATHER_T ||
                      op == PTC_PROBE_RHS_T)
                         ? 6
                     : op == PTC_PROBE_FEATURES ? 16
                                                : 1;
    double *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(double) * (nin + nout) * n));
    hipError_t e = hipMemcpy(d, in, sizeof(double) * nin * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_probe(f->d, sp, op, n, d, d + nin * n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(out, d + nin * n, sizeof(double) * nout * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "ptc_probe");
    ok();
    return PTC_OK;
}

// ---------------------------------------------------------------------------
// step-size controller modes
// ---------------------------------------------------------------------------
int ptc_particles_record(ptc_particles *p, double rate, int64_t capacity)
{
    if (!p) return fail(PTC_EARG, "ptc_particles_record: null handle");
    if (!(rate > 0.0 && rate <= 1.0) && capacity > 0)
        return fail(PTC_EARG, "ptc_particles_record: rate must be in (0, 1] (got %g)", rate);
    if (capacity < 0 || capacity > ((int64_t)1 << 32))
        return fail(PTC_EARG, "ptc_particles_record: bad capacity %lld", (long long)capacity);
    if (capacity > 0 && p->fields->d.nsnap > 0)
        return fail(PTC_ESTATE, "ptc_particles_record: needs a static field (not a series)");
    HIP_TRY(hipStreamSynchronize(p->stream));
    if (p->rec_block) {
        HIP_TRY(hipFree(p->rec_block));
        p->rec_block = nullptr;
    }
    p->d.pend = nullptr;
    p->d.pend_h = nullptr;
    p->d.pend_rej = nullptr;
    p->d.rec = nullptr;
    p->d.rec_count = nullptr;
    p->d.rec_cap = 0;
    if (p->ctrl == CTRL_RECORD) p->ctrl = CTRL_PI;
    if (capacity == 0) { ok(); return PTC_OK; }
    if (p->fields->d.layout != LAYOUT_AOS)
        return fail(PTC_ESTATE, "ptc_particles_record: needs the aos field layout");
    const int64_t n = p->d.np > 0 ? p->d.np : 1;
    const size_t bytes = sizeof(double) * n + sizeof(int64_t) * n + 256 +
                         sizeof(float) * kMlpIn * n + sizeof(float) * kRecord * (size_t)capacity;
    HIP_TRY(hipMalloc(&p->rec_block, bytes));
    char *c = (char *)p->rec_block;
    p->d.pend_h = (double *)c;
    c += sizeof(double) * n;
    p->d.pend_rej = (int64_t *)c;
    c += sizeof(int64_t) * n;
    p->d.rec_count = (unsigned long long *)c;
    c += 256;
    p->d.pend = (float *)c;
    c += sizeof(float) * kMlpIn * n;
    p->d.rec = (float *)c;
    p->d.rec_cap = capacity;
    p->d.rec_thresh = (uint64_t)(rate * 4294967296.0);
    p->ctrl = CTRL_RECORD;
    HIP_TRY(launch_zero_i64((int64_t *)p->d.rec_count, 1, p->stream));
    if (p->d.np > 0) HIP_TRY(launch_clear_flag(p->d.flags, p->d.np, FLAG_PENDING, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    ok();
    return PTC_OK;
}

int ptc_particles_records(const ptc_particles *p, int64_t *count, float *host_out,
                          const float **device_records)
{
    if (!p || !count) return fail(PTC_EARG, "ptc_particles_records: null argument");
    *count = 0;
    if (device_records) *device_records = p->d.rec;
    if (!p->rec_block) { ok(); return PTC_OK; }
    unsigned long long claimed = 0;
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(&claimed, p->d.rec_count, sizeof claimed, hipMemcpyDeviceToHost));
    const int64_t k = claimed < (unsigned long long)p->d.rec_cap ? (int64_t)claimed : p->d.rec_cap;
    *count = k;
    if (host_out && k > 0)
        HIP_TRY(hipMemcpy(host_out, p->d.rec, sizeof(float) * kRecord * (size_t)k,
                          hipMemcpyDeviceToHost));
    ok();
    return PTC_OK;
}

int ptc_particles_use_mlp(ptc_particles *p, const ptc_mlp *net, double shift)
{
    if (!p) return fail(PTC_EARG, "ptc_particles_use_mlp: null handle");
    if (!std::isfinite(shift)) return fail(PTC_EARG, "ptc_particles_use_mlp: shift not finite");
    if (!net) {
        if (p->ctrl == CTRL_LEARNED) p->ctrl = CTRL_PI;
        p->d.net = nullptr;
        ok();
        return PTC_OK;
    }
    const DeviceFields &fd = p->fields->d;
    const bool rows_ok = fd.layout == LAYOUT_AOS || (fd.prec == PTC_FP32 && rows_layout(fd.layout));
    if (!rows_ok || fd.nsnap > 0)
        return fail(PTC_ESTATE, "ptc_particles_use_mlp: needs a static field in the aos layout "
                                "(or FP32 pair, quad or column rows)");
    if (!net->trained)
        return fail(PTC_ESTATE, "ptc_particles_use_mlp: the network still has its random "
                                "initialisation (train it or set its parameters)");
    p->d.net = net->packed;
    p->d.ctrl_shift = (float)shift;
    p->ctrl = CTRL_LEARNED;
    ok();
    return PTC_OK;
}

// ---------------------------------------------------------------------------
// step-size network
// ---------------------------------------------------------------------------
namespace {

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// He-uniform hidden layers, small output layer, zero biases, identity
// normalisation
void init_params(uint64_t seed, std::vector<float> &w)
{
    w.assign(kMlpParams, 0.0f);
    uint64_t s = seed;
    auto uni = [&](float a) {
        const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * a);
    };
    for (int i = 0; i < kMlpIn * kMlpHidden; i++) w[kMlpW1 + i] = uni(std::sqrt(6.0f / kMlpIn));
    for (int i = 0; i < kMlpHidden * kMlpHidden; i++)
        w[kMlpW2 + i] = uni(std::sqrt(6.0f / kMlpHidden));
    for (int i = 0; i < kMlpHidden; i++) w[kMlpW3 + i] = uni(0.1f * std::sqrt(6.0f / kMlpHidden));
    for (int i = 0; i < kMlpIn; i++) w[kMlpSig + i] = 1.0f;
}

int check_params(const float *params)
{
    for (int i = 0; i < kMlpParams; i++)
        if (!std::isfinite(params[i])) return fail(PTC_EARG, "network parameter %d not finite", i);
    for (int i = 0; i < kMlpIn; i++)
        if (!(params[kMlpSig + i] > 0.0f))
            return fail(PTC_EARG, "network input scale sig[%d] must be > 0", i);
    return PTC_OK;
}

// device copy of host records (or the device pointer itself)
struct RecordBuf {
    const float *d = nullptr;
    float *owned = nullptr;
    ~RecordBuf() { if (owned) (void)hipFree(owned); }
    hipError_t set(int64_t n, const float *records, int on_device)
    {
        if (on_device) { d = records; return hipSuccess; }
        hipError_t e = hipMalloc(&owned, sizeof(float) * kRecord * (size_t)n);
        if (e != hipSuccess) return e;
        d = owned;
        return hipMemcpy(owned, records, sizeof(float) * kRecord * (size_t)n, hipMemcpyHostToDevice);
    }
};

}  // namespace

int ptc_mlp_create(const float *params, uint64_t seed, ptc_mlp **out)
{
    if (!out) return fail(PTC_EARG, "ptc_mlp_create: out is null");
    *out = nullptr;
    std::vector<float> w;
    if (params) {
        int rc = check_params(params);
        if (rc) return rc;
        w.assign(params, params + kMlpParams);
    } else {
        init_params(seed, w);
    }
    ptc_mlp *net = new ptc_mlp();
    net->trained = params != nullptr;
    const size_t fl = (size_t)kMlpParams + 2 * (size_t)kMlpTrainable +
                      (size_t)kMlpMaxGradBlocks * (kMlpTrainable + 1) + (kMlpTrainable + 1) + 64;
    void *blk = nullptr;
    hipError_t e = hipMalloc(&blk, sizeof(float) * fl);
    if (e != hipSuccess) { delete net; return hip_fail(e, "ptc_mlp_create"); }
    float *c = (float *)blk;
    net->params = c;
    net->m1 = c + kMlpParams;
    net->m2 = net->m1 + kMlpTrainable;
    net->partials = net->m2 + kMlpTrainable;
    net->grad = net->partials + (size_t)kMlpMaxGradBlocks * (kMlpTrainable + 1);
    e = hipMalloc(&net->packed, sizeof(MlpPacked));
    if (e == hipSuccess) e = hipMalloc(&net->stats, sizeof(double) * 2 * kMlpIn * 1024);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpy(net->params, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(net->m1, 0, sizeof(float) * 2 * kMlpTrainable);
    if (e == hipSuccess) e = hipMemset(net->packed, 0, sizeof(MlpPacked));
    if (e == hipSuccess) e = mlp_pack(net->params, net->packed, net->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
    if (e != hipSuccess) {
        (void)hipFree(blk);
        if (net->packed) (void)hipFree(net->packed);
        if (net->stats) (void)hipFree(net->stats);
        delete net;
        return hip_fail(e, "ptc_mlp_create");
    }
    *out = net;
    ok();
    return PTC_OK;
}

int ptc_mlp_destroy(ptc_mlp *net)
{
    if (!net) return PTC_OK;
    (void)hipStreamSynchronize(net->stream);
    (void)hipStreamDestroy(net->stream);
    hipError_t e = hipFree(net->params);
    (void)hipFree(net->packed);
    (void)hipFree(net->stats);
    delete net;
    if (e != hipSuccess) return hip_fail(e, "ptc_mlp_destroy");
    ok();
    return PTC_OK;
}

int ptc_mlp_get_params(const ptc_mlp *net, float *params)
{
    if (!net || !params) return fail(PTC_EARG, "ptc_mlp_get_params: null argument");
    HIP_TRY(hipStreamSynchronize(net->stream));
    HIP_TRY(hipMemcpy(params, net->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    ok();
    return PTC_OK;
}

int ptc_mlp_set_params(ptc_mlp *net, const float *params)
{
    if (!net || !params) return fail(PTC_EARG, "ptc_mlp_set_params: null argument");
    int rc = check_params(params);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(net->stream));
    HIP_TRY(hipMemcpy(net->params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(net->m1, 0, sizeof(float) * 2 * kMlpTrainable));
    net->t = 0;
    net->trained = true;
    HIP_TRY(mlp_pack(net->params, net->packed, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    ok();
    return PTC_OK;
}

static int mlp_forward_host(const ptc_mlp *net, int64_t n, const float *features, float *out,
                            int groups, const char *who)
{
    if (!net) return fail(PTC_EARG, "%s: null handle", who);
    if (n < 0) return fail(PTC_EARG, "%s: negative count", who);
    if (n == 0) { ok(); return PTC_OK; }
    if (!features || !out) return fail(PTC_EARG, "%s: null array", who);
    float *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(float) * (kMlpIn + 1) * (size_t)n));
    hipError_t e = hipMemcpy(d, features, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = mlp_forward(net->packed, n, d, kMlpIn, d + kMlpIn * n, net->stream, groups);
    if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d + kMlpIn * n, sizeof(float) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, who);
    ok();
    return PTC_OK;
}

int ptc_mlp_forward(const ptc_mlp *net, int64_t n, const float *features, float *out)
{
    return mlp_forward_host(net, n, features, out, 0, "ptc_mlp_forward");
}

int ptc_mlp_forward_groups(const ptc_mlp *net, int64_t n, const float *features, float *out)
{
    return mlp_forward_host(net, n, features, out, 1, "ptc_mlp_forward_groups");
}

int ptc_mlp_gradient(const ptc_mlp *net, int64_t n, const float *records, float alpha,
                     float *grad, double *loss)
{
    if (!net || !records || !grad) return fail(PTC_EARG, "ptc_mlp_gradient: null argument");
    if (n <= 0) return fail(PTC_EARG, "ptc_mlp_gradient: need at least one record");
    RecordBuf rb;
    HIP_TRY(rb.set(n, records, 0));
    HIP_TRY(mlp_gradient(net->packed, rb.d, n, n, 0, 0, alpha, -4.0f, 4.0f, net->partials,
                         net->stream));
    HIP_TRY(mlp_reduce(net->partials, mlp_grad_blocks(n), net->grad, 0, 0.0f, 1.0f, 1.0f, nullptr,
                       nullptr, nullptr, nullptr, nullptr, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    std::vector<float> g(kMlpTrainable + 1);
    HIP_TRY(hipMemcpy(g.data(), net->grad, sizeof(float) * (kMlpTrainable + 1), hipMemcpyDeviceToHost));
    for (int i = 0; i < kMlpTrainable; i++) grad[i] = g[i];
    if (loss) *loss = (double)g[kMlpTrainable] / (double)n;
    ok();
    return PTC_OK;
}

int ptc_mlp_normalize(ptc_mlp *net, int64_t n, const float *records, int on_device)
{
    if (!net || !records) return fail(PTC_EARG, "ptc_mlp_normalize: null argument");
    if (n <= 1) return fail(PTC_EARG, "ptc_mlp_normalize: need at least two records");
    RecordBuf rb;
    HIP_TRY(rb.set(n, records, on_device));
    int64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    HIP_TRY(mlp_stats(rb.d, n, net->stats, (int)blocks, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    std::vector<double> h(2 * kMlpIn * (size_t)blocks);
    HIP_TRY(hipMemcpy(h.data(), net->stats, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    std::vector<float> w(kMlpParams);
    HIP_TRY(hipMemcpy(w.data(), net->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    for (int i = 0; i < kMlpIn; i++) {
        double s = 0.0, q = 0.0;
        for (int64_t b = 0; b < blocks; b++) {
            s += h[b * 2 * kMlpIn + i];
            q += h[b * 2 * kMlpIn + kMlpIn + i];
        }
        const double mu = s / (double)n;
        const double var = q / (double)n - mu * mu;
        const double sd = var > 1e-6 ? std::sqrt(var) : 1e-3;
        w[kMlpMu + i] = (float)mu;
        w[kMlpSig + i] = (float)sd;
    }
    HIP_TRY(hipMemcpy(net->params, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    HIP_TRY(mlp_pack(net->params, net->packed, net->stream));
    HIP_TRY(hipStreamSynchronize(net->stream));
    ok();
    return PTC_OK;
}

int ptc_mlp_train(ptc_mlp *net, int64_t n, const float *records, int on_device, int steps,
                  int64_t batch, float lr, float alpha, uint64_t seed, float *losses)
{
    if (!net || !records) return fail(PTC_EARG, "ptc_mlp_train: null argument");
    if (n <= 0 || steps < 0 || batch <= 0 || !(lr > 0.0f) || !std::isfinite(alpha))
        return fail(PTC_EARG, "ptc_mlp_train: bad n/steps/batch/lr/alpha");
    if (seed == 0) seed = 1;
    RecordBuf rb;
    HIP_TRY(rb.set(n, records, on_device));
    float *d_loss = nullptr;
    if (steps > 0) HIP_TRY(hipMalloc(&d_loss, sizeof(float) * steps));
    const int blocks = mlp_grad_blocks(batch);
    hipError_t e = hipSuccess;
    for (int s = 0; s < steps && e == hipSuccess; s++) {
        net->t++;
        const float bc1 = 1.0f - (float)std::pow(0.9, (double)net->t);
        const float bc2 = 1.0f - (float)std::pow(0.999, (double)net->t);
        e = mlp_gradient(net->packed, rb.d, n, batch, seed, (uint64_t)net->t, alpha, -4.0f, 4.0f,
                         net->partials, net->stream);
        if (e == hipSuccess)
            e = mlp_reduce(net->partials, blocks, nullptr, 1, lr, bc1, bc2, net->params, net->m1,
                           net->m2, net->packed, d_loss + s, net->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
    if (e == hipSuccess && losses && steps > 0) {
        e = hipMemcpy(losses, d_loss, sizeof(float) * steps, hipMemcpyDeviceToHost);
        for (int s = 0; s < steps; s++) losses[s] /= (float)batch;
    }
    if (d_loss) (void)hipFree(d_loss);
    if (e != hipSuccess) return hip_fail(e, "ptc_mlp_train");
    if (steps > 0) net->trained = true;
    ok();
    return PTC_OK;
}

int ptc_mlp_probe_gemm(const float *a, const float *bt, float *c)
{
    if (!a || !bt || !c) return fail(PTC_EARG, "ptc_mlp_probe_gemm: null array");
    float *d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(float) * 3 * 4096));
    hipError_t e = hipMemcpy(d, a, sizeof(float) * 4096, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 4096, bt, sizeof(float) * 4096, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = mlp_gemm_probe(d, d + 4096, d + 8192, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(c, d + 8192, sizeof(float) * 4096, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "ptc_mlp_probe_gemm");
    ok();
    return PTC_OK;
}

int ptc_integrate(const ptc_fields *f, int method, int64_t np, int m, const double *params,
                  int nparams, const double *time, double *state, int64_t *counter, int32_t *status)
{
    ptc_particles *p = nullptr;
    int rc = ptc_particles_create(f, method, np, m, params, nparams, time, state, &p);
    if (rc) return rc;
    // advance snapshot interval by snapshot interval (at most 64 segments):
    // the same result as one launch, and from the second segment on the
    // work queue is ordered by each particle's cost in the previous one
    const int nseg = m - 1 < 64 ? m - 1 : 64;
    for (int k = 1; k < nseg && !rc; k++) {
        const int n = 1 + (int)((int64_t)k * (m - 1) / nseg);
        if (time[n - 1] < params[2]) rc = ptc_particles_advance(p, time[n - 1]);
    }
    if (!rc) rc = ptc_particles_advance(p, INFINITY);
    if (!rc) rc = ptc_particles_download(p, state, counter, status);
    std::string keep = g_err;
    int rc2 = ptc_particles_destroy(p);
    if (rc) { g_err = keep; return rc; }
    return rc2;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// allocation-free integration (ptc_integrate_noalloc)
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kNoAllocMaxParticles = 65536;
// per particle: x[6], t, dt, errold (double); cnt[3], nrhs (int64);
// nout, flags, status (int32)
constexpr size_t kNoAllocPerParticle =
    9 * sizeof(double) + 4 * sizeof(int64_t) + 3 * sizeof(int32_t);
// header: the queue heads (3 int32) at byte 0, the cell-load counter at 64
constexpr size_t kNoAllocHeader = 256;
constexpr size_t kNoAllocBytes =
    kNoAllocHeader + (size_t)kNoAllocMaxParticles * kNoAllocPerParticle;

}  // namespace

// The batch state of ptc_integrate_noalloc (7.3 MiB): each device holds its
// own copy until the process ends, and every call on a device reuses it
// (the call's first kernel re-initialises what it reads).  Host code takes
// its device address from g_scratch_ptr, never from this name.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[256];  // device address of `scratch`, per device id

// The device address of the current device's `scratch`, looked up once per
// device (with that device current).
int scratch_address(char **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ptc_integrate_noalloc: hipGetDevice");
    if (dev < 0 || dev >= 256) return fail(PTC_EARG, "ptc_integrate_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "ptc_integrate_noalloc: hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = ptr;
    }
    *out = (char *)g_scratch_ptr[dev];
    return PTC_OK;
}

// Initial batch state from the caller's state(8, m, np) (rows 1:6 of
// snapshot 1), as ptc_particles_create sets it up; zeroes the queue heads
// and the cell-load counter.
__global__ void noalloc_start_kernel(const DeviceParticles S, int m)
{
    if (blockIdx.x == 0 && threadIdx.x < 3) S.qcount[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *S.loads = 0ull;
    const int64_t N = S.np;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double *s = S.out + (size_t)8 * m * i;
#pragma unroll
        for (int c = 0; c < 6; c++) S.x[c][i] = s[c];
        S.t[i] = 0.0;
        S.dt[i] = 0.0;
        S.errold[i] = 1.0e-04;  // integrations.F90:349
        S.nout[i] = 2;
        S.flags[i] = 0;
        S.status[i] = PTC_STATUS_RUNNING;
        S.cnt[i] = 0;
        S.cnt[N + i] = 0;
        S.cnt[2 * N + i] = 0;
        S.nrhs[i] = 0;
    }
}

// The step counters into the caller's counter(3, np), the status into
// status(np) (if given).
__global__ void noalloc_finish_kernel(const DeviceParticles S, int64_t *counter, int32_t *status)
{
    const int64_t N = S.np;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x) {
        counter[3 * i] = S.cnt[i];
        counter[3 * i + 1] = S.cnt[N + i];
        counter[3 * i + 2] = S.cnt[2 * N + i];
        if (status) status[i] = S.status[i];
    }
}

}  // namespace

extern "C" {

int ptc_integrate_noalloc(const ptc_fields *f, int method, int64_t np, int m,
                          const double *params, int nparams, const double *time, double *state,
                          int64_t *counter, int32_t *status, void *stream)
{
    if (!f) return fail(PTC_EARG, "ptc_integrate_noalloc: fields handle is null");
    if (method != PTC_RK5 && method != PTC_DOP853) return fail(PTC_EARG, "bad method %d", method);
    if (np < 0 || np > kNoAllocMaxParticles)
        return fail(PTC_EARG, "ptc_integrate_noalloc: %lld particles; its fixed workspace holds "
                              "0 to %lld (ptc_integrate takes any number)",
                    (long long)np, (long long)kNoAllocMaxParticles);
    if (m < 1) return fail(PTC_EARG, "need at least one snapshot (m = %d)", m);
    if (np > 0 && (!time || !state || !counter))
        return fail(PTC_EARG, "ptc_integrate_noalloc: null time, state or counter");
    StepParams sp;
    int rc = build_step_params(f, m, params, nparams, sp);
    if (rc) return rc;
    if (np == 0) { ok(); return PTC_OK; }
    char *base = nullptr;
    rc = scratch_address(&base);
    if (rc) return rc;

    const int64_t n = kNoAllocMaxParticles;  // plane capacity; planes are np apart
    DeviceParticles d{};
    d.np = np;
    double *dx = (double *)(base + kNoAllocHeader);
    for (int k = 0; k < 6; k++) d.x[k] = dx + k * np;
    d.t = dx + 6 * np;
    d.dt = dx + 7 * np;
    d.errold = dx + 8 * np;
    d.cnt = (int64_t *)(dx + 9 * n);
    d.nrhs = d.cnt + 3 * np;
    d.nout = (int32_t *)(d.cnt + 4 * n);
    d.flags = d.nout + np;
    d.status = d.nout + 2 * np;
    d.qcount = (int32_t *)base;
    d.loads = (unsigned long long *)(base + 64);
    d.out = state;
    d.time = time;
    d.order = nullptr;

    Options o = options();
    sp.max_steps = o.max_steps;
    const int cache = f->d.layout != LAYOUT_SOA && f->d.layout != LAYOUT_BRICK ? o.cell_cache : 0;
    LaunchConfig lc{o.block, o.grid_blocks, cache, CTRL_PI,
                    choose_tail(o.tail, f->d, method, np, CTRL_PI),
                    choose_lanes(o.lanes, f->d, method, np)};
    hipStream_t s = (hipStream_t)stream;
    const int blocks = (int)((np + 255) / 256);
    hipLaunchKernelGGL(noalloc_start_kernel, dim3(blocks), dim3(256), 0, s, d, m);
    HIP_TRY(hipGetLastError());
    HIP_TRY(launch_integrate(method, f->d, sp, d, INFINITY, lc, s));
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3(blocks), dim3(256), 0, s, d, counter, status);
    HIP_TRY(hipGetLastError());
    ok();
    return PTC_OK;
}

}  // extern "C"
