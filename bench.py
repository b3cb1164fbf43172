"""Benchmark: accepted particle-steps/s of the device integrator on one MI355X.

Workload (default): the DOP853 turbulent-box configuration -- 256^3 periodic
grid of side 1, B = z + dB and V built from 256 random solenoidal Fourier
modes (|k| in [2 pi, 16 pi], amplitude ~ |k|^(-5/3), rms 0.5 and 0.1), force
q/m (v - V) x B, 100,000 protons (q/m = 1, c = 1) with gamma log-uniform in
[1, 1e3] and isotropic directions, atol = rtol = 1e-8, snapshots every
100 gyro-periods of B0 up to 1000.  One benchmark step advances every
particle by one gyro-period of B0 (T0 = 2 pi); the integration continues
across steps exactly as one long run would.  Inputs are synthetic (seeded)
and resident in HBM before the timed region.

After the timed plain run, the same workload runs (untimed for `value`) in
mixed precision (FP32 fields, weights and RHS; FP64 state, sums and error
norm) and with the learned step-size proposal in mixed precision (network
trained on the GPU from records of a PI-controlled run of a separate particle
set); the `variants` object reports their rates, rejected-step fractions and
RHS counts, their physical-time throughput relative to the plain run, and
their deviation from the plain FP64 result.

Prints one JSON line.  `roofline` is the integrator kernel's algorithmic
gather traffic (8 corners x 6 components x 8 B = 384 B per RHS evaluation,
times the RHS evaluations the kernel counted) over its duration measured with
HIP events on the batch's stream, against the MI355X HBM peak.
With `--snapshots N` the field is time-dependent: N device-built snapshots
of the same turbulent recipe, its modes travelling in time, spread over the
benchmarked time and blended linearly in time at every stage; each RHS then
reads two snapshots (768 B per evaluation in FP64), the second kept in LDS
by the cell cache.  The reference has no
time-dependent field, so no CPU baseline or variants are run for it.

`cpu_baseline` times the reference's own FP64 OpenMP integrator (compiled
from its sources into oracle/_ref) -- or the C oracle where that library is
absent -- on the host cores, on the first particles of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_RHS_FP64 = 8 * 6 * 8


def build_workload(args):
    """(make_fields(prec) -> Fields, host fields (uu, bb) or None, params, time
    grid, initial state, method, config)."""
    from particles_amd import workloads as W
    gr = None
    device = False
    if args.workload == "turbulent256":
        dm, npart, method, tol, gr = 256, 100_000, "dop853", 1e-8, (1.0, 1.0e3)
        tmax, nsnap, recipe = 1000 * W.T0, 10, "turbulent"
    elif args.workload == "turbulent64":
        dm, npart, method, tol = 64, 1_000, "rk5", 1e-6
        tmax, nsnap, recipe = 100 * W.T0, 10, "turbulent"
    elif args.workload == "reconnection512":
        dm, npart, method, tol, gr = 512, 1_000_000, "rk5", 1e-8, (1.0, 1.0e4)
        tmax, nsnap, recipe, device = 10 * W.T0, 10, "reconnection", True
    elif args.workload == "reconnection1024":
        dm, npart, method, tol, gr = 1024, 4_000_000, "dop853", 1e-8, (1.0, 1.0e4)
        tmax, nsnap, recipe, device = 10 * W.T0, 10, "reconnection", True
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    if args.particles:
        npart = args.particles
    if args.grid:
        dm = args.grid
    if args.method:
        method = args.method
    uu = bb = None
    if args.snapshots > 0:
        if recipe != "turbulent":
            raise SystemExit("--snapshots applies to the turbulent workloads")
        # a snapshot series every `snapshot_every` T0 from t = 0, every mode
        # travelling at 0.01 box lengths per unit time, built on the device
        times = np.arange(args.snapshots) * args.snapshot_every * W.T0

        def make_fields(prec):
            return W.turbulent_vb_series_device(dm, seed=args.seed, times=times, prec=prec,
                                                speed=0.01)
        desc = (f"{dm}^3 VB turbulent field (B0=1, dB rms 0.5, V rms 0.1, 256 modes) as "
                f"{args.snapshots} snapshots {args.snapshot_every:g} T0 apart (modes "
                "travelling at 0.01 per unit time), blended linearly in time at each stage, "
                "built on the device")
    elif recipe == "turbulent" and not device:
        uu, bb = W.turbulent_vb(dm, seed=args.seed)

        def make_fields(prec):
            import particles_amd as P
            return P.Fields(uu, bb, prec=prec)
        desc = (f"{dm}^3 VB turbulent field (B0=1, dB rms 0.5, V rms 0.1, 256 modes)")
    else:
        def make_fields(prec):
            return W.reconnection_eb_device(dm, seed=args.seed, prec=prec)
        desc = (f"{dm}^3 EB reconnection field built on the device (B_x = sheet(y), width 0.02, "
                f"guide 0.1, 256-mode noise rms 0.05; E = -V x B + 1e-3 curl B, V_y = -0.1 sheet(y))")
    p = W.params_vector(qom=1.0, vun=1.0, tmax=tmax, dtmax=W.T0, atol=tol, rtol=tol)
    tg = W.time_grid(nsnap, tmax)
    if gr is None:
        st0 = W.initial_state(npart, len(tg), seed=args.seed + 1, gamma=10.0)
    else:
        st0 = W.initial_state(npart, len(tg), seed=args.seed + 1, gamma_range=gr)
    cfg = {"workload": f"{args.workload}: {method} tol={tol:g}, {desc}, {npart} protons, "
                       f"gamma {'10' if gr is None else 'log-uniform [%g, %g]' % gr}, "
                       f"step = advance all particles by 1 gyro-period of B0",
           "grid": dm, "particles": npart, "method": method, "tol": tol,
           "snapshots_every_T0": tmax / W.T0 / nsnap, "layout": args.layout,
           "cell_cache": args.cell_cache, "field_precision": args.prec, "sort": args.sort,
           "order": args.order, "field_snapshots": args.snapshots, "tail": args.tail,
           "f32_layout": args.f32_layout}
    if args.snapshots > 0:
        cfg["snapshot_every_T0"] = args.snapshot_every
    return make_fields, uu, bb, p, tg, st0, method, cfg


def cpu_baseline(uu, bb, p, tg, st0, method, target_s):
    """Reference FP64 OpenMP integrator on the host, first particles of the
    workload, horizon scaled so the timed run takes about target_s."""
    from particles_amd import workloads as W
    try:
        from oracle.oracle import REF_SO, Oracle, Reference
    except ImportError:
        return None
    if os.path.exists(REF_SO):
        impl, kind = Reference(), "reference"
        threads = impl.num_threads()
        run = lambda q, s: impl.integrate(method, uu, bb, q, np.array([0.0, 1e300]), s)
    else:
        impl, kind = Oracle(), "port"
        threads = impl.max_threads()
        run = lambda q, s: impl.integrate(method, uu, bb, q, np.array([0.0, 1e300]), s)[:2]
    n = min(st0.shape[2], 64 * threads)
    s0 = np.asfortranarray(st0[:, :1, :n].repeat(2, axis=1))
    s0[:, 1, :] = 0.0
    horizon = 0.05 * W.T0
    for _ in range(6):
        q = p.copy(); q[2] = horizon
        t0 = time.perf_counter()
        _, cnt = run(q, s0)
        el = time.perf_counter() - t0
        if el > 0.3 * target_s:
            break
        horizon *= min(8.0, max(1.5, target_s / max(el, 1e-3)))
    acc = int(cnt[1].sum())
    return {"value": acc / el, "unit": "accepted particle-steps/s", "cores": int(threads),
            "kind": kind,
            "sample": f"first {n} particles of the workload, t = 0 .. {horizon / W.T0:.3g} T0 "
                      f"({acc} accepted steps in {el:.2f} s, {threads} OpenMP threads)"}


def traffic_from_profile(cfg):
    """HBM bytes per integrator launch of this workload from a committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (profiles/*.json, scaled by a
    copy-kernel calibration), or None when no summary matches the
    configuration."""
    import glob
    keys = ("grid", "particles", "method", "tol", "layout", "cell_cache", "field_precision",
            "sort", "order", "field_snapshots", "snapshot_every_T0", "tail", "f32_layout")
    defaults = {"field_snapshots": 0, "tail": "0", "f32_layout": "nodes"}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.json"))):
        try:
            with open(path) as fh:
                prof = json.load(fh)
        except (OSError, ValueError):
            continue
        pc = prof.get("config", {})
        if all(_norm(pc, k, defaults) == _norm(cfg, k, defaults) for k in keys):
            return prof["hbm_bytes_per_launch"]
    return None


def _norm(cfg, key, defaults):
    """A configuration value for matching; the tail setting only selects
    kernels for node rows ("aos"), elsewhere it is moot."""
    if key == "tail" and cfg.get("layout") != "aos":
        return None
    if key == "f32_layout" and cfg.get("field_precision") != "fp32":
        return None  # FP64 fields ignore it
    return cfg.get(key, defaults.get(key))


def run_variant(fields, method, p, tg, st, steps, warmup, net=None, shift=0.0):
    """Advances a fresh batch warmup + steps gyro-periods; returns the timed
    part's counters and wall time, and the snapshots."""
    from particles_amd import workloads as W
    import particles_amd as P
    with P.Particles(fields, method, p, tg, st) as b:
        if net is not None:
            b.use_mlp(net, shift)
        t = 0.0
        for _ in range(warmup):
            t += W.T0
            b.advance(t)
        tot0 = b.totals()
        kms = 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            t += W.T0
            kms += b.advance(t)
        wall = time.perf_counter() - t0
        tot1 = b.totals()
        state = b.download()[0]
    acc = tot1["accepted"] - tot0["accepted"]
    rej = tot1["rejected"] - tot0["rejected"]
    rhs = tot1["rhs"] - tot0["rhs"]
    return dict(acc=acc, rej=rej, rhs=rhs, wall=wall, kernel_ms=kms, t_end=t), state


def variants(args, fields64, make_fields, p, tg, st0, method, plain):
    """Mixed precision and learned-h (mixed) on the same workload, then their
    accuracy next to the plain run's on the first particles over one T0."""
    import particles_amd as P
    from particles_amd import workloads as W

    out = {}
    time_per_s_plain = st0.shape[2] * args.steps / plain["wall"]
    f32 = make_fields("fp32")
    r, _ = run_variant(f32, method, p, tg, st0, args.steps, args.warmup)
    out["mixed"] = summary(r, st0.shape[2], args.steps, time_per_s_plain)
    print(f"mixed: {out['mixed']['value']:.3e} accepted/s", file=sys.stderr, flush=True)
    # network trained on records of a separate particle set under the PI
    # controller, its log-step offset calibrated on that set against PI
    from particles_amd.learned import train_controller
    st_train = W.initial_state(args.train_particles, 2, seed=args.seed + 101,
                               gamma_range=tuple(args.train_gamma))
    ctl = train_controller(f32, method, p, tg, st_train, fields_fp64=fields64, horizon=W.T0,
                           steps=args.train_steps, seed=args.seed)
    if args.shift is not None:
        ctl.shift = args.shift
    net = ctl.net
    r, _ = run_variant(f32, method, p, tg, st0, args.steps, args.warmup, net, ctl.shift)
    out["learned_h"] = summary(r, st0.shape[2], args.steps, time_per_s_plain)
    out["learned_h"]["controller"] = dict(
        network="16-64-64-1 GELU, BF16 MFMA, FP32 accumulation", shift=ctl.shift,
        record_run=f"{args.train_particles} particles (separate seed), 1 T0, PI controller",
        **ctl.info)
    print(f"learned-h: {out['learned_h']['value']:.3e} accepted/s (shift {ctl.shift})",
          file=sys.stderr, flush=True)

    # accuracy: one T0 from the initial state, snapshot at the end, against
    # the FP64 integrator at atol = rtol = 1e-12
    n = min(args.accuracy_particles, st0.shape[2])
    q = p.copy()
    q[2] = W.T0
    tga = np.array([0.0, W.T0])
    sa = np.asfortranarray(st0[:, :1, :n].repeat(2, axis=1))
    tight = q.copy()
    tight[5] = tight[6] = 1e-12
    ref = P.integrate(fields64, method, tight, tga, sa)[0]

    def err(state):
        d = np.abs(state[:6, 1, :] - ref[:6, 1, :]).max(0)
        return {"median": float(np.median(d)), "p90": float(np.percentile(d, 90))}

    s_plain = P.integrate(fields64, method, q, tga, sa)[0]
    s_mixed = P.integrate(f32, method, q, tga, sa)[0]
    with P.Particles(f32, method, q, tga, sa) as b:
        b.use_mlp(net, ctl.shift)
        b.advance()
        s_learned = b.download()[0]

    def dev_plain(state):
        # the same-seed FP64 run at the same tolerance: position/momentum and
        # relative energy (gamma) deviation
        d = np.abs(state[:6, 1, :] - s_plain[:6, 1, :]).max(0)
        dg = np.abs(state[6, 1, :] - s_plain[6, 1, :]) / s_plain[6, 1, :]
        return {"median": float(np.median(d)), "p90": float(np.percentile(d, 90)),
                "gamma_rel_median": float(np.median(dg)), "gamma_rel_max": float(dg.max())}

    out["plain_error_vs_fp64_tol1e-12"] = err(s_plain)
    out["mixed"]["error_vs_fp64_tol1e-12"] = err(s_mixed)
    out["learned_h"]["error_vs_fp64_tol1e-12"] = err(s_learned)
    out["mixed"]["deviation_vs_plain_fp64"] = dev_plain(s_mixed)
    out["learned_h"]["deviation_vs_plain_fp64"] = dev_plain(s_learned)
    out["accuracy_sample"] = (f"first {n} particles, t = 0 .. 1 T0, max over x, u of "
                              "|state - state(FP64, atol = rtol = 1e-12)|")
    net.close()
    f32.close()
    return out


def summary(r, npart, steps, time_per_s_plain):
    """One variant (FP32 node rows: 192 algorithmic bytes per RHS; the
    learned controller's features come from the cell already gathered)."""
    tps = npart * steps / r["wall"]
    bpr = BYTES_PER_RHS_FP64 // 2
    achieved = bpr * r["rhs"] / (r["kernel_ms"] * 1e-3) / 1e9
    return {"value": r["acc"] / r["wall"], "unit": "accepted particle-steps/s",
            "particle_T0_per_s": tps, "time_throughput_vs_plain": tps / time_per_s_plain,
            "accepted": r["acc"], "rejected": r["rej"],
            "rejected_per_accepted": r["rej"] / max(r["acc"], 1),
            "rhs_per_accepted": r["rhs"] / max(r["acc"], 1), "kernel_ms": r["kernel_ms"],
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "bytes_per_rhs": bpr}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="turbulent256")
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--method", default="", choices=["", "rk5", "dop853"])
    ap.add_argument("--layout", default="aos", choices=["aos", "soa", "brick"])
    ap.add_argument("--f32-layout", default="nodes", choices=["nodes", "pairs"],
                    help="FP32 fields (--prec fp32 and the mixed / learned-h variants): "
                         "32-byte node rows or 64-byte rows of a node and its +x neighbour "
                         "(read by the group kernel only)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cell-cache", type=int, default=1, choices=[0, 1])
    ap.add_argument("--tail", default="auto",
                    help="FP64 static runs: hand a wave's last particles (at most this many) to "
                         "the group kernel once the queue is empty; 'group': that kernel only")
    ap.add_argument("--grid-blocks", type=int, default=0)
    ap.add_argument("--sort", type=int, default=1,
                    help="re-order the work queue every N steps (0 = never)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune block size and queue order during the warm-up steps")
    ap.add_argument("--order", default="cost", choices=["spatial", "cost"],
                    help="cost: most work in the previous step first; spatial: Morton order")
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32"],
                    help="fp32 = mixed precision (FP32 fields, weights and RHS; FP64 state)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the mixed-precision and learned-h runs")
    ap.add_argument("--train-particles", type=int, default=20000)
    ap.add_argument("--accuracy-particles", type=int, default=20000)
    ap.add_argument("--train-gamma", type=float, nargs=2, default=(1.0, 1.0e3))
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--snapshots", type=int, default=0,
                    help="time-dependent field: this many snapshots from t = 0 (0 = static)")
    ap.add_argument("--snapshot-every", type=float, default=2.0,
                    help="time-dependent field: snapshot spacing in gyro-periods of B0")
    ap.add_argument("--shift", type=float, default=None,
                    help="added to the network's log step factor (default: calibrated)")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="only time device-to-device copies and full reductions of a 4 GiB buffer "
                         "(bytes per second: practical HBM figures next to the nominal 8 TB/s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 1 GiB float64 device tensor three times (the known-size "
                         "copy that scales the rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_bandwidth:
        import torch
        a = torch.ones(1 << 29, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3
        gbps = 20 * 2 * a.numel() * 8 / sec / 1e9
        # read-only: a full reduction reads every byte once
        r = a.sum()
        e0.record()
        for _ in range(20):
            r = a.sum()
        e1.record()
        torch.cuda.synchronize()
        rsec = e0.elapsed_time(e1) / 1e3
        print(json.dumps({"bytes": a.numel() * 8, "passes": 20,
                          "copy_read_plus_write_GBps": gbps,
                          "sum_read_GBps": 20 * a.numel() * 8 / rsec / 1e9,
                          "check": float(r)}), flush=True)
        return
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copied 3 x 1 GiB", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the particle batch would "
              "shard by contiguous particle ranges with the fields replicated, and no such run "
              "is made here.", file=sys.stderr)
        sys.exit(2)

    import particles_amd as P
    from particles_amd import workloads as W

    t_setup = time.perf_counter()
    make_fields, uu, bb, p, tg, st0, method, cfg = build_workload(args)
    P.set_option("layout", args.layout)
    P.set_option("f32_layout", args.f32_layout)
    P.set_option("cell_cache", args.cell_cache)
    P.set_option("grid_blocks", args.grid_blocks)
    P.set_option("sort", args.sort)
    P.set_option("order", args.order)
    P.set_option("tail", args.tail)
    fields = make_fields(args.prec)
    batch = P.Particles(fields, method, p, tg, st0)
    print(f"setup {time.perf_counter() - t_setup:.1f}s: {cfg['workload']}", file=sys.stderr,
          flush=True)

    dt_step = W.T0
    t = 0.0
    tuned = None
    for i in range(args.warmup):
        t += dt_step
        if args.autotune and i == args.warmup - 1:
            tuned = batch.autotune(t)
            print(f"warmup {i + 1}/{args.warmup}: autotuned {tuned}", file=sys.stderr, flush=True)
            continue
        ms = batch.advance(t)
        print(f"warmup {i + 1}/{args.warmup}: kernel {ms:.1f} ms", file=sys.stderr, flush=True)
    tot0 = batch.totals()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        t += dt_step
        ms = batch.advance(t)
        kernel_ms += ms
        print(f"step {i + 1}/{args.steps}: kernel {ms:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    tot1 = batch.totals()
    acc = tot1["accepted"] - tot0["accepted"]
    rej = tot1["rejected"] - tot0["rejected"]
    rhs = tot1["rhs"] - tot0["rhs"]
    # algorithmic gather bytes per RHS: 8 corners x 6 components x 8 B (FP64)
    # or x 4 B (FP32 fields)
    bytes_per_rhs = BYTES_PER_RHS_FP64 if args.prec == "fp64" else BYTES_PER_RHS_FP64 // 2
    if args.snapshots > 1:
        bytes_per_rhs *= 2  # the two bracketing snapshots
    achieved = bytes_per_rhs * rhs / (kernel_ms * 1e-3) / 1e9
    batch.close()

    var = None
    if not args.no_variants and args.prec == "fp64" and args.snapshots == 0:
        print("variants ...", file=sys.stderr, flush=True)
        var = variants(args, fields, make_fields, p, tg, st0, method, dict(acc=acc, wall=wall))

    cpu = None
    if not args.no_cpu and uu is not None:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(uu, bb, p, tg, st0, method, args.cpu_seconds)

    out = {
        "metric": "accepted particle-steps/s on 1 MI355X at tol=1e-8, plain vs learned-h",
        "value": acc / wall,
        "unit": "accepted particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded random-mode fields and particles)",
        "config": cfg,
        "variant": ("plain FP64, bit-identical to the reference" if uu is not None else
                    "plain FP64, time-dependent field (bit-identical to the oracle; the "
                    "reference's field is static)" if args.snapshots > 0 else
                    "plain FP64 (EB fields: bit-identical to the oracle; the reference has "
                    "no E-field form)"),
        "accepted": acc, "rejected": rej, "rejected_per_accepted": rej / max(acc, 1),
        "rhs_per_accepted": rhs / max(acc, 1), "kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(cfg),
                     "bytes_per_rhs": bytes_per_rhs, "rhs_per_launch": rhs / args.steps},
        "cpu_baseline": cpu,
        "variants": var,
        "autotuned": tuned,
    }
    if args.prec == "fp32":
        out["dtype"] = "f64 state, f32 fields and RHS"
        out["variant"] = "mixed precision"
    print(json.dumps(out), flush=True)
    fields.close()


if __name__ == "__main__":
    main()
