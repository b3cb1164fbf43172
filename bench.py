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

Prints one JSON line.  `roofline` is the integrator kernel's algorithmic
gather traffic (8 corners x 6 components x 8 B = 384 B per RHS evaluation,
times the RHS evaluations the kernel counted) over its duration measured with
HIP events on the batch's stream, against the MI355X HBM peak.
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
    from particles_amd import workloads as W
    if args.workload == "turbulent256":
        dm, npart, method, tol, gr = 256, 100_000, "dop853", 1e-8, (1.0, 1.0e3)
        tmax, nsnap = 1000 * W.T0, 10
    elif args.workload == "turbulent64":
        dm, npart, method, tol, gr = 64, 1_000, "rk5", 1e-6, None
        tmax, nsnap = 100 * W.T0, 10
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    if args.particles:
        npart = args.particles
    if args.grid:
        dm = args.grid
    if args.method:
        method = args.method
    uu, bb = W.turbulent_vb(dm, seed=args.seed)
    p = W.params_vector(qom=1.0, vun=1.0, tmax=tmax, dtmax=W.T0, atol=tol, rtol=tol)
    tg = W.time_grid(nsnap, tmax)
    if gr is None:
        st0 = W.initial_state(npart, len(tg), seed=args.seed + 1, gamma=10.0)
    else:
        st0 = W.initial_state(npart, len(tg), seed=args.seed + 1, gamma_range=gr)
    cfg = {"workload": f"{args.workload}: {method} tol={tol:g}, {dm}^3 VB turbulent field "
                       f"(B0=1, dB rms 0.5, V rms 0.1, 256 modes), {npart} protons, "
                       f"gamma {'10' if gr is None else 'log-uniform [%g, %g]' % gr}, "
                       f"step = advance all particles by 1 gyro-period of B0",
           "grid": dm, "particles": npart, "method": method, "tol": tol,
           "snapshots_every_T0": tmax / W.T0 / nsnap, "layout": args.layout,
           "cell_cache": args.cell_cache, "field_precision": args.prec}
    return uu, bb, p, tg, st0, method, cfg


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
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cell-cache", type=int, default=1, choices=[0, 1])
    ap.add_argument("--grid-blocks", type=int, default=0)
    ap.add_argument("--prec", default="fp64", choices=["fp64", "fp32"],
                    help="fp32 = mixed precision (FP32 fields, weights and RHS; FP64 state)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the particle batch would "
              "shard by contiguous particle ranges with the fields replicated, and no such run "
              "is made here.", file=sys.stderr)
        sys.exit(2)

    import particles_amd as P
    from particles_amd import workloads as W

    t_setup = time.perf_counter()
    uu, bb, p, tg, st0, method, cfg = build_workload(args)
    P.set_option("layout", args.layout)
    P.set_option("cell_cache", args.cell_cache)
    P.set_option("grid_blocks", args.grid_blocks)
    fields = P.Fields(uu, bb, prec=args.prec)
    batch = P.Particles(fields, method, p, tg, st0)
    print(f"setup {time.perf_counter() - t_setup:.1f}s: {cfg['workload']}", file=sys.stderr,
          flush=True)

    dt_step = W.T0
    t = 0.0
    for i in range(args.warmup):
        t += dt_step
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
    achieved = BYTES_PER_RHS_FP64 * rhs / (kernel_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
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
        "variant": "plain FP64, bit-identical to the reference",
        "accepted": acc, "rejected": rej, "rejected_per_accepted": rej / max(acc, 1),
        "rhs_per_accepted": rhs / max(acc, 1), "kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_rhs": BYTES_PER_RHS_FP64, "rhs_per_launch": rhs / args.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    fields.close()


if __name__ == "__main__":
    main()
