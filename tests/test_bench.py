"""bench.py's contract on the CPU: the workload configurations it reports
and the committed rocprofv3 traffic figures it attaches to them (a figure
only for the configuration it was measured on), and the refusal of more than
one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _cfg(argv):
    import argparse
    ap = argparse.ArgumentParser()
    for a, d in (("--workload", "turbulent256"), ("--layout", "aos"), ("--order", "cost"), ("--tail", "auto"),
                 ("--prec", "fp64"), ("--method", ""), ("--f32-layout", "nodes"),
                 ("--group-lanes", "6")):
        ap.add_argument(a, default=d)
    for a, d in (("--particles", 0), ("--grid", 0), ("--seed", 1), ("--cell-cache", 1),
                 ("--sort", 1), ("--snapshots", 0), ("--steps", 10), ("--warmup", 2)):
        ap.add_argument(a, type=int, default=d)
    ap.add_argument("--snapshot-every", type=float, default=2.0)
    args = ap.parse_args(argv)
    if args.workload == "turbulent256" and args.snapshots == 0:
        # the static default builds its field on the host: skip that here
        cfg = {"grid": 256, "particles": 100000, "method": "dop853", "tol": 1e-8,
               "layout": args.layout, "cell_cache": args.cell_cache, "field_precision": args.prec,
               "sort": args.sort, "order": args.order, "field_snapshots": 0, "tail": args.tail,
               "group_lanes": args.group_lanes}
        return cfg
    return bench.build_workload(args)[-1]


def test_profiles_name_their_configuration():
    names = sorted(f for f in os.listdir(os.path.join(ROOT, "profiles")) if f.endswith(".json"))
    assert names, "no traffic figures committed"
    for n in names:
        with open(os.path.join(ROOT, "profiles", n)) as fh:
            prof = json.load(fh)
        assert prof["hbm_bytes_per_launch"] == pytest.approx(prof["read_bytes"] + prof["write_bytes"])
        assert prof["config"]["grid"] > 0 and "source" in prof


def test_traffic_matches_only_its_configuration():
    static = bench.traffic_from_profile(_cfg([]))
    series = bench.traffic_from_profile(_cfg(["--snapshots", "8"]))
    assert static is not None and series is not None and series > static
    assert bench.traffic_from_profile(_cfg(["--snapshots", "8", "--snapshot-every", "1"])) is None
    assert bench.traffic_from_profile(_cfg(["--layout", "soa"])) is None
    assert bench.traffic_from_profile(_cfg(["--order", "spatial"])) is None
    assert bench.traffic_from_profile(_cfg(["--prec", "fp32"])) is None
    assert bench.traffic_from_profile(_cfg(["--tail", "16"])) is None
    assert bench.traffic_from_profile(_cfg(["--tail", "0"])) is None
    # the static figure was measured with six lanes per particle; a series
    # never runs the trio kernel, so the lanes setting does not split it
    assert bench.traffic_from_profile(_cfg(["--group-lanes", "3"])) is None
    assert bench.traffic_from_profile(_cfg(["--snapshots", "8", "--group-lanes", "3"])) == series


def test_default_workload_traffic_follows_the_field_storage():
    """The 1024^3 figures: column rows (the default storage) and pair /
    quad rows each get their own profile; the FP64 storage does not split
    FP32 runs, nor the FP32 storage FP64 runs."""
    base = ["--workload", "reconnection1024", "--order", "auto", "--group-lanes", "auto"]
    cols64 = bench.traffic_from_profile(_cfg(base + ["--layout", "columns",
                                                     "--f32-layout", "columns"]))
    pairs64 = bench.traffic_from_profile(_cfg(base + ["--layout", "pairs"]))
    cols32 = bench.traffic_from_profile(_cfg(base + ["--prec", "fp32", "--f32-layout", "columns"]))
    quads32 = bench.traffic_from_profile(_cfg(base + ["--prec", "fp32", "--f32-layout", "quads"]))
    assert None not in (cols64, pairs64, cols32, quads32)
    assert cols64 < pairs64 and cols32 < pairs64
    assert bench.traffic_from_profile(_cfg(base + ["--layout", "columns",
                                                   "--f32-layout", "nodes"])) == cols64
    assert bench.traffic_from_profile(_cfg(base + ["--prec", "fp32", "--layout", "aos",
                                                   "--f32-layout", "columns"])) == cols32
    assert bench.traffic_from_profile(_cfg(base + ["--layout", "columns",
                                                   "--group-lanes", "6"])) is None


def test_secondary_workloads_have_traffic_figures():
    """The 256^3 and 512^3 workloads' default configurations (FP64 column
    rows; FP32 node rows at 256^3, column rows at 512^3) each get their
    profile's traffic, in both precisions."""
    for wl, f32 in (("turbulent256", "nodes"), ("reconnection512", "columns")):
        for prec in ("fp64", "fp32"):
            cfg = _cfg(["--workload", wl, "--layout", "columns", "--f32-layout", f32,
                        "--order", "auto", "--group-lanes", "auto", "--prec", prec])
            if wl == "turbulent256":
                cfg["f32_layout"] = f32
            assert bench.traffic_from_profile(cfg) is not None, (wl, prec)


def test_launches_past_the_horizon_are_refused():
    """Every launch advances 1 T0; the 512^3 workload's horizon is 10 T0, so
    2 warm-up and 10 timed launches are refused before any device work."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--workload",
                        "reconnection512", "--warmup", "2", "--steps", "10"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "horizon of 10 T0" in r.stderr


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
