"""The device path at the size of each benchmark configuration.

The parity tests elsewhere run small fields; here every configuration runs
at its own grid and particle count (over a short horizon), and what can be
compared is compared:

* 256^3 turbulent VB field, 100,000 protons, DOP853 at tol 1e-8 (the
  turbulent-box DOP853 configuration): every particle of a sample of 1,024
  bit-identical to the oracle on the same host field, with the field in
  node rows and in column rows;
* 512^3 EB reconnection field built on the device, 1,000,000 protons,
  Cash-Karp at tol 1e-8: a sample bit-identical to the oracle on the field
  downloaded from the device, in node rows and in column rows;
* 1024^3 EB reconnection field, 4,000,000 protons, DOP853 at tol 1e-8: the
  FP64 run (column rows, as the benchmark stores it) bit-identical to the
  oracle on a sample (51.5 GB field downloaded), the mixed-precision run
  (FP32 field) within the stated bound of the FP64 run from the same seed,
  every status 0;
* the learned step-size controller on that workload: every status 0 and an
  error against a tight-tolerance FP64 run no larger than the PI
  controller's.
"""
import math

import numpy as np
import pytest

import particles_amd as P
from particles_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def exact(sa, ca, sb, cb):
    return np.array([np.array_equal(sa[..., i], sb[..., i]) and np.array_equal(ca[:, i], cb[:, i])
                     for i in range(sa.shape[2])])


def deviation(a, b):
    """Per-particle max over snapshots of |a - b| / max(|b|, 1) for position
    and momentum."""
    d = a[:6] - b[:6]
    pos = np.sqrt((d[:3] ** 2).sum(0)) / np.maximum(np.sqrt((b[:3] ** 2).sum(0)), 1.0)
    mom = np.sqrt((d[3:6] ** 2).sum(0)) / np.maximum(np.sqrt((b[3:6] ** 2).sum(0)), 1.0)
    return np.maximum(pos, mom).max(0)


SAMPLE = 1024


def _sample(np_):
    return np.linspace(0, np_ - 1, SAMPLE).astype(np.int64)


@pytest.fixture(params=["aos", "columns"])
def layout(request):
    """The FP64 field storage: node rows, and column rows (the benchmark's)."""
    old = P.get_option("layout")
    P.set_option("layout", request.param)
    yield request.param
    P.set_option("layout", old)


def test_turbulent256_dop853_100k(oracle, layout):
    dm, npart = 256, 100_000
    say("[c1] 256^3 turbulent field on the host")
    uu, bb = W.turbulent_vb(dm, seed=1)
    p = W.params_vector(tmax=0.25 * W.T0, dtmax=W.T0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(npart, len(tg), seed=2, gamma_range=(1.0, 1.0e3))
    with P.Fields(uu, bb) as f:
        say(f"[c1] device, 100,000 particles, 0.25 T0, layout {layout}")
        st, cnt, status = P.integrate(f, "dop853", p, tg, st0)
    assert (status == 0).all()
    idx = _sample(npart)
    so, co, sto = oracle.integrate("dop853", uu, bb, p, tg, np.asfortranarray(st0[:, :, idx]))
    ok = exact(st[:, :, idx], cnt[:, idx], so, co)
    say(f"[c1] {ok.sum()}/{len(idx)} sampled particles bit-identical, "
        f"{int(cnt[1].sum())} accepted steps")
    assert ok.all() and (sto == 0).all()


def test_reconnection512_rk5_1m(oracle, layout):
    dm, npart = 512, 1_000_000
    say("[c2] 512^3 reconnection field on the device")
    with W.reconnection_eb_device(dm, seed=1) as f:
        p = W.params_vector(tmax=0.1 * W.T0, dtmax=W.T0, atol=1e-8, rtol=1e-8)
        tg = W.time_grid(2, p[2])
        st0 = W.initial_state(npart, len(tg), seed=2, gamma_range=(1.0, 1.0e4))
        say(f"[c2] device, 1,000,000 particles, 0.1 T0, layout {layout}")
        st, cnt, status = P.integrate(f, "rk5", p, tg, st0)
        say("[c2] download the field")
        ee, bb = f.download()
    assert (status == 0).all()
    idx = _sample(npart)
    so, co, sto = oracle.integrate("rk5", ee, bb, p, tg, np.asfortranarray(st0[:, :, idx]),
                                   mode="eb")
    ok = exact(st[:, :, idx], cnt[:, idx], so, co)
    say(f"[c2] {ok.sum()}/{len(idx)} sampled particles bit-identical, "
        f"{int(cnt[1].sum())} accepted steps")
    assert ok.all() and (sto == 0).all()


@pytest.fixture(scope="module")
def reconnection1024():
    """The FP64 field in column rows (192 GiB, the benchmark's storage),
    and FP32 node rows converted from it on the device (32 GiB)."""
    dm = 1024
    say("[c3] 1024^3 reconnection field on the device: FP64 column rows, FP32 node rows")
    old = P.get_option("layout")
    P.set_option("layout", "columns")
    try:
        f64 = W.reconnection_eb_device(dm, seed=1)
    finally:
        P.set_option("layout", old)
    f32 = f64.convert("fp32")
    yield f64, f32
    f32.close()
    f64.close()


def test_reconnection1024_fp64_and_mixed_4m(oracle, reconnection1024):
    f64, f32 = reconnection1024
    npart = 4_000_000
    p = W.params_vector(tmax=0.1 * W.T0, dtmax=W.T0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(npart, len(tg), seed=2, gamma_range=(1.0, 1.0e4))
    say("[c3] FP64, 4,000,000 particles, 0.1 T0")
    s64, c64, t64 = P.integrate(f64, "dop853", p, tg, st0)
    say("[c3] mixed precision")
    s32, c32, t32 = P.integrate(f32, "dop853", p, tg, st0)
    assert (t64 == 0).all() and (t32 == 0).all()
    dev = deviation(s32, s64)
    say(f"[c3] mixed vs FP64: median {np.median(dev):.2e}, p99 {np.percentile(dev, 99):.2e}, "
        f"max {dev.max():.2e}")
    # FP32 fields and slopes: the median deviation after 0.1 T0 is FP32
    # rounding grown by the orbit; a few particles near the current sheets
    # (field gradients over 0.02 box lengths) diverge further
    assert np.median(dev) < 1e-6 and dev.max() < 1e-2
    say("[c3] download the FP64 field (51.5 GB)")
    ee, bb = f64.download()
    idx = _sample(npart)[:256]
    so, co, sto = oracle.integrate("dop853", ee, bb, p, tg, np.asfortranarray(st0[:, :, idx]),
                                   mode="eb")
    del ee, bb
    ok = exact(s64[:, :, idx], c64[:, idx], so, co)
    say(f"[c3] FP64: {ok.sum()}/{len(idx)} sampled particles bit-identical")
    assert ok.all() and (sto == 0).all()


def test_reconnection1024_learned_4m(reconnection1024):
    f64, f32 = reconnection1024
    p = W.params_vector(tmax=0.2 * W.T0, dtmax=W.T0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(1, p[2])
    train = W.initial_state(20_000, len(tg), seed=12, gamma_range=(1.0, 1.0e4))
    test = W.initial_state(4_000_000, len(tg), seed=2, gamma_range=(1.0, 1.0e4))
    say("[c4] records from the PI controller, 20,000 particles, 0.2 T0")
    with P.Particles(f32, "dop853", p, tg, train) as b:
        b.record(1.0, 20_000_000)
        b.advance()
        recs = b.records(device=True)
        net = P.Mlp(seed=1)
        net.normalize(recs)
        losses = net.train(recs, steps=1500, batch=16384, lr=2e-3, alpha=1.0, seed=7)
    try:
        say(f"[c4] {len(recs)} records, loss {losses[:50].mean():.3f} -> "
            f"{losses[-50:].mean():.3f}; learned controller on 4,000,000 particles")
        with P.Particles(f32, "dop853", p, tg, test) as b:
            b.use_mlp(net, -0.3)
            ms_l = b.advance()
            t_l = b.totals()
            s_l, _, st_l = b.download()
    finally:
        net.close()
    with P.Particles(f32, "dop853", p, tg, test) as b:
        ms_pi = b.advance()
        t_pi = b.totals()
        s_pi, _, st_pi = b.download()
    assert (st_l == 0).all() and (st_pi == 0).all()
    # error against FP64 at atol = rtol = 1e-12 on a sample
    idx = _sample(4_000_000)
    tight = p.copy()
    tight[5] = tight[6] = 1e-12
    ref, _, st_r = P.integrate(f64, "dop853", tight, tg, np.asfortranarray(test[:, :, idx]))
    assert (st_r == 0).all()

    def err(s):
        return float(np.median(np.abs(s[:6, -1, idx] - ref[:6, -1, :]).max(0)))

    e_l, e_pi = err(s_l), err(s_pi)
    say(f"[c4] learned: {ms_l:.0f} ms, rejected/accepted {t_l['rejected'] / t_l['accepted']:.3f}, "
        f"error {e_l:.2e}; PI: {ms_pi:.0f} ms, {t_pi['rejected'] / t_pi['accepted']:.3f}, "
        f"error {e_pi:.2e}")
    assert e_l <= e_pi
