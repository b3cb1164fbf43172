"""Device integrator (HIP, gfx950) against the reference's golden vectors and
the CPU oracle on the same seeded inputs.

Bar: the device path runs the reference's FP64 operation order without FMA
contraction, and its step-size powers reproduce the host libm's pow, so
snapshots and step counters are bit-identical to the reference's.  Each test
also reports the deviation against a tolerance of 1e-8 relative in position
and momentum, which bit-identity implies.
"""
import math

import numpy as np
import pytest

import particles_amd as P
from conftest import golden_cases, load_golden
from particles_amd import workloads as W

pytestmark = pytest.mark.gpu

PARITY_TOL = 1e-8


def say(*a):
    print(*a, flush=True)


def deviation(a, b):
    """Per-particle max over snapshots of |a - b| / max(|b|, 1), for the
    position and the momentum vector (state rows 1:3 and 4:6)."""
    d = a[:6] - b[:6]
    pos = np.sqrt((d[:3] ** 2).sum(0)) / np.maximum(np.sqrt((b[:3] ** 2).sum(0)), 1.0)
    mom = np.sqrt((d[3:6] ** 2).sum(0)) / np.maximum(np.sqrt((b[3:6] ** 2).sum(0)), 1.0)
    return np.maximum(pos, mom).max(0)


def exact_particles(sa, ca, sb, cb):
    return np.array([np.array_equal(sa[..., i], sb[..., i]) and np.array_equal(ca[:, i], cb[:, i])
                     for i in range(sa.shape[2])])


@pytest.fixture(params=["aos", "soa", "brick"])
def layout(request):
    old = P.get_option("layout")
    P.set_option("layout", request.param)
    yield request.param
    P.set_option("layout", old)


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ia = a.view(np.int64).astype(object)
    ib = b.view(np.int64).astype(object)
    return np.array([abs(x - y) for x, y in zip(ia.ravel(), ib.ravel())], dtype=np.float64)


def test_device_probes_match_reference_golden(layout):
    """gather and RHS are bit-identical to the reference's interpolate_lin /
    acceleration, including positions on cell faces and far periodic images."""
    g = load_golden("probe_dm9x12x8")
    say(f"[probe] layout={layout} n={len(g['pos'])}")
    with P.Fields(g["uu"], g["bb"]) as f:
        fb = P.probe(f, "gather", g["pos"], g["params"])
        acc = P.probe(f, "rhs", g["vec"], g["params"])
        d5 = P.probe(f, "dti5", g["vec"], g["params"])
        d8 = P.probe(f, "dti8", g["vec"], g["params"])
    nu = int((fb[:, :3] != g["u"]).any(1).sum() + (fb[:, 3:] != g["b"]).any(1).sum())
    na = int((acc != g["acc"]).any(1).sum())
    say(f"  gather mismatches {nu}, rhs mismatches {na}, dti5 {int((d5 != g['dti5']).sum())}, "
        f"dti8 {int((d8 != g['dti8']).sum())}, max dti ulps {max(ulps(d5, g['dti5']).max(), ulps(d8, g['dti8']).max())}")
    assert np.array_equal(fb[:, :3], g["u"]) and np.array_equal(fb[:, 3:], g["b"])
    assert np.array_equal(acc, g["acc"])
    assert np.array_equal(d5, g["dti5"]) and np.array_equal(d8, g["dti8"])


def test_device_pow_against_host_libm():
    """The step-size factor x**y on the device equals the host libm's pow
    (what the reference's real ** real calls) bit for bit."""
    import ctypes as C
    libm = C.CDLL("libm.so.6")
    libm.pow.argtypes = [C.c_double, C.c_double]
    libm.pow.restype = C.c_double
    rng = np.random.default_rng(0)
    n = 200000
    x = np.exp(rng.uniform(-700, 700, size=n))
    x[: n // 2] = np.exp(rng.uniform(-20, 20, size=n // 2))
    y = rng.choice([-0.2, -0.25, -0.125, -0.117, 0.04, 0.2, 0.125], size=n)
    uu = np.zeros((4, 4, 4, 3), order="F")
    with P.Fields(uu, uu) as f:
        dev = P.probe(f, "pow", np.stack([x, y], 1), W.params_vector())
    host = np.array([libm.pow(a, b) for a, b in zip(x, y)])
    nbad = int((dev.view(np.int64) != host.view(np.int64)).sum())
    say(f"[pow] device vs host libm mismatches {nbad}/{n}")
    assert nbad == 0


@pytest.mark.parametrize("case", golden_cases())
def test_device_matches_reference_golden(case, layout):
    g = load_golden(case)
    say(f"[golden] {case} layout={layout} np={g['state0'].shape[2]}")
    with P.Fields(g["uu"], g["bb"]) as f:
        st, cnt, status = P.integrate(f, g["method"], g["params"], g["time"], g["state0"])
    assert (status == 0).all()
    ex = exact_particles(st, cnt, g["state"], g["counter"])
    dev = deviation(st, g["state"])
    say(f"  bit-exact particles {ex.sum()}/{len(ex)}, max deviation {dev.max():.3e}")
    assert dev.max() <= PARITY_TOL
    assert ex.all(), "snapshots and step counters are bit-identical to the reference"


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_device_matches_oracle_turbulent_10_gyroperiods(oracle, method):
    """Turbulent-box recipe (B0 = 1, dB rms 0.5, V rms 0.1, 256 modes with
    |k| in [2 pi, 16 pi]) on a 32^3 grid, 10 gyro-periods of B0 at tol 1e-8."""
    uu, bb = W.turbulent_vb(32, seed=1)
    p = W.params_vector(tmax=10 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(10, p[2])
    st0 = W.initial_state(192, len(tg), seed=2, gamma_range=(1.0, 1.0e3))
    say(f"[oracle] turbulent 32^3 {method} np=192 t=10 T0")
    so, co, _ = oracle.integrate(method, uu, bb, p, tg, st0)
    with P.Fields(uu, bb) as f:
        st, cnt, status = P.integrate(f, method, p, tg, st0)
    assert (status == 0).all()
    ex = exact_particles(st, cnt, so, co)
    dev = deviation(st, so)
    say(f"  bit-exact {ex.sum()}/{len(ex)}, max deviation {dev.max():.3e}, "
        f"accepted {cnt[1].sum()} vs {co[1].sum()}")
    assert dev.max() <= PARITY_TOL
    assert ex.all()


def test_cpu_config_full_size_matches_oracle(oracle):
    """The reference's CPU-runnable case at its full size: 1,000 protons
    (gamma 10, isotropic, uniform in the unit box) in the 64^3 turbulent
    field (B0 = 1, dB rms 0.5, V rms 0.1, 256 modes), Cash-Karp RK5(4) at
    atol = rtol = 1e-6 to t = 100 gyro-periods of B0, ten snapshots.  Every
    snapshot and step counter bit-identical to the oracle (bound: 1e-8
    relative in position and momentum); where the compiled reference is
    present, its first 64 particles bit-identical to it as well."""
    import os
    from oracle.oracle import REF_SO, Reference
    uu, bb = W.turbulent_vb(64, seed=1)
    p = W.params_vector(tmax=100 * W.T0, dtmax=W.T0, atol=1e-6, rtol=1e-6)
    tg = W.time_grid(10, p[2])
    st0 = W.initial_state(1000, len(tg), seed=2, gamma=10.0)
    say("[config 0] turbulent 64^3 rk5 tol 1e-6 np=1000 t=100 T0")
    with P.Fields(uu, bb) as f:
        st, cnt, status = P.integrate(f, "rk5", p, tg, st0)
    so, co, _ = oracle.integrate("rk5", uu, bb, p, tg, st0)
    ex = exact_particles(st, cnt, so, co)
    dev = deviation(st, so)
    say(f"  bit-exact {ex.sum()}/{len(ex)}, max deviation {dev.max():.3e}, "
        f"accepted {cnt[1].sum()}, rejected {cnt[2].sum()}")
    assert (status == 0).all()
    assert dev.max() <= PARITY_TOL
    assert ex.all()
    if os.path.exists(REF_SO):
        sr, cr = Reference().integrate("rk5", uu, bb, p, tg, st0[:, :, :64])
        assert np.array_equal(st[:, :, :64], sr) and np.array_equal(cnt[:, :64], cr)
        say("  first 64 particles bit-identical to the compiled reference")


def test_large_field_offsets(oracle):
    """A 512^3 field: its node rows span 6.4 GB (FP64) and 4.3 GB (FP32),
    beyond 32-bit byte offsets.  Gathers and RHS at positions over the whole
    box, the far corner and its periodic faces included, are bit-identical to
    the oracle in FP64 and within 1e-6 of the field scale in FP32, and a
    short DOP853 run of 64 particles is bit-identical to the oracle."""
    n = 512
    rng = np.random.default_rng(31)
    uu = np.empty((n, n, n, 3), dtype=np.float64, order="F")
    bb = np.empty((n, n, n, 3), dtype=np.float64, order="F")
    for arr, off in ((uu, 0.0), (bb, 1.0)):
        for c in range(3):
            # a sum of three random profiles, one per axis: every node distinct
            ax, ay, az = (0.1 * rng.normal(size=n) for _ in range(3))
            arr[..., c] = ax[:, None, None] + ay[None, :, None]
            arr[..., c] += az[None, None, :] + (off if c == 2 else 0.0)
    pos = rng.uniform(0.0, 1.0, size=(3000, 3))
    pos[:1000] = 1.0 - rng.uniform(0.0, 2.0 / n, size=(1000, 3))  # the last cells
    pos[1000:1100] = np.round(pos[1000:1100] * n) / n              # nodes and faces
    vec = np.hstack([pos, rng.normal(size=(3000, 3))])
    p = W.params_vector(tmax=0.2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(64, len(tg), seed=32, gamma_range=(1.0, 100.0))
    st0[0:3, 0, :32] = 1.0 - rng.uniform(0.0, 2.0 / n, size=(3, 32))
    gu = np.array([np.concatenate(oracle.interpolate_lin(uu, bb, x)) for x in pos])
    ga = np.array([oracle.acceleration(uu, bb, v) for v in vec])
    so, co, _ = oracle.integrate("dop853", uu, bb, p, tg, st0)
    say(f"[large] {n}^3 fp64: {len(pos)} probes, 64 particles")
    with P.Fields(uu, bb) as f:
        g = P.probe(f, "gather", pos, p)
        a = P.probe(f, "rhs", vec, p)
        st, cnt, status = P.integrate(f, "dop853", p, tg, st0)
    assert np.array_equal(g, gu) and np.array_equal(a, ga)
    assert (status == 0).all() and np.array_equal(st, so) and np.array_equal(cnt, co)
    say(f"[large] {n}^3 fp32 probes")
    with P.Fields(uu, bb, prec="fp32") as f:
        g32 = P.probe(f, "gather", pos, p)
    err = np.abs(g32 - gu).max() / np.abs(gu).max()
    say(f"  fp32 gather max error {err:.2e} of the field scale")
    assert err < 1e-6


def test_eb_mode_matches_oracle(oracle):
    ee, bb = W.reconnection_eb(32, seed=3, nmodes=64)
    p = W.params_vector(tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(128, len(tg), seed=4, gamma_range=(1.0, 1.0e2))
    for method in ("rk5", "dop853"):
        say(f"[eb] reconnection 32^3 {method}")
        so, co, _ = oracle.integrate(method, ee, bb, p, tg, st0, mode="eb")
        with P.Fields(ee, bb, mode="eb") as f:
            st, cnt, status = P.integrate(f, method, p, tg, st0)
        dev = deviation(st, so)
        say(f"  max deviation {dev.max():.3e}")
        assert (status == 0).all() and dev.max() <= PARITY_TOL
        assert np.array_equal(st, so) and np.array_equal(cnt, co)


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_segments_equal_one_call(method):
    """Advancing in snapshot-interval segments continues each trajectory
    exactly: bit-identical to one uninterrupted integration."""
    uu, bb = W.turbulent_vb(16, seed=7, nmodes=64)
    p = W.params_vector(tmax=3 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(6, p[2])
    st0 = W.initial_state(300, len(tg), seed=8, gamma_range=(1.0, 100.0))
    say(f"[segments] {method}")
    with P.Fields(uu, bb) as f:
        s1, c1, _ = P.integrate(f, method, p, tg, st0)
        with P.Particles(f, method, p, tg, st0) as b:
            for t_stop in np.linspace(0, p[2], 8)[1:-1]:
                b.advance(t_stop)
                assert b.totals()["running"] > 0
            b.advance()
            tot = b.totals()
            s2, c2, st2 = b.download()
    assert np.array_equal(s1, s2) and np.array_equal(c1, c2)
    assert (st2 == 0).all() and tot["running"] == 0
    assert tot["accepted"] == c1[1].sum() and tot["rejected"] == c1[2].sum()
    assert tot["all"] == c1[0].sum()


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_cell_cache_modes_identical(oracle, method):
    """The per-particle cell cache changes where the corner values come from,
    never their values or summation order."""
    uu, bb = W.turbulent_vb(24, seed=12, nmodes=64)
    p = W.params_vector(tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(500, len(tg), seed=13, gamma_range=(1.0, 1000.0))
    so, co, _ = oracle.integrate(method, uu, bb, p, tg, st0)
    try:
        P.set_option("tail", 0)  # the lane-per-particle kernel's cache
        for mode in (0, 1):
            P.set_option("cell_cache", mode)
            say(f"[cache] {method} cell_cache={mode}")
            with P.Fields(uu, bb) as f:
                st, cnt, status = P.integrate(f, method, p, tg, st0)
            assert (status == 0).all()
            assert np.array_equal(st, so) and np.array_equal(cnt, co), mode
    finally:
        P.set_option("cell_cache", 1)
        P.set_option("tail", "auto")


def test_launch_shape_does_not_change_results():
    """The work queue hands particles to lanes in a launch-dependent order;
    results must not depend on it."""
    uu, bb = W.turbulent_vb(16, seed=9, nmodes=64)
    p = W.params_vector(tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(3, p[2])
    st0 = W.initial_state(1000, len(tg), seed=10, gamma_range=(1.0, 1000.0))
    outs = []
    try:
        P.set_option("tail", 0)  # the lane-per-particle kernel's launch shape
        for block, grid in ((256, 0), (64, 3), (128, 1)):
            P.set_option("block", block)
            P.set_option("grid_blocks", grid)
            say(f"[launch] block={block} grid_blocks={grid}")
            with P.Fields(uu, bb) as f:
                outs.append(P.integrate(f, "dop853", p, tg, st0))
    finally:
        P.set_option("block", 256)
        P.set_option("grid_blocks", 0)
        P.set_option("tail", "auto")
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])


def test_nonfinite_field_and_caps():
    """Failure handling forced from the host: a NaN field node retires the
    particles that meet it (status 2), a step cap retires with status 1,
    others finish normally; every launch returns."""
    dm = (8, 8, 8)
    uu = np.zeros(dm + (3,), order="F")
    bb = np.zeros(dm + (3,), order="F")
    bb[..., 2] = 1.0
    bb[2, 4, 6, 0] = np.nan
    p = W.params_vector(tmax=20.0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = np.zeros((8, len(tg), 3), order="F")
    st0[:6, 0, 0] = [0.27, 0.52, 0.77, 0.5, 0.0, 0.0]   # next to the NaN node
    st0[:6, 0, 1] = [0.9, 0.1, 0.1, 0.1, 0.0, 0.0]      # far away, small orbit
    st0[:6, 0, 2] = [0.9, 0.1, 0.2, 0.1, 0.0, 0.0]
    say("[failure] NaN node")
    with P.Fields(uu, bb) as f:
        _, cnt, status = P.integrate(f, "dop853", p, tg, st0)
        assert status.tolist() == [2, 0, 0]
        try:
            P.set_option("max_steps", 25)
            _, cnt, status = P.integrate(f, "rk5", p, tg, st0)
        finally:
            P.set_option("max_steps", 0)
        assert status[1] == 1 and cnt[0, 1] == 25
        bad = p.copy()
        bad[4] = 0.0  # dtmax <= 0 would never terminate in the reference
        with pytest.raises(P.PtcError) as e:
            P.integrate(f, "rk5", bad, tg, st0)
        assert e.value.code == 1


def test_empty_and_single_snapshot():
    uu, bb = W.turbulent_vb(8, seed=1, nmodes=16)
    p = W.params_vector(tmax=1.0, dtmax=0.5)
    with P.Fields(uu, bb) as f:
        st, cnt, status = P.integrate(f, "rk5", p, np.zeros(1), np.zeros((8, 1, 0), order="F"))
        assert st.shape == (8, 1, 0)
        st0 = W.initial_state(5, 1, seed=3, gamma=3.0)
        st, cnt, status = P.integrate(f, "dop853", p, np.zeros(1), st0)
        assert (status == 0).all() and (cnt[1] > 0).all()
        np.testing.assert_allclose(st[6, 0], 3.0, rtol=1e-14)


def test_energy_conserved_without_electric_field(oracle):
    """U = 0 (E = 0): gamma is invariant.  The snapshots come from the cubic
    dense output, so they hold it to the tolerance level only; the device
    run must show exactly the oracle's (reference's) drift and stay small."""
    _, bb = W.turbulent_vb(24, seed=2)
    uu = np.zeros_like(bb, order="F")
    p = W.params_vector(tmax=10 * W.T0, dtmax=0.5, atol=1e-10, rtol=1e-10)
    tg = W.time_grid(10, p[2])
    st0 = W.initial_state(256, len(tg), seed=5, gamma_range=(2.0, 100.0))
    with P.Fields(uu, bb) as f:
        st, _, status = P.integrate(f, "dop853", p, tg, st0)
    so, _, _ = oracle.integrate("dop853", uu, bb, p, tg, st0)
    g0 = st[6, 0, :]
    rel = np.abs(st[6, 1:, :] - g0) / g0
    rel_o = np.abs(so[6, 1:, :] - so[6, 0, :]) / so[6, 0, :]
    say(f"[energy] max relative gamma drift {rel.max():.3e} (oracle {rel_o.max():.3e})")
    assert (status == 0).all() and np.array_equal(rel, rel_o) and rel.max() < 1e-6


# ---------------------------------------------------------------------------
# mixed precision (FP32 fields, weights and RHS; FP64 state, sums, error norm)
# ---------------------------------------------------------------------------

def test_mixed_probes_close_to_reference():
    """FP32 node storage and FP32 weights: the gather and the RHS agree with
    the reference's FP64 values to FP32 accuracy (tolerance 1e-6 relative to
    the field scale)."""
    g = load_golden("probe_dm9x12x8")
    with P.Fields(g["uu"], g["bb"], prec="fp32") as f:
        fb = P.probe(f, "gather", g["pos"], g["params"])
        acc = P.probe(f, "rhs", g["vec"], g["params"])
    scale_u = np.abs(g["uu"]).max()
    scale_b = np.abs(g["bb"]).max()
    eu = np.abs(fb[:, :3] - g["u"]).max() / scale_u
    eb = np.abs(fb[:, 3:] - g["b"]).max() / scale_b
    ea = np.abs(acc - g["acc"]).max(0) / np.maximum(np.abs(g["acc"]).max(0), 1e-300)
    say(f"[mixed probe] gather rel err U {eu:.2e} B {eb:.2e}; rhs rel err {ea.max():.2e}")
    assert eu < 1e-6 and eb < 1e-6 and ea.max() < 1e-6


@pytest.mark.parametrize("mode", ["vb", "eb"])
@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_mixed_precision_error_against_oracle(oracle, method, tail, mode):
    """Mixed precision differs from the FP64 oracle by FP32 rounding in the
    RHS; over 0.1 gyro-periods the median relative deviation in position
    and momentum stays below 1e-6 and every particle below 1e-3 (the bound
    is set by chaotic growth of FP32-level differences, which the FP64 path
    does not have), with one lane per particle, with the group kernel and
    with the hand-over between them, in the turbulent VB field and in the
    EB reconnection field (whose FP32 force has its own E + v x B branch)."""
    if mode == "vb":
        uu, bb = W.turbulent_vb(32, seed=1)
    else:
        uu, bb = W.reconnection_eb(32, seed=1)
    p = W.params_vector(tmax=0.1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(512, len(tg), seed=2, gamma_range=(1.0, 1.0e3))
    so, co, _ = oracle.integrate(method, uu, bb, p, tg, st0, mode=mode)
    with P.Fields(uu, bb, prec="fp32", mode=mode) as f:
        st, cnt, status = P.integrate(f, method, p, tg, st0)
    dev = deviation(st, so)
    say(f"[mixed] {method} {mode} tail={tail}: deviation median {np.median(dev):.2e} "
        f"max {dev.max():.2e}")
    assert (status == 0).all()
    assert np.median(dev) < 1e-6 and dev.max() < 1e-3


def test_mixed_precision_energy_without_electric_field():
    _, bb = W.turbulent_vb(24, seed=2)
    uu = np.zeros_like(bb, order="F")
    p = W.params_vector(tmax=5 * W.T0, dtmax=0.5, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(5, p[2])
    st0 = W.initial_state(256, len(tg), seed=5, gamma_range=(2.0, 100.0))
    with P.Fields(uu, bb, prec="fp32") as f:
        st, _, status = P.integrate(f, "dop853", p, tg, st0)
    rel = np.abs(st[6, 1:, :] - st[6, 0, :]) / st[6, 0, :]
    say(f"[mixed energy] max relative gamma drift {rel.max():.3e}")
    assert (status == 0).all() and rel.max() < 1e-6


def test_fortran_caller_integrates():
    """The Fortran demo (ptc_binding) runs the device integrator: ten
    particles gyrating in uniform B, one gyro-period."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "particles_amd", "fortran", "ptc_demo")
    if not os.path.exists(exe):
        pytest.skip("Fortran demo not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    say(r.stdout[-400:])
    assert r.returncode == 0
    rows = [l.split() for l in r.stdout.splitlines() if l.strip().startswith("particle")]
    assert len(rows) == 10 and all(int(row[4]) > 0 for row in rows)


# ---------------------------------------------------------------------------
# device field synthesis (large-grid inputs)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("recipe", ["turbulent_vb", "reconnection_eb"])
def test_device_synthesis_matches_host_generator(recipe):
    """ptc_fields_synthesize builds the same fields as the host generator
    (inverse FFT of the same mode tables; sheet, guide field and E = -V x B
    + eta curl B): gathering both at random points agrees to 1e-12."""
    dm = 24
    if recipe == "turbulent_vb":
        f1, bb = W.turbulent_vb(dm, seed=3, nmodes=64)
        dev = W.turbulent_vb_device(dm, seed=3, nmodes=64)
        mode = "vb"
    else:
        f1, bb = W.reconnection_eb(dm, seed=3, nmodes=64)
        dev = W.reconnection_eb_device(dm, seed=3, nmodes=64)
        mode = "eb"
    pos = np.random.default_rng(1).uniform(0, 1, (2000, 3))
    p = W.params_vector()
    with P.Fields(f1, bb, mode=mode) as host, dev:
        a = P.probe(host, "gather", pos, p)
        b = P.probe(dev, "gather", pos, p)
    scale = max(np.abs(f1).max(), np.abs(bb).max())
    err = np.abs(a - b).max() / scale
    say(f"[synthesis {recipe}] max relative difference {err:.2e}")
    assert err < 1e-12


def test_spatial_order_keeps_results():
    """Handing particles to lanes in Morton order of their cells or most
    work first (re-sorted before every segment or every other) changes no
    snapshot and no counter."""
    uu, bb = W.turbulent_vb(24, seed=3, nmodes=64)
    p = W.params_vector(tmax=2.0 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(3000, len(tg), seed=4, gamma_range=(1.0, 100.0))
    out = []
    for sort, order in ((0, "spatial"), (1, "spatial"), (2, "spatial"), (1, "cost")):
        P.set_option("sort", sort)
        P.set_option("order", order)
        try:
            with P.Fields(uu, bb) as f, P.Particles(f, "dop853", p, tg, st0) as b:
                for k in range(1, 5):
                    b.advance(k * 0.5 * W.T0)
                out.append(b.download())
        finally:
            P.set_option("sort", 1)
            P.set_option("order", "auto")
    for s, c, status in out[1:]:
        assert np.array_equal(s, out[0][0]) and np.array_equal(c, out[0][1])
        assert (status == 0).all()


def test_autotune_keeps_results():
    """The autotuner's probe segments are real integration: the batch ends
    exactly where an untuned one does."""
    uu, bb = W.turbulent_vb(16, seed=2, nmodes=32)
    p = W.params_vector(tmax=1.0 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(3, p[2])
    st0 = W.initial_state(2000, len(tg), seed=6, gamma_range=(1.0, 100.0))
    with P.Fields(uu, bb) as f:
        s_ref, c_ref, _ = P.integrate(f, "rk5", p, tg, st0)
        with P.Particles(f, "rk5", p, tg, st0) as b:
            chosen = b.autotune(0.5 * W.T0)
            b.advance()
            s, c, status = b.download()
    say(f"[autotune] chose {chosen}")
    assert chosen["block"] in (64, 128, 256)
    assert chosen["tail"] in ("0", "group", "32768")
    assert chosen["group_lanes"] in (3, 6)
    assert np.array_equal(s, s_ref) and np.array_equal(c, c_ref) and (status == 0).all()


@pytest.mark.parametrize("layout", ["aos", "columns"])
def test_graph_advances_equal_eager_advances(layout):
    """Snapshot intervals submitted as one captured HIP graph give the same
    snapshots and counters as eager advances; two graphs in a row, with
    eager work and synchronisation between them; node rows (the lane
    kernel with its hand-over) and column rows (the trio kernel)."""
    uu, bb = W.turbulent_vb(16, seed=2, nmodes=32)
    p = W.params_vector(tmax=2.0 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(8, p[2])
    st0 = W.initial_state(3000, len(tg), seed=7, gamma_range=(1.0, 100.0))
    old = P.get_option("layout")
    P.set_option("layout", layout)
    try:
        f = P.Fields(uu, bb)
    finally:
        P.set_option("layout", old)
    with f:
        with P.Particles(f, "dop853", p, tg, st0) as b:
            for t in tg[1:]:
                b.advance(t)
            ref = b.download()
        with P.Particles(f, "dop853", p, tg, st0) as b:
            b.advance_graph(tg[1:4])
            mid = b.totals()
            assert mid["accepted"] > 0
            b.advance_graph(tg[4:])
            got = b.download()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert (got[2] == 0).all()


@pytest.mark.parametrize("tail,sort", [("0", 1), ("group", 0), ("600", 1)])
def test_graph_is_kept_and_replayed_with_new_stop_times(tail, sort, lanes):
    """advance_graph instantiates its graph once per batch and replays it
    with the stop times replaced while n and the launch settings stay; a
    change of either captures a new one.  After every call the batch equals,
    bit for bit, one advanced eagerly through the same stop times (eager
    downloads and synchronisation between the replays), with the group
    kernel of six lanes or the trio kernel where a group kernel runs."""
    if tail == "0" and lanes == "3":
        pytest.skip("one lane per particle: no group kernel runs")
    uu, bb = W.turbulent_vb(16, seed=2, nmodes=32)
    p = W.params_vector(tmax=2.0 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(10, p[2])
    st0 = W.initial_state(3000, len(tg), seed=8, gamma_range=(1.0, 100.0))
    calls = [tg[1:3], tg[3:5], tg[5:7], tg[7:8], tg[8:]]
    # instantiated / replayed after each call: n = 2, 2, 2, then 1, 2
    want = [(1, 0), (1, 1), (1, 2), (2, 2), (3, 2)]
    P.set_option("tail", tail)
    P.set_option("sort", sort)
    try:
        with P.Fields(uu, bb) as f, P.Particles(f, "dop853", p, tg, st0) as eager, \
                P.Particles(f, "dop853", p, tg, st0) as b:
            for k, stops in enumerate(calls):
                for t in stops:
                    eager.advance(t)
                b.advance_graph(stops)
                ref, got = eager.download(), b.download()
                st = b.graph_stats()
                say(f"[graph tail {tail} sort {sort} lanes {lanes}] call {k}: {len(stops)} "
                    f"intervals, {st}")
                assert (st["instantiated"], st["replayed"]) == want[k]
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
                assert np.array_equal(got[2], ref[2])
            assert (got[2] == 0).all() and got[1][1].sum() > 0
            # another setting captures anew
            P.set_option("block", 128)
            b.advance_graph(tg[-1:])
            assert b.graph_stats()["instantiated"] == 4
            # so does the other group width
            P.set_option("group_lanes", "6" if lanes == "3" else "3")
            b.advance_graph(tg[-1:])
            assert b.graph_stats()["instantiated"] == 5
    finally:
        P.set_option("tail", "auto")
        P.set_option("sort", 1)
        P.set_option("block", 256)


@pytest.fixture(params=["0", "group", "600", "100"])
def tail(request):
    """Which kernel integrates: "0" one lane per particle, "group" the group
    kernel (six lanes per particle) alone, "N" the former handing every
    particle still in flight to the latter once the work queue is empty and
    at most N are left."""
    P.set_option("tail", request.param)
    yield request.param
    P.set_option("tail", "auto")


@pytest.fixture(params=["6", "3"])
def lanes(request):
    """Lanes per particle where a group kernel runs: "6" the group kernel,
    "3" the trio kernel (two components per lane)."""
    P.set_option("group_lanes", request.param)
    yield request.param
    P.set_option("group_lanes", "auto")


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_group_kernel_matches_oracle(oracle, method, tail, lanes):
    """The group kernel forms every per-component operation as the lane-per-
    particle kernel does and every sum over components in the same order on
    every lane: snapshots and counters bit-identical to the oracle, in the
    unit box (power-of-two geometry) and in a shifted box of size 2.5 x 3 x
    1.75 with velocity unit 0.7, in VB and EB modes, in one call or in
    segments."""
    uu, bb = W.turbulent_vb(32, seed=41, nmodes=64)
    p = W.params_vector(tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(700, len(tg), seed=42, gamma_range=(1.0, 1000.0))
    if tail == "0" and lanes == "3":
        pytest.skip("one lane per particle: no group kernel runs")
    say(f"[group] {method} tail={tail} lanes={lanes} unit box")
    so, co, _ = oracle.integrate(method, uu, bb, p, tg, st0)
    with P.Fields(uu, bb) as f:
        st, cnt, status = P.integrate(f, method, p, tg, st0)
        assert (status == 0).all()
        assert np.array_equal(st, so) and np.array_equal(cnt, co)
        with P.Particles(f, method, p, tg, st0) as b:
            for t_stop in (0.3 * W.T0, 0.9 * W.T0, 1.4 * W.T0):
                b.advance(t_stop)
            b.advance()
            s2, c2, st2 = b.download()
        assert np.array_equal(s2, so) and np.array_equal(c2, co) and (st2 == 0).all()
    g = W.params_vector(vun=0.7, tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8,
                        xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
    tg = W.time_grid(3, g[2])
    sg = W.initial_state(300, len(tg), seed=43, gamma_range=(1.0, 100.0),
                         xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
    say(f"[group] {method} tail={tail} general box")
    so, co, _ = oracle.integrate(method, uu, bb, g, tg, sg)
    with P.Fields(uu, bb) as f:
        st, cnt, status = P.integrate(f, method, g, tg, sg)
    assert np.array_equal(st, so) and np.array_equal(cnt, co)
    ee, be = W.reconnection_eb(32, seed=44, nmodes=64)
    say(f"[group] {method} tail={tail} EB")
    so, co, _ = oracle.integrate(method, ee, be, p, W.time_grid(4, p[2]), st0, mode="eb")
    with P.Fields(ee, be, mode="eb") as f:
        st, cnt, status = P.integrate(f, method, p, W.time_grid(4, p[2]), st0)
    assert np.array_equal(st, so) and np.array_equal(cnt, co)


def test_group_kernel_failure_handling(lanes):
    """Status codes from the group and trio kernels: a NaN field node retires
    the particle that meets it (2), a step cap retires with status 1 after
    exactly that many attempts, the others finish (0)."""
    dm = (8, 8, 8)
    uu = np.zeros(dm + (3,), order="F")
    bb = np.zeros(dm + (3,), order="F")
    bb[..., 2] = 1.0
    bb[2, 4, 6, 0] = np.nan
    p = W.params_vector(tmax=20.0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = np.zeros((8, len(tg), 3), order="F")
    st0[:6, 0, 0] = [0.27, 0.52, 0.77, 0.5, 0.0, 0.0]
    st0[:6, 0, 1] = [0.9, 0.1, 0.1, 0.1, 0.0, 0.0]
    st0[:6, 0, 2] = [0.9, 0.1, 0.2, 0.1, 0.0, 0.0]
    try:
        P.set_option("tail", "group")
        with P.Fields(uu, bb) as f:
            say(f"[group failure] NaN node, lanes {lanes}")
            _, cnt, status = P.integrate(f, "dop853", p, tg, st0)
            assert status.tolist() == [2, 0, 0]
            P.set_option("max_steps", 25)
            _, cnt, status = P.integrate(f, "rk5", p, tg, st0)
            assert status[1] == 1 and cnt[0, 1] == 25
    finally:
        P.set_option("max_steps", 0)
        P.set_option("tail", "auto")


def test_handover_counts_retired_particles(oracle, lanes):
    """Particles already retired when a launch fetches them (here 3 of 4,
    non-finite initial momenta, status 2 after the first segment) count as
    finished for the hand-over threshold, so the running ones are handed to
    the group kernel with a threshold far below the retired count; every
    snapshot and counter still equals the oracle's."""
    uu, bb = W.turbulent_vb(32, seed=51, nmodes=64)
    p = W.params_vector(tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(3, p[2])
    st0 = W.initial_state(4000, len(tg), seed=52, gamma_range=(1.0, 1000.0))
    st0[3, 0, 1::4] = np.nan
    st0[4, 0, 2::4] = np.inf
    st0[5, 0, 3::4] = np.nan
    so, co, sto = oracle.integrate("dop853", uu, bb, p, tg, st0)
    assert (sto[0::4] == 0).all() and (sto[1::4] == 2).all()
    try:
        P.set_option("tail", "256")
        with P.Fields(uu, bb) as f, P.Particles(f, "dop853", p, tg, st0) as b:
            say(f"[hand-over] 3000 retired of 4000, tail 256, lanes {lanes}")
            for t_stop in (0.2 * W.T0, 0.6 * W.T0):
                b.advance(t_stop)
            b.advance()
            s2, c2, st2 = b.download()
    finally:
        P.set_option("tail", "auto")
    assert np.array_equal(st2, sto)
    assert np.array_equal(c2, co) and np.array_equal(s2, so, equal_nan=True)


@pytest.mark.parametrize("form", ["aos", "soa", "brick", "pairs", "columns", "fp32 nodes",
                                  "fp32 pairs", "fp32 quads", "fp32 columns", "series"])
def test_fields_download_returns_stored_values(form):
    """ptc_fields_download gives back what each storage form holds: FP64
    layouts exactly the uploaded arrays, FP32 ones the arrays rounded to
    float, a series the chosen snapshot."""
    rng = np.random.default_rng(8)
    dm = (6, 5, 7)
    uu = np.asfortranarray(rng.normal(size=dm + (3,)))
    bb = np.asfortranarray(rng.normal(size=dm + (3,)))
    old = P.get_option("layout")
    try:
        if form in ("aos", "soa", "brick", "pairs", "columns"):
            P.set_option("layout", form)
            with P.Fields(uu, bb) as f:
                u2, b2 = f.download()
            assert np.array_equal(u2, uu) and np.array_equal(b2, bb)
        elif form.startswith("fp32"):
            with _fields_fp32(uu, bb, form.split()[1]) as f:
                u2, b2 = f.download()
            assert np.array_equal(u2, uu.astype(np.float32).astype(np.float64))
            assert np.array_equal(b2, bb.astype(np.float32).astype(np.float64))
        else:
            with P.Fields(uu, bb) as f0, P.Fields(2 * uu, -bb) as f1:
                with P.Fields.series([f0, f1], [0.0, 1.0]) as fs:
                    u2, b2 = fs.download(1)
                    u0, _ = fs.download(0)
            assert np.array_equal(u2, 2 * uu) and np.array_equal(b2, -bb)
            assert np.array_equal(u0, uu)
    finally:
        P.set_option("layout", old)


@pytest.mark.parametrize("target", ["fp32 nodes", "fp32 pairs", "fp32 quads", "fp32 columns",
                                    "fp64 pairs", "fp64 columns", "fp64 aos"])
@pytest.mark.parametrize("source", ["aos", "pairs", "columns", "soa"])
def test_fields_convert_equals_direct_creation(source, target):
    """ptc_fields_convert builds, on the device, the same stored values as
    creating the target form from the host arrays (FP32: the values rounded
    to float), and integrating on it gives the same snapshots and counters
    as the directly created field (group kernel)."""
    uu, bb = W.turbulent_vb(16, seed=61, nmodes=32)
    prec, form = target.split()
    opt = "f32_layout" if prec == "fp32" else "layout"
    p = W.params_vector(tmax=0.3 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(400, len(tg), seed=62, gamma_range=(1.0, 300.0))
    old = P.get_option("layout"), P.get_option("f32_layout")
    try:
        P.set_option("layout", source)
        src = P.Fields(uu, bb)
        P.set_option("layout", "aos")
        P.set_option(opt, form)
        with src, src.convert(prec) as conv, P.Fields(uu, bb, prec=prec) as direct:
            assert all(np.array_equal(a, b) for a, b in zip(conv.download(), direct.download()))
            P.set_option("tail", "group")
            a = P.integrate(conv, "dop853", p, tg, st0)
            b = P.integrate(direct, "dop853", p, tg, st0)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and (a[2] == 0).all()
    finally:
        P.set_option("layout", old[0])
        P.set_option("f32_layout", old[1])
        P.set_option("tail", "auto")


@pytest.mark.parametrize("kernel", ["0", "group"])
def test_cell_loads_counted(kernel):
    """The cell-load counter: a particle at rest in uniform fields never
    leaves its cell, so each launch loads its cell once; with the cell cache
    off every RHS evaluation loads; moving particles load at most once per
    evaluation."""
    uu = np.zeros((8, 8, 8, 3), order="F")
    bb = np.zeros((8, 8, 8, 3), order="F")
    bb[..., 2] = 1.0
    p = W.params_vector(tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = np.zeros((8, len(tg), 1), order="F")
    st0[:3, 0, 0] = [0.3, 0.4, 0.5]
    try:
        P.set_option("tail", kernel)
        with P.Fields(uu, bb) as f:
            with P.Particles(f, "dop853", p, tg, st0) as b:
                b.advance(W.T0)
                b.advance()
                tot = b.totals()
            say(f"[loads] tail={kernel} at rest: {tot['cell_loads']} loads, {tot['rhs']} RHS")
            assert tot["cell_loads"] == 2 and tot["rhs"] > 2
            ut, bt = W.turbulent_vb(16, seed=3, nmodes=32)
            sm = W.initial_state(300, len(tg), seed=4, gamma_range=(1.0, 100.0))
            with P.Fields(ut, bt) as g, P.Particles(g, "dop853", p, tg, sm) as b:
                b.advance()
                tot = b.totals()
            say(f"[loads] tail={kernel} moving: {tot['cell_loads']} loads, {tot['rhs']} RHS")
            assert 0 < tot["cell_loads"] < tot["rhs"]
        if kernel == "0":
            P.set_option("cell_cache", 0)
            with P.Fields(ut, bt) as g, P.Particles(g, "dop853", p, tg, sm) as b:
                b.advance()
                tot = b.totals()
            assert tot["cell_loads"] == tot["rhs"]
    finally:
        P.set_option("cell_cache", 1)
        P.set_option("tail", "auto")


def _fields_fp32(uu, bb, layout, **kw):
    P.set_option("f32_layout", layout)
    try:
        return P.Fields(uu, bb, prec="fp32", **kw)
    finally:
        P.set_option("f32_layout", "nodes")


def _fields_rows(uu, bb, form, **kw):
    """Fields in one of the group kernel's row forms: "nodes" / "fp64
    nodes" (node rows), "pairs" (FP32 pair rows), "quads" (FP32 quad rows),
    "columns" (FP32 quad rows stored z fastest), "fp64 pairs", "fp64 columns"
    (FP64 quad rows stored z fastest)."""
    if form.startswith("fp64"):
        P.set_option("layout", form.split()[1].replace("nodes", "aos"))
        try:
            return P.Fields(uu, bb, **kw)
        finally:
            P.set_option("layout", "aos")
    return _fields_fp32(uu, bb, form, **kw)


@pytest.mark.parametrize("form", ["pairs", "quads", "columns", "fp64 pairs", "fp64 columns"])
@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_pair_rows_match_node_rows(method, form):
    """Pair rows (a node and its +x neighbour in one row: 64 bytes in FP32,
    128 in FP64), FP32 quad rows (a node and its +x, +y, +x+y neighbours
    in one 128-byte row) and column rows (those rows stored z fastest; in
    FP64 192-byte rows) hold the same values as node rows,
    and the group kernel reads a right corner from the left corner's row
    only where the right index is the left one's periodic successor:
    snapshots and counters identical to node rows with the group kernel
    (FP64: to the oracle), in the unit box (power-of-two geometry), on a
    12 x 10 x 8 grid in a shifted box, in EB mode, with particles started
    exactly on x, y and z cell faces (where the separately rounded right
    index can differ) and across the periodic boundaries; every "tail"
    setting runs the group kernel on these rows."""
    uu, bb = W.turbulent_vb(32, seed=51, nmodes=64)
    p = W.params_vector(tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(3, p[2])
    st0 = W.initial_state(1200, len(tg), seed=52, gamma_range=(1.0, 1000.0))
    rng = np.random.default_rng(53)
    st0[0, 0, :200] = rng.integers(0, 32, 200) / 32.0  # on x faces
    st0[0, 0, 200:260] = 1.0 - 1e-12  # last cell in x
    st0[1, 0, 260:460] = rng.integers(0, 32, 200) / 32.0  # on y faces
    st0[1, 0, 460:520] = 1.0 - 1e-12  # last cell in y
    st0[:2, 0, 520:560] = 1.0 - 1e-12  # last cell in x and y
    st0[2, 0, 560:760] = rng.integers(0, 32, 200) / 32.0  # on z faces
    st0[2, 0, 760:800] = 1.0 - 1e-12  # last cell in z
    st0[:3, 0, 800:840] = 1.0 - 1e-12  # last cell in x, y and z
    cases = [("unit box", uu, bb, {}, p, st0)]
    ug, bg = W.turbulent_vb(32, seed=54, nmodes=64)
    ug, bg = np.asfortranarray(ug[:12, :10, :8]), np.asfortranarray(bg[:12, :10, :8])
    g = W.params_vector(vun=0.7, tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8,
                        xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
    sg = W.initial_state(800, len(tg), seed=55, gamma_range=(1.0, 100.0),
                         xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
    sg[0, 0, :200] = -1.0 + 2.5 * rng.integers(0, 12, 200) / 12.0
    sg[1, 0, 200:400] = 0.5 + 3.0 * rng.integers(0, 10, 200) / 10.0
    sg[2, 0, 400:600] = 2.0 + 1.75 * rng.integers(0, 8, 200) / 8.0
    cases.append(("12x10x8 shifted box", ug, bg, {}, g, sg))
    ee, be = W.reconnection_eb(32, seed=56, nmodes=64)
    cases.append(("EB", ee, be, {"mode": "eb"}, p, st0))
    node_form = "fp64 nodes" if form.startswith("fp64") else "nodes"
    try:
        for name, a, b, kw, q, s0 in cases:
            say(f"[rows] {method} {form} {name}")
            P.set_option("tail", "group")
            with _fields_rows(a, b, node_form, **kw) as f:
                ref = P.integrate(f, method, q, tg, s0)
            for tail in ("group", "0", "auto"):
                P.set_option("tail", tail)
                with _fields_rows(a, b, form, **kw) as f:
                    got = P.integrate(f, method, q, tg, s0)
                assert (got[2] == 0).all()
                assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), (name, tail)
    finally:
        P.set_option("tail", "auto")


@pytest.mark.parametrize("dims", [(2, 3, 2), (5, 2, 3), (3, 4, 2)])
def test_row_forms_on_the_smallest_grids(oracle, dims):
    """Grids of two to five nodes per axis, where the periodic +x and +y
    neighbours wrap within one or two cells: FP64 pair and column rows with
    the six-lane and trio kernels bit-identical to the oracle, FP32 pair,
    quad and column rows identical to FP32 node rows (group kernel)."""
    uu, bb = W.turbulent_vb(8, seed=91, nmodes=16)
    uu = np.asfortranarray(uu[:dims[0], :dims[1], :dims[2]])
    bb = np.asfortranarray(bb[:dims[0], :dims[1], :dims[2]])
    p = W.params_vector(tmax=0.5 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(300, len(tg), seed=92, gamma_range=(1.0, 30.0))
    so, co, _ = oracle.integrate("dop853", uu, bb, p, tg, st0)
    try:
        P.set_option("tail", "group")
        with _fields_rows(uu, bb, "nodes") as f:
            ref32 = P.integrate(f, "dop853", p, tg, st0)
        for form in ("fp64 pairs", "fp64 columns", "pairs", "quads", "columns"):
            for lanes in ("6", "3"):
                P.set_option("group_lanes", lanes)
                say(f"[small grid] {dims} {form} group_lanes {lanes}")
                with _fields_rows(uu, bb, form) as f:
                    st, cnt, status = P.integrate(f, "dop853", p, tg, st0)
                assert (status == 0).all(), (form, lanes)
                want = (so, co) if form.startswith("fp64") else ref32[:2]
                assert np.array_equal(st, want[0]) and np.array_equal(cnt, want[1]), (form, lanes)
    finally:
        P.set_option("tail", "auto")
        P.set_option("group_lanes", "auto")


@pytest.mark.parametrize("form", ["fp64 nodes", "fp64 pairs", "fp64 columns", "nodes", "pairs",
                                  "quads", "columns"])
@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_trio_kernel_matches_group_kernel(method, form):
    """The trio kernel (three lanes per particle, group_lanes 3) against
    the group kernel (six lanes, group_lanes 6) on the
    same fields and particles, with the trio or group kernel running the
    whole launch ("tail" group) and taking the lane-per-particle kernel's
    last particles (a hand-over, node rows): snapshots and counters
    bit-identical, in FP64 (both restate the reference's operations per
    component) and in FP32 (the mixed-precision unit contracts only within
    expressions, which both kernels write alike).  VB in the unit box, EB in a shifted box with
    particles on x and y cell faces and in the last cells, so that both
    the pair and quad rows' successor rows and the per-corner fallback are
    read (column rows: the z-right rows as well)."""
    uu, bb = W.turbulent_vb(24, seed=71, nmodes=48)
    p = W.params_vector(tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(3, p[2])
    st0 = W.initial_state(1000, len(tg), seed=72, gamma_range=(1.0, 1000.0))
    rng = np.random.default_rng(73)
    st0[0, 0, :100] = rng.integers(0, 24, 100) / 24.0
    st0[1, 0, 100:200] = rng.integers(0, 24, 100) / 24.0
    st0[:2, 0, 200:240] = 1.0 - 1e-12
    st0[2, 0, 240:340] = rng.integers(0, 24, 100) / 24.0
    st0[:3, 0, 340:360] = 1.0 - 1e-12
    ee, be = W.reconnection_eb(24, seed=74, nmodes=48)
    g = W.params_vector(vun=0.5, tmax=1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8,
                        xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
    sg = W.initial_state(1000, len(tg), seed=75, gamma_range=(1.0, 100.0),
                         xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
    sg[0, 0, :100] = -1.0 + 2.5 * rng.integers(0, 24, 100) / 24.0
    sg[1, 0, 100:200] = 0.5 + 3.0 * rng.integers(0, 24, 100) / 24.0
    cases = [("VB unit box", uu, bb, {}, p, st0), ("EB shifted box", ee, be, {"mode": "eb"}, g, sg)]
    tails = ["group"] + (["300"] if form == "fp64 nodes" or form == "nodes" else [])
    try:
        for name, a, b, kw, q, s0 in cases:
            for tail in tails:
                P.set_option("tail", tail)
                got = {}
                for lanes in ("6", "3"):
                    say(f"[trio] {method} {form} {name} tail {tail} group_lanes {lanes}")
                    P.set_option("group_lanes", lanes)
                    with _fields_rows(a, b, form, **kw) as f:
                        got[lanes] = P.integrate(f, method, q, tg, s0)
                    assert (got[lanes][2] == 0).all(), (name, tail, lanes)
                (s6, c6, _), (s3, c3, _) = got["6"], got["3"]
                assert np.array_equal(s3, s6) and np.array_equal(c3, c6), (name, tail)
    finally:
        P.set_option("tail", "auto")
        P.set_option("group_lanes", "auto")


@pytest.mark.parametrize("form", ["pairs", "quads", "columns", "fp64 pairs", "fp64 columns"])
def test_pair_rows_refused_where_node_rows_are_read(form):
    """Probes and record mode read node rows: they refuse pair, quad and
    column rows with an error; snapshot series take column rows and refuse
    pair and quad rows; the learned controller runs on FP32 pair, quad and
    column rows (group kernel) with the results it gives on node rows, and
    refuses FP64 pair and column rows (it runs FP64 fields one lane per
    particle)."""

    def series_check(f):
        if form.endswith("columns"):
            with P.Fields.series([f, f], [0.0, 1.0]) as s:
                assert s.download(1)[0].shape == uu.shape
        else:
            with pytest.raises(P.PtcError):
                P.Fields.series([f, f], [0.0, 1.0])

    from oracle import mlp_oracle as M
    uu, bb = W.turbulent_vb(16, seed=57, nmodes=32)
    p = W.params_vector(tmax=0.5 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(1, p[2])
    st0 = W.initial_state(300, len(tg), seed=58, gamma_range=(1.0, 100.0))
    zero = np.zeros(P.MLP_PARAMS, dtype=np.float32)
    zero[M.SIG:] = 1.0  # output 0: keep h, within the PI floor
    if form.startswith("fp64"):
        with _fields_rows(uu, bb, form) as f, P.Mlp(zero) as net:
            with pytest.raises(P.PtcError):
                P.probe(f, "gather", np.full((4, 3), 0.3), p)
            series_check(f)
            with P.Particles(f, "dop853", p, tg, st0) as b:
                with pytest.raises(P.PtcError):
                    b.record(1.0, 1000)
                with pytest.raises(P.PtcError):
                    b.use_mlp(net)
        return
    with _fields_fp32(uu, bb, form) as f, P.Mlp(zero) as net:
        with pytest.raises(P.PtcError):
            P.probe(f, "gather", np.full((4, 3), 0.3), p)
        with P.Particles(f, "dop853", p, tg, st0) as b:
            with pytest.raises(P.PtcError):
                b.record(1.0, 1000)
        series_check(f)
        P.set_option("tail", "group")
        try:
            with P.Particles(f, "dop853", p, tg, st0) as b:
                b.use_mlp(net)
                b.advance()
                got = b.download()
            with _fields_fp32(uu, bb, "nodes") as g:
                with P.Particles(g, "dop853", p, tg, st0) as b:
                    b.use_mlp(net)
                    b.advance()
                    ref = b.download()
        finally:
            P.set_option("tail", "auto")
    assert (got[2] == 0).all()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_cell_gather_diagnostic():
    """The cell-load diagnostic counts the loads it makes (ten groups per
    wave, four waves per block, three blocks per CU) and reports a positive
    rate; it refuses an unknown pattern and a buffer under 1 MiB."""
    for pat in ("lines", "span"):
        r = P.diag_cell_gather(256 << 20, pat, cells=64)
        say(f"[diag] cell loads, {pat}: {r['cell_loads_per_s']:.3e}/s")
        assert r["cell_loads"] > 0 and r["cell_loads"] % (3 * 4 * 10 * 64) == 0
        assert r["cell_loads_per_s"] > 0 and r["ms"] > 0
    with pytest.raises(KeyError):
        P.diag_cell_gather(256 << 20, "rows")
    with pytest.raises(P.PtcError):
        P.diag_cell_gather(1024, "span")


def test_line_gather_diagnostic():
    """The line-gather diagnostic reads what it says it reads and reports a
    positive rate below the nominal HBM peak."""
    r = P.diag_line_gather(256 << 20, 16)
    say(f"[diag] 256 MiB buffer: {r['GBps']:.0f} GB/s")
    # 8 blocks of 256 threads per CU, 8 lanes a line (32 lines per block and
    # load), 8 loads in flight, 16 rounds: a whole number of CUs' worth
    per_cu = 8 * (256 // 8) * 8 * 16 * 128
    assert r["bytes"] > 0 and r["bytes"] % per_cu == 0
    assert 0 < r["GBps"] < 2.0e4 and r["ms"] > 0
    with pytest.raises(P.PtcError):
        P.diag_line_gather(1024, 16)
