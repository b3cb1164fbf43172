"""Time-dependent fields: a series of snapshots blended linearly in time at
every Runge-Kutta stage (ptc_fields_series).  The reference has no such
path (its field is static and its RHS takes no time, sources/fields.F90:51-78,
sources/accelerations.F90:59), so the oracle's series functions are pinned
here by what the mathematics fixes:
  * the stage nodes are the closed-form values of the published tableaux
    and the row sums of the stage matrices the integrators use;
  * at a snapshot time the blend is that snapshot bit for bit, before the
    first / after the last time it is the end snapshot, and for fields
    linear in time it is the field at that time;
  * a series of identical snapshots reproduces the static integration bit
    for bit;
  * in a uniform E(t) with B = 0 the momentum is u0 + q/m int E dt, which
    both methods integrate exactly (to rounding) while E is linear in time,
    and to the tolerance across the kinks of a piecewise-linear E(t);
  * a time-dependent B with E = 0 does no work: gamma stays constant.
The device path is then bit-identical to the oracle (FP64) and within the
mixed-precision bound of it (FP32)."""
import math

import numpy as np
import pytest

from particles_amd import workloads as W


def say(*a):
    print(*a, flush=True)


TIMES = np.array([0.0, 2.5, 6.0, 10.0]) * W.T0  # non-uniform spacing


def uniform_field(dm, vec):
    a = np.zeros((dm, dm, dm, 3), dtype=np.float64, order="F")
    a[...] = np.asarray(vec, dtype=np.float64)
    return a


def stack(arrs):
    return np.asfortranarray(np.stack(arrs, axis=-1))


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_stage_nodes_closed_forms_and_row_sums(oracle, method):
    c, rs = oracle.stage_nodes(method)
    if method == "rk5":
        exact = [0.0, 1 / 5, 3 / 10, 3 / 5, 1.0, 7 / 8]
    else:
        s6 = math.sqrt(6.0)
        c4 = (6.0 - s6) / 30.0
        exact = [0.0, 4.0 * c4 / 9.0, 2.0 * c4 / 3.0, c4, (6.0 + s6) / 30.0, 1 / 3, 1 / 4,
                 4 / 13, 127 / 195, 3 / 5, 6 / 7, 1.0]
    np.testing.assert_allclose(c, exact, rtol=4e-16, atol=0)
    # row sums carry the rounding of the large DOP853 entries (|a| < 50)
    np.testing.assert_allclose(rs, c, rtol=0, atol=4e-14)


def test_series_gather_ends_and_snapshot_times(oracle):
    rng = np.random.default_rng(1)
    uu, bb = W.turbulent_vb_series(8, seed=3, times=TIMES, nmodes=32)
    for _ in range(20):
        pos = rng.uniform(-0.5, 1.5, 3)
        for k, tk in enumerate(TIMES):
            u, b = oracle.interpolate_series(uu, bb, TIMES, pos, tk)
            us, bs = oracle.interpolate_lin(uu[..., k].copy(order="F"), bb[..., k].copy(order="F"), pos)
            assert np.array_equal(u, us) and np.array_equal(b, bs)
        for t, k in ((-1.0, 0), (TIMES[-1] + 5.0, -1), (math.nan, 0)):
            u, b = oracle.interpolate_series(uu, bb, TIMES, pos, t)
            us, bs = oracle.interpolate_lin(uu[..., k].copy(order="F"), bb[..., k].copy(order="F"), pos)
            assert np.array_equal(u, us) and np.array_equal(b, bs)


def test_series_gather_linear_in_time(oracle):
    """Snapshots f0 + t_k d: the blend at any t is the trilinear interpolant
    of f0 + t d (both maps are linear; rounding only)."""
    rng = np.random.default_rng(2)
    u0, b0 = W.turbulent_vb(8, seed=4, nmodes=32)
    du, db = W.turbulent_vb(8, seed=5, nmodes=32)
    uu = stack([u0 + t * du for t in TIMES])
    bb = stack([b0 + t * db for t in TIMES])
    worst = 0.0
    for _ in range(200):
        pos = rng.uniform(0, 1, 3)
        t = rng.uniform(TIMES[0], TIMES[-1])
        u, b = oracle.interpolate_series(uu, bb, TIMES, pos, t)
        us, bs = oracle.interpolate_lin(np.asfortranarray(u0 + t * du),
                                        np.asfortranarray(b0 + t * db), pos)
        worst = max(worst, np.abs(u - us).max(), np.abs(b - bs).max())
    say(f"[series] linear-in-time blend vs direct: max abs diff {worst:.2e}")
    assert worst < 1e-12


@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("mode", ["vb", "eb"])
def test_identical_snapshots_reproduce_static(oracle, method, mode):
    uu, bb = W.turbulent_vb(16, seed=6, nmodes=64)
    if mode == "eb":
        uu = 0.05 * uu
    p = W.params_vector(tmax=3 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(64, len(tg), seed=7, gamma_range=(1.0, 1e3))
    ss, cs, _ = oracle.integrate(method, uu, bb, p, tg, st0, mode=mode)
    tt = np.array([0.0, 1.0, 2.5]) * W.T0
    s2, c2, _ = oracle.integrate_series(method, stack([uu] * 3), stack([bb] * 3), tt, p, tg, st0,
                                        mode=mode)
    assert np.array_equal(ss, s2) and np.array_equal(cs, c2)


def ramp_case(kinks: bool):
    """Uniform E(t), B = 0 (EB mode): snapshots of E at TIMES; returns the
    fields and u(t) - u0 = q/m int_0^t E (q/m = 1), exact for piecewise
    linear E."""
    e = np.array([[0.02, -0.01, 0.03], [0.05, 0.02, -0.04], [0.06, 0.025, -0.05],
                  [0.01, -0.03, 0.02]])
    if not kinks:
        # E linear in time through the first and last values
        e = e[0] + (e[-1] - e[0]) * (TIMES / TIMES[-1])[:, None]
    ee = stack([uniform_field(4, v) for v in e])
    bb = stack([uniform_field(4, (0.0, 0.0, 0.0))] * len(TIMES))

    def du(t):
        out = np.zeros(3)
        for k in range(len(TIMES) - 1):
            a, b = TIMES[k], TIMES[k + 1]
            s = min(max(t, a), b)
            if s > a:
                ea = e[k]
                eb = e[k] + (e[k + 1] - e[k]) * (s - a) / (b - a)
                out += 0.5 * (ea + eb) * (s - a)
        return out
    return ee, bb, du


@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("kinks", [False, True])
def test_uniform_e_ramp_closed_form(oracle, method, kinks):
    """Every stage of a step is exact here, so a wrong stage node (b_j dc_j
    ~ 0.01) shows as an error of about T dt b_j dc_j |E'| ~ 1e-7 relative
    over the run.  The snapshots are interpolated: DOP853's cubic Hermite
    uses the end slope f(t + dt) and is exact for the quadratic u(t); RK5's
    takes its end slope from stage 6 at t + 7/8 dt (as the reference's,
    integrations.F90:156-166), which biases it by up to 0.02 dt^2 |E'|
    (2e-11 relative at dtmax = 0.005), and 12,600 steps accumulate about
    1e-12 of rounding.  Bounds: 1e-11 (DOP853) and 1e-10 (RK5) for linear E,
    1e-8 across kinks (measured 3e-10); an oracle built with one node off
    by 0.01 to 0.03 (RK5 c6, DOP853 c7 or c8) errs by 1e-7 to 2e-6 here."""
    ee, bb, du = ramp_case(kinks)
    tol = 1e-10
    p = W.params_vector(tmax=TIMES[-1], dtmax=0.005, atol=tol, rtol=tol)
    tg = W.time_grid(12, p[2])
    st0 = W.initial_state(16, len(tg), seed=8, gamma_range=(1.0, 10.0))
    st, _, status = oracle.integrate_series(method, ee, bb, TIMES, p, tg, st0, mode="eb")
    assert (status == 0).all()
    ref = np.stack([du(t) for t in tg], 1)[:, :, None] + st0[3:6, :1, :]
    err = np.abs(st[3:6] - ref).max() / np.abs(ref).max()
    say(f"[series ramp] {method} kinks={kinks}: max relative momentum error {err:.2e}")
    assert err < (1e-8 if kinks else (1e-11 if method == "dop853" else 1e-10))


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_time_dependent_b_does_no_work(oracle, method):
    """E = 0, B changing in time: gamma is conserved up to the integrator's
    own error, which on this coarse 8^3 field (cell-face kinks of the
    trilinear B) matches the static field's: median drift about 1e-8 (RK5,
    whose error norm is atol-only) and 1e-10 (DOP853) at tol 1e-10.  A blend
    that leaked B into E would change gamma at order one."""
    _, bb = W.turbulent_vb_series(8, seed=9, times=TIMES, speed=0.02, nmodes=32)
    ee = np.zeros_like(bb, order="F")
    p = W.params_vector(tmax=TIMES[-1], dtmax=1.0, atol=1e-10, rtol=1e-10)
    tg = W.time_grid(10, p[2])
    st0 = W.initial_state(64, len(tg), seed=10, gamma_range=(1.0, 100.0))
    st, _, status = oracle.integrate_series(method, ee, bb, TIMES, p, tg, st0, mode="eb")
    assert (status == 0).all()
    rel = np.abs(st[6, 1:] - st[6, :1]) / st[6, :1]
    say(f"[series E=0] {method}: gamma drift median {np.median(rel):.2e} max {rel.max():.2e}")
    assert np.median(rel) < (1e-7 if method == "rk5" else 1e-9) and rel.max() < 1e-4


# ---------------------------------------------------------------------------
# device path (GPU)
# ---------------------------------------------------------------------------

def device_series(P, uu, bb, times, mode="vb", prec="fp64"):
    snaps = [P.Fields(uu[..., k].copy(order="F"), bb[..., k].copy(order="F"), mode=mode, prec=prec)
             for k in range(len(times))]
    try:
        return P.Fields.series(snaps, times)
    finally:
        for f in snaps:
            f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("mode", ["vb", "eb"])
@pytest.mark.parametrize("cache", [0, 1])
def test_device_series_matches_oracle(oracle, method, mode, cache):
    """Three propagating turbulent snapshots; the run goes past the last
    snapshot time (held constant there).  One lane per particle ("tail 0");
    with the cell cache (second snapshot in LDS) also at 64 threads per
    block."""
    import particles_amd as P
    tt = TIMES[:3]
    uu, bb = W.turbulent_vb_series(24, seed=11, times=tt, speed=0.02, nmodes=128)
    if mode == "eb":
        uu = np.asfortranarray(0.05 * uu)
    p = W.params_vector(tmax=8 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(8, p[2])
    st0 = W.initial_state(256, len(tg), seed=12, gamma_range=(1.0, 1e3))
    say(f"[series] {method} {mode} cache {cache} 24^3 x3 np=256 t=8 T0")
    so, co, sto = oracle.integrate_series(method, uu, bb, tt, p, tg, st0, mode=mode)
    old = P.get_option("cell_cache"), P.get_option("block")
    try:
        P.set_option("tail", 0)
        P.set_option("cell_cache", cache)
        runs = []
        with device_series(P, uu, bb, tt, mode=mode) as f:
            for block in ((256, 64) if cache else (256,)):
                P.set_option("block", block)
                runs.append((block, P.integrate(f, method, p, tg, st0)))
    finally:
        P.set_option("cell_cache", old[0])
        P.set_option("block", old[1])
        P.set_option("tail", "auto")
    for block, (st, cnt, status) in runs:
        assert (status == sto).all() and (status == 0).all()
        same = np.array([np.array_equal(st[..., i], so[..., i]) and
                         np.array_equal(cnt[:, i], co[:, i]) for i in range(st.shape[2])])
        say(f"  block {block}: bit-exact {same.sum()}/{len(same)}; accepted {cnt[1].sum()} "
            f"vs {co[1].sum()}")
        assert same.all()


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("mode", ["vb", "eb"])
def test_group_kernel_series_matches_oracle(oracle, method, mode):
    """The group kernel (six lanes per particle, "tail group") in a snapshot
    series: FP64 bit-identical to the oracle's restatement (three
    propagating snapshots, past the last snapshot time; and a shifted box of
    non-power-of-two size with velocity unit 0.9), FP32 within the static
    mixed-precision bound (median relative deviation 1e-6, every particle
    1e-3 over 0.1 gyro-periods)."""
    import particles_amd as P
    tt = TIMES[:3]
    uu, bb = W.turbulent_vb_series(24, seed=11, times=tt, speed=0.02, nmodes=128)
    if mode == "eb":
        uu = np.asfortranarray(0.05 * uu)
    p = W.params_vector(tmax=8 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(8, p[2])
    st0 = W.initial_state(256, len(tg), seed=12, gamma_range=(1.0, 1e3))
    xl, xs = (-0.5, 0.2, 0.1), (1.5, 2.0, 0.75)
    g = W.params_vector(qom=1.0, vun=0.9, tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8,
                        xl=xl, xs=xs)
    tgg = W.time_grid(3, g[2])
    sg = W.initial_state(256, len(tgg), seed=26, gamma_range=(1.0, 1e3), xl=xl, xs=xs)
    q = W.params_vector(tmax=0.1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tgq = W.time_grid(2, q[2])
    say(f"[group series] {method} {mode}")
    so, co, _ = oracle.integrate_series(method, uu, bb, tt, p, tg, st0, mode=mode)
    sgo, cgo, _ = oracle.integrate_series(method, uu, bb, tt, g, tgg, sg, mode=mode)
    sqo, _, _ = oracle.integrate_series(method, uu, bb, tt, q, tgq, st0[:, :len(tgq), :], mode=mode)
    try:
        P.set_option("tail", "group")
        with device_series(P, uu, bb, tt, mode=mode) as f:
            st, cnt, status = P.integrate(f, method, p, tg, st0)
            stg, cntg, statusg = P.integrate(f, method, g, tgg, sg)
        with device_series(P, uu, bb, tt, mode=mode, prec="fp32") as f:
            s32, _, status32 = P.integrate(f, method, q, tgq, st0[:, :len(tgq), :])
    finally:
        P.set_option("tail", "auto")
    assert (status == 0).all() and np.array_equal(st, so) and np.array_equal(cnt, co)
    assert (statusg == 0).all() and np.array_equal(stg, sgo) and np.array_equal(cntg, cgo)
    d = s32[:6] - sqo[:6]
    dev = np.maximum(np.sqrt((d[:3] ** 2).sum(0)) / np.maximum(np.sqrt((sqo[:3] ** 2).sum(0)), 1),
                     np.sqrt((d[3:6] ** 2).sum(0)) / np.maximum(np.sqrt((sqo[3:6] ** 2).sum(0)), 1)).max(0)
    say(f"  fp32 deviation median {np.median(dev):.2e} max {dev.max():.2e}")
    assert (status32 == 0).all() and np.median(dev) < 1e-6 and dev.max() < 1e-3


def device_series_columns(P, uu, bb, times, mode="vb", prec="fp64"):
    """device_series with every snapshot in column rows (FP64 layout or FP32
    f32_layout "columns")."""
    opt = "layout" if prec == "fp64" else "f32_layout"
    old = P.get_option(opt)
    P.set_option(opt, "columns")
    try:
        return device_series(P, uu, bb, times, mode=mode, prec=prec)
    finally:
        P.set_option(opt, old)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("mode", ["vb", "eb"])
def test_column_row_series_matches_node_rows(oracle, method, mode):
    """A snapshot series in column rows (each snapshot's quad rows stored z
    fastest; the six-lane group kernel reads both bracketing snapshots with
    the column loads): FP64 bit-identical to the oracle's restatement, in
    the unit box with particles on x, y and z cell faces and in the last
    cells, and in a shifted box of non-power-of-two size; FP32 identical to
    the FP32 series in node rows with the group kernel; the download gives
    each snapshot's values back."""
    import particles_amd as P
    tt = TIMES[:3]
    uu, bb = W.turbulent_vb_series(24, seed=11, times=tt, speed=0.02, nmodes=128)
    if mode == "eb":
        uu = np.asfortranarray(0.05 * uu)
    p = W.params_vector(tmax=4 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(400, len(tg), seed=12, gamma_range=(1.0, 1e3))
    rng = np.random.default_rng(27)
    for a in range(3):
        st0[a, 0, 100 * a:100 * a + 60] = rng.integers(0, 24, 60) / 24.0  # on faces
        st0[a, 0, 100 * a + 60:100 * a + 80] = 1.0 - 1e-12  # last cell
    st0[:3, 0, 300:320] = 1.0 - 1e-12
    xl, xs = (-0.5, 0.2, 0.1), (1.5, 2.0, 0.75)
    g = W.params_vector(qom=1.0, vun=0.9, tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8,
                        xl=xl, xs=xs)
    tgg = W.time_grid(3, g[2])
    sg = W.initial_state(256, len(tgg), seed=26, gamma_range=(1.0, 1e3), xl=xl, xs=xs)
    q = W.params_vector(tmax=0.2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tgq = W.time_grid(2, q[2])
    say(f"[column series] {method} {mode}")
    so, co, _ = oracle.integrate_series(method, uu, bb, tt, p, tg, st0, mode=mode)
    sgo, cgo, _ = oracle.integrate_series(method, uu, bb, tt, g, tgg, sg, mode=mode)
    with device_series_columns(P, uu, bb, tt, mode=mode) as f:
        st, cnt, status = P.integrate(f, method, p, tg, st0)
        stg, cntg, statusg = P.integrate(f, method, g, tgg, sg)
        u1, b1 = f.download(1)
    assert np.array_equal(u1, uu[..., 1]) and np.array_equal(b1, bb[..., 1])
    assert (status == 0).all() and np.array_equal(st, so) and np.array_equal(cnt, co)
    assert (statusg == 0).all() and np.array_equal(stg, sgo) and np.array_equal(cntg, cgo)
    try:
        P.set_option("tail", "group")
        with device_series(P, uu, bb, tt, mode=mode, prec="fp32") as f:
            ref32 = P.integrate(f, method, q, tgq, st0[:, :len(tgq), :])
    finally:
        P.set_option("tail", "auto")
    with device_series_columns(P, uu, bb, tt, mode=mode, prec="fp32") as f:
        got32 = P.integrate(f, method, q, tgq, st0[:, :len(tgq), :])
    assert (got32[2] == 0).all()
    assert np.array_equal(got32[0], ref32[0]) and np.array_equal(got32[1], ref32[1])


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("mode", ["vb", "eb"])
@pytest.mark.parametrize("rows", ["nodes", "columns"])
def test_trio_series_matches_group_series(oracle, method, mode, rows):
    """A snapshot series in the trio kernel (three lanes per particle,
    group_lanes 3) against the six-lane group kernel, with the group or trio
    kernel running the whole launch: FP64 bit-identical to each other and to
    the oracle's restatement, FP32 identical to each other, on node and
    column rows, with particles on cell faces and past the last snapshot
    time."""
    import particles_amd as P
    tt = TIMES[:3]
    uu, bb = W.turbulent_vb_series(24, seed=11, times=tt, speed=0.02, nmodes=128)
    if mode == "eb":
        uu = np.asfortranarray(0.05 * uu)
    p = W.params_vector(tmax=4 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(400, len(tg), seed=12, gamma_range=(1.0, 1e3))
    rng = np.random.default_rng(28)
    for a in range(3):
        st0[a, 0, 100 * a:100 * a + 60] = rng.integers(0, 24, 60) / 24.0
    st0[:3, 0, 300:320] = 1.0 - 1e-12
    q = W.params_vector(tmax=0.2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tgq = W.time_grid(2, q[2])
    make = device_series_columns if rows == "columns" else device_series
    so, co, _ = oracle.integrate_series(method, uu, bb, tt, p, tg, st0, mode=mode)
    got = {}
    try:
        P.set_option("tail", "group")
        for lanes in ("6", "3"):
            say(f"[trio series] {method} {mode} {rows} group_lanes {lanes}")
            P.set_option("group_lanes", lanes)
            with make(P, uu, bb, tt, mode=mode) as f:
                r64 = P.integrate(f, method, p, tg, st0)
            with make(P, uu, bb, tt, mode=mode, prec="fp32") as f:
                r32 = P.integrate(f, method, q, tgq, st0[:, :len(tgq), :])
            got[lanes] = (r64, r32)
    finally:
        P.set_option("tail", "auto")
        P.set_option("group_lanes", "auto")
    for lanes in ("6", "3"):
        (st, cnt, status), (s32, c32, st32) = got[lanes]
        assert (status == 0).all() and (st32 == 0).all(), lanes
        assert np.array_equal(st, so) and np.array_equal(cnt, co), lanes
    assert np.array_equal(got["3"][1][0], got["6"][1][0])
    assert np.array_equal(got["3"][1][1], got["6"][1][1])


@pytest.mark.gpu
def test_column_row_series_refuses_mixed_layouts():
    import particles_amd as P
    uu, bb = W.turbulent_vb(8, seed=22, nmodes=16)
    old = P.get_option("layout")
    try:
        with P.Fields(uu, bb) as f:
            P.set_option("layout", "columns")
            with P.Fields(uu, bb) as c:
                with pytest.raises(P.PtcError, match="differs"):
                    P.Fields.series([c, f], [0.0, 1.0])
    finally:
        P.set_option("layout", old)


@pytest.mark.gpu
def test_device_series_many_snapshots_matches_oracle(oracle):
    """33 snapshots over the run: lanes that pick up a new particle start
    from an interval far from their last one (the lookup's bisection)."""
    import particles_amd as P
    tt = np.linspace(0.0, 6.0, 33) * W.T0
    uu, bb = W.turbulent_vb_series(12, seed=23, times=tt, speed=0.05, nmodes=48)
    p = W.params_vector(tmax=6 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(6, p[2])
    st0 = W.initial_state(512, len(tg), seed=24, gamma_range=(1.0, 1e3))
    so, co, sto = oracle.integrate_series("dop853", uu, bb, tt, p, tg, st0)
    with device_series(P, uu, bb, tt) as f:
        st, cnt, status = P.integrate(f, "dop853", p, tg, st0)
    assert (status == sto).all() and (status == 0).all()
    assert np.array_equal(st, so) and np.array_equal(cnt, co)


@pytest.mark.gpu
def test_device_series_probes_match_oracle(oracle):
    import particles_amd as P
    rng = np.random.default_rng(13)
    uu, bb = W.turbulent_vb_series(12, seed=14, times=TIMES, nmodes=64)
    p = W.params_vector()
    n = 400
    pos = rng.uniform(-0.5, 1.5, (n, 3))
    t = rng.uniform(-W.T0, TIMES[-1] + W.T0, n)
    t[:4] = TIMES
    t[4] = math.nan
    vec = np.concatenate([pos, rng.normal(0, 5, (n, 3))], 1)
    with device_series(P, uu, bb, TIMES) as f:
        g = P.probe(f, "gather_t", np.concatenate([pos, t[:, None]], 1), p)
        a = P.probe(f, "rhs_t", np.concatenate([vec, t[:, None]], 1), p)
    for i in range(n):
        u, b = oracle.interpolate_series(uu, bb, TIMES, pos[i], t[i])
        assert np.array_equal(g[i], np.concatenate([u, b])), i
        acc = oracle.acceleration_series(uu, bb, TIMES, vec[i], t[i])
        assert np.array_equal(a[i], acc), i


@pytest.mark.gpu
def test_identical_snapshot_series_equals_static_device():
    import particles_amd as P
    uu, bb = W.turbulent_vb(16, seed=15, nmodes=64)
    p = W.params_vector(tmax=3 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(4, p[2])
    st0 = W.initial_state(256, len(tg), seed=16, gamma_range=(1.0, 1e3))
    with P.Fields(uu, bb) as f:
        s1, c1, _ = P.integrate(f, "dop853", p, tg, st0)
        with P.Fields.series([f, f], [0.0, W.T0]) as fs:
            s2, c2, _ = P.integrate(fs, "dop853", p, tg, st0)
        with P.Fields.series([f], [0.0]) as f1:  # one snapshot through the series kernels
            s3, c3, _ = P.integrate(f1, "dop853", p, tg, st0)
    assert np.array_equal(s1, s2) and np.array_equal(c1, c2)
    assert np.array_equal(s1, s3) and np.array_equal(c1, c3)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_series_general_geometry(oracle, prec):
    """A domain that is neither at the origin nor of power-of-two size (the
    general, dividing geometry kernels) and a non-unit velocity unit: FP64
    bit-identical to the oracle, FP32 within the mixed-precision bound."""
    import particles_amd as P
    tt = TIMES[:3]
    uu, bb = W.turbulent_vb_series(16, seed=25, times=tt, speed=0.02, nmodes=64)
    xl, xs = (-0.5, 0.2, 0.1), (1.5, 2.0, 0.75)
    p = W.params_vector(qom=1.0, vun=0.9, tmax=(2 if prec == "fp64" else 0.1) * W.T0,
                        dtmax=1.0, atol=1e-8, rtol=1e-8, xl=xl, xs=xs)
    tg = W.time_grid(3, p[2])
    st0 = W.initial_state(256, len(tg), seed=26, gamma_range=(1.0, 1e3), xl=xl, xs=xs)
    so, co, sto = oracle.integrate_series("dop853", uu, bb, tt, p, tg, st0)
    with device_series(P, uu, bb, tt, prec=prec) as f:
        st, cnt, status = P.integrate(f, "dop853", p, tg, st0)
    assert (status == 0).all() and (sto == 0).all()
    if prec == "fp64":
        assert np.array_equal(st, so) and np.array_equal(cnt, co)
    else:
        d = st[:6] - so[:6]
        dev = np.maximum(np.sqrt((d[:3] ** 2).sum(0)) / np.maximum(np.sqrt((so[:3] ** 2).sum(0)), 1),
                         np.sqrt((d[3:6] ** 2).sum(0)) / np.maximum(np.sqrt((so[3:6] ** 2).sum(0)), 1)).max(0)
        say(f"[mixed series, general geometry] deviation median {np.median(dev):.2e} max {dev.max():.2e}")
        assert np.median(dev) < 1e-6 and dev.max() < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["rk5", "dop853"])
@pytest.mark.parametrize("tail", ["0", "group"])
def test_mixed_series_close_to_oracle(oracle, method, tail):
    """FP32 series, one lane per particle and the group kernel: the bound of
    the static mixed-precision test (median relative deviation 1e-6, every
    particle 1e-3 over 0.1 gyro-periods), and the ramp closed form to FP32
    accuracy."""
    import particles_amd as P
    P.set_option("tail", tail)
    try:
        _mixed_series_checks(P, oracle, method)
    finally:
        P.set_option("tail", "auto")


def _mixed_series_checks(P, oracle, method):
    tt = np.array([0.0, 0.04, 0.1]) * W.T0
    uu, bb = W.turbulent_vb_series(24, seed=17, times=tt, speed=1.0, nmodes=128)
    p = W.params_vector(tmax=0.1 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st0 = W.initial_state(512, len(tg), seed=18, gamma_range=(1.0, 1e3))
    so, _, _ = oracle.integrate_series(method, uu, bb, tt, p, tg, st0)
    with device_series(P, uu, bb, tt, prec="fp32") as f:
        st, _, status = P.integrate(f, method, p, tg, st0)
    d = st[:6] - so[:6]
    dev = np.maximum(np.sqrt((d[:3] ** 2).sum(0)) / np.maximum(np.sqrt((so[:3] ** 2).sum(0)), 1),
                     np.sqrt((d[3:6] ** 2).sum(0)) / np.maximum(np.sqrt((so[3:6] ** 2).sum(0)), 1)).max(0)
    say(f"[mixed series] {method}: deviation median {np.median(dev):.2e} max {dev.max():.2e}")
    assert (status == 0).all() and np.median(dev) < 1e-6 and dev.max() < 1e-3

    ee, bz, du = ramp_case(False)
    p = W.params_vector(tmax=TIMES[-1], dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(12, p[2])
    st0 = W.initial_state(64, len(tg), seed=19, gamma_range=(1.0, 10.0))
    with device_series(P, ee, bz, TIMES, mode="eb", prec="fp32") as f:
        st, _, status = P.integrate(f, method, p, tg, st0)
    ref = np.stack([du(t) for t in tg], 1)[:, :, None] + st0[3:6, :1, :]
    err = np.abs(st[3:6] - ref).max() / np.abs(ref).max()
    say(f"[mixed series ramp] {method}: max relative momentum error {err:.2e}")
    assert (status == 0).all() and err < 1e-6


@pytest.mark.gpu
def test_series_graph_segments_and_queue_orders_keep_results():
    import particles_amd as P
    tt = TIMES[:3]
    uu, bb = W.turbulent_vb_series(16, seed=20, times=tt, speed=0.02, nmodes=64)
    p = W.params_vector(tmax=6 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(6, p[2])
    st0 = W.initial_state(2048, len(tg), seed=21, gamma_range=(1.0, 1e3))
    old = P.get_option("sort"), P.get_option("order")
    try:
        with device_series(P, uu, bb, tt) as f:
            P.set_option("sort", "0")  # index order, one call
            ref, rc, _ = P.integrate(f, "dop853", p, tg, st0)
            P.set_option("sort", "1")
            P.set_option("order", "auto")
            with P.Particles(f, "dop853", p, tg, st0) as b:
                b.advance_graph(tg[1:])
                b.advance()
                s2, c2, _ = b.download()
    finally:
        P.set_option("sort", old[0])
        P.set_option("order", old[1])
    assert np.array_equal(ref, s2) and np.array_equal(rc, c2)


@pytest.mark.gpu
def test_series_argument_errors():
    import particles_amd as P
    uu, bb = W.turbulent_vb(8, seed=22, nmodes=16)
    with P.Fields(uu, bb) as f, P.Fields(uu[:, :, :4].copy(order="F"),
                                         bb[:, :, :4].copy(order="F")) as g, \
            P.Fields(uu, bb, prec="fp32") as h:
        for snaps, times, what in (([f, f], [1.0, 1.0], "increase"),
                                   ([f, f], [2.0, 1.0], "increase"),
                                   ([f, f], [0.0, math.inf], "finite"),
                                   ([f, g], [0.0, 1.0], "differs"),
                                   ([f, h], [0.0, 1.0], "differs"),
                                   ([], [], "n >= 1")):
            with pytest.raises(P.PtcError, match=what):
                P.Fields.series(snaps, times)
        with P.Fields.series([f, f], [0.0, 1.0]) as s:
            with pytest.raises(P.PtcError, match="series"):
                P.Fields.series([s, f], [0.0, 1.0])
            p = W.params_vector(tmax=1.0)
            tg = W.time_grid(2, 1.0)
            st0 = W.initial_state(4, len(tg), seed=1)
            with P.Particles(s, "dop853", p, tg, st0) as b:
                with pytest.raises(P.PtcError, match="static"):
                    b.record(0.5, 100)
                zero = np.zeros(P.MLP_PARAMS, dtype=np.float32)
                zero[-len(P.FEATURES):] = 1.0  # the input scales (last 16) must be > 0
                with P.Mlp(zero) as net:
                    with pytest.raises(P.PtcError, match="static"):
                        b.use_mlp(net)
    old = P.get_option("layout")
    try:
        P.set_option("layout", "soa")
        with P.Fields(uu, bb) as f:
            with pytest.raises(P.PtcError, match="aos"):
                P.Fields.series([f], [0.0])
    finally:
        P.set_option("layout", old)
