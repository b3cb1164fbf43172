"""Learned step-size proposal around the integrator: the features (oracle
closed forms on the CPU, device probe against the oracle), record mode
(training records from the PI controller without changing its results) and
the learned controller itself (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from particles_amd import workloads as W


def _uniform(dm, b=(0.0, 0.0, 2.0), u=(0.0, 0.0, 0.0)):
    uu = np.zeros(tuple(dm) + (3,), order="F")
    bb = np.zeros(tuple(dm) + (3,), order="F")
    uu[...] = u
    bb[...] = b
    return uu, bb


# ---------------------------------------------------------------------------
# features: closed forms of the oracle (CPU)
# ---------------------------------------------------------------------------
def test_features_uniform_field():
    uu, bb = _uniform((8, 8, 8), b=(0.0, 0.0, 2.0))
    p = W.params_vector(qom=0.5, vun=1.0, tmax=1.0, dtmax=1.0)
    u = np.array([3.0, 0.0, 4.0])             # |u| = 5, gamma = sqrt(26)
    vec = np.array([0.5 / 8, 0.3, 0.9, *u])  # x at the middle of cell 0
    f = M.features(uu, bb, vec, 0.01, 0.25, 0.3, True, p)
    g = np.sqrt(26.0)
    assert abs(f[0] - np.log(2.0)) < 1e-14
    assert f[1] == -30.0                      # no E field: clamped
    assert np.all(np.abs(f[2:8]) < 1e-14)     # uniform |B|: faces equal the centre
    assert abs(f[8] - np.log(g)) < 1e-14
    assert abs(f[9] - 0.8) < 1e-14            # cos pitch = u_z / |u|
    assert abs(f[10] - np.log(0.01 * 0.5 * 2.0 / g)) < 1e-14
    assert abs(f[11] - np.log(0.25)) < 1e-14
    # x: half a cell to the face at velocity 3/g cells... in cells per time
    # vel_x = u_x / (g vun) * dm / xs = 24 / g; z: 4/g*8 cells per time,
    # z = 0.9 -> fraction 0.2 of cell 7 -> 0.8 cells to go
    t_x = 0.5 / (24 / g)
    t_z = 0.8 / (32 / g)
    assert abs(f[12] - np.log(min(t_x, t_z) / 0.01)) < 1e-12
    assert f[13] == 0.3 and f[14] == 1.0
    assert abs(f[15] - np.log(3.0 / (0.5 * 2.0) * 8)) < 1e-12  # u_perp / (qom B) / dx


def test_features_field_gradient_and_e_field():
    dm = (8, 8, 8)
    uu, bb = _uniform(dm, b=(0.0, 0.0, 1.0), u=(0.1, 0.0, 0.0))
    x = np.arange(8) / 8.0
    bb[:, :, :, 2] = (1.0 + x)[:, None, None]  # |B| grows along x
    p = W.params_vector(tmax=1.0, dtmax=1.0)
    vec = np.array([2.5 / 8, 0.5, 0.5, 0.0, 1.0, 0.0])
    f = M.features(uu, bb, vec, 0.01, 0.5, 0.0, False, p)
    bl, br = 1.0 + 2 / 8, 1.0 + 3 / 8
    bc = 0.5 * (bl + br)
    assert abs(f[0] - np.log(bc)) < 1e-14
    assert abs(f[2] - np.log(bl / bc)) < 1e-14 and abs(f[3] - np.log(br / bc)) < 1e-14
    assert np.all(np.abs(f[4:8]) < 1e-14)
    # VB mode: E = -U x B, |E| = |U| |B| for U perpendicular to B
    assert abs(f[1] - np.log(0.1)) < 1e-12
    assert f[14] == 0.0


def test_calibration_falls_back_to_pi():
    """The calibration keeps the network only where a shift runs faster than
    the PI controller at no larger error; otherwise the PI controller is the
    result and the reason is stated (CPU: the decision on measured
    candidates)."""
    from particles_amd.learned import choose_shift
    cand = [{"shift": -0.3, "kernel_ms": 30.0, "error": 0.8e-5},
            {"shift": -0.1, "kernel_ms": 20.0, "error": 1.2e-5},
            {"shift": 0.1, "kernel_ms": 15.0, "error": 2.0e-5}]
    # the fastest shift at no larger error beats PI: kept
    best, cal, beats, why = choose_shift(cand, pi_ms=25.0, pi_error=1.3e-5)
    assert best["shift"] == -0.1 and cal and beats and why == ""
    # qualifying shifts exist but none is faster than PI: PI
    best, cal, beats, why = choose_shift(cand, pi_ms=19.0, pi_error=1.3e-5)
    assert best is None and cal and not beats and "no faster" in why
    # equal time is not faster
    best, _, beats, _ = choose_shift(cand, pi_ms=20.0, pi_error=1.3e-5)
    assert best is None and not beats
    # no shift reaches PI's error: PI, never the least inaccurate network
    best, cal, beats, why = choose_shift(cand, pi_ms=100.0, pi_error=0.5e-5)
    assert best is None and not cal and not beats and "error" in why


# ---------------------------------------------------------------------------
# device (GPU)
# ---------------------------------------------------------------------------
def _feature_inputs(n, seed, xs=1.0):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 10))
    x[:, :3] = rng.uniform(0, xs, (n, 3))
    g = np.exp(rng.uniform(0, np.log(100), n))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x[:, 3:6] = d * np.sqrt(g * g - 1)[:, None]
    x[:, 6] = np.exp(rng.uniform(-8, -2, n))
    x[:, 7] = np.exp(rng.uniform(-6, 0, n))
    x[:, 8] = rng.uniform(-1, 1, n)
    x[:, 9] = rng.integers(0, 2, n)
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["vb", "eb"])
def test_feature_probe_matches_oracle(mode):
    import particles_amd as P
    if mode == "vb":
        f1, bb = W.turbulent_vb(16, seed=4, nmodes=64)
    else:
        f1, bb = W.reconnection_eb(16, seed=4)
    p = W.params_vector(tmax=1.0, dtmax=1.0)
    x = _feature_inputs(400, 5)
    with P.Fields(f1, bb, mode=mode) as fl:
        got = P.probe(fl, "features", x, p)
    want = np.array([M.features(f1, bb, r[:6], r[6], r[7], r[8], r[9] != 0, p, mode) for r in x])
    d = np.abs(got - want)
    print(f"feature probe max abs diff per feature: {np.round(d.max(0), 6)}")
    assert np.all(d < 5e-4 * np.maximum(1.0, np.abs(want)))


def _small_run(method="dop853", np_=1500, T=0.5, seed=2, prec="fp64"):
    uu, bb = W.turbulent_vb(32, seed=1, nmodes=128)
    p = W.params_vector(tmax=T * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(2, p[2])
    st = W.initial_state(np_, len(tg), seed=seed, gamma_range=(1.0, 100.0))
    return uu, bb, p, tg, st


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_record_mode_keeps_results_and_writes_pairs(method):
    import particles_amd as P
    uu, bb, p, tg, st = _small_run(method)
    with P.Fields(uu, bb) as f:
        with P.Particles(f, method, p, tg, st) as b:
            b.advance()
            s0, c0, _ = b.download()
        with P.Particles(f, method, p, tg, st) as b:
            b.record(1.0, 10_000_000)
            b.advance()
            s1, c1, status = b.download()
            r = b.records()
    assert np.array_equal(s0, s1) and np.array_equal(c0, c1)
    assert (status == 0).all()
    # one record per pair of consecutive accepted steps
    assert len(r) == int(c1[1].sum()) - st.shape[2]
    nrej = r[:, 18]
    # without a rejection in between, the next step is the PI proposal itself
    assert np.array_equal(r[nrej == 0, 16], r[nrej == 0, 13])
    assert np.all(r[nrej > 0, 16] < r[nrej > 0, 13])
    assert int(nrej.sum()) <= int(c1[2].sum())
    q = 5.0 if method == "rk5" else 8.0
    assert np.all(r[:, 19] == q) and np.all(r[:, 17] <= 0.0)
    assert np.all(np.isfinite(r))


@pytest.mark.gpu
def test_record_sampling_rate_and_capacity():
    import particles_amd as P
    uu, bb, p, tg, st = _small_run()
    with P.Fields(uu, bb) as f:
        with P.Particles(f, "dop853", p, tg, st) as b:
            b.record(0.25, 10_000_000)
            b.advance()
            n_quarter = len(b.records())
            _, c, _ = b.download()
        with P.Particles(f, "dop853", p, tg, st) as b:
            b.record(1.0, 1000)
            b.advance()
            r = b.records()
            dev = b.records(device=True)
    full = int(c[1].sum()) - st.shape[2]
    assert abs(n_quarter / full - 0.25) < 0.02
    assert len(r) == 1000 and len(dev) == 1000


@pytest.mark.gpu
def test_learned_mode_needs_a_trained_network_and_the_cell_cache():
    import particles_amd as P
    uu, bb, p, tg, st = _small_run(np_=64)
    zero = np.zeros(P.MLP_PARAMS, dtype=np.float32)
    zero[M.SIG:] = 1.0  # output 0: keep h, within the PI floor
    with P.Fields(uu, bb) as f, P.Mlp(seed=1) as untrained, P.Mlp(zero) as net:
        with P.Particles(f, "dop853", p, tg, st) as b:
            with pytest.raises(P.PtcError):
                b.use_mlp(untrained)
            b.use_mlp(net)
            P.set_option("cell_cache", 0)
            try:
                with pytest.raises(P.PtcError):
                    b.advance()
            finally:
                P.set_option("cell_cache", 1)
            b.advance()
            _, c, status = b.download()
    assert (status == 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["group", "trio", "0"])
@pytest.mark.parametrize("method", ["dop853", "rk5"])
def test_trained_network_gains_over_pi(method, kernel):
    """Record with the PI controller on one particle set, train the network
    on the GPU, integrate a different set: the learned controller rejects
    far fewer steps and needs fewer RHS evaluations, at an error against a
    tight-tolerance FP64 run no larger than the PI controller's (mixed
    precision, tol 1e-8, turbulent 32^3 field, gamma in [1, 1e3]), in the
    six-lanes-per-particle kernel ("group"), in the three-lanes-per-particle
    kernel ("trio") and in one lane per particle ("0"), the PI runs in the
    same kernel."""
    import particles_amd as P
    P.set_option("tail", "0" if kernel == "0" else "group")
    P.set_option("group_lanes", "3" if kernel == "trio" else "6")
    try:
        _gains_over_pi(P, method)
    finally:
        P.set_option("tail", "auto")
        P.set_option("group_lanes", "auto")


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["dop853", "rk5"])
def test_learned_trio_matches_group(method):
    """The learned controller in the trio kernel (three lanes per particle,
    21 samples per network tile) and in the group kernel (six lanes, ten
    samples): the same features and network per particle from the same
    values, the same FP32 arithmetic; every particle finishes, and the two
    runs agree to the integration error (median relative deviation of
    position and momentum after 0.5 T0 below 1e-4; the step sequences are
    equal where both kernels round every feature alike)."""
    import particles_amd as P
    uu, bb = W.turbulent_vb(32, seed=11, nmodes=64)
    p = W.params_vector(tmax=0.5 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(1, p[2])
    train = W.initial_state(5000, len(tg), seed=12, gamma_range=(1.0, 1e3))
    test = W.initial_state(3000, len(tg), seed=13, gamma_range=(1.0, 1e3))
    got = {}
    try:
        P.set_option("tail", "group")
        with P.Fields(uu, bb, prec="fp32") as f:
            with P.Particles(f, method, p, tg, train) as b:
                b.record(1.0, 5_000_000)
                b.advance()
                r = b.records()
            with P.Mlp(seed=1) as net:
                net.normalize(r)
                net.train(r, steps=500, batch=8192, lr=2e-3, alpha=1.0, seed=7)
                for lanes in ("6", "3"):
                    P.set_option("group_lanes", lanes)
                    with P.Particles(f, method, p, tg, test) as b:
                        b.use_mlp(net, -0.2)
                        b.advance()
                        got[lanes] = b.download()
    finally:
        P.set_option("tail", "auto")
        P.set_option("group_lanes", "auto")
    (s6, c6, st6), (s3, c3, st3) = got["6"], got["3"]
    assert (st6 == 0).all() and (st3 == 0).all()
    d = s3[:6, -1, :] - s6[:6, -1, :]
    rel = np.maximum(np.sqrt((d[:3] ** 2).sum(0)) / np.maximum(np.sqrt((s6[:3, -1] ** 2).sum(0)), 1.0),
                     np.sqrt((d[3:] ** 2).sum(0)) / np.maximum(np.sqrt((s6[3:6, -1] ** 2).sum(0)), 1.0))
    same = float((c3 == c6).all(0).mean())
    print(f"{method}: counters equal for {same:.3f} of the particles, deviation median "
          f"{np.median(rel):.2e} max {rel.max():.2e}; accepted trio {c3[1].sum()} group "
          f"{c6[1].sum()}")
    assert np.median(rel) < 1e-4
    assert abs(c3[1].sum() / c6[1].sum() - 1.0) < 0.05


def _gains_over_pi(P, method):
    uu, bb = W.turbulent_vb(32, seed=1, nmodes=128)
    p = W.params_vector(tmax=0.5 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(1, p[2])
    train = W.initial_state(20000, len(tg), seed=2, gamma_range=(1.0, 1e3))
    test = W.initial_state(20000, len(tg), seed=3, gamma_range=(1.0, 1e3))
    tight = p.copy()
    tight[5] = tight[6] = 1e-12
    with P.Fields(uu, bb) as f64:
        ref, _, _ = P.integrate(f64, "dop853", tight, tg, test)

    def err(s):
        return float(np.median(np.abs(s[:6, -1, :] - ref[:6, -1, :]).max(0)))

    with P.Fields(uu, bb, prec="fp32") as f:
        with P.Particles(f, method, p, tg, train) as b:
            b.record(1.0, 20_000_000)
            b.advance()
            r = b.records()
        with P.Particles(f, method, p, tg, test) as b:
            b.advance()
            t_pi = b.totals()
            e_pi = err(b.download()[0])
        with P.Mlp(seed=1) as net:
            net.normalize(r)
            losses = net.train(r, steps=3000, batch=16384, lr=2e-3, alpha=1.0, seed=7)
            with P.Particles(f, method, p, tg, test) as b:
                b.use_mlp(net, -0.35)
                b.advance()
                t_l = b.totals()
                s_l, _, status = b.download()
    e_l = err(s_l)
    rho_pi = t_pi["rejected"] / t_pi["accepted"]
    rho_l = t_l["rejected"] / t_l["accepted"]
    print(f"{method}: {len(r)} records, loss {losses[:50].mean():.3f} -> {losses[-50:].mean():.3f}; "
          f"rejected/accepted PI {rho_pi:.3f} learned {rho_l:.3f}; RHS learned/PI "
          f"{t_l['rhs'] / t_pi['rhs']:.3f}; median error PI {e_pi:.2e} learned {e_l:.2e}")
    assert (status == 0).all()
    assert losses[-50:].mean() < 0.5 * losses[:50].mean()
    assert rho_l < (0.6 if method == "dop853" else 0.7) * rho_pi
    assert t_l["rhs"] < 0.85 * t_pi["rhs"]
    assert e_l <= 1.2 * e_pi
