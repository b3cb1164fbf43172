"""The step-size network: the float64 oracle pinned by finite differences and
closed forms (CPU), then the device's BF16 matrix-core forward, gradient and
training against it (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _records(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 20))
    r[:, :16] = rng.normal(0, 1.5 * scale, (n, 16))
    r[:, 16] = rng.normal(0, 0.5, n)
    r[:, 17] = -np.abs(rng.normal(2.0, 1.0, n))
    r[:, 18] = rng.integers(0, 3, n)
    r[:, 19] = 8.0
    return r


# ---------------------------------------------------------------------------
# oracle (CPU)
# ---------------------------------------------------------------------------
def test_bf16_rounding_known_values():
    one = 1.0
    assert M.bf16(one) == 1.0
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0                 # tie -> even (mantissa 0)
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie -> even (up)
    assert M.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert M.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    rel = np.abs(M.bf16(x) - x) / np.abs(x)
    assert rel.max() <= 2.0 ** -8 + 1e-12


def test_gelu_closed_forms():
    assert M.gelu(0.0) == 0.0
    assert abs(M.gelu(10.0) - 10.0) < 1e-12 and abs(M.gelu(-10.0)) < 1e-12
    # tanh form at z = 1: 0.5 (1 + tanh(sqrt(2/pi) 1.044715))
    assert abs(M.gelu(1.0) - 0.5 * (1 + np.tanh(np.sqrt(2 / np.pi) * 1.044715))) < 1e-15
    z = np.linspace(-6, 6, 101)
    num = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.max(np.abs(num - M.gelu_grad(z))) < 1e-8


def test_oracle_gradient_matches_finite_differences():
    p = M.init_params(3)
    r = _records(40, 4)
    for alpha in (0.0, 1.0):
        loss, g = M.gradient(p, r, alpha, emulate=False)
        rng = np.random.default_rng(5)
        idx = np.concatenate([rng.choice(M.NTRAIN, 150, replace=False),
                              [M.W1, M.B1, M.W2, M.B2, M.W3, M.B3, M.B3 - 1]])
        for i in idx:
            e = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += e
            pm[i] -= e
            num = (M.gradient(pp, r, alpha, emulate=False)[0]
                   - M.gradient(pm, r, alpha, emulate=False)[0]) / (2 * e)
            assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)), (alpha, i, num, g[i])


def test_oracle_target_and_loss():
    r = _records(10, 1)
    t0 = M.targets(r, 0.0)
    t1 = M.targets(r, 1.0)
    assert np.allclose(t0, np.clip(r[:, 16], -4, 4))
    assert np.allclose(t1, np.clip(r[:, 16] - r[:, 17] / 8.0, -4, 4))
    assert np.all(t1 >= t0)  # err <= 1 extrapolates to larger steps
    p = np.zeros(M.NPARAMS)
    p[M.SIG:] = 1.0
    p[M.B3] = 0.25
    loss, g = M.gradient(p, r, 0.0, emulate=False)
    assert abs(loss - np.mean((0.25 - t0) ** 2)) < 1e-15
    assert abs(g[M.B3] - 2 * np.mean(0.25 - t0)) < 1e-15


def test_emulation_is_close_to_float64():
    p = M.init_params(1)
    f = _records(200, 2)[:, :16]
    y64 = M.forward(p, f, emulate=False)
    y16 = M.forward(p, f, emulate=True)
    assert 1e-5 < np.max(np.abs(y64 - y16)) < 0.2


# ---------------------------------------------------------------------------
# device (GPU)
# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_gemm_operand_layout():
    import particles_amd as P
    rng = np.random.default_rng(0)
    a = rng.normal(size=(64, 64)).astype(np.float32)
    bt = rng.normal(size=(64, 64)).astype(np.float32)
    c = P.probe_gemm(a, bt)
    want = M.bf16(a) @ M.bf16(bt).T
    assert np.max(np.abs(c - want)) < 1e-4 * np.max(np.abs(want))
    # a row/column permutation would not pass: every entry is checked
    assert np.allclose(c, want, rtol=1e-5, atol=1e-4)


@pytest.mark.gpu
def test_forward_matches_oracle():
    import particles_amd as P
    p = M.init_params(7)
    f = _records(4096 + 17, 8)[:, :16]
    with P.Mlp(p.astype(np.float32)) as net:
        y = net.forward(f)
    p32 = p.astype(np.float32).astype(np.float64)
    want = M.forward(p32, f.astype(np.float32), emulate=True)
    exact = M.forward(p32, f.astype(np.float32), emulate=False)
    d = np.abs(y - want)
    print(f"forward: median {np.median(d):.2e} max {d.max():.2e}; "
          f"vs unrounded float64 max {np.max(np.abs(y - exact)):.2e}")
    assert np.median(d) < 1e-5
    assert d.max() < 1e-2
    assert np.max(np.abs(y - exact)) > 10 * np.median(d)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 10, 4096 + 17])
def test_group_forward_matches_wave_forward(n):
    """The forward pass of the six-lanes-per-particle integrator (ten
    samples per wavefront in one 32-sample tile) gives the per-wavefront
    forward's outputs exactly, for whole and ragged last waves."""
    import particles_amd as P
    p = M.init_params(7)
    f = _records(n, 8)[:, :16]
    with P.Mlp(p.astype(np.float32)) as net:
        y = net.forward(f)
        yg = net.forward(f, groups=True)
    want = M.forward(p.astype(np.float32).astype(np.float64), f.astype(np.float32), emulate=True)
    assert np.array_equal(yg, y)
    assert np.max(np.abs(yg - want)) < 1e-2


@pytest.mark.gpu
def test_gradient_matches_oracle():
    import particles_amd as P
    p = M.init_params(11).astype(np.float32)
    r = _records(3000, 12).astype(np.float32)
    with P.Mlp(p) as net:
        for alpha in (0.0, 1.0):
            loss, g = net.gradient(r, alpha)
            want_loss, want = M.gradient(p.astype(np.float64), r.astype(np.float64), alpha)
            assert abs(loss - want_loss) < 1e-4 * want_loss
            rel = np.linalg.norm(g - want) / np.linalg.norm(want)
            print(f"alpha {alpha}: loss {loss:.6f} vs {want_loss:.6f}; gradient rel err {rel:.2e}")
            assert rel < 2e-3
            for name, lo, hi in (("w1", M.W1, M.B1), ("b1", M.B1, M.W2), ("w2", M.W2, M.B2),
                                 ("b2", M.B2, M.W3), ("w3", M.W3, M.B3), ("b3", M.B3, M.MU)):
                e = np.linalg.norm(g[lo:hi] - want[lo:hi]) / np.linalg.norm(want[lo:hi])
                assert e < 1e-2, (name, e)


@pytest.mark.gpu
def test_training_fits_a_smooth_target():
    import particles_amd as P
    rng = np.random.default_rng(21)
    n = 40000
    r = np.zeros((n, 20), dtype=np.float32)
    r[:, :16] = rng.normal(0, 1, (n, 16)) * np.linspace(0.5, 3, 16) + np.linspace(-2, 2, 16)
    z = (r[:, :16] - np.linspace(-2, 2, 16)) / np.linspace(0.5, 3, 16)
    r[:, 16] = 0.6 * np.tanh(z[:, 0]) - 0.4 * z[:, 13] * 0.3 + 0.3 * np.sin(2 * z[:, 9])
    r[:, 19] = 8.0
    with P.Mlp(seed=3) as net:
        net.normalize(r)
        params = net.params
        assert np.allclose(params[M.MU:M.SIG], r[:, :16].mean(0), atol=1e-3)
        assert np.allclose(params[M.SIG:], r[:, :16].std(0), rtol=1e-3)
        losses = net.train(r, steps=600, batch=4096, lr=3e-3, seed=5)
        test = np.zeros((2000, 20), dtype=np.float32)
        test[:, :16] = rng.normal(0, 1, (2000, 16)) * np.linspace(0.5, 3, 16) + np.linspace(-2, 2, 16)
        zt = (test[:, :16] - np.linspace(-2, 2, 16)) / np.linspace(0.5, 3, 16)
        yt = 0.6 * np.tanh(zt[:, 0]) - 0.12 * zt[:, 13] + 0.3 * np.sin(2 * zt[:, 9])
        pred = net.forward(test[:, :16])
        # the trained parameters evaluated by the oracle give the same outputs
        want = M.forward(net.params.astype(np.float64), test[:, :16], emulate=True)
    mse = float(np.mean((pred - yt) ** 2))
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.5f}; held-out mse {mse:.5f} (target var "
          f"{np.var(yt):.4f})")
    assert losses[-1] < 0.05 * losses[0]
    assert mse < 0.05 * np.var(yt)
    assert np.median(np.abs(pred - want)) < 1e-4


@pytest.mark.gpu
def test_training_is_deterministic():
    import particles_amd as P
    r = _records(5000, 30).astype(np.float32)
    out = []
    for _ in range(2):
        with P.Mlp(seed=9) as net:
            net.normalize(r)
            net.train(r, steps=20, batch=2048, lr=1e-3, seed=4)
            out.append(net.params)
    assert np.array_equal(out[0], out[1])
