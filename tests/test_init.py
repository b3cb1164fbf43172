"""Device initial states: the Philox4x32-10 restatement against the
published known-answer vectors (CPU), and the device generator against the
restatement and its distributions (GPU)."""
import numpy as np
import pytest

from oracle import init_oracle as I
from particles_amd import workloads as W


def test_philox_known_answers():
    """Known-answer vectors of Philox4x32-10 from the Random123 distribution
    (kat_vectors): zero counter/key, all-ones, and the pi-digit case."""
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        got = I.philox4x32_10(np.array([ctr], np.uint32), np.array([key], np.uint32))[0]
        assert tuple(int(x) for x in got) == want


def test_uniforms_cover_the_interval():
    d = I.draws(100000, 12345)
    assert d.min() >= 0.0 and d.max() < 1.0
    assert np.allclose(d.mean(0), 0.5, atol=0.005)
    assert np.allclose(d.var(0), 1 / 12, atol=0.002)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,a,b", [("maxwell", 0.01, 1.0), ("gamma", 10.0, 1.0),
                                      ("gamma_log", 1.0, 1.0e3)])
def test_device_generator_matches_restatement(kind, a, b):
    import particles_amd as P
    n = 20001
    uu, bb = W.turbulent_vb(8, seed=1, nmodes=8)
    p = W.params_vector(tmax=1.0, dtmax=1.0)
    tg = W.time_grid(1, 1.0)
    lo, hi = (-0.5, 0.25, 1.0), (0.5, 0.75, 3.0)
    with P.Fields(uu, bb) as f, P.Particles(f, "rk5", p, tg, n) as batch:
        batch.generate(kind, 987654321, lo, hi, a, b)
        st = batch.download()[0]
        mom = batch.moments()
    want = I.initial_state(n, 987654321, kind, lo, hi, a, b)
    got = st[:6, 0, :]
    assert np.array_equal(got[:3], want[:3])          # positions: exact
    # momenta: the device's and NumPy's log/exp/cos may differ by an ulp or
    # two, relative to the particle's |u|
    umag = np.sqrt((want[3:] ** 2).sum(0))
    ratio = np.abs(got[3:] - want[3:]).max(0) / np.maximum(umag, 1e-300)
    k = int(np.argmax(ratio))
    print(f"{kind}: worst momentum difference {ratio[k]:.2e} of |u| at particle {k}: "
          f"{got[3:, k]} vs {want[3:, k]}, gamma {np.sqrt(1 + umag[k] ** 2)}")
    # |u| = sqrt(gamma^2 - 1) amplifies a one-ulp difference in gamma by
    # gamma^2 / (gamma^2 - 1) near gamma = 1
    g2 = 1.0 + umag ** 2
    assert np.all(ratio <= 1e-14 * (1.0 + g2 / np.maximum(g2 - 1.0, 1e-300)))
    # moments as the reference defines them, against NumPy on the same data
    for c in range(6):
        x = got[c]
        m, s = x.mean(), x.std()
        z = (x - m) / s
        assert np.allclose(mom[c], [m, s, (z ** 3).mean(), (z ** 4).mean() - 3.0],
                           rtol=1e-9, atol=1e-12)
    if kind == "maxwell":
        assert np.allclose(mom[3:, 1], a, rtol=0.02) and np.all(np.abs(mom[3:, 2]) < 0.05)
        assert np.all(np.abs(mom[3:, 3]) < 0.1)
    else:
        g = np.sqrt(1.0 + (got[3:] ** 2).sum(0))
        if kind == "gamma":
            assert np.allclose(g, a, rtol=1e-12)
        else:
            lg = np.log(g)
            assert abs(lg.mean() - 0.5 * np.log(b)) < 0.05 and lg.min() >= -1e-12
        # isotropy: direction components uniform in [-1, 1]
        uz = got[5] / np.sqrt((got[3:] ** 2).sum(0))
        assert abs(uz.mean()) < 0.02 and abs((uz ** 2).mean() - 1 / 3) < 0.02
    assert np.all(got[0] >= lo[0]) and np.all(got[0] < hi[0])


@pytest.mark.gpu
def test_generate_only_before_advance():
    import particles_amd as P
    uu, bb = W.turbulent_vb(8, seed=1, nmodes=8)
    p = W.params_vector(tmax=0.1, dtmax=1.0)
    tg = W.time_grid(1, 0.1)
    with P.Fields(uu, bb) as f, P.Particles(f, "rk5", p, tg, 16) as batch:
        batch.generate("gamma", 1, a=2.0)
        batch.advance()
        with pytest.raises(P.PtcError):
            batch.generate("gamma", 1, a=2.0)
