"""Long-run ensemble statistics: after 50 gyro-periods the energy (gamma)
and pitch-angle distributions of the mixed-precision and learned-h runs
match the FP64 run's (two-sample Kolmogorov-Smirnov, p > 0.01), and in a
field without E the FP64 run keeps gamma: median relative drift below 1e-9
at atol = rtol = 1e-8 and below 1e-11 at 1e-10 (DOP853).  The FP64 run
equals the reference bit for bit (test_gpu_parity), so these are also the
reference's statistics."""
import numpy as np
import pytest
from scipy import stats

from particles_amd import workloads as W

pytestmark = pytest.mark.gpu


def _final(P, f, method, p, tg, st0, net=None, shift=0.0):
    with P.Particles(f, method, p, tg, st0) as b:
        if net is not None:
            b.use_mlp(net, shift)
        b.advance()
        st, _, status = b.download()
    assert (status == 0).all()
    return st[:6, -1, :]


def _gamma_pitch(P, f, s, p):
    b = P.probe(f, "gather", s[:3].T, p)[:, 3:]
    u = s[3:].T
    g = np.sqrt(1.0 + (u * u).sum(1))
    cp = (u * b).sum(1) / (np.linalg.norm(u, axis=1) * np.linalg.norm(b, axis=1))
    return g, cp


def test_mixed_and_learned_distributions_match_fp64():
    import particles_amd as P
    from particles_amd.learned import train_controller
    uu, bb = W.turbulent_vb(32, seed=11, nmodes=128)
    p = W.params_vector(tmax=50 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    tg = W.time_grid(5, p[2])
    st0 = W.initial_state(4000, len(tg), seed=12, gamma_range=(1.0, 100.0))
    with P.Fields(uu, bb) as f64, P.Fields(uu, bb, prec="fp32") as f32:
        s64 = _final(P, f64, "dop853", p, tg, st0)
        s32 = _final(P, f32, "dop853", p, tg, st0)
        ctl = train_controller(f32, "dop853", p, tg,
                               W.initial_state(4000, 2, seed=13, gamma_range=(1.0, 100.0)),
                               fields_fp64=f64, horizon=W.T0, steps=1500, fallback=False)
        sl = _final(P, f32, "dop853", p, tg, st0, ctl.net, ctl.shift)
        ctl.close()
        g64, c64 = _gamma_pitch(P, f64, s64, p)
        for name, s in (("mixed", s32), ("learned", sl)):
            g, c = _gamma_pitch(P, f64, s, p)
            pg = stats.ks_2samp(g64, g).pvalue
            pc = stats.ks_2samp(c64, c).pvalue
            dev = np.median(np.abs(s - s64).max(0))
            print(f"{name}: KS p gamma {pg:.3f}, pitch {pc:.3f}; median state deviation "
                  f"{dev:.2e} after 50 T0")
            assert pg > 0.01 and pc > 0.01


def test_energy_conserved_without_electric_field():
    import particles_amd as P
    _, bb = W.turbulent_vb(32, seed=2, nmodes=128)
    uu = np.zeros_like(bb, order="F")
    with P.Fields(uu, bb) as f:
        for tol, bound in ((1e-8, 1e-9), (1e-10, 1e-11)):
            p = W.params_vector(tmax=10 * W.T0, dtmax=0.5, atol=tol, rtol=tol)
            tg = W.time_grid(10, p[2])
            st0 = W.initial_state(2000, len(tg), seed=5, gamma_range=(1.0, 1e3))
            st, _, status = P.integrate(f, "dop853", p, tg, st0)
            assert (status == 0).all()
            rel = np.abs(st[6, 1:, :] - st[6, 0, :]) / st[6, 0, :]
            print(f"tol {tol:g}: gamma drift median {np.median(rel):.2e} p99 "
                  f"{np.percentile(rel, 99):.2e} max {rel.max():.2e}")
            assert np.median(rel) < bound
