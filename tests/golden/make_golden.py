"""Writes the golden vectors in this folder from the reference's own
hot-path modules (interpolations.F90, accelerations.F90, integrations.F90 of
gkowal/particles), compiled from its sources into oracle/_ref by
``make -C oracle ref``.  Run in the build container only (the reference is
not present elsewhere):

    python tests/golden/make_golden.py

Each case file holds inputs (fields, params, time grid, initial state) and the
reference's outputs (state snapshots, step counters, single-call probes).
The reference has no tests or fixtures of its own; these vectors are the
pinned record of its behaviour that the oracle and the device path are
checked against.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Reference  # noqa: E402
from particles_amd import workloads as W  # noqa: E402


def turbulent(dm, seed, brms=0.5, vrms=0.1, b0=1.0, nmodes=48):
    mb = W.random_modes(seed, nmodes=nmodes, kmax=3.0, rms=brms, stream=1)
    mv = W.random_modes(seed, nmodes=nmodes, kmax=3.0, rms=vrms, stream=2)
    bb = W.eval_modes(mb, dm)
    bb[..., 2] += b0
    uu = W.eval_modes(mv, dm)
    return uu, bb


def integration_case(ref, name, method, dm, seed, np_, params, tg, gamma_range):
    uu, bb = turbulent(dm, seed)
    xl, xs = params[11:14], params[14:17]
    st0 = W.initial_state(np_, len(tg), seed + 100, gamma_range=gamma_range, xl=xl, xs=xs)
    st, cnt = ref.integrate(method, uu, bb, params, tg, st0)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), method=np.array(method), uu=uu, bb=bb,
                        params=params, time=tg, state0=st0, state=st, counter=cnt)
    print(f"{name}: {method} dm={dm} np={np_} accepted={cnt[1].sum()} rejected={cnt[2].sum()}")


def probe_case(ref, name, dm, seed):
    """Single interpolate_lin / acceleration / dti_guess calls, including
    positions on cell faces, on periodic images far from the box and at
    negative coordinates, where the reference's separate right-index floor
    (interpolations.F90:141-143) matters."""
    uu, bb = turbulent(dm, seed)
    rng = np.random.Generator(np.random.Philox(key=seed))
    xl = np.array([-0.5, 0.25, 1.0])
    xs = np.array([1.0, 2.0, 0.75])
    pts = [rng.uniform(-5, 5, size=3) for _ in range(200)]
    # exact nodes and faces, and one ulp either side
    for _ in range(100):
        i = rng.integers(-3 * max(dm), 3 * max(dm), size=3)
        p = xl + xs * i / np.asarray(dm)
        pts.append(p)
        pts.append(np.nextafter(p, np.inf))
        pts.append(np.nextafter(p, -np.inf))
    pos = np.array(pts)
    u = np.zeros_like(pos); b = np.zeros_like(pos)
    for k, p in enumerate(pos):
        u[k], b[k] = ref.interpolate_lin(uu, bb, p, xl=xl, xs=xs)
    vec = np.concatenate([pos, rng.normal(size=pos.shape) * np.exp(rng.uniform(0, 7, size=(len(pos), 1)))], 1)
    acc = np.array([ref.acceleration(uu, bb, v, qom=1.7, vun=0.9, xl=xl, xs=xs) for v in vec])
    params = W.params_vector(qom=1.7, vun=0.9, dtmax=0.3, atol=1e-7, rtol=1e-6, xl=xl, xs=xs)
    dti5 = np.array([ref.dti_guess(5, uu, bb, params, v) for v in vec])
    dti8 = np.array([ref.dti_guess(8, uu, bb, params, v) for v in vec])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), uu=uu, bb=bb, xl=xl, xs=xs, pos=pos,
                        u=u, b=b, vec=vec, acc=acc, params=params, dti5=dti5, dti8=dti8)
    print(f"{name}: {len(pos)} probes")


def main():
    ref = Reference()
    probe_case(ref, "probe_dm9x12x8", (9, 12, 8), seed=11)
    p = W.params_vector(tmax=2 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
    integration_case(ref, "rk5_tol1e-8", "rk5", (12, 12, 12), 21, 24, p, W.time_grid(4, p[2]),
                     (1.0, 100.0))
    integration_case(ref, "dop853_tol1e-8", "dop853", (12, 12, 12), 22, 24, p,
                     W.time_grid(4, p[2]), (1.0, 100.0))
    # reference defaults for the controller, non-unit units and domain,
    # logarithmic snapshot grid, PI controller on (beta > 0)
    p2 = W.params_vector(qom=2.5, vun=0.8, tmax=3.0, dtmax=0.05, atol=1e-6, rtol=1e-6, beta=0.04,
                         xl=(-1.0, 0.5, 2.0), xs=(2.0, 1.5, 0.5))
    integration_case(ref, "rk5_defaults_log", "rk5", (8, 10, 6), 23, 16, p2,
                     W.time_grid(3, p2[2], "log", tmin=1e-2), (1.0, 20.0))
    integration_case(ref, "dop853_pi_log", "dop853", (8, 10, 6), 24, 16, p2,
                     W.time_grid(3, p2[2], "log", tmin=1e-2), (1.0, 20.0))


if __name__ == "__main__":
    main()
