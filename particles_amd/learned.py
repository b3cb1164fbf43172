"""Training workflow of the learned step-size proposal.

The reference sizes every step with its PI controller
(sources/integrations.F90:177,186 for RK5, :466-490 for DOP853).  Here the
PI controller stays the fallback (after every rejection) and the floor of
every proposal, and the embedded error estimate still accepts or rejects
each step:

  1. ``record``: the PI-controlled integrator runs a training particle set
     and writes (features, next accepted step, error ratio) records on the
     device (``Particles.record``);
  2. ``train``: the network (16-64-64-1, BF16 matrix cores) is normalised to
     the records' feature statistics and trained on the GPU by the library's
     own forward/backward kernels, reading the records where they lie;
  3. ``calibrate``: the network's log-step offset (``shift``) is chosen on the
     training particles: among the candidates, the fastest one (integrator
     kernel time) whose error against a tight-tolerance FP64 run is no larger
     than the PI controller's own error on the same particles.  The result
     says whether that candidate also ran faster than PI (``beats_pi``); where
     it does not, the result is the PI controller alone (``net`` None, the
     reason in ``info["fallback"]``).

Everything runs through the C-ABI; nothing here is on the integration path.
"""
from __future__ import annotations

import time as _time
from dataclasses import dataclass, field

import numpy as np

from . import Fields, Mlp, Particles, integrate


@dataclass
class Controller:
    """A calibrated step-size controller.  ``net`` is None when the PI
    controller is the better choice for the workload (the network ran no
    faster than PI at no larger error, or it never reached PI's error):
    ``Particles.use_mlp(None)`` then keeps the PI rule; ``info["fallback"]``
    says why."""
    net: Mlp | None
    shift: float
    beats_pi: bool = False  # the calibrated shift ran faster than PI at no larger error
    info: dict = field(default_factory=dict)

    def close(self):
        if self.net is not None:
            self.net.close()
            self.net = None


def choose_shift(cand, pi_ms: float, pi_error: float):
    """The calibration's decision, from the candidates' measured kernel time
    and error (dicts with "shift", "kernel_ms", "error") and PI's on the same
    particles: (chosen candidate or None, calibrated, beats_pi, reason).
    A candidate qualifies when its error is no larger than PI's; the fastest
    qualifying one is chosen when it is also faster than PI, else PI is."""
    ok = [c for c in cand if c["error"] <= pi_error]
    if not ok:
        return None, False, False, (f"no shift reached the PI controller's error "
                                    f"({pi_error:.3g}); the lowest was "
                                    f"{min(c['error'] for c in cand):.3g}")
    best = min(ok, key=lambda c: c["kernel_ms"])
    if not best["kernel_ms"] < pi_ms:
        return None, True, False, (f"the fastest shift at no larger error ({best['shift']:g}, "
                                   f"{best['kernel_ms']:.2f} ms) ran no faster than the PI "
                                   f"controller ({pi_ms:.2f} ms)")
    return best, True, True, ""


def _final_error(state, ref):
    d = np.abs(state[:6, -1, :] - ref[:6, -1, :]).max(0)
    return float(np.median(d))


def record_and_train(fields: Fields, method: str, params, state_train, *,
                     horizon: float | None = None, steps: int = 3000, batch: int = 16384,
                     lr: float = 2e-3, alpha: float = 1.0, capacity: int = 60_000_000,
                     seed: int = 1):
    """Steps 1 and 2: records of the PI controller on `fields` (FP32 node
    rows: record mode runs one lane per particle on node rows) over
    [0, horizon], then the network normalised to them and trained on the
    GPU.  Returns (network, info)."""
    params = np.asarray(params, dtype=np.float64)
    q = params.copy()
    if horizon is not None:
        q[2] = horizon
    tg = np.array([0.0, q[2]])
    st = np.asfortranarray(np.asarray(state_train, dtype=np.float64)[:, :1, :].repeat(2, axis=1))
    info = {}
    t0 = _time.perf_counter()
    net = Mlp(seed=seed)
    with Particles(fields, method, q, tg, st) as b:
        b.record(1.0, capacity)
        b.advance()
        recs = b.records(device=True)
        info["records"] = len(recs)
        net.normalize(recs)
        losses = net.train(recs, steps=steps, batch=batch, lr=lr, alpha=alpha, seed=seed + 6)
    info.update(loss_first=float(losses[:50].mean()), loss_last=float(losses[-50:].mean()),
                train_steps=steps, batch=batch, alpha=alpha,
                train_seconds=_time.perf_counter() - t0)
    return net, info


def reference_states(fields_fp64: Fields, method: str, params, state, horizon: float | None = None):
    """The FP64 states at atol = rtol = 1e-12 over [0, horizon] from
    state[:, 0, :], the calibration's accuracy reference."""
    q = np.asarray(params, dtype=np.float64).copy()
    if horizon is not None:
        q[2] = horizon
    q[5] = q[6] = 1e-12
    tg = np.array([0.0, q[2]])
    st = np.asfortranarray(np.asarray(state, dtype=np.float64)[:, :1, :].repeat(2, axis=1))
    return integrate(fields_fp64, method, q, tg, st)[0]


def calibrate(net: Mlp, fields: Fields, method: str, params, state_cal, ref, *,
              horizon: float | None = None,
              shifts=(-0.45, -0.4, -0.35, -0.3, -0.25, -0.2, -0.15, -0.1, 0.0, 0.1),
              fallback: bool = True, info: dict | None = None) -> Controller:
    """Step 3: each candidate shift of the network's log step on `fields`
    (the precision and layout it will run with) against the PI controller,
    both timed and compared with `ref` (reference_states of the same
    particles); choose_shift decides.  Where the network does not beat PI,
    the result is the PI controller (net None, closed here) unless
    `fallback` is False, which keeps the network at the shift closest to
    qualifying (for measurements)."""
    info = dict(info or {})
    q = np.asarray(params, dtype=np.float64).copy()
    if horizon is not None:
        q[2] = horizon
    tg = np.array([0.0, q[2]])
    st = np.asfortranarray(np.asarray(state_cal, dtype=np.float64)[:, :1, :].repeat(2, axis=1))
    with Particles(fields, method, q, tg, st) as b:
        ms_pi = b.advance()
        e_pi = _final_error(b.download()[0], ref)
    cand = []
    for s in shifts:
        with Particles(fields, method, q, tg, st) as b:
            b.use_mlp(net, s)
            ms = b.advance()
            e = _final_error(b.download()[0], ref)
        cand.append({"shift": float(s), "kernel_ms": ms, "error": e})
    best, calibrated, beats, why = choose_shift(cand, ms_pi, e_pi)
    info.update(calibration=cand, pi_kernel_ms=ms_pi, pi_error=e_pi,
                calibrated=calibrated, beats_pi=beats, particles=int(st.shape[2]),
                horizon=float(q[2]))
    if best is not None:
        return Controller(net=net, shift=best["shift"], beats_pi=True, info=info)
    if fallback:
        net.close()
        info["fallback"] = "PI controller: " + why
        return Controller(net=None, shift=0.0, beats_pi=False, info=info)
    ok = [c for c in cand if c["error"] <= e_pi]
    near = min(ok, key=lambda c: c["kernel_ms"]) if ok else min(cand, key=lambda c: c["error"])
    info["not_beating_pi"] = why
    return Controller(net=net, shift=near["shift"], beats_pi=False, info=info)


def train_controller(fields: Fields, method: str, params, time, state_train, *,
                     fields_fp64: Fields | None = None, horizon: float | None = None,
                     steps: int = 3000, batch: int = 16384, lr: float = 2e-3, alpha: float = 1.0,
                     shifts=(-0.45, -0.4, -0.35, -0.3, -0.25, -0.2, -0.15, -0.1, 0.0, 0.1),
                     capacity: int = 60_000_000,
                     seed: int = 1, fallback: bool = True) -> Controller:
    """Records, trains and calibrates a learned controller for `fields` (the
    precision it will run with, in node rows).  `state_train` is (8, m, n)
    like every state; records are taken over [0, horizon] (default
    params[2]).  The calibration runs the same particles over the same
    horizon, with `fields_fp64` (default: `fields`) at atol = rtol = 1e-12
    as reference (calibrate)."""
    net, info = record_and_train(fields, method, params, state_train, horizon=horizon,
                                 steps=steps, batch=batch, lr=lr, alpha=alpha,
                                 capacity=capacity, seed=seed)
    ref = reference_states(fields_fp64 if fields_fp64 is not None else fields, method, params,
                           state_train, horizon)
    return calibrate(net, fields, method, params, state_train, ref, horizon=horizon,
                     shifts=shifts, fallback=fallback, info=info)
