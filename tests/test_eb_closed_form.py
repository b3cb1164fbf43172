"""EB field mode and the reconnection recipe, pinned by closed forms.

The reference has one force, q/m (v - U) x B with the plasma velocity U
stored (sources/accelerations.F90:84-92).  The EB mode stores E instead and
evaluates q/m (E + v x B); the reconnection workload builds
E = -V x B + eta curl B on the grid.  Neither has a reference counterpart,
so the checks below fix them by what the physics and the mathematics fix:

* with E = -U x B in uniform fields, the EB force equals the reference's VB
  force (the compiled reference where it is built, else the oracle, which
  the golden vectors pin to it), and whole trajectories agree;
* a particle moving at exactly E x B / B^2 feels no force (straight line);
* with E = 0 the orbit is the closed-form gyration (sense of rotation and
  radius); with E perpendicular to B the guiding centre drifts at
  E x B / B^2;
* the central-difference curl equals its closed form on trigonometric
  fields, and the recipe's E at the nodes equals -V x B + eta curl B built
  independently from the mode table (direct sums, not the FFT).

The device kernels are compared bitwise with the oracle in EB mode by the
GPU tests, so pinning the oracle pins them.  The last test compiles copies
of the oracle with a planted mistake in the EB force (wrong sign of v x B,
two components swapped, wrong sign of E) and shows that the checks catch
each one.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from particles_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def uniform(dm, v):
    return np.asfortranarray(np.broadcast_to(np.asarray(v, float), tuple(dm) + (3,)).copy())


# ---------------------------------------------------------------------------
# checks (each raises AssertionError when the oracle is wrong)
# ---------------------------------------------------------------------------

def check_rhs_equals_vb(oracle, vb_rhs):
    """E = -U x B in uniform fields: q/m (E + v x B) = q/m (v - U) x B to
    rounding, for random U, B, momenta, q/m and velocity unit."""
    rng = np.random.default_rng(11)
    dm = (4, 5, 3)
    for _ in range(40):
        u_pl = rng.normal(size=3) * 0.3
        b = rng.normal(size=3)
        uu, bb = uniform(dm, u_pl), uniform(dm, b)
        ee = uniform(dm, -np.cross(u_pl, b))
        g = math.exp(rng.uniform(0.0, math.log(1e3)))
        d = rng.normal(size=3)
        vec = np.concatenate([rng.uniform(-2, 2, size=3), d / np.linalg.norm(d) * math.sqrt(g * g - 1)])
        qom, vun = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        a_eb = oracle.acceleration(ee, bb, vec, qom=qom, vun=vun, mode="eb")
        a_vb = vb_rhs(uu, bb, vec, qom, vun)
        scale = qom * (1.0 + np.linalg.norm(u_pl)) * np.linalg.norm(b)
        assert np.array_equal(a_eb[:3], a_vb[:3])
        assert np.abs(a_eb[3:] - a_vb[3:]).max() <= 1e-14 * scale, (a_eb, a_vb)


def check_trajectory_equals_vb(oracle, method):
    """Whole DOP853 / RK5 trajectories in E = -U x B and in (U, B) agree to
    the rounding of the two force forms (uniform fields: no chaos)."""
    dm = (4, 4, 4)
    u_pl, b = np.array([0.03, -0.01, 0.02]), np.array([0.2, 0.1, 1.1])
    uu, bb = uniform(dm, u_pl), uniform(dm, b)
    ee = uniform(dm, -np.cross(u_pl, b))
    p = W.params_vector(tmax=3 * W.T0, dtmax=0.5, atol=1e-11, rtol=1e-11)
    tg = W.time_grid(6, p[2])
    st0 = W.initial_state(6, len(tg), seed=3, gamma_range=(1.0, 20.0))
    s_eb, _, st_eb = oracle.integrate(method, ee, bb, p, tg, st0, mode="eb")
    s_vb, _, st_vb = oracle.integrate(method, uu, bb, p, tg, st0, mode="vb")
    assert (st_eb == 0).all() and (st_vb == 0).all()
    d = np.abs(s_eb[:6] - s_vb[:6]) / np.maximum(np.abs(s_vb[:6]), 1.0)
    assert d.max() < 1e-8, d.max()


def check_drift_frame_straight_line(oracle, method="dop853"):
    """E = (0, e0, 0), B = (0, 0, b0): a particle with v = E x B / B^2 =
    (e0 / b0) x feels E + v x B = 0 and moves on a straight line with
    constant momentum."""
    dm = (4, 4, 4)
    b0, e0 = 1.3, 0.026
    vd = e0 / b0
    ee, bb = uniform(dm, (0.0, e0, 0.0)), uniform(dm, (0.0, 0.0, b0))
    p = W.params_vector(tmax=40.0, dtmax=1.0, atol=1e-10, rtol=1e-10)
    tg = W.time_grid(8, p[2])
    st0 = np.zeros((8, len(tg), 1), order="F")
    gd = 1.0 / math.sqrt(1.0 - vd * vd)
    st0[:6, 0, 0] = [0.1, 0.2, 0.3, gd * vd, 0.0, 0.0]
    st, _, status = oracle.integrate(method, ee, bb, p, tg, st0, mode="eb")
    assert status[0] == 0
    np.testing.assert_allclose(st[0, :, 0], 0.1 + vd * tg, rtol=0, atol=1e-12)
    np.testing.assert_allclose(st[1, :, 0], 0.2, rtol=0, atol=1e-12)
    np.testing.assert_allclose(st[3, :, 0], gd * vd, rtol=1e-13)
    np.testing.assert_allclose(st[4:6, :, 0], 0.0, rtol=0, atol=1e-13)


def check_gyration(oracle, method):
    """E = 0, B = b0 z: u rotates clockwise about +z (positive charge) at
    Omega = q/m b0 / gamma on a circle of radius |u_perp| / (q/m b0)."""
    b0, qom, gam = 1.3, 1.0, 5.0
    dm = (4, 4, 4)
    ee, bb = uniform(dm, (0, 0, 0)), uniform(dm, (0, 0, b0))
    om = qom * b0 / gam
    p = W.params_vector(qom=qom, tmax=2 * 2 * math.pi / om, dtmax=0.02 / om, atol=1e-12,
                        rtol=1e-12)
    tg = W.time_grid(10, p[2])
    up = math.sqrt(gam ** 2 - 1)
    st0 = np.zeros((8, len(tg), 1), order="F")
    st0[:6, 0, 0] = [0.2, 0.3, 0.4, up, 0.0, 0.0]
    st, _, status = oracle.integrate(method, ee, bb, p, tg, st0, mode="eb")
    assert status[0] == 0
    t = tg[1:]
    r = up / (qom * b0)
    tol = 1e-4 if method == "rk5" else 1e-8  # dense-output accuracy (test_oracle)
    np.testing.assert_allclose(st[3, 1:, 0], up * np.cos(om * t), atol=tol * up)
    np.testing.assert_allclose(st[4, 1:, 0], -up * np.sin(om * t), atol=tol * up)
    np.testing.assert_allclose(st[0, 1:, 0], 0.2 + r * np.sin(om * t), atol=tol * r)
    np.testing.assert_allclose(st[1, 1:, 0], 0.3 - r * (1 - np.cos(om * t)), atol=tol * r)


def check_exb_drift(oracle, method):
    """E = (0, e0, 0) perpendicular to B = (0, 0, b0): starting at rest the
    guiding centre drifts at E x B / B^2 = (e0 / b0) x; over whole gyrations
    the mean velocity is that drift, with no net motion along y or z."""
    b0, e0 = 1.0, 0.02
    vd = e0 / b0
    dm = (4, 4, 4)
    ee, bb = uniform(dm, (0.0, e0, 0.0)), uniform(dm, (0.0, 0.0, b0))
    # gyro-period at gamma ~ 1 + O(vd^2): 2 pi / b0; 25 of them
    p = W.params_vector(tmax=25 * 2 * math.pi / b0, dtmax=0.5, atol=1e-11, rtol=1e-11)
    tg = W.time_grid(25, p[2])
    st0 = np.zeros((8, len(tg), 1), order="F")
    st0[:3, 0, 0] = [0.5, 0.5, 0.5]
    st, _, status = oracle.integrate(method, ee, bb, p, tg, st0, mode="eb")
    assert status[0] == 0
    dx = st[0, -1, 0] - st[0, 0, 0]
    assert abs(dx / p[2] - vd) < 0.01 * vd, dx / p[2]
    # the cycloid's y excursion is 2 vd / Omega; it never drifts along E
    assert -1e-9 <= st[1, :, 0].min() - 0.5 and st[1, :, 0].max() - 0.5 <= 2.01 * vd / b0
    assert np.abs(st[2, :, 0] - 0.5).max() < 1e-12


# ---------------------------------------------------------------------------
# the checks on the oracle
# ---------------------------------------------------------------------------

def _vb_rhs(oracle, reference):
    if reference is not None:
        return lambda uu, bb, vec, qom, vun: reference.acceleration(uu, bb, vec, qom=qom, vun=vun)
    return lambda uu, bb, vec, qom, vun: oracle.acceleration(uu, bb, vec, qom=qom, vun=vun)


@pytest.fixture(scope="module")
def reference_or_none():
    from oracle.oracle import REF_SO, Reference
    return Reference() if os.path.exists(REF_SO) else None


def test_eb_rhs_equals_reference_vb_with_motional_field(oracle, reference_or_none):
    check_rhs_equals_vb(oracle, _vb_rhs(oracle, reference_or_none))


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_eb_trajectory_equals_vb_with_motional_field(oracle, method):
    check_trajectory_equals_vb(oracle, method)


def test_eb_drift_frame_straight_line(oracle):
    # DOP853 only: with momentum components that stay zero to rounding,
    # Cash-Karp's atol-only error norm relative to |si| + |k1|
    # (integrations.F90:147-148) divides rounding noise by ~tiny(1d0) and
    # never settles on a step
    check_drift_frame_straight_line(oracle, "dop853")


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_eb_gyration_closed_form(oracle, method):
    check_gyration(oracle, method)


@pytest.mark.parametrize("method", ["rk5", "dop853"])
def test_eb_exb_drift(oracle, method):
    check_exb_drift(oracle, method)


# ---------------------------------------------------------------------------
# the recipe: central-difference curl and E = -V x B + eta J
# ---------------------------------------------------------------------------

def test_curl_central_trig_closed_form():
    """On f = (sin 2 pi (a y + b z), cos 2 pi (c z + d x), sin 2 pi (e x + g y))
    with integer wave numbers the periodic central difference of
    sin(2 pi k s) along s with spacing h is cos(2 pi k s) sin(2 pi k h) / h;
    the curl assembled from those closed forms equals curl_central."""
    dm = (16, 12, 10)
    h = tuple(1.0 / d for d in dm)
    a, b, c, d, e, g = 2, 1, 3, 1, 2, 3
    x, y, z = np.meshgrid(*(np.arange(n) / n for n in dm), indexing="ij")
    tp = 2 * math.pi
    f = np.empty(dm + (3,), order="F")
    f[..., 0] = np.sin(tp * (a * y + b * z))
    f[..., 1] = np.cos(tp * (c * z + d * x))
    f[..., 2] = np.sin(tp * (e * x + g * y))

    def dsin(k, hh):  # factor of the central difference of sin / cos
        return math.sin(tp * k * hh) / hh

    ph0, ph1, ph2 = tp * (a * y + b * z), tp * (c * z + d * x), tp * (e * x + g * y)
    d2_dy = np.cos(ph2) * dsin(g, h[1])
    d1_dz = -np.sin(ph1) * dsin(c, h[2])
    d0_dz = np.cos(ph0) * dsin(b, h[2])
    d2_dx = np.cos(ph2) * dsin(e, h[0])
    d1_dx = -np.sin(ph1) * dsin(d, h[0])
    d0_dy = np.cos(ph0) * dsin(a, h[1])
    want = np.stack([d2_dy - d1_dz, d0_dz - d2_dx, d1_dx - d0_dy], -1)
    got = W.curl_central(f, h=h)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_reconnection_e_equals_independent_construction():
    """workloads.reconnection_eb's E at every node equals -V x B + eta J
    assembled here from the mode table by direct sums: B from
    eval_modes_at plus the sheet and guide field, J from the closed-form
    central differences of every mode and of the sheet profile, V x B by
    numpy's cross product."""
    dm, seed, delta, guide, vin, eta, nm = 12, 7, 0.05, 0.1, 0.1, 0.3, 24
    ee, bb = W.reconnection_eb(dm, seed=seed, delta=delta, guide=guide, noise_rms=0.05, vin=vin,
                               eta=eta, nmodes=nm)
    modes = W.random_modes(seed, nmodes=nm, rms=0.05, stream=1)
    h = 1.0 / dm
    idx = np.stack(np.meshgrid(*(np.arange(dm),) * 3, indexing="ij"), -1).reshape(-1, 3)
    xs = idx / dm
    ph = 2 * math.pi * (xs @ modes.n.T) + modes.phase[None, :]
    bnode = (np.cos(ph) * modes.amp) @ modes.pol
    prof = W.sheet_profile(xs[:, 1], delta)
    bnode[:, 0] += prof
    bnode[:, 2] += guide
    # d/ds_a of amp pol cos(phase) by central differences: -amp pol sin(phase) sin(2 pi n_a h) / h
    grad = np.empty((len(xs), 3, 3))  # grad[:, component, axis]
    for ax in range(3):
        fac = modes.amp * np.sin(2 * math.pi * modes.n[:, ax] * h) / h
        grad[:, :, ax] = -(np.sin(ph) * fac) @ modes.pol
    # the profile is sampled at the nodes of the periodic grid, so the
    # neighbours of the edge nodes are the nodes across the boundary
    dprof = (W.sheet_profile((xs[:, 1] + h) % 1.0, delta) -
             W.sheet_profile((xs[:, 1] - h) % 1.0, delta)) / (2 * h)
    grad[:, 0, 1] += dprof
    jj = np.stack([grad[:, 2, 1] - grad[:, 1, 2], grad[:, 0, 2] - grad[:, 2, 0],
                   grad[:, 1, 0] - grad[:, 0, 1]], -1)
    vv = np.zeros_like(bnode)
    vv[:, 1] = -vin * prof
    enode = -np.cross(vv, bnode) + eta * jj
    got_b = bb[idx[:, 0], idx[:, 1], idx[:, 2]]
    got_e = ee[idx[:, 0], idx[:, 1], idx[:, 2]]
    np.testing.assert_allclose(got_b, bnode, rtol=0, atol=1e-12)
    np.testing.assert_allclose(got_e, enode, rtol=0, atol=1e-11)
    # the motional part dominates where the inflow crosses the guide field
    assert np.abs(eta * jj).max() > 1e-3 and np.abs(np.cross(vv, bnode)).max() > 1e-3


# ---------------------------------------------------------------------------
# planted mistakes in the oracle's EB force are caught
# ---------------------------------------------------------------------------

PLANTS = {1: "v x B with the wrong sign", 2: "two components of v x B swapped",
          3: "E with the wrong sign"}


def _fails(fn, *a):
    try:
        fn(*a)
    except AssertionError:
        return True
    return False


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_eb_mistakes_are_caught(tmp_path, oracle, reference_or_none, plant):
    from oracle.oracle import Oracle
    so = str(tmp_path / f"libptc_oracle_plant{plant}.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", f"-DORC_PLANT={plant}", "-shared", "-o", so,
                    os.path.join(ROOT, "oracle", "ptc_oracle.c"), "-lm"], check=True)
    bad = Oracle(so)
    vb = _vb_rhs(oracle, reference_or_none)
    caught = {
        "rhs": _fails(check_rhs_equals_vb, bad, vb),
        "trajectory": _fails(check_trajectory_equals_vb, bad, "dop853"),
        "straight line": _fails(check_drift_frame_straight_line, bad, "dop853"),
        "gyration": _fails(check_gyration, bad, "dop853"),
        "drift": _fails(check_exb_drift, bad, "dop853"),
    }
    print(PLANTS[plant], caught)
    assert caught["rhs"] and caught["trajectory"] and caught["straight line"] and caught["drift"]
    # with E = 0 only the magnetic term acts: a mistake there changes the orbit
    assert caught["gyration"] == (plant in (1, 2))
