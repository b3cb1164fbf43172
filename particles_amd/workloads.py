"""Synthetic inputs for the integration path: parameter vectors, snapshot
time grids, random-Fourier-mode fields and initial particle states.

Everything here is host-side NumPy and seeded (``numpy.random.Philox``), so
the same seed gives the same inputs on any machine.  The device-resident
field synthesis used for large grids lives in the C-ABI library
(``ptc_fields_synthesize``) and evaluates the same mode tables.

Conventions follow the reference driver:
  * params: the 32-entry vector of sources/particles.F90:317-333
    (index 0 = q/m, 1 = velocity unit, 2 = tmax, 3 = dtini (unused),
    4 = dtmax, 5 = atol, 6 = rtol, 7 = facmin, 8 = facmax, 9 = safe,
    10 = beta, 11..13 = domain origin, 14..16 = domain size);
  * time grid: sources/particles.F90:240-249,269-280;
  * fields: Fortran-ordered (dm1, dm2, dm3, 3) arrays on a node-centred
    periodic grid, node i at origin + i * size / dm (sources/interpolations.F90:140-149).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

TWO_PI = 2.0 * math.pi
# gyro-period of B0 = 1 for q/m = 1 in code units (c = 1)
T0 = TWO_PI


def params_vector(qom=1.0, vun=1.0, tmax=1.0, dtini=1.0e-4, dtmax=1.0e-3, atol=1.0e-6,
                  rtol=1.0e-6, facmin=3.33e-1, facmax=6.0, safe=0.9, beta=0.0,
                  xl=(0.0, 0.0, 0.0), xs=(1.0, 1.0, 1.0)) -> np.ndarray:
    """The driver's params vector (sources/particles.F90:317-333); defaults
    are the driver's own (sources/particles.F90:53-84)."""
    p = np.zeros(32, dtype=np.float64)
    p[0], p[1], p[2], p[3], p[4] = qom, vun, tmax, dtini, dtmax
    p[5], p[6], p[7], p[8], p[9], p[10] = atol, rtol, facmin, facmax, safe, beta
    p[11:14] = xl
    p[14:17] = xs
    return p


def time_grid(nsteps: int, tmax: float, tscale: str = "linear", tmin: float = 1.0e-3) -> np.ndarray:
    """Snapshot times (sources/particles.F90:240-249 adjusts nsteps,
    :269-280 fills the grid).  `nsteps` is the params.in value."""
    if tscale in ("lin", "linear"):
        m = nsteps + 1
        t = np.zeros(m)
        for n in range(2, m + 1):
            t[n - 1] = (tmax * (n - 1)) / (m - 1)
        return t
    if tscale in ("log", "logarithmic", "log10"):
        m = int(math.log10(tmax) - math.log10(tmin)) * nsteps + 2
        t = np.zeros(m)
        ltmn, ltmx = math.log10(tmin), math.log10(tmax)
        for n in range(2, m + 1):
            t[n - 1] = 10.0 ** (((m - n) * ltmn + (n - 2) * ltmx) / (m - 2))
        return t
    raise ValueError(f"unknown time scale {tscale!r}")


@dataclass
class Modes:
    """A sum of real Fourier modes  sum_m amp_m * pol_m * cos(2 pi n_m . x + phase_m)
    on the unit periodic box.  Integer wave vectors keep every mode periodic."""
    n: np.ndarray      # (M, 3) int
    amp: np.ndarray    # (M,)
    pol: np.ndarray    # (M, 3) unit vectors, pol . n = 0 (solenoidal)
    phase: np.ndarray  # (M,)

    def rms(self) -> float:
        """Exact box rms (distinct wave vectors, none the negative of another)."""
        return float(np.sqrt(0.5 * np.sum(self.amp ** 2)))


def random_modes(seed: int, nmodes: int = 256, kmin: float = 1.0, kmax: float = 8.0,
                 slope: float = -5.0 / 3.0, rms: float = 0.5, stream: int = 0) -> Modes:
    """`nmodes` distinct integer wave vectors with kmin <= |n| <= kmax (|k| = 2 pi |n|,
    so |k| in [2 pi, 16 pi] by default), amplitude ~ |k|^slope rescaled to the
    given rms, random phases and solenoidal polarisation (pol . k = 0)."""
    rng = np.random.Generator(np.random.Philox(key=seed, counter=[stream, 0, 0, 0]))
    r = int(math.floor(kmax))
    g = np.arange(-r, r + 1)
    cand = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    kk = np.sqrt((cand ** 2).sum(1))
    cand = cand[(kk >= kmin) & (kk <= kmax)]
    # one representative of each +-n pair: first non-zero component positive
    first = np.where(cand[:, 0] != 0, cand[:, 0], np.where(cand[:, 1] != 0, cand[:, 1], cand[:, 2]))
    cand = cand[first > 0]
    if nmodes > len(cand):
        raise ValueError(f"only {len(cand)} distinct wave vectors in the shell")
    pick = rng.choice(len(cand), size=nmodes, replace=False)
    n = cand[np.sort(pick)].astype(np.int64)
    kn = np.sqrt((n ** 2).sum(1))
    amp = kn ** slope
    # random unit vector orthogonal to n
    v = rng.normal(size=(nmodes, 3))
    nh = n / kn[:, None]
    v -= (v * nh).sum(1)[:, None] * nh
    pol = v / np.sqrt((v ** 2).sum(1))[:, None]
    phase = rng.uniform(0.0, TWO_PI, size=nmodes)
    m = Modes(n=n, amp=amp, pol=pol, phase=phase)
    m.amp = m.amp * (rms / m.rms())
    return m


def eval_modes(modes: Modes, dm) -> np.ndarray:
    """Modes on the node-centred grid x_i = i / dm (i = 0 .. dm-1 per axis),
    returned Fortran-ordered (dm1, dm2, dm3, 3).  Exact inverse FFT of the
    sparse spectrum."""
    dm = tuple(int(d) for d in dm)
    spec = np.zeros((3,) + dm, dtype=np.complex128)
    coef = 0.5 * modes.amp * np.exp(1j * modes.phase)
    for c in range(3):
        for sgn in (1, -1):
            idx = tuple((sgn * modes.n[:, a]) % dm[a] for a in range(3))
            val = coef * modes.pol[:, c] if sgn == 1 else np.conj(coef) * modes.pol[:, c]
            np.add.at(spec[c], idx, val)
    out = np.empty(dm + (3,), dtype=np.float64, order="F")
    size = dm[0] * dm[1] * dm[2]
    for c in range(3):
        out[..., c] = np.real(np.fft.ifftn(spec[c])) * size
    return out


def eval_modes_at(modes: Modes, x: np.ndarray) -> np.ndarray:
    """Direct evaluation at arbitrary points x (N, 3) in box units (checker)."""
    ph = TWO_PI * (x @ modes.n.T) + modes.phase[None, :]
    return (np.cos(ph) * modes.amp[None, :]) @ modes.pol


# --------------------------------------------------------------------------
# field recipes
# --------------------------------------------------------------------------

def turbulent_vb(dm: int, seed: int, b0: float = 1.0, brms: float = 0.5, vrms: float = 0.1,
                 nmodes: int = 256):
    """Turbulent-box recipe: B = B0 z + dB, V built the same way; the force is
    q/m (v - V) x B at the particle (VB mode, the reference's formulation).
    Returns (uu, bb) Fortran arrays."""
    mb = random_modes(seed, nmodes=nmodes, rms=brms, stream=1)
    mv = random_modes(seed, nmodes=nmodes, rms=vrms, stream=2)
    bb = eval_modes(mb, (dm, dm, dm))
    bb[..., 2] += b0
    uu = eval_modes(mv, (dm, dm, dm))
    return uu, bb


def propagate(modes: Modes, t: float, speed: float) -> Modes:
    """The modes as waves travelling along their wave vectors at `speed`
    (box lengths per time unit): phase_m - 2 pi |n_m| speed t."""
    kn = np.sqrt((modes.n ** 2).sum(1))
    return Modes(n=modes.n, amp=modes.amp, pol=modes.pol,
                 phase=modes.phase - TWO_PI * kn * speed * t)


def turbulent_vb_series(dm: int, seed: int, times, speed: float = 0.01, b0: float = 1.0,
                        brms: float = 0.5, vrms: float = 0.1, nmodes: int = 256):
    """Snapshots of turbulent_vb at `times` with every mode propagating
    (propagate); returns (uu, bb) Fortran (dm, dm, dm, 3, len(times))."""
    mb = random_modes(seed, nmodes=nmodes, rms=brms, stream=1)
    mv = random_modes(seed, nmodes=nmodes, rms=vrms, stream=2)
    ns = len(times)
    uu = np.empty((dm, dm, dm, 3, ns), dtype=np.float64, order="F")
    bb = np.empty_like(uu, order="F")
    for k, t in enumerate(times):
        bb[..., k] = eval_modes(propagate(mb, t, speed), (dm, dm, dm))
        bb[..., 2, k] += b0
        uu[..., k] = eval_modes(propagate(mv, t, speed), (dm, dm, dm))
    return uu, bb


def sheet_profile(y: np.ndarray, delta: float) -> np.ndarray:
    """Periodic double current sheet: tanh((y-1/4)/d) - tanh((y-3/4)/d) - 1,
    which is tanh(y'/d) about the sheet at y = 1/4 and continuous across the
    periodic boundary."""
    return np.tanh((y - 0.25) / delta) - np.tanh((y - 0.75) / delta) - 1.0


def curl_central(f: np.ndarray, h=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Periodic central-difference curl of a Fortran (dm1,dm2,dm3,3) field."""
    def d(a, axis):
        return (np.roll(a, -1, axis) - np.roll(a, 1, axis)) / (2.0 * h[axis])
    out = np.empty_like(f, order="F")
    out[..., 0] = d(f[..., 2], 1) - d(f[..., 1], 2)
    out[..., 1] = d(f[..., 0], 2) - d(f[..., 2], 0)
    out[..., 2] = d(f[..., 1], 0) - d(f[..., 0], 1)
    return out


def reconnection_eb(dm: int, seed: int, delta: float = 0.02, guide: float = 0.1,
                    noise_rms: float = 0.05, vin: float = 0.1, eta: float = 1.0e-3,
                    nmodes: int = 256):
    """Reconnection-like recipe: B_x = sheet(y), B_z = guide, plus solenoidal
    random-mode noise; V_y = -vin * sheet(y); J = curl B (central
    differences); E = -V x B + eta J (EB mode).  Returns (ee, bb)."""
    y = np.arange(dm) / dm
    prof = sheet_profile(y, delta)
    bb = eval_modes(random_modes(seed, nmodes=nmodes, rms=noise_rms, stream=1), (dm, dm, dm))
    bb[..., 0] += prof[None, :, None]
    bb[..., 2] += guide
    vy = -vin * prof[None, :, None]
    jj = curl_central(bb, h=(1.0 / dm,) * 3)
    ee = np.empty_like(bb, order="F")
    # -V x B with V = (0, vy, 0)
    ee[..., 0] = -(vy * bb[..., 2]) + eta * jj[..., 0]
    ee[..., 1] = eta * jj[..., 1]
    ee[..., 2] = vy * bb[..., 0] + eta * jj[..., 2]
    return ee, bb


def turbulent_vb_device(dm: int, seed: int, prec: str = "fp64", b0: float = 1.0,
                        brms: float = 0.5, vrms: float = 0.1, nmodes: int = 256):
    """turbulent_vb built on the device (particles_amd.Fields), same modes."""
    from . import Fields
    mb = random_modes(seed, nmodes=nmodes, rms=brms, stream=1)
    mv = random_modes(seed, nmodes=nmodes, rms=vrms, stream=2)
    return Fields.synthesize((dm, dm, dm), "turbulent_vb", (b0,), mb, mv, prec=prec)


def turbulent_vb_series_device(dm: int, seed: int, times, prec: str = "fp64",
                               speed: float = 0.01, b0: float = 1.0, brms: float = 0.5,
                               vrms: float = 0.1, nmodes: int = 256):
    """turbulent_vb_series built on the device: one synthesised snapshot per
    time, gathered into a series (particles_amd.Fields.series)."""
    from . import Fields
    mb = random_modes(seed, nmodes=nmodes, rms=brms, stream=1)
    mv = random_modes(seed, nmodes=nmodes, rms=vrms, stream=2)
    snaps = []
    try:
        for t in times:
            snaps.append(Fields.synthesize((dm, dm, dm), "turbulent_vb", (b0,),
                                           propagate(mb, t, speed), propagate(mv, t, speed),
                                           prec=prec))
        return Fields.series(snaps, times)
    finally:
        for f in snaps:
            f.close()


def reconnection_eb_device(dm: int, seed: int, prec: str = "fp64", delta: float = 0.02,
                           guide: float = 0.1, noise_rms: float = 0.05, vin: float = 0.1,
                           eta: float = 1.0e-3, nmodes: int = 256):
    """reconnection_eb built on the device (particles_amd.Fields), same modes."""
    from . import Fields
    mb = random_modes(seed, nmodes=nmodes, rms=noise_rms, stream=1)
    return Fields.synthesize((dm, dm, dm), "reconnection_eb", (delta, guide, vin, eta), mb,
                             None, prec=prec)


# --------------------------------------------------------------------------
# initial states
# --------------------------------------------------------------------------

def initial_state(np_: int, m: int, seed: int, gamma=10.0, gamma_range=None,
                  xl=(0.0, 0.0, 0.0), xs=(1.0, 1.0, 1.0)) -> np.ndarray:
    """state(8, m, np) Fortran array with uniform positions in the box and
    isotropic momenta u = gamma v (c = 1).  Either a fixed `gamma` or a
    log-uniform `gamma_range` = (gmin, gmax)."""
    rng = np.random.Generator(np.random.Philox(key=seed, counter=[7, 0, 0, 0]))
    st = np.zeros((8, m, np_), dtype=np.float64, order="F")
    pos = rng.uniform(size=(np_, 3))
    st[0:3, 0, :] = (np.asarray(xl)[None, :] + pos * np.asarray(xs)[None, :]).T
    if gamma_range is not None:
        lg = rng.uniform(math.log(gamma_range[0]), math.log(gamma_range[1]), size=np_)
        g = np.exp(lg)
    else:
        g = np.full(np_, float(gamma))
    umag = np.sqrt(np.maximum(g * g - 1.0, 0.0))
    d = rng.normal(size=(np_, 3))
    d /= np.sqrt((d ** 2).sum(1))[:, None]
    st[3:6, 0, :] = (d * umag[:, None]).T
    return st
