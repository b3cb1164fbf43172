"""The Fortran host driver (particles_amd/fortran/particles_host): the
reference program's parameter file, unit algebra, time grid and output
files around the device integrator.  CPU: a dry run (no GPU) checks the
parsed parameters and derived quantities against the reference's formulas
(sources/particles.F90:163-214,240-280,317-333), and the initial state it
writes equals the reference program's own draw (its random module, compiled
from the reference sources into oracle/_ref/ref_init).  GPU: a full run
writes counter.fits, time.fits and state.fits, and its snapshots equal the
device integrator's and the reference's compiled integrator's on the initial
states it wrote; with field_snapshots = 3 (a
time-dependent field, no counterpart in the reference) they equal the
library's series integration of the same snapshots."""
import os
import re
import subprocess

import numpy as np
import pytest

from particles_amd import fits
from particles_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "particles_amd", "fortran", "particles_host")
C_LIGHT, KB, MP, ME, QE = 2.99792458e8, 1.38064852e-23, 1.672621898e-27, 9.10938356e-31, 1.6021766208e-19


def _setup(tmp, params_text, dm=12, gz=False):
    uu, bb = W.turbulent_vb(dm, seed=5, nmodes=24)
    data = os.path.join(tmp, "data")
    os.makedirs(data, exist_ok=True)
    names = ("velx", "vely", "velz", "magx", "magy", "magz")
    for k, n in enumerate(names):
        arr = uu[..., k] if k < 3 else bb[..., k - 3]
        path = os.path.join(data, n + ".fits")
        fits.write_f64(path, arr)
        if gz:
            import gzip
            with open(path, "rb") as s, gzip.open(path + ".gz", "wb") as d:
                d.write(s.read())
            os.remove(path)
    with open(os.path.join(tmp, "params.in"), "w") as fh:
        fh.write(params_text)
    return uu, bb


def _run(tmp, *args, timeout=300):
    if not os.path.exists(EXE):
        pytest.skip("host driver not built")
    return subprocess.run([EXE, "-i", "params.in", *args], cwd=tmp, capture_output=True,
                          text=True, timeout=timeout)


def _floats(text, label):
    m = re.search(re.escape(label) + r"\s*=\s*\n((?:\s+[-+0-9.E]+)+)", text)
    return np.array([float(x) for x in m.group(1).split()])


PARAMS = """# unit test
method      = 'dop853'      # DOP853
particle_type = electron
nparticles  = 9
nsteps      = 3
time_scale  = log
tmin        = 1.0e-2
tmax        = 10.0
tunit       = 2.0e-6
vunit       = 3.0e5
temperature = 5.0e4
dtmax       = 0.5
atol        = 1.0e-9
rtol        = 1.0e-9
zmax        = 0.25
xmin_domain = -0.5
xmax_domain = 0.5
datadir     = "data"
"""


def test_dry_run_parameters_and_units(tmp_path):
    tmp = str(tmp_path)
    _setup(tmp, PARAMS, gz=True)
    r = _run(tmp, "--dry-run")
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    vunit = 3.0e5 / C_LIGHT
    qom = -1.0e-4 * QE * 2.0e-6 / ME
    vth = np.sqrt(KB * 5.0e4 / MP) / C_LIGHT
    nsteps = int(np.log10(10.0) - np.log10(1.0e-2)) * 3 + 2
    assert re.search(r"nsteps\s*=\s*%d\b" % nsteps, out)
    p = _floats(out, "params(1:17)")
    want = W.params_vector(qom=qom, vun=vunit, tmax=10.0, dtmax=0.5, atol=1e-9, rtol=1e-9,
                           xl=(-0.5, 0, 0), xs=(1.0, 1.0, 1.0))[:17]
    assert np.allclose(p, want, rtol=1e-14, atol=0)
    t = _floats(out, "time(1:nsteps)")
    assert np.allclose(t, W.time_grid(3, 10.0, "log", 1.0e-2), rtol=1e-14, atol=0)
    got_vth = float(re.search(r"thermal speed\s*=\s*([-+0-9.E]+)", out).group(1))
    got_qom = float(re.search(r"q/m\s*=\s*([-+0-9.E]+)", out).group(1))
    assert abs(got_vth - vth) < 1e-14 * vth and abs(got_qom - qom) < 1e-14 * abs(qom)
    # seeding region: x in [max(0, -0.5), min(1, 0.5)], z in [0, 0.25]
    mean = {m.group(1): float(m.group(2)) for m in
            re.finditer(r"(pos[xyz]):\s+([-+0-9.E]+)", out)}
    assert 0.0 < mean["posx"] < 0.5 and 0.0 < mean["posz"] < 0.25
    # momenta ~ normal(0, vth): 9 particles, odd count handled
    assert re.search(r"momz:\s+[-+0-9.E]+\s+[0-9.E+-]+", out)


REF_INIT = os.path.join(ROOT, "oracle", "_ref", "ref_init")


@pytest.mark.parametrize("ptype,npart,temp", [("electron", 1000, 5.0e4), ("proton", 64, 1.0e6)])
def test_dry_run_initial_state_is_the_reference_programs(tmp_path, ptype, npart, temp):
    """initial_state.fits of a dry run equals, bit for bit, the state the
    reference program draws with its own random module (sources/random.F90
    uniform and normal, called as in particles.F90:289-296) in a fresh
    process, and params.fits holds the printed params(1:17).  The region is
    the reference's clamp of the parameter file's bounds to the domain
    (particles.F90:240-280): x in [0, 0.5], y in [0, 1], z in [0, 0.25].
    (The reference's normal() fills an even count, so npart is even.)"""
    if not os.path.exists(REF_INIT):
        pytest.skip("reference random module (oracle/_ref/ref_init) not built")
    tmp = str(tmp_path)
    text = PARAMS.replace("nparticles  = 9", f"nparticles  = {npart}").replace(
        "particle_type = electron", f"particle_type = {ptype}").replace(
        "temperature = 5.0e4", f"temperature = {temp:.6e}")
    _setup(tmp, text)
    r = _run(tmp, "--dry-run")
    assert r.returncode == 0, r.stdout + r.stderr
    path = os.path.join(tmp, "initial_state.fits")
    state = fits.read(path)
    assert state.shape == (6, npart) and float(fits.header(path)["TEMP"]) == temp
    ref = os.path.join(tmp, "ref_init.bin")
    subprocess.run([REF_INIT, str(npart), "0", "0.5", "0", "1", "0", "0.25", repr(temp), ref],
                   check=True, timeout=60)
    expect = np.fromfile(ref, dtype=np.float64).reshape(npart, 6).T
    assert np.array_equal(state, expect)
    params = fits.read(os.path.join(tmp, "params.fits"))
    # printed with 16 significant digits
    assert np.allclose(params, _floats(r.stdout, "params(1:17)"), rtol=1e-15, atol=0)
    assert float(fits.header(os.path.join(tmp, "params.fits"))["TUNIT"]) == 2.0e-6


def test_dry_run_errors(tmp_path):
    tmp = str(tmp_path)
    _setup(tmp, PARAMS.replace("'dop853'", "'euler'"))
    r = _run(tmp, "--dry-run")
    assert r.returncode != 0 and "unknown integration method" in r.stdout
    _setup(tmp, PARAMS.replace('"data"', '"nowhere"'))
    r = _run(tmp, "--dry-run")
    assert r.returncode != 0 and "missing field file" in r.stdout


def test_dry_run_kernel_choice(tmp_path):
    tmp = str(tmp_path)
    _setup(tmp, PARAMS)
    r = _run(tmp, "--dry-run")
    assert r.returncode == 0 and re.search(r"kernels\s*=\s*auto", r.stdout)
    r = _run(tmp, "--kernels", "group", "--dry-run")
    assert r.returncode == 0 and re.search(r"kernels\s*=\s*group", r.stdout)
    r = _run(tmp, "--kernels", "fastest", "--dry-run")
    assert r.returncode != 0 and "--kernels" in r.stdout and "tail" in r.stdout
    r = _run(tmp, "--layout", "columns", "--dry-run")
    assert r.returncode == 0 and re.search(r"field storage\s*=\s*columns", r.stdout)
    r = _run(tmp, "--mixed", "--f32-layout", "columns", "--dry-run")
    assert r.returncode == 0 and re.search(r"field storage\s*=\s*columns", r.stdout)
    r = _run(tmp, "--layout", "column", "--dry-run")
    assert r.returncode != 0 and "--layout" in r.stdout


@pytest.mark.gpu
def test_full_run_kernel_choices_agree(tmp_path):
    """--kernels lanes (one lane per particle) and --kernels group (six
    lanes per particle) write bit-identical state and counter files."""
    out = {}
    for k in ("lanes", "group"):
        tmp = os.path.join(str(tmp_path), k)
        os.makedirs(tmp)
        _setup(tmp, PARAMS.replace("nparticles  = 9", "nparticles  = 40"))
        r = _run(tmp, "--kernels", k)
        assert r.returncode == 0, r.stdout + r.stderr
        out[k] = (fits.read(os.path.join(tmp, "state.fits")),
                  fits.read(os.path.join(tmp, "counter.fits")))
    assert np.array_equal(out["lanes"][0], out["group"][0])
    assert np.array_equal(out["lanes"][1], out["group"][1])


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_full_run_field_storage_choices_agree(tmp_path, mixed):
    """--layout columns (FP64) and --mixed --f32-layout columns (FP32) store
    the fields in column rows, read by the trio kernel; the state and counter
    files equal those of the default node rows bit for bit."""
    out = {}
    for lay in (("nodes", "columns") if mixed else ("aos", "columns")):
        tmp = os.path.join(str(tmp_path), lay)
        os.makedirs(tmp)
        _setup(tmp, PARAMS.replace("nparticles  = 9", "nparticles  = 40"))
        args = ("--mixed", "--f32-layout", lay) if mixed else ("--layout", lay)
        r = _run(tmp, *args)
        assert r.returncode == 0, r.stdout + r.stderr
        out[lay] = (fits.read(os.path.join(tmp, "state.fits")),
                    fits.read(os.path.join(tmp, "counter.fits")))
    a, b = out.values()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.gpu
def test_full_run_writes_reference_files(tmp_path):
    import particles_amd as P
    tmp = str(tmp_path)
    text = PARAMS.replace("particle_type = electron", "particle_type = proton").replace(
        "tunit       = 2.0e-6", "tunit       = 1.0e-4").replace("nparticles  = 9", "nparticles  = 64")
    uu, bb = _setup(tmp, text)
    r = _run(tmp)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stdout + r.stderr
    state = fits.read(os.path.join(tmp, "state.fits"))
    counter = fits.read(os.path.join(tmp, "counter.fits"))
    time = fits.read(os.path.join(tmp, "time.fits"))
    assert state.shape == (8, len(time), 64) and counter.shape == (3, 64)
    assert float(fits.header(os.path.join(tmp, "time.fits"))["TUNIT"]) == 1.0e-4
    assert float(fits.header(os.path.join(tmp, "state.fits"))["TEMP"]) == 5.0e4
    assert fits.dims(os.path.join(tmp, "counter.fits"))[1] == 64
    # the same integration through the library on the written initial state
    vunit = 3.0e5 / C_LIGHT
    qom = 1.0e-4 * QE * 1.0e-4 / MP
    p = W.params_vector(qom=qom, vun=vunit, tmax=10.0, dtmax=0.5, atol=1e-9, rtol=1e-9,
                        xl=(-0.5, 0, 0), xs=(1.0, 1.0, 1.0))
    st0 = np.zeros_like(state, order="F")
    st0[:, 0, :] = state[:, 0, :]
    st0[6:, 0, :] = 0.0
    with P.Fields(vunit * uu, 1.0 * bb) as f:
        st, cnt, status = P.integrate(f, "dop853", p, time, st0)
    assert (status == 0).all()
    assert np.array_equal(st[:, 1:, :], state[:, 1:, :])
    assert np.array_equal(cnt.astype(np.float64), counter)
    # and through the reference's own integrator (its modules compiled from
    # the reference sources, CPU) on the same initial state and parameters
    from oracle.oracle import REF_SO, Reference
    if os.path.exists(REF_SO):
        sr, cr = Reference().integrate("dop853", vunit * uu, 1.0 * bb, p, time, st0)
        assert np.array_equal(sr[:, 1:, :], state[:, 1:, :])
        assert np.array_equal(cr.astype(np.float64), counter)
        print("[driver] state.fits and counter.fits equal the reference integrator's", flush=True)


# ---- field snapshot series (field_snapshots / field_dt) ----

SERIES_KEYS = "field_snapshots = 3\nfield_dt = 4.0\n"


def _setup_series(tmp, params_text, dm=12):
    times = np.arange(3) * 4.0
    uu, bb = W.turbulent_vb_series(dm, seed=6, times=times, speed=0.05, nmodes=24)
    data = os.path.join(tmp, "data")
    os.makedirs(data, exist_ok=True)
    names = ("velx", "vely", "velz", "magx", "magy", "magz")
    for s in range(3):
        for k, n in enumerate(names):
            arr = uu[..., k, s] if k < 3 else bb[..., k - 3, s]
            fits.write_f64(os.path.join(data, f"{n}_{s + 1:04d}.fits"), np.asfortranarray(arr))
    with open(os.path.join(tmp, "params.in"), "w") as fh:
        fh.write(params_text)
    return uu, bb, times


def test_dry_run_series_checks_every_snapshot(tmp_path):
    tmp = str(tmp_path)
    _setup_series(tmp, PARAMS + SERIES_KEYS)
    r = _run(tmp, "--dry-run")
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"field snaps\s*=\s*3 every\s+4\.0", r.stdout)
    # a missing file of the third snapshot
    os.remove(os.path.join(tmp, "data", "magy_0003.fits"))
    r = _run(tmp, "--dry-run")
    assert r.returncode != 0 and "missing field file data/magy_0003.fits" in r.stdout
    # a snapshot of another shape
    fits.write_f64(os.path.join(tmp, "data", "magy_0003.fits"), np.zeros((12, 12, 6), order="F"))
    r = _run(tmp, "--dry-run")
    assert r.returncode != 0 and "does not match the first" in r.stdout
    # several snapshots need a spacing
    with open(os.path.join(tmp, "params.in"), "w") as fh:
        fh.write(PARAMS + "field_snapshots = 3\n")
    r = _run(tmp, "--dry-run")
    assert r.returncode != 0 and "field_dt > 0" in r.stdout


@pytest.mark.gpu
def test_full_run_with_snapshot_series(tmp_path):
    import particles_amd as P
    tmp = str(tmp_path)
    text = PARAMS.replace("particle_type = electron", "particle_type = proton").replace(
        "tunit       = 2.0e-6", "tunit       = 1.0e-4").replace(
        "nparticles  = 9", "nparticles  = 64") + SERIES_KEYS
    uu, bb, times = _setup_series(tmp, text)
    r = _run(tmp)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stdout + r.stderr
    state = fits.read(os.path.join(tmp, "state.fits"))
    counter = fits.read(os.path.join(tmp, "counter.fits"))
    time = fits.read(os.path.join(tmp, "time.fits"))
    vunit = 3.0e5 / C_LIGHT
    qom = 1.0e-4 * QE * 1.0e-4 / MP
    p = W.params_vector(qom=qom, vun=vunit, tmax=10.0, dtmax=0.5, atol=1e-9, rtol=1e-9,
                        xl=(-0.5, 0, 0), xs=(1.0, 1.0, 1.0))
    st0 = np.zeros_like(state, order="F")
    st0[:, 0, :] = state[:, 0, :]
    st0[6:, 0, :] = 0.0
    snaps = [P.Fields(np.asfortranarray(vunit * uu[..., s]), np.asfortranarray(bb[..., s]))
             for s in range(3)]
    try:
        with P.Fields.series(snaps, times) as f:
            st, cnt, status = P.integrate(f, "dop853", p, time, st0)
    finally:
        for g in snaps:
            g.close()
    assert (status == 0).all()
    assert np.array_equal(st[:, 1:, :], state[:, 1:, :])
    assert np.array_equal(cnt.astype(np.float64), counter)
