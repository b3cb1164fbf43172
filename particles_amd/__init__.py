"""particles_amd -- MI355X-native integrator for the test-particle hot path
of gkowal/particles.

Python host mirror of the reference's integration interface, over the C-ABI
library ``libparticles_amd.so`` (include/particles_amd.h):

  * ``Fields``      -- initialize_fields / finalize_fields
                       (reference sources/fields.F90:51-101): one field
                       snapshot uploaded to HBM once;
  * ``integrate``   -- the particle loop of sources/particles.F90:337-356
                       calling integrate_rk5 / integrate_dp853
                       (sources/integrations.F90:57,224) per particle,
                       with the same arguments and array layouts;
  * ``Particles``   -- the same integration as a device-resident batch that
                       advances in segments (snapshot intervals).

Arrays follow the reference's Fortran layouts: fields (dm1, dm2, dm3, 3),
state (8, m, np), counter (3, np) with (all, accepted, rejected) steps.
The library must be built (``__graft_entry__.build()``); importing the
package without it raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

__all__ = ["Fields", "Particles", "Mlp", "integrate", "set_option", "get_option", "PtcError",
           "METHODS", "MODES", "STATUS", "FEATURES", "library_path"]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libparticles_amd.so")

# method strings of sources/particles.F90:338,347
METHODS = {"rk5": 0, "dop853": 1, "dp853": 1, "rk8": 1}
MODES = {"vb": 0, "eb": 1}
PRECS = {"fp64": 0, "fp32": 1}
STATUS = {-1: "running", 0: "done", 1: "max_steps", 2: "nonfinite", 3: "dt_small"}
# inputs of the step-size network (PTC_PROBE_FEATURES), in order
FEATURES = ("log|B|", "log(|E|/|B|)", "log(Bx-/B)", "log(Bx+/B)", "log(By-/B)", "log(By+/B)",
            "log(Bz-/B)", "log(Bz+/B)", "log(gamma)", "cos(pitch)", "log(h*Omega)", "log(err)",
            "log(t_face/h)", "log(h_PI/h)", "rejected", "log(r_L/dx)")
MLP_PARAMS, MLP_TRAINABLE, MLP_RECORD = 5345, 5313, 20


class PtcError(RuntimeError):
    """A nonzero return code of the C-ABI, with ptc_last_error()'s message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[ptc error {code}] {msg}")
        self.code = code


def _load():
    if not os.path.exists(library_path):
        raise ImportError(f"{library_path} is not built; run __graft_entry__.build() "
                          "(make -C particles_amd/csrc)")
    lib = C.CDLL(library_path)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "ptc_fields_create": ([vp, vp, vp, i32, i32, C.POINTER(vp)], i32),
        "ptc_fields_destroy": ([vp], i32),
        "ptc_fields_series": ([vp, vp, i32, C.POINTER(vp)], i32),
        "ptc_fields_synthesize": ([vp, i32, vp, i32, vp, vp, vp, vp, i32, vp, vp, vp, vp, i32,
                                   C.POINTER(vp)], i32),
        "ptc_integrate": ([vp, i32, i64, i32, vp, i32, vp, vp, vp, vp], i32),
        "ptc_particles_create": ([vp, i32, i64, i32, vp, i32, vp, vp, C.POINTER(vp)], i32),
        "ptc_particles_advance": ([vp, dbl], i32),
        "ptc_particles_download": ([vp, vp, vp, vp], i32),
        "ptc_particles_totals": ([vp, vp], i32),
        "ptc_particles_last_kernel_ms": ([vp, C.POINTER(dbl)], i32),
        "ptc_particles_cell_loads": ([vp, C.POINTER(i64)], i32),
        "ptc_fields_download": ([vp, i32, vp, vp], i32),
        "ptc_fields_convert": ([vp, i32, C.POINTER(vp)], i32),
        "ptc_particles_destroy": ([vp], i32),
        "ptc_probe": ([vp, i32, i64, vp, i32, vp, vp], i32),
        "ptc_set_option": ([C.c_char_p, C.c_char_p], i32),
        "ptc_get_option": ([C.c_char_p, C.c_char_p, i32], i32),
        "ptc_particles_generate": ([vp, i32, C.c_uint64, vp, vp, dbl, dbl], i32),
        "ptc_particles_moments": ([vp, vp], i32),
        "ptc_particles_autotune": ([vp, dbl, vp], i32),
        "ptc_particles_advance_graph": ([vp, i32, vp], i32),
        "ptc_particles_graph_stats": ([vp, vp], i32),
        "ptc_particles_record": ([vp, dbl, i64], i32),
        "ptc_particles_records": ([vp, C.POINTER(i64), vp, C.POINTER(vp)], i32),
        "ptc_particles_use_mlp": ([vp, vp, dbl], i32),
        "ptc_mlp_create": ([vp, C.c_uint64, C.POINTER(vp)], i32),
        "ptc_mlp_destroy": ([vp], i32),
        "ptc_mlp_get_params": ([vp, vp], i32),
        "ptc_mlp_set_params": ([vp, vp], i32),
        "ptc_mlp_forward": ([vp, i64, vp, vp], i32),
        "ptc_mlp_forward_groups": ([vp, i64, vp, vp], i32),
        "ptc_mlp_gradient": ([vp, i64, vp, C.c_float, vp, C.POINTER(dbl)], i32),
        "ptc_mlp_normalize": ([vp, i64, vp, i32], i32),
        "ptc_mlp_train": ([vp, i64, vp, i32, i32, i64, C.c_float, C.c_float, C.c_uint64, vp], i32),
        "ptc_mlp_probe_gemm": ([vp, vp, vp], i32),
        "ptc_last_error": ([], C.c_char_p),
        "ptc_version": ([], C.c_char_p),
        "ptc_diag_line_gather": ([i64, i32, vp], i32),
        "ptc_diag_cell_gather": ([i64, i32, i32, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = _load()


def _check(rc: int):
    if rc != 0:
        raise PtcError(rc, _lib.ptc_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def version() -> str:
    return _lib.ptc_version().decode()


def diag_line_gather(nbytes: int, rounds: int = 4096) -> dict:
    """Random aligned 128-byte lines read from a device buffer of nbytes
    (ptc_diag_line_gather): the rate that bounds gathers from fields far
    larger than the caches."""
    r = np.zeros(3, dtype=np.float64)
    _check(_lib.ptc_diag_line_gather(int(nbytes), int(rounds), _ptr(r)))
    return {"GBps": float(r[0]), "ms": float(r[1]), "bytes": float(r[2])}


def diag_cell_gather(nbytes: int, pattern: str, cells: int = 8192) -> dict:
    """Cell loads as the group kernels make them from a device buffer of
    nbytes (ptc_diag_cell_gather): pattern "lines", four random 128-byte
    lines per load (pair rows), or "span", one 384-byte span at a random
    192-byte-aligned offset (column rows)."""
    pat = {"lines": 0, "span": 1}[pattern]
    r = np.zeros(3, dtype=np.float64)
    _check(_lib.ptc_diag_cell_gather(int(nbytes), pat, int(cells), _ptr(r)))
    return {"cell_loads_per_s": float(r[0]), "ms": float(r[1]), "cell_loads": float(r[2])}


def set_option(key: str, value) -> None:
    """Process-wide tuning switch (block, grid_blocks, layout, max_steps)."""
    _check(_lib.ptc_set_option(key.encode(), str(value).encode()))


def get_option(key: str) -> str:
    buf = C.create_string_buffer(128)
    _check(_lib.ptc_get_option(key.encode(), buf, 128))
    return buf.value.decode()


class Fields:
    """One field snapshot resident in HBM (reference fields.F90 module
    arrays uu, bb).  `f1` is U (mode "vb", the reference's formulation) or E
    (mode "eb"); arrays are (dm1, dm2, dm3, 3) float64."""

    def __init__(self, f1: np.ndarray, bb: np.ndarray, mode: str = "vb", prec: str = "fp64"):
        f1 = np.asfortranarray(f1, dtype=np.float64)
        bb = np.asfortranarray(bb, dtype=np.float64)
        if f1.ndim != 4 or f1.shape[3] != 3 or f1.shape != bb.shape:
            raise ValueError(f"fields must both be (dm1, dm2, dm3, 3), got {f1.shape} and {bb.shape}")
        self.dm = np.ascontiguousarray(f1.shape[:3], dtype=np.int32)
        self.mode, self.prec = mode, prec
        h = C.c_void_p()
        _check(_lib.ptc_fields_create(_ptr(self.dm), _ptr(f1), _ptr(bb), MODES[mode], PRECS[prec],
                                      C.byref(h)))
        self._h = h

    @classmethod
    def synthesize(cls, dm, recipe: str, scalars, bmodes=None, vmodes=None, prec: str = "fp64"):
        """A field built on the device (ptc_fields_synthesize): recipe
        "turbulent_vb" (scalars = (b0,)) or "reconnection_eb" (scalars =
        (delta, guide, vin, eta)); bmodes / vmodes are workloads.Modes."""
        self = cls.__new__(cls)
        self.dm = np.ascontiguousarray(dm, dtype=np.int32)
        self.mode = "vb" if recipe == "turbulent_vb" else "eb"
        self.prec = prec
        code = {"turbulent_vb": 0, "reconnection_eb": 1}[recipe]
        sc = np.zeros(8, dtype=np.float64)
        sc[:len(scalars)] = scalars
        keep = []

        def table(m):
            if m is None:
                return 0, None, None, None, None
            t = (np.ascontiguousarray(m.n, dtype=np.int32), np.ascontiguousarray(m.amp, dtype=np.float64),
                 np.ascontiguousarray(m.pol, dtype=np.float64),
                 np.ascontiguousarray(m.phase, dtype=np.float64))
            keep.append(t)
            return len(m.amp), *(_ptr(a) for a in t)

        nb, bn, ba, bp, bph = table(bmodes)
        nv, vn, va, vpo, vph = table(vmodes)
        h = C.c_void_p()
        _check(_lib.ptc_fields_synthesize(_ptr(self.dm), code, _ptr(sc), nb, bn, ba, bp, bph, nv,
                                          vn, va, vpo, vph, PRECS[prec], C.byref(h)))
        self._h = h
        return self

    @classmethod
    def series(cls, snapshots, times):
        """Time-dependent fields (ptc_fields_series): the snapshots (Fields
        of equal shape, mode and precision) at strictly increasing times,
        blended linearly in time at every stage.  The device arrays are
        copied, so the snapshots may be closed afterwards."""
        snaps = list(snapshots)
        times = np.ascontiguousarray(times, dtype=np.float64)
        if len(snaps) != times.size:
            raise ValueError(f"{len(snaps)} snapshots but {times.size} times")
        self = cls.__new__(cls)
        self.dm = snaps[0].dm.copy() if snaps else np.zeros(3, np.int32)
        self.mode = snaps[0].mode if snaps else "vb"
        self.prec = snaps[0].prec if snaps else "fp64"
        self.times = times
        arr = (C.c_void_p * max(len(snaps), 1))(*[f.handle for f in snaps])
        h = C.c_void_p()
        _check(_lib.ptc_fields_series(arr, _ptr(times), len(snaps), C.byref(h)))
        self._h = h
        return self

    def convert(self, prec: str = "fp32") -> "Fields":
        """A second device copy in precision `prec` and the layout the
        options name for it (ptc_fields_convert), built on the device."""
        out = Fields.__new__(Fields)
        out.dm = self.dm.copy()
        out.mode, out.prec = self.mode, prec
        h = C.c_void_p()
        _check(_lib.ptc_fields_convert(self._h, PRECS[prec], C.byref(h)))
        out._h = h
        return out

    def download(self, snapshot: int = 0):
        """(f1, bb): the stored values back on the host as Fortran
        (dm1, dm2, dm3, 3) float64 arrays (FP32 storage widened), the form
        the constructor takes (ptc_fields_download)."""
        shape = tuple(int(d) for d in self.dm) + (3,)
        f1 = np.empty(shape, dtype=np.float64, order="F")
        bb = np.empty(shape, dtype=np.float64, order="F")
        _check(_lib.ptc_fields_download(self._h, int(snapshot), _ptr(f1), _ptr(bb)))
        return f1, bb

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _check(_lib.ptc_fields_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


_PROBES = {"gather": (0, 3, 6), "rhs": (1, 6, 6), "dti5": (2, 6, 1), "dti8": (3, 6, 1),
           "pow": (4, 2, 1), "features": (5, 10, 16), "gather_t": (6, 4, 6), "rhs_t": (7, 7, 6)}


def probe(fields: Fields, op: str, inputs, params) -> np.ndarray:
    """Evaluates one device building block per input row: "gather" (pos ->
    U or E, B; interpolate_lin), "rhs" (vec -> d vec/dt; acceleration),
    "dti5"/"dti8" (vec -> initial step; dti_guess), "pow" ((x, y) -> x**y as
    the step controller computes it), "features" ((vec, h, err, pi_ratio,
    rejected) -> the step-size network's 16 inputs, FEATURES), "gather_t" /
    "rhs_t" ((pos, t) / (vec, t) in a snapshot series).  inputs is
    (n, k); returns (n, 6), (n, 16) or (n,)."""
    code, nin, nout = _PROBES[op]
    x = np.ascontiguousarray(inputs, dtype=np.float64).reshape(-1, nin)
    params = np.ascontiguousarray(params, dtype=np.float64)
    out = np.zeros((x.shape[0], nout), dtype=np.float64)
    _check(_lib.ptc_probe(fields.handle, code, x.shape[0], _ptr(params), params.size, _ptr(x),
                          _ptr(out)))
    return out if nout > 1 else out[:, 0]


def _prep(params, time, state):
    params = np.ascontiguousarray(params, dtype=np.float64)
    time = np.ascontiguousarray(time, dtype=np.float64)
    state = np.asfortranarray(state, dtype=np.float64)
    if state.ndim != 3 or state.shape[0] != 8:
        raise ValueError(f"state must be (8, m, np), got {state.shape}")
    if state.shape[1] != time.shape[0]:
        raise ValueError(f"state has {state.shape[1]} snapshots but time has {time.shape[0]}")
    return params, time, state


def integrate(fields: Fields, method: str, params, time, state):
    """The particle loop of sources/particles.F90:337-356: integrates every
    particle from t = 0 to params[2] with integrate_rk5 ('rk5') or
    integrate_dp853 ('dop853', 'dp853', 'rk8').  Returns (state, counter,
    status): state (8, m, np) with the snapshots filled in, counter (3, np)
    int64 = (all, accepted, rejected) steps, status (np,) int32."""
    params, time, state = _prep(params, time, state)
    state = state.copy(order="F")
    np_ = state.shape[2]
    counter = np.zeros((3, np_), dtype=np.int64, order="F")
    status = np.zeros(np_, dtype=np.int32)
    _check(_lib.ptc_integrate(fields.handle, METHODS[method], np_, state.shape[1], _ptr(params),
                              params.size, _ptr(time), _ptr(state), _ptr(counter), _ptr(status)))
    return state, counter, status


INIT_KINDS = {"maxwell": 0, "gamma": 1, "gamma_log": 2}


class Particles:
    """Device-resident particle batch advanced in segments; the result of
    advancing to t1 < t2 < ... < tmax equals one uninterrupted integrate().
    `state` may be an int (the particle count) to start from zeros and fill
    the batch with generate()."""

    def __init__(self, fields: Fields, method: str, params, time, state):
        if isinstance(state, (int, np.integer)):
            params = np.ascontiguousarray(params, dtype=np.float64)
            time = np.ascontiguousarray(time, dtype=np.float64)
            self.fields = fields
            self.m, self.np = time.shape[0], int(state)
            h = C.c_void_p()
            _check(_lib.ptc_particles_create(fields.handle, METHODS[method], self.np, self.m,
                                             _ptr(params), params.size, _ptr(time), None,
                                             C.byref(h)))
            self._h = h
            return
        params, time, state = _prep(params, time, state)
        self.fields = fields  # keeps the field handle alive
        self.m, self.np = state.shape[1], state.shape[2]
        h = C.c_void_p()
        _check(_lib.ptc_particles_create(fields.handle, METHODS[method], self.np, self.m,
                                         _ptr(params), params.size, _ptr(time), _ptr(state),
                                         C.byref(h)))
        self._h = h

    def advance(self, t_stop: float = math.inf) -> float:
        """Integrates until t >= min(t_stop, tmax); returns the integrator
        kernel's duration in ms (HIP events on the batch's stream)."""
        _check(_lib.ptc_particles_advance(self._h, float(t_stop)))
        ms = C.c_double()
        _check(_lib.ptc_particles_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def totals(self):
        """dict of summed counters, reduced on the device."""
        t = np.zeros(5, dtype=np.int64)
        _check(_lib.ptc_particles_totals(self._h, _ptr(t)))
        loads = C.c_int64()
        _check(_lib.ptc_particles_cell_loads(self._h, C.byref(loads)))
        return dict(all=int(t[0]), accepted=int(t[1]), rejected=int(t[2]), rhs=int(t[3]),
                    running=int(t[4]), cell_loads=int(loads.value))

    def generate(self, kind: str, seed: int, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                 a: float = 1.0, b: float = 1.0) -> None:
        """Initial states on the device (ptc_particles_generate): kind
        "maxwell" (momenta normal(0, a)), "gamma" (gamma = a) or "gamma_log"
        (gamma log-uniform in [a, b]); positions uniform in [lo, hi]."""
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        _check(_lib.ptc_particles_generate(self._h, INIT_KINDS[kind], int(seed), _ptr(lo),
                                           _ptr(hi), float(a), float(b)))

    def advance_graph(self, t_stops) -> float:
        """Advances through every stop time in one captured HIP graph;
        returns the graph's duration in ms."""
        t = np.ascontiguousarray(t_stops, dtype=np.float64)
        _check(_lib.ptc_particles_advance_graph(self._h, t.size, _ptr(t)))
        ms = C.c_double()
        _check(_lib.ptc_particles_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value

    def graph_stats(self) -> dict:
        """Graphs instantiated by advance_graph for this batch, and calls
        that replayed the kept graph with new stop times."""
        st = np.zeros(2, dtype=np.int64)
        _check(_lib.ptc_particles_graph_stats(self._h, _ptr(st)))
        return dict(instantiated=int(st[0]), replayed=int(st[1]))

    def autotune(self, t_end: float) -> dict:
        """Integrates to t_end while timing candidate launch settings, keeps
        the fastest for this batch (ptc_particles_autotune)."""
        c = np.zeros(5, dtype=np.int32)
        _check(_lib.ptc_particles_autotune(self._h, float(t_end), _ptr(c)))
        return {"block": int(c[0]), "sort": int(c[1]),
                "order": {1: "spatial", 2: "cost"}.get(int(c[2]), "index"),
                "tail": "group" if int(c[3]) < 0 else str(int(c[3])),
                "group_lanes": int(c[4])}

    def moments(self) -> np.ndarray:
        """(6, 4): mean, standard deviation, skewness, excess kurtosis of each
        component of state(1:6,1,:)."""
        m = np.zeros(24, dtype=np.float64)
        _check(_lib.ptc_particles_moments(self._h, _ptr(m)))
        return m.reshape(6, 4)

    def record(self, rate: float, capacity: int) -> None:
        """Record mode: keep the PI controller and write a training record for
        a fraction `rate` of accepted steps into a device buffer of
        `capacity` records; capacity 0 leaves record mode."""
        _check(_lib.ptc_particles_record(self._h, float(rate), int(capacity)))

    def records(self, device: bool = False):
        """The records written so far: a (count, 20) float32 array, or with
        device=True a DeviceRecords view of the device buffer (no copy)."""
        n = C.c_int64()
        if device:
            d = C.c_void_p()
            _check(_lib.ptc_particles_records(self._h, C.byref(n), None, C.byref(d)))
            return DeviceRecords(d.value, n.value, self)
        _check(_lib.ptc_particles_records(self._h, C.byref(n), None, None))
        out = np.zeros((n.value, MLP_RECORD), dtype=np.float32)
        if n.value:
            _check(_lib.ptc_particles_records(self._h, C.byref(n), _ptr(out), None))
        return out

    def use_mlp(self, net, shift: float = 0.0) -> None:
        """Learned mode: the network proposes the step after each accepted
        step (log factor + shift, clamped to [facmin, facmax]); None returns
        to the PI controller."""
        self._net = net  # keeps the network alive while in use
        _check(_lib.ptc_particles_use_mlp(self._h, net.handle if net is not None else None,
                                          float(shift)))

    def download(self):
        state = np.zeros((8, self.m, self.np), dtype=np.float64, order="F")
        counter = np.zeros((3, self.np), dtype=np.int64, order="F")
        status = np.zeros(self.np, dtype=np.int32)
        _check(_lib.ptc_particles_download(self._h, _ptr(state), _ptr(counter), _ptr(status)))
        return state, counter, status

    def close(self):
        if self._h:
            _check(_lib.ptc_particles_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceRecords:
    """Training records left in device memory by a Particles batch in
    record mode (valid while the batch exists and stays in record mode)."""

    def __init__(self, ptr: int, count: int, owner):
        self.ptr, self.count, self.owner = ptr, count, owner

    def __len__(self):
        return self.count


def _records_arg(records):
    if isinstance(records, DeviceRecords):
        return C.c_void_p(records.ptr), records.count, 1, None
    r = np.ascontiguousarray(records, dtype=np.float32)
    if r.ndim != 2 or r.shape[1] != MLP_RECORD:
        raise ValueError(f"records must be (n, {MLP_RECORD}), got {r.shape}")
    return _ptr(r), r.shape[0], 0, r


class Mlp:
    """The step-size network (16 -> 64 -> 64 -> 1, GELU, BF16 MFMA with FP32
    accumulation) with its device training state.  params is the packed
    PTC_MLP_PARAMS float vector (None: random initialisation from seed)."""

    def __init__(self, params=None, seed: int = 0):
        h = C.c_void_p()
        if params is None:
            _check(_lib.ptc_mlp_create(None, int(seed), C.byref(h)))
        else:
            p = np.ascontiguousarray(params, dtype=np.float32)
            if p.size != MLP_PARAMS:
                raise ValueError(f"need {MLP_PARAMS} parameters, got {p.size}")
            _check(_lib.ptc_mlp_create(_ptr(p), 0, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        _check(_lib.ptc_mlp_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, params) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != MLP_PARAMS:
            raise ValueError(f"need {MLP_PARAMS} parameters, got {p.size}")
        _check(_lib.ptc_mlp_set_params(self._h, _ptr(p)))

    def forward(self, features, groups: bool = False) -> np.ndarray:
        """Network outputs for rows of 16 features; groups=True evaluates
        them as the six-lanes-per-particle integrator does (same values)."""
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 16)
        out = np.zeros(f.shape[0], dtype=np.float32)
        fn = _lib.ptc_mlp_forward_groups if groups else _lib.ptc_mlp_forward
        _check(fn(self._h, f.shape[0], _ptr(f), _ptr(out)))
        return out

    def gradient(self, records, alpha: float = 0.0):
        """(loss, grad) of mean((out - target)^2) over host records."""
        ptr, n, on_dev, keep = _records_arg(records)
        if on_dev:
            raise ValueError("gradient takes host records")
        g = np.zeros(MLP_TRAINABLE, dtype=np.float32)
        loss = C.c_double()
        _check(_lib.ptc_mlp_gradient(self._h, n, ptr, float(alpha), _ptr(g), C.byref(loss)))
        return loss.value, g

    def normalize(self, records) -> None:
        """Input normalisation from the records' feature statistics."""
        ptr, n, on_dev, keep = _records_arg(records)
        _check(_lib.ptc_mlp_normalize(self._h, n, ptr, on_dev))

    def train(self, records, steps: int, batch: int = 16384, lr: float = 1e-3,
              alpha: float = 0.0, seed: int = 1) -> np.ndarray:
        """`steps` Adam steps on minibatches drawn from the records (host
        array or DeviceRecords); returns each minibatch's loss."""
        ptr, n, on_dev, keep = _records_arg(records)
        losses = np.zeros(max(int(steps), 1), dtype=np.float32)
        _check(_lib.ptc_mlp_train(self._h, n, ptr, on_dev, int(steps), int(batch), float(lr),
                                  float(alpha), int(seed), _ptr(losses)))
        return losses[:int(steps)]

    def close(self):
        if self._h:
            _check(_lib.ptc_mlp_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def probe_gemm(a, bt) -> np.ndarray:
    """c = a bt^T (64 x 64 each) through the device's BF16 wave product."""
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(64, 64)
    bt = np.ascontiguousarray(bt, dtype=np.float32).reshape(64, 64)
    c = np.zeros((64, 64), dtype=np.float32)
    _check(_lib.ptc_mlp_probe_gemm(_ptr(a), _ptr(bt), _ptr(c)))
    return c
