"""ctypes front end of the CPU checkers (test infrastructure only).

`Oracle` wraps ``oracle/libptc_oracle.so`` (the plain-C FP64 restatement of
the reference hot path, ptc_oracle.c).  `Reference` wraps
``oracle/_ref/libptc_ref.so``, the reference's own Fortran modules compiled
from its sources with the harness ref_bindc.F90; it exists only where those
sources were present at build time.

Only tests/, ``__graft_entry__.smoke`` and the CPU-baseline leg of
``bench.py`` import this module; the product never does.

Array conventions (both libraries): fields ``uu``/``bb`` are Fortran-ordered
(dm1, dm2, dm3, 3) float64; ``state`` is Fortran (8, m, np); ``cnt`` is
Fortran (3, np) int64; ``params`` has 32 entries (reference
sources/particles.F90:95,317-333).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libptc_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libptc_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="F_CONTIGUOUS")

METHODS = {"rk5": 0, "dop853": 1, "dp853": 1, "rk8": 1}
MODES = {"vb": 0, "eb": 1}


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _dm(uu):
    return np.ascontiguousarray(uu.shape[:3], dtype=np.int32)


class Oracle:
    """The C restatement (ptc_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        lib = C.CDLL(path)
        lib.orc_interpolate_lin.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.orc_acceleration.argtypes = [C.c_int, C.c_double, C.c_double, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.orc_dti_guess.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp]
        lib.orc_dti_guess.restype = C.c_double
        lib.orc_integrate.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int, _ip, _dp, _dp, _dp, _dp,
                                      _dp, _lp, C.c_void_p, C.c_int64, C.c_int]
        lib.orc_max_threads.restype = C.c_int
        lib.orc_interpolate_series.argtypes = [_ip, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp,
                                               C.c_double, _dp, _dp]
        lib.orc_acceleration_series.argtypes = [C.c_int, C.c_double, C.c_double, _ip, _dp, _dp,
                                                C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp]
        lib.orc_integrate_series.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int, _ip, _dp, _dp,
                                             C.c_int, _dp, _dp, _dp, _dp, _lp, C.c_void_p,
                                             C.c_int64, C.c_int]
        lib.orc_stage_nodes.argtypes = [C.c_int, _dp, _dp]
        lib.orc_stage_nodes.restype = C.c_int
        self.lib = lib

    # ---- time series: uu / bb are (dm1, dm2, dm3, 3, ns), times (ns,) ----
    def interpolate_series(self, uu, bb, times, pos, t, xl=(0, 0, 0), xs=(1, 1, 1)):
        u = np.zeros(3, order="F"); b = np.zeros(3, order="F")
        self.lib.orc_interpolate_series(_dm(uu), _f64(xl), _f64(xs), uu.shape[4], _f64(times),
                                        _f64(uu), _f64(bb), _f64(pos), t, u, b)
        return u, b

    def acceleration_series(self, uu, bb, times, vec, t, qom=1.0, vun=1.0, xl=(0, 0, 0),
                            xs=(1, 1, 1), mode="vb"):
        acc = np.zeros(6, order="F")
        self.lib.orc_acceleration_series(MODES[mode], qom, vun, _dm(uu), _f64(xl), _f64(xs),
                                         uu.shape[4], _f64(times), _f64(uu), _f64(bb), _f64(vec),
                                         t, acc)
        return acc

    def integrate_series(self, method, uu, bb, times, params, time, state, mode="vb",
                         max_steps=0, nthreads=0):
        state = _f64(state).copy(order="F")
        m, np_ = state.shape[1], state.shape[2]
        cnt = np.zeros((3, np_), dtype=np.int64, order="F")
        status = np.zeros(np_, dtype=np.int32)
        self.lib.orc_integrate_series(METHODS[method], MODES[mode], np_, m, _dm(uu), _f64(params),
                                      _f64(time), uu.shape[4], _f64(times), _f64(uu), _f64(bb),
                                      state, cnt, status.ctypes.data, max_steps, nthreads)
        return state, cnt, status

    def stage_nodes(self, method):
        """(c, row sums of the stage matrix) as the oracle holds them."""
        c = np.zeros(12, order="F"); r = np.zeros(12, order="F")
        k = self.lib.orc_stage_nodes(METHODS[method], c, r)
        return c[:k].copy(), r[:k].copy()

    def interpolate_lin(self, uu, bb, pos, xl=(0, 0, 0), xs=(1, 1, 1)):
        u = np.zeros(3, order="F"); b = np.zeros(3, order="F")
        self.lib.orc_interpolate_lin(_dm(uu), _f64(xl), _f64(xs), uu, bb, _f64(pos), u, b)
        return u, b

    def acceleration(self, uu, bb, vec, qom=1.0, vun=1.0, xl=(0, 0, 0), xs=(1, 1, 1), mode="vb"):
        acc = np.zeros(6, order="F")
        self.lib.orc_acceleration(MODES[mode], qom, vun, _dm(uu), _f64(xl), _f64(xs), uu, bb,
                                  _f64(vec), acc)
        return acc

    def dti_guess(self, iord, uu, bb, params, si, mode="vb"):
        return self.lib.orc_dti_guess(iord, MODES[mode], _dm(uu), _f64(params), uu, bb, _f64(si))

    def integrate(self, method, uu, bb, params, time, state, mode="vb", max_steps=0, nthreads=0):
        """Integrates every particle; returns (state, cnt, status).  `state`
        is (8, m, np) Fortran with state[:6, 0, :] the initial conditions."""
        state = _f64(state).copy(order="F")
        m, np_ = state.shape[1], state.shape[2]
        cnt = np.zeros((3, np_), dtype=np.int64, order="F")
        status = np.zeros(np_, dtype=np.int32)
        self.lib.orc_integrate(METHODS[method], MODES[mode], np_, m, _dm(uu), _f64(params),
                               _f64(time), uu, bb, state, cnt, status.ctypes.data, max_steps,
                               nthreads)
        return state, cnt, status

    def max_threads(self):
        return self.lib.orc_max_threads()


class Reference:
    """The reference's own modules (VB mode only), compiled into _ref/."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        lib = C.CDLL(path)
        lib.ref_interpolate_lin.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.ref_acceleration.argtypes = [C.c_double, C.c_double, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.ref_dti_guess.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp]
        lib.ref_dti_guess.restype = C.c_double
        lib.ref_integrate.argtypes = [C.c_int, C.c_int64, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _lp]
        lib.ref_num_threads.restype = C.c_int
        self.lib = lib

    def interpolate_lin(self, uu, bb, pos, xl=(0, 0, 0), xs=(1, 1, 1)):
        u = np.zeros(3, order="F"); b = np.zeros(3, order="F")
        self.lib.ref_interpolate_lin(_dm(uu), _f64(xl), _f64(xs), uu, bb, _f64(pos), u, b)
        return u, b

    def acceleration(self, uu, bb, vec, qom=1.0, vun=1.0, xl=(0, 0, 0), xs=(1, 1, 1)):
        acc = np.zeros(6, order="F")
        self.lib.ref_acceleration(qom, vun, _dm(uu), _f64(xl), _f64(xs), uu, bb, _f64(vec), acc)
        return acc

    def dti_guess(self, iord, uu, bb, params, si):
        return self.lib.ref_dti_guess(iord, _dm(uu), _f64(params), uu, bb, _f64(si))

    def integrate(self, method, uu, bb, params, time, state):
        state = _f64(state).copy(order="F")
        m, np_ = state.shape[1], state.shape[2]
        cnt = np.zeros((3, np_), dtype=np.int64, order="F")
        self.lib.ref_integrate(METHODS[method], np_, m, _dm(uu), _f64(params), _f64(time), uu, bb,
                               state, cnt)
        return state, cnt

    def num_threads(self):
        return self.lib.ref_num_threads()
