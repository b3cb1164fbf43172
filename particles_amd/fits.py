"""FITS primary-image I/O of the library (include/particles_amd_fits.h) from
Python: the field files the reference reads and the counter / time / state
files it writes (reference sources/fitsio.F90).  Arrays are returned and
taken in Fortran order, shape (NAXIS1, NAXIS2, ...), like the reference's."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import library_path

_lib = C.CDLL(library_path)
_lib.ptc_fits_dims.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_int)]
_lib.ptc_fits_read.argtypes = [C.c_char_p, C.c_int64, C.c_void_p]
_lib.ptc_fits_write_f64.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
_lib.ptc_fits_write_i64.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p]
_lib.ptc_fits_last_error.restype = C.c_char_p
for _f in ("ptc_fits_dims", "ptc_fits_read", "ptc_fits_write_f64", "ptc_fits_write_i64"):
    getattr(_lib, _f).restype = C.c_int


class FitsError(RuntimeError):
    pass


def _check(rc):
    if rc:
        raise FitsError(_lib.ptc_fits_last_error().decode())


def dims(path: str):
    """(shape, bitpix) of the primary image."""
    n = C.c_int()
    d = np.zeros(8, dtype=np.int64)
    bp = C.c_int()
    _check(_lib.ptc_fits_dims(path.encode(), C.byref(n), d.ctypes.data, C.byref(bp)))
    return tuple(int(x) for x in d[:n.value]), bp.value


def read(path: str) -> np.ndarray:
    """The primary image as FP64, Fortran order."""
    shape, _ = dims(path)
    out = np.zeros(shape, dtype=np.float64, order="F")
    _check(_lib.ptc_fits_read(path.encode(), out.size, out.ctypes.data))
    return out


def write_f64(path: str, data, keys: dict | None = None, comments: dict | None = None) -> None:
    a = np.asfortranarray(data, dtype=np.float64)
    d = np.asarray(a.shape, dtype=np.int64)
    keys = keys or {}
    names = [k.encode() for k in keys]
    vals = np.asarray([float(v) for v in keys.values()], dtype=np.float64)
    com = [(comments or {}).get(k, "").encode() for k in keys]
    karr = (C.c_char_p * max(len(names), 1))(*names)
    carr = (C.c_char_p * max(len(com), 1))(*com)
    _check(_lib.ptc_fits_write_f64(path.encode(), a.ndim, d.ctypes.data, a.ctypes.data, len(names),
                                   karr, vals.ctypes.data if len(vals) else None, carr))


def write_i64(path: str, data) -> None:
    a = np.asfortranarray(data, dtype=np.int64)
    d = np.asarray(a.shape, dtype=np.int64)
    _check(_lib.ptc_fits_write_i64(path.encode(), a.ndim, d.ctypes.data, a.ctypes.data))


def header(path: str) -> dict:
    """Header cards of the primary HDU as {key: value string} (for checks)."""
    import gzip
    opener = gzip.open if path.endswith(".gz") else open
    out = {}
    with opener(path, "rb") as fh:
        while True:
            card = fh.read(80).decode("ascii")
            if len(card) < 80 or card.startswith("END"):
                return out
            if card[8:10] == "= ":
                out[card[:8].strip()] = card[10:].split("/")[0].strip()
