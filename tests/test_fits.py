"""FITS primary-image I/O of the host driver (CPU only): round trips, every
BITPIX the reader accepts, gzip, header keys, and the reference's file
layouts (counter int64 3 x np, time FP64 1-D with TUNIT, state FP64
8 x m x np with TEMP)."""
import gzip
import os

import numpy as np
import pytest

from particles_amd import fits


def _numpy_fits(path, arr, bitpix, extra=()):
    """A FITS file written independently of the library (big-endian, 2880-byte
    blocks), Fortran order."""
    cards = [f"{'SIMPLE':8}= {'T':>20}", f"{'BITPIX':8}= {bitpix:>20}",
             f"{'NAXIS':8}= {arr.ndim:>20}"]
    cards += [f"{'NAXIS%d' % (i + 1):8}= {n:>20}" for i, n in enumerate(arr.shape)]
    cards += list(extra) + ["END"]
    hdr = "".join(c.ljust(80) for c in cards)
    hdr += " " * ((2880 - len(hdr) % 2880) % 2880)
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    data = np.asfortranarray(arr).astype(dt).tobytes(order="F")
    data += b"\0" * ((2880 - len(data) % 2880) % 2880)
    with open(path, "wb") as fh:
        fh.write(hdr.encode("ascii") + data)


@pytest.mark.parametrize("bitpix", [8, 16, 32, 64, -32, -64])
def test_reader_every_bitpix(tmp_path, bitpix):
    rng = np.random.default_rng(bitpix + 100)
    a = rng.uniform(0, 100, (5, 4, 3))
    if bitpix > 0:
        a = np.round(a)
    elif bitpix == -32:
        a = a.astype(np.float32).astype(np.float64)
    p = str(tmp_path / "f.fits")
    _numpy_fits(p, a, bitpix)
    assert fits.dims(p) == ((5, 4, 3), bitpix)
    assert np.array_equal(fits.read(p), a)


def test_reader_applies_bscale_bzero(tmp_path):
    a = np.arange(12, dtype=np.float64).reshape(3, 4, order="F")
    p = str(tmp_path / "s.fits")
    _numpy_fits(p, a, 16, extra=(f"{'BSCALE':8}= {'0.5':>20}", f"{'BZERO':8}= {'10.0':>20}"))
    assert np.array_equal(fits.read(p), 10.0 + 0.5 * a)


def test_fp64_field_is_read_at_full_precision(tmp_path):
    """A BITPIX -64 field keeps every bit (the reference reads FP64 fields
    through the FP32 call, sources/fitsio.F90:255-260)."""
    a = np.random.default_rng(0).normal(size=(7, 6, 5))
    p = str(tmp_path / "magx.fits")
    _numpy_fits(p, a, -64)
    assert np.array_equal(fits.read(p), a)
    assert not np.array_equal(a.astype(np.float32).astype(np.float64), a)


def test_gzip_files(tmp_path):
    a = np.random.default_rng(1).normal(size=(4, 4, 4))
    p = str(tmp_path / "velx.fits")
    _numpy_fits(p, a, -64)
    with open(p, "rb") as src, gzip.open(p + ".gz", "wb") as dst:
        dst.write(src.read())
    assert np.array_equal(fits.read(p + ".gz"), a)


def test_writer_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(2)
    state = np.asfortranarray(rng.normal(size=(8, 5, 7)))
    counter = np.asfortranarray(rng.integers(0, 2 ** 40, size=(3, 7)))
    time = np.linspace(0, 1, 5)
    ps, pc, pt = (str(tmp_path / n) for n in ("state.fits", "counter.fits", "time.fits"))
    fits.write_f64("!" + ps, state, {"TEMP": 1.0e3}, {"TEMP": "Initial thermal temperature"})
    fits.write_i64("!" + pc, counter)
    fits.write_f64("!" + pt, time, {"TUNIT": 2.5}, {"TUNIT": "Time unit in seconds"})
    for p in (ps, pc, pt):
        assert os.path.getsize(p) % 2880 == 0
    assert np.array_equal(fits.read(ps), state)
    assert np.array_equal(fits.read(pc), counter.astype(np.float64))
    assert np.array_equal(fits.read(pt), time)
    assert fits.dims(pc) == ((3, 7), 64) and fits.dims(ps) == ((8, 5, 7), -64)
    h = fits.header(pt)
    assert float(h["TUNIT"]) == 2.5 and h["BITPIX"] == "-64" and h["NAXIS1"] == "5"
    assert float(fits.header(ps)["TEMP"]) == 1.0e3
    # data written once: file = header block(s) + exactly the data blocks
    assert os.path.getsize(pt) == 2880 + 2880
    # big-endian on disk
    raw = open(pc, "rb").read()
    first = int.from_bytes(raw[2880:2888], "big", signed=True)
    assert first == counter[0, 0]


def test_errors(tmp_path):
    with pytest.raises(fits.FitsError):
        fits.read(str(tmp_path / "missing.fits"))
    p = str(tmp_path / "bad.fits")
    open(p, "wb").write(b"NOT FITS" * 400)
    with pytest.raises(fits.FitsError):
        fits.read(p)


# ---- interoperability with CFITSIO, the library the reference program reads
# and writes its files with (sources/fitsio.F90:188-191,248-265,300-308) ----

def _cfitsio():
    import ctypes as C
    for name in ("/opt/conda/lib/libcfitsio.so", "libcfitsio.so", "libcfitsio.so.10",
                 "libcfitsio.so.9", "libcfitsio.so.8"):
        try:
            lib = C.CDLL(name)
        except OSError:
            continue
        ll = C.POINTER(C.c_longlong)
        pi = C.POINTER(C.c_int)
        lib.ffopen.argtypes = [C.POINTER(C.c_void_p), C.c_char_p, C.c_int, pi]
        lib.ffinit.argtypes = [C.POINTER(C.c_void_p), C.c_char_p, pi]
        lib.ffgidm.argtypes = [C.c_void_p, pi, pi]
        lib.ffgidt.argtypes = [C.c_void_p, pi, pi]
        lib.ffgiszll.argtypes = [C.c_void_p, C.c_int, ll, pi]
        lib.ffgpvd.argtypes = [C.c_void_p, C.c_long, C.c_longlong, C.c_longlong, C.c_double,
                               C.c_void_p, pi, pi]
        lib.ffgpvjj.argtypes = [C.c_void_p, C.c_long, C.c_longlong, C.c_longlong, C.c_longlong,
                                C.c_void_p, pi, pi]
        lib.ffgkyd.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_char_p, pi]
        lib.ffcrimll.argtypes = [C.c_void_p, C.c_int, C.c_int, ll, pi]
        lib.ffpprd.argtypes = [C.c_void_p, C.c_long, C.c_longlong, C.c_longlong, C.c_void_p, pi]
        lib.ffpkyd.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_int, C.c_char_p, pi]
        lib.ffclos.argtypes = [C.c_void_p, pi]
        return lib
    pytest.skip("CFITSIO not installed")


def _cf_read(lib, path, key=None):
    """(array in Fortran order, bitpix, key value) read by CFITSIO."""
    import ctypes as C
    f, st = C.c_void_p(), C.c_int(0)
    lib.ffopen(C.byref(f), path.encode(), 0, C.byref(st))
    assert st.value == 0, f"ffopen status {st.value}"
    nax, bp = C.c_int(), C.c_int()
    lib.ffgidm(f, C.byref(nax), C.byref(st))
    lib.ffgidt(f, C.byref(bp), C.byref(st))
    dims = (C.c_longlong * nax.value)()
    lib.ffgiszll(f, nax.value, dims, C.byref(st))
    shape = tuple(int(d) for d in dims)
    n = int(np.prod(shape))
    anynul = C.c_int()
    if bp.value == 64:
        out = np.zeros(n, dtype=np.int64)
        lib.ffgpvjj(f, 0, 1, n, 0, out.ctypes.data, C.byref(anynul), C.byref(st))
    else:
        out = np.zeros(n, dtype=np.float64)
        lib.ffgpvd(f, 0, 1, n, 0.0, out.ctypes.data, C.byref(anynul), C.byref(st))
    val = None
    if key is not None:
        v = C.c_double()
        com = C.create_string_buffer(81)
        lib.ffgkyd(f, key.encode(), C.byref(v), com, C.byref(st))
        val = v.value
    lib.ffclos(f, C.byref(st))
    assert st.value == 0, f"CFITSIO status {st.value}"
    return out.reshape(shape, order="F"), bp.value, val


def test_cfitsio_reads_what_the_library_writes(tmp_path):
    """counter (int64 3 x np), time (FP64, TUNIT) and state (FP64 8 x m x np,
    TEMP) files written by the library, read back by CFITSIO as the
    reference's users read them."""
    lib = _cfitsio()
    rng = np.random.default_rng(3)
    state = np.asfortranarray(rng.normal(size=(8, 5, 7)))
    counter = np.asfortranarray(rng.integers(-2 ** 62, 2 ** 62, size=(3, 7)))
    time = np.linspace(0, 1, 5)
    ps, pc, pt = (str(tmp_path / n) for n in ("state.fits", "counter.fits", "time.fits"))
    fits.write_f64("!" + ps, state, {"TEMP": 1.0e3}, {"TEMP": "Initial thermal temperature"})
    fits.write_i64("!" + pc, counter)
    fits.write_f64("!" + pt, time, {"TUNIT": 2.5e-6}, {"TUNIT": "Time unit in seconds"})
    s, bp, temp = _cf_read(lib, ps, "TEMP")
    assert bp == -64 and s.shape == state.shape and np.array_equal(s, state) and temp == 1.0e3
    c, bp, _ = _cf_read(lib, pc)
    assert bp == 64 and c.shape == counter.shape and np.array_equal(c, counter)
    t, bp, tu = _cf_read(lib, pt, "TUNIT")
    assert bp == -64 and np.array_equal(t, time) and tu == 2.5e-6


def test_library_reads_what_cfitsio_writes(tmp_path):
    """A 3-D FP64 field file written by CFITSIO (as a simulation code would
    write velx.fits ...) read by the library: identical values and shape."""
    import ctypes as C
    lib = _cfitsio()
    a = np.asfortranarray(np.random.default_rng(4).normal(size=(6, 5, 4)))
    p = str(tmp_path / "magx.fits")
    f, st = C.c_void_p(), C.c_int(0)
    lib.ffinit(C.byref(f), p.encode(), C.byref(st))
    dims = (C.c_longlong * 3)(*a.shape)
    lib.ffcrimll(f, -64, 3, dims, C.byref(st))
    flat = np.ascontiguousarray(a.ravel(order="F"))
    lib.ffpprd(f, 0, 1, flat.size, flat.ctypes.data, C.byref(st))
    lib.ffpkyd(f, b"BUNIT", 1.5, 10, b"field unit", C.byref(st))
    lib.ffclos(f, C.byref(st))
    assert st.value == 0
    assert fits.dims(p) == (a.shape, -64)
    assert np.array_equal(fits.read(p), a)
    assert float(fits.header(p)["BUNIT"]) == 1.5
