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
