"""ptc_integrate_noalloc through the C-ABI: the allocation-free form of
ptc_integrate (device pointers in and out, the batch state in the library's
fixed per-device workspace) gives ptc_integrate's results bit for bit, on
two calls in a row with the second smaller (the workspace keeps the first
call's contents, which the second must not see), and refuses a batch larger
than its workspace."""
import ctypes as C

import numpy as np
import pytest

import particles_amd as P
from particles_amd import workloads as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    for f in (lib.hipMalloc, lib.hipFree, lib.hipMemcpy, lib.hipDeviceSynchronize):
        f.restype = C.c_int
    return lib


def _noalloc():
    lib = C.CDLL(P.library_path)
    fn = lib.ptc_integrate_noalloc
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    lib.ptc_last_error.restype = C.c_char_p
    return fn, lib


class DeviceArray:
    """A device copy of a host array (the caller's buffer)."""

    def __init__(self, hip, host):
        self.hip, self.shape, self.dtype = hip, host.shape, host.dtype
        self.nbytes = host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.nbytes, 8)) == 0
        h = np.asfortranarray(host)
        assert hip.hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _run_noalloc(fields, method, p, tg, st0, with_status=True):
    hip = _hip()
    fn, _ = _noalloc()
    np_ = st0.shape[2]
    bufs = [DeviceArray(hip, np.ascontiguousarray(tg, dtype=np.float64)),
            DeviceArray(hip, np.asfortranarray(st0, dtype=np.float64)),
            DeviceArray(hip, np.full((3, np_), -7, dtype=np.int64)),
            DeviceArray(hip, np.full(np_, -7, dtype=np.int32))]
    pp = np.ascontiguousarray(p, dtype=np.float64)
    try:
        rc = fn(fields.handle, P.METHODS[method], np_, len(tg), pp.ctypes.data, pp.size,
                bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr if with_status else None, None)
        assert rc == 0, P._lib.ptc_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        return bufs[1].get(), bufs[2].get(), bufs[3].get()
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("layout", ["aos", "columns"])
def test_noalloc_matches_integrate_twice_then_smaller(layout):
    """Node rows (one lane per particle with the hand-over) and column rows
    (the trio kernel)."""
    uu, bb = W.turbulent_vb(16, seed=21, nmodes=64)
    old = P.get_option("layout")
    P.set_option("layout", layout)
    try:
        f = P.Fields(uu, bb)
    finally:
        P.set_option("layout", old)
    with f:
        # first call: DOP853, 3,000 particles, four snapshots
        p = W.params_vector(tmax=0.5 * W.T0, dtmax=1.0, atol=1e-8, rtol=1e-8)
        tg = W.time_grid(3, p[2])
        st0 = W.initial_state(3000, len(tg), seed=22, gamma_range=(1.0, 1000.0))
        s_ref, c_ref, t_ref = P.integrate(f, "dop853", p, tg, st0)
        s, c, t = _run_noalloc(f, "dop853", p, tg, st0)
        print(f"noalloc {layout} dop853 3000: {int(c[1].sum())} accepted steps", flush=True)
        assert (t_ref == 0).all()
        assert np.array_equal(s, s_ref) and np.array_equal(c, c_ref) and np.array_equal(t, t_ref)
        # second call, smaller: Cash-Karp, 700 other particles, another box,
        # status not requested
        g = W.params_vector(vun=0.7, tmax=0.4 * W.T0, dtmax=1.0, atol=1e-7, rtol=1e-7,
                            xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
        tg2 = W.time_grid(2, g[2])
        st2 = W.initial_state(700, len(tg2), seed=23, gamma_range=(1.0, 100.0),
                              xl=(-1.0, 0.5, 2.0), xs=(2.5, 3.0, 1.75))
        s_ref, c_ref, t_ref = P.integrate(f, "rk5", g, tg2, st2)
        s, c, t = _run_noalloc(f, "rk5", g, tg2, st2, with_status=False)
        print(f"noalloc {layout} rk5 700: {int(c[1].sum())} accepted steps", flush=True)
        assert np.array_equal(s, s_ref) and np.array_equal(c, c_ref)
        assert (t == -7).all()  # status not written when not requested


def test_noalloc_refuses_batches_beyond_its_workspace():
    fn, lib = _noalloc()
    uu, bb = W.turbulent_vb(8, seed=1, nmodes=16)
    p = np.ascontiguousarray(W.params_vector(tmax=1.0, dtmax=1.0), dtype=np.float64)
    with P.Fields(uu, bb) as f:
        rc = fn(f.handle, 1, 65537, 2, p.ctypes.data, p.size, None, None, None, None, None)
        assert rc == 1 and b"65536" in lib.ptc_last_error()
        rc = fn(f.handle, 1, -1, 2, p.ctypes.data, p.size, None, None, None, None, None)
        assert rc == 1
        # nothing to do: no pointers needed
        rc = fn(f.handle, 1, 0, 2, p.ctypes.data, p.size, None, None, None, None, None)
        assert rc == 0
