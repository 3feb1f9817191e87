"""The allocation-free CFL entry point (fv_cfl_dt_noalloc) against fv_dev_cfl_dt:
the same kernels on caller-owned device arrays, bit for bit, on two calls in a
row (the second on a smaller mesh, so that stale partial minima of the first
call would show), plus the size limit."""
import ctypes as C

import numpy as np
import pytest

import forestfv_public_amd as F
from cases import bubble_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class HipBuf:
    """Caller-owned device memory through the HIP runtime the library itself uses."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0   # host to device

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0              # device to host
        return out

    def free(self):
        self.hip().hipFree(self.ptr)


def noalloc_dt(p, m, S, cfl):
    pk = m.pack()
    S_dev = HipBuf(np.ascontiguousarray(S.T))          # [NEQ][n], the context's layout
    info_dev = HipBuf(pk.arrays()["info"])
    dt_dev = HipBuf(np.full(1, -1.0))
    rc = F.lib().fv_cfl_dt_noalloc(C.byref(p), C.byref(pk.pk.geo), m.n, S_dev.ptr, info_dev.ptr, cfl, dt_dev.ptr,
                                   None)
    assert rc == F.FV_OK, F.lib().fv_dev_last_error()
    d = float(dt_dev.read()[0])
    for b in (S_dev, info_dev, dt_dev):
        b.free()
    return d


def context_dt(p, m, S, cfl):
    dev = F.Device(p, m)
    dev.set_state(S)
    d = dev.cfl_dt(cfl)
    dev.close()
    return d


def test_noalloc_cfl_equals_the_context_entry_point_on_two_calls():
    for level, cfl in ((6, 0.4), (4, 0.7)):
        p, m, S = bubble_case(level=level)
        say(f"[noalloc] level={level} n={m.n} cfl={cfl}")
        want = context_dt(p, m, S, cfl)
        got = noalloc_dt(p, m, S, cfl)
        say(f"[noalloc] context {want!r} noalloc {got!r}")
        assert got == want


def test_noalloc_cfl_three_materials():
    from cases import three_material_case
    p, m, S = three_material_case(level=4)
    assert noalloc_dt(p, m, S, 0.5) == context_dt(p, m, S, 0.5)


def test_noalloc_cfl_rejects_sizes_beyond_its_scratch():
    p, m, S = bubble_case(level=3)
    pk = m.pack()
    buf = C.c_double(0.0)
    rc = F.lib().fv_cfl_dt_noalloc(C.byref(p), C.byref(pk.pk.geo), (1 << 26) + 1, C.c_void_p(C.addressof(buf)),
                                   C.c_void_p(C.addressof(buf)), 0.4, C.c_void_p(C.addressof(buf)), None)
    assert rc == F.FV_ERR_ARG
    assert b"exceed" in F.lib().fv_dev_last_error()
