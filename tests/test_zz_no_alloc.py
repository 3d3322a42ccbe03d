"""kspace_field_noalloc against kspace_field (the reference-named drop-in).

Two calls in a row through the C ABI on device buffers (allocated here by
the test through the HIP runtime the library itself uses, never by the
entry point): the 648-site water box, then a smaller 81-site subset of it,
each compared with kspace_field on host arrays to 1e-12 relative (both FP64;
the sums over wave vectors run in another order).  Outputs accumulate (+=)
into what the buffers held.  Sizes beyond the array's limits are refused
before anything is queued.
"""
import ctypes as C

import numpy as np
import pytest

from molmodel_amd import waterbox as W
from molmodel_amd.engine import lib

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()   # the library loads its HIP runtime first; the same one is used here
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, a):
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def recip(boxl):
    a = 2 * np.pi / boxl
    return np.array([a, 0, 0, 0, a, 0, 0, 0, a], dtype=np.float64)


def dropin(s, kc, eta):
    L = lib()
    L.kspace_new.restype = C.c_void_p
    L.kspace_new.argtypes = [C.c_double, C.c_int]
    L.kspace_field.argtypes = [C.c_void_p] + [C.c_void_p] * 3 + [C.c_double, C.c_double,
                                                                 C.c_void_p, C.c_void_p]
    L.kspace_delete.argtypes = [C.c_void_p]
    k = L.kspace_new(kc, s.nsite)
    phi = np.zeros(s.nsite)
    ev = np.zeros((s.nsite, 3))
    rl = recip(s.boxl)
    r = np.ascontiguousarray(s.r)
    L.kspace_field(k, r.ctypes.data, s.q.ctypes.data, rl.ctypes.data, s.boxl ** 3, eta,
                   phi.ctypes.data, ev.ctypes.data)
    L.kspace_delete(k)
    return phi, ev


def noalloc(s, kc, eta, phi0, ev0):
    L = lib()
    L.kspace_field_noalloc.argtypes = [C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
    r, q, phi, ev = DevBuf(s.r), DevBuf(s.q), DevBuf(phi0), DevBuf(ev0)
    rl = recip(s.boxl)
    rc = L.kspace_field_noalloc(kc, s.nsite, r.p, q.p, rl.ctypes.data, s.boxl ** 3, eta, phi.p,
                                ev.p, None)
    assert rc == 0, L.mm_last_error()
    assert hip().hipDeviceSynchronize() == 0
    out = phi.get(), ev.get()
    for b in (r, q, phi, ev):
        b.free()
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_noalloc_matches_kspace_field_twice(gpu_lib):
    big, p = W.config("water648_ewald")
    small = W.subset(big, np.arange(27))
    kc, eta = p["kc"], p["eta"]
    for tag, s in (("648 sites", big), ("81 sites", small)):
        print(f"[noalloc] {tag}", flush=True)
        phi_ref, ev_ref = dropin(s, kc, eta)
        phi0 = np.full(s.nsite, 1.5)                  # += onto what the buffers hold
        ev0 = np.full((s.nsite, 3), -0.25)
        phi, ev = noalloc(s, kc, eta, phi0, ev0)
        ep, ee = rel(phi - phi0, phi_ref), rel(ev - ev0, ev_ref)
        print(f"[noalloc] {tag}: phi rel {ep:.2e}, evec rel {ee:.2e}", flush=True)
        assert ep <= 1e-12 and ee <= 1e-12


def test_noalloc_refuses_sizes_beyond_its_array(gpu_lib):
    L = lib()
    L.kspace_field_noalloc.argtypes = [C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                       C.c_void_p]
    rl = recip(20.0)
    # checked on the host before anything is queued: no device pointer is read
    assert L.kspace_field_noalloc(7.0, 65537, None, None, rl.ctypes.data, 8000.0, 0.35, None,
                                  None, None) == -1
    assert b"nsite" in L.mm_last_error()
    assert L.kspace_field_noalloc(12.5, 10, None, None, rl.ctypes.data, 8000.0, 0.35, None,
                                  None, None) == -1
    assert b"wave vectors" in L.mm_last_error()
