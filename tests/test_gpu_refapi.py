"""GPU parity of the reference-named drop-ins (include/molmodel_amd_ref.h):
RSpaceEwald, kspace_new/kspace_field, p3m_new/p3m_init/p3m_field, called
exactly as ElecLJ calls them (eleclj.cpp:932-950), against the FP64 oracle."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import engine as E
from molmodel_amd import waterbox as W

pytestmark = pytest.mark.gpu
P = C.c_void_p
D = C.c_double
I = C.c_int


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def lib():
    L = E.lib()
    L.PairBinary_new.restype = P
    L.PairBinary_new.argtypes = [I]
    L.PairBinary_add.argtypes = [P, I, I]
    L.PairBinary_init.argtypes = [P]
    L.PairBinary_delete.argtypes = [P]
    L.RSpaceChebyshev_new.restype = P
    L.RSpaceChebyshev_new.argtypes = [D, D, P, P]
    L.RSpaceEwald.argtypes = [I, I, P, P, I, P, P, P, P, P, D, D, P, P, P, P, D, P, P, P, P, I,
                              P, P, P, D, D]
    L.kspace_new.restype = P
    L.kspace_new.argtypes = [D, I]
    L.kspace_field.argtypes = [P, P, P, P, D, D, P, P]
    L.kspace_delete.argtypes = [P]
    L.kspace_number_of_wavevectors.argtypes = [P]
    L.p3m_new.restype = P
    L.p3m_new.argtypes = [I, P, D, I, I, I, P]
    L.p3m_init.argtypes = [P, D, P]
    L.p3m_field.argtypes = [P, P, P, P, P]
    L.p3m_delete.argtypes = [P]
    return L


def test_rspace_ewald_dropin(gpu_lib):
    L = lib()
    s, p = W.config("water648_ewald")
    gr = O.first_sites(s)
    pairs = O.nl_pairs(gr, s.boxl, p["rc"], p["skin"])
    # reference-shaped neighbour rows
    rows = [[] for _ in range(s.ngroup)]
    for a, b in pairs:
        rows[a].append(b)
    arrs = [np.array(r if r else [0], dtype=np.int32) for r in rows]
    ptrs = (C.POINTER(C.c_int) * s.ngroup)(*[a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs])
    nn = np.array([len(r) for r in rows], dtype=np.int32)
    excl = L.PairBinary_new(s.nsite)
    for a, b in s.excl:
        L.PairBinary_add(excl, int(a), int(b))
    L.PairBinary_init(excl)
    gexcl = L.PairBinary_new(s.ngroup)
    L.PairBinary_init(gexcl)
    cheb = L.RSpaceChebyshev_new(p["rc"], p["eta"], None, None)
    latv = np.array([s.boxl, 0, 0, 0, s.boxl, 0, 0, 0, s.boxl])
    phi = np.zeros(s.nsite); ev = np.zeros((s.nsite, 3)); ljf = np.zeros((s.nsite, 3))
    lje = np.zeros(1)
    grp = np.ascontiguousarray(s.group)
    L.RSpaceEwald(s.nsite, s.ngroup, grp.ctypes.data_as(P), s.r.ctypes.data_as(P), 1,
                  latv.ctypes.data_as(P), C.cast(ptrs, P), nn.ctypes.data_as(P), gexcl, excl,
                  (p["rc"] - p["smoothing_width"]) ** 2, p["rc"] ** 2,
                  s.site_mask.ctypes.data_as(P), s.q.ctypes.data_as(P), phi.ctypes.data_as(P),
                  ev.ctypes.data_as(P), p["eta"], cheb, None, s.sig.ctypes.data_as(P),
                  s.eps.ctypes.data_as(P), 0, lje.ctypes.data_as(P), ljf.ctypes.data_as(P), None,
                  1.0, 1.0)
    ophi, oev, oljf, olje = O.rspace_parts(s, p)
    print("dropin rspace", rel(phi, ophi), rel(ev, oev), rel(ljf, oljf), abs(lje[0] - olje) / abs(olje))
    assert rel(phi, ophi) < 2e-6 and rel(ev, oev) < 5e-6 and rel(ljf, oljf) < 5e-6
    assert abs(lje[0] - olje) < 1e-5 * abs(olje)


def test_rspace_ewald_phi_dropin(gpu_lib):
    """RSpaceEwaldPhi with molecules straddling the box edge (the excluded
    correction without the minimum image) against the reference's output."""
    L = lib()
    L.RSpaceEwaldPhi.argtypes = [I, I, P, P, I, P, P, P, P, P, D, D, P, P, P, D, P, P, P, D]
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_water648_phi.npz"))
    s, p = W.config("water648_ewald")
    r = np.ascontiguousarray(g["r"])
    gr = O.first_sites(s, r)
    pairs = O.nl_pairs(gr, s.boxl, p["rc"], p["skin"])
    rows = [[] for _ in range(s.ngroup)]
    for a, b in pairs:
        rows[a].append(b)
    arrs = [np.array(x if x else [0], dtype=np.int32) for x in rows]
    ptrs = (C.POINTER(C.c_int) * s.ngroup)(*[a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs])
    nn = np.array([len(x) for x in rows], dtype=np.int32)
    excl = L.PairBinary_new(s.nsite)
    for a, b in s.excl:
        L.PairBinary_add(excl, int(a), int(b))
    L.PairBinary_init(excl)
    gexcl = L.PairBinary_new(s.ngroup)
    L.PairBinary_init(gexcl)
    latv = np.array([s.boxl, 0, 0, 0, s.boxl, 0, 0, 0, s.boxl])
    phi = np.zeros(s.nsite)
    grp = np.ascontiguousarray(s.group)
    L.RSpaceEwaldPhi(s.nsite, s.ngroup, grp.ctypes.data_as(P), r.ctypes.data_as(P), 1,
                     latv.ctypes.data_as(P), C.cast(ptrs, P), nn.ctypes.data_as(P), gexcl, excl,
                     (p["rc"] - p["smoothing_width"]) ** 2, p["rc"] ** 2,
                     s.site_mask.ctypes.data_as(P), s.q.ctypes.data_as(P), phi.ctypes.data_as(P),
                     p["eta"], None, None, None, 1.0)
    L.PairBinary_delete(excl)
    L.PairBinary_delete(gexcl)
    print("RSpaceEwaldPhi rel", rel(phi, g["phi_only"]), "straddling sites",
          rel(phi[:30], g["phi_only"][:30]))
    assert rel(phi, g["phi_only"]) < 2e-6
    assert rel(phi[:30], g["phi_only"][:30]) < 2e-6


def test_phi_at_gridpoints_dropin(gpu_lib):
    L = lib()
    L.RSpaceEwaldPhiAtGridpoints.argtypes = [I, I, P, P, I, P, D, D, P, P, P, D, P, P]
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_esp_grid.npz"))
    s, p = W.config("water648_ewald")
    grid = np.ascontiguousarray(g["grid"])
    latv = np.array([s.boxl, 0, 0, 0, s.boxl, 0, 0, 0, s.boxl])
    phi = np.zeros(len(grid))
    L.RSpaceEwaldPhiAtGridpoints(s.nsite, len(grid), s.r.ctypes.data_as(P), grid.ctypes.data_as(P),
                                 1, latv.ctypes.data_as(P), (p["rc"] - p["smoothing_width"]) ** 2,
                                 p["rc"] ** 2, s.site_mask.ctypes.data_as(P),
                                 s.q.ctypes.data_as(P), phi.ctypes.data_as(P), p["eta"], None, None)
    err = np.abs(phi - g["phi"]).max() / np.abs(g["phi"]).max()
    print("phi at gridpoints max err / max |phi|", err, "rel", rel(phi, g["phi"]))
    assert err < 2e-6


def test_kspace_dropin(gpu_lib):
    L = lib()
    s, p = W.config("water648_ewald")
    k = L.kspace_new(p["kc"], s.nsite)
    assert L.kspace_number_of_wavevectors(k) == 709
    rl = O.recip_cubic(s.boxl)
    phi = np.zeros(s.nsite); ev = np.zeros((s.nsite, 3))
    L.kspace_field(k, s.r.ctypes.data_as(P), s.q.ctypes.data_as(P), rl.ctypes.data_as(P),
                   s.boxl ** 3, p["eta"], phi.ctypes.data_as(P), ev.ctypes.data_as(P))
    L.kspace_delete(k)
    ophi, oev, _, _ = O.kspace_parts(s, p["kc"], p["eta"])
    assert rel(phi, ophi) < 1e-11 and rel(ev, oev) < 1e-11


def test_p3m_dropin(gpu_lib):
    L = lib()
    s, p = W.config("water12k_p3m")
    latv = np.array([s.boxl, 0, 0, 0, s.boxl, 0, 0, 0, s.boxl])
    ng = np.zeros(3, dtype=np.int32)
    h = s.boxl / 48
    pp = L.p3m_new(s.nsite, latv.ctypes.data_as(P), h, 2, 0, 5, ng.ctypes.data_as(P))
    assert list(ng) == [48, 48, 48]
    chi = np.zeros(1)
    L.p3m_init(pp, p["eta"], chi.ctypes.data_as(P))
    _, ochi = O.green((48, 48, 48), s.boxl, p["eta"], 5)
    assert abs(chi[0] - ochi) < 1e-10 * ochi
    phi = np.zeros(s.nsite); ev = np.zeros((s.nsite, 3))
    L.p3m_field(pp, s.r.ctypes.data_as(P), s.q.ctypes.data_as(P), phi.ctypes.data_as(P),
                ev.ctypes.data_as(P))
    L.p3m_delete(pp)
    ophi, oev, _, _ = O.p3m_parts(s, p)
    print("dropin p3m", rel(phi, ophi), rel(ev, oev))
    assert rel(phi, ophi) < 1e-5 and rel(ev, oev) < 1e-5
