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


def np_lib():
    L = lib()
    L.RSpaceNonPeriodic.argtypes = [I, I, P, P, P, P, P, P, P, P, P, P, P, I, P, P, P, D, D]
    L.RSpaceNonPeriodicPhi.argtypes = [I, I, P, P, P, P, P, P, P, P, P, D]
    L.RSpaceNonPeriodicPhiAtGridpoints.argtypes = [I, I, P, P, P, P, P, P]
    return L


def test_nonperiodic_dropins(gpu_lib):
    """RSpaceNonPeriodic, RSpaceNonPeriodicPhi and
    RSpaceNonPeriodicPhiAtGridpoints (FP64 on the device) against the oracle,
    which equals the compiled reference bit for bit on this cluster: 1e-12
    relative (summation order differs); outputs accumulate (+=)."""
    L = np_lib()
    s, rs = O.cluster_system()
    grid = O.cluster_grid(s)
    n, ng = s.nsite, len(grid)
    r = np.ascontiguousarray(s.r)
    group = np.ascontiguousarray(s.group)
    phi = np.full(n, 0.5); ev = np.zeros((n, 3)); ljf = np.zeros((n, 3)); lje = np.full(1, 2.0)
    phi_only = np.zeros(n); gphi = np.full(ng, -1.0)
    print("[refapi] non-periodic drop-ins on a 124-site cluster", flush=True)
    L.RSpaceNonPeriodic(n, s.ngroup, group.ctypes.data_as(P), r.ctypes.data_as(P), None, None,
                        s.site_mask.ctypes.data_as(P), s.q.ctypes.data_as(P),
                        phi.ctypes.data_as(P), ev.ctypes.data_as(P), rs.ctypes.data_as(P),
                        s.sig.ctypes.data_as(P), s.eps.ctypes.data_as(P), 0,
                        lje.ctypes.data_as(P), ljf.ctypes.data_as(P), None, 1.0, 1.0)
    L.RSpaceNonPeriodicPhi(n, s.ngroup, group.ctypes.data_as(P), r.ctypes.data_as(P), None,
                           None, s.site_mask.ctypes.data_as(P), s.q.ctypes.data_as(P),
                           phi_only.ctypes.data_as(P), rs.ctypes.data_as(P), None, 1.0)
    g = np.ascontiguousarray(grid)
    L.RSpaceNonPeriodicPhiAtGridpoints(n, ng, r.ctypes.data_as(P), g.ctypes.data_as(P),
                                       s.site_mask.ctypes.data_as(P), s.q.ctypes.data_as(P),
                                       gphi.ctypes.data_as(P), rs.ctypes.data_as(P))
    o = O.nonperiodic_parts(s, rs, grid)
    errs = [rel(phi - 0.5, o[0]), rel(ev, o[1]), rel(ljf, o[2]), abs(lje[0] - 2.0 - o[3]) / abs(o[3]),
            rel(phi_only, o[4]), rel(gphi + 1.0, o[5])]
    print("[refapi] non-periodic rel errors", errs)
    assert max(errs) <= 1e-12


def test_kspace_phi_dropin(gpu_lib):
    """kspace_phi (kspace.c:176-251): the potential of kspace_field alone."""
    L = lib()
    L.kspace_phi.argtypes = [P, P, P, P, D, D, P]
    s, p = W.config("water648_ewald")
    k = L.kspace_new(p["kc"], s.nsite)
    rl = O.recip_cubic(s.boxl)
    phi = np.zeros(s.nsite)
    L.kspace_phi(k, s.r.ctypes.data_as(P), s.q.ctypes.data_as(P), rl.ctypes.data_as(P),
                 s.boxl ** 3, p["eta"], phi.ctypes.data_as(P))
    L.kspace_delete(k)
    ophi, _, _, _ = O.kspace_parts(s, p["kc"], p["eta"])
    assert rel(phi, ophi) < 1e-11


def p3m_handle(L, s, mesh, eta, order=5, ik=0):
    latv = np.array([s.boxl, 0, 0, 0, s.boxl, 0, 0, 0, s.boxl])
    ng = np.zeros(3, dtype=np.int32)
    pp = L.p3m_new(s.nsite, latv.ctypes.data_as(P), s.boxl / mesh, 2, ik, order,
                   ng.ctypes.data_as(P))
    assert list(ng) == [mesh] * 3
    chi = np.zeros(1)
    L.p3m_init(pp, eta, chi.ctypes.data_as(P))
    return pp, chi[0]


def p3m_run(L, pp, s, r=None, evec=True):
    r = np.ascontiguousarray(s.r if r is None else r)
    phi = np.zeros(s.nsite)
    ev = np.zeros((s.nsite, 3)) if evec else None
    L.p3m_field(pp, r.ctypes.data_as(P), s.q.ctypes.data_as(P), phi.ctypes.data_as(P),
                ev.ctypes.data_as(P) if evec else None)
    return phi, ev


def test_p3m_field_phi_only(gpu_lib):
    """p3m_field(..., evec = NULL) gives the potential alone (pppm.h:48-51)."""
    L = lib()
    s, p = W.config("water648_ewald")
    pp, _ = p3m_handle(L, s, 24, 0.35)
    phi, _ = p3m_run(L, pp, s, evec=False)
    phi2, ev2 = p3m_run(L, pp, s)
    L.p3m_delete(pp)
    assert np.array_equal(phi, phi2)
    ophi, oev, _, _ = O.p3m_parts(s, dict(mesh=(24, 24, 24), eta=0.35, order=5))
    assert rel(phi, ophi) < 1e-5 and rel(ev2, oev) < 1e-5


@pytest.mark.parametrize("ik", [0, 1])
def test_p3m_scale_and_copy(gpu_lib, ik):
    """p3m_scale (pppm.c:643-664): the box times f and G / f is the P3M of
    the scaled box with eta / f (G is homogeneous in k and eta), so the field
    at scaled positions equals the oracle's for that box and eta.
    p3m_copy (pppm.c:294-311) to fewer sites keeps the influence function."""
    L = lib()
    L.p3m_scale.argtypes = [P, D]
    L.p3m_copy.restype = P
    L.p3m_copy.argtypes = [P, I]
    s, p = W.config("water648_ewald")
    eta, f = 0.35, 1.03
    pp, _ = p3m_handle(L, s, 24, eta, ik=ik)
    L.p3m_scale(pp, f)
    t = W.System(**{k: getattr(s, k) for k in s.__dataclass_fields__})
    t.boxl, t.r = s.boxl * f, s.r * f
    phi, ev = p3m_run(L, pp, t)
    ophi, oev, _, _ = O.p3m_parts(t, dict(mesh=(24, 24, 24), eta=eta / f, order=5, ik=ik))
    print("[refapi] p3m_scale rel", rel(phi, ophi), rel(ev, oev))
    assert rel(phi, ophi) < 1e-5 and rel(ev, oev) < 1e-5
    # copy to the first 27 waters, at the unscaled box again
    small = W.subset(s, np.arange(27))
    L.p3m_scale(pp, 1 / f)
    cp = L.p3m_copy(pp, small.nsite)
    L.p3m_delete(pp)
    phi_c, ev_c = p3m_run(L, cp, small)
    L.p3m_delete(cp)
    sphi, sev, _, _ = O.p3m_parts(small, dict(mesh=(24, 24, 24), eta=eta, order=5, ik=ik))
    print("[refapi] p3m_copy rel", rel(phi_c, sphi), rel(ev_c, sev))
    assert rel(phi_c, sphi) < 1e-5 and rel(ev_c, sev) < 1e-5


def test_p3m_optimize_finds_the_minimum(gpu_lib):
    """p3m_optimize (pppm.c:667-716) minimises chi_r^2 + chi_p3m^2 over eta
    (the reference uses GSL's simplex, absent here): the returned eta is
    within 1e-3 relative of the minimum of a fine scan of the same function
    computed with the oracle's influence function, its value is the scan's
    minimum to 1e-4 (the simplex stops at an eta step of 1e-4, where the
    function is flat to second order), and the returned estimates are those
    of the returned eta."""
    L = lib()
    L.p3m_optimize.argtypes = [P, D, P, P, P]
    s, p = W.config("water648_ewald")
    rc, mesh = 9.0, 20
    pp, _ = p3m_handle(L, s, mesh, 0.3)
    eta, cr, ck = np.zeros(1), np.zeros(1), np.zeros(1)
    L.p3m_optimize(pp, rc, eta.ctypes.data_as(P), cr.ctypes.data_as(P), ck.ctypes.data_as(P))
    L.p3m_delete(pp)
    V = s.boxl ** 3

    def f(x):
        chi_r = 2 / np.sqrt(rc / V ** (1 / 3)) * np.exp(-(rc * x) ** 2)
        _, chi_k = O.green((mesh,) * 3, s.boxl, x, 5)
        return chi_r ** 2 + chi_k ** 2, chi_r, chi_k

    xs = np.linspace(0.8 * eta[0], 1.2 * eta[0], 81)
    vals = [f(x)[0] for x in xs]
    k = int(np.argmin(vals))
    assert 0 < k < len(xs) - 1, "the scan brackets the minimum"
    fine = np.linspace(xs[k - 1], xs[k + 1], 41)
    fv = [f(x)[0] for x in fine]
    x_min = fine[int(np.argmin(fv))]
    fo, cro, cko = f(eta[0])
    print(f"[refapi] p3m_optimize eta {eta[0]:.6f}, scan minimum {x_min:.6f}")
    assert abs(eta[0] - x_min) <= 1e-3 * x_min
    assert fo <= min(fv) * (1 + 1e-4)
    assert abs(cr[0] - cro) <= 1e-12 * cro and abs(ck[0] - cko) <= 1e-9 * cko
