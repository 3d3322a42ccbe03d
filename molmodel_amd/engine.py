"""Python access to libmolmodel_amd.so through its C ABI (include/molmodel_amd.h).

This is plumbing for tests, benchmarks and scripting; all computation runs in
the HIP library.  Importing this module never falls back to anything else:
if the library is missing or no GPU is present, creating an Engine raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MOLMODEL_AMD_LIB selects another build of the same library (e.g. one
# compiled with different kernel parameters); there is no other fallback
LIB_PATH = os.environ.get("MOLMODEL_AMD_LIB", os.path.join(HERE, "libmolmodel_amd.so"))

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_long


class SystemDesc(C.Structure):
    _fields_ = [("nsite", _I), ("ngroup", _I), ("group", _P), ("q", _P), ("sigma", _P),
                ("epsilon", _P), ("site_mask", _P), ("mass", _P), ("nexcl", _I), ("excl", _P),
                ("boxl", _D), ("rc", _D), ("smoothing_width", _D), ("skin", _D), ("eta", _D),
                ("kmode", _I), ("kc", _D), ("mesh", _I * 3), ("order", _I), ("nalias", _I),
                ("ik", _I), ("include_lj_correction", _I), ("geometric_sigma", _I)]


NE = 8
E_ELEC, E_LJ, E_TAIL, E_TOTAL, E_KINETIC, E_KSPACE = range(6)
VIRIAL = 1

_lib = None


def lib():
    """Load the built library (fails loudly if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run molmodel_amd.build.build() first")
        L = C.CDLL(LIB_PATH)
        L.mm_last_error.restype = C.c_char_p
        L.mm_engine_create.argtypes = [_I, _P, C.POINTER(_P)]
        L.mm_engine_destroy.argtypes = [_P]
        for name in ("mm_set_positions", "mm_get_positions", "mm_set_velocities",
                     "mm_get_velocities", "mm_get_forces", "mm_get_potential",
                     "mm_get_energies", "mm_get_virial"):
            getattr(L, name).argtypes = [_P, _P]
        L.mm_compute.argtypes = [_P, _I]
        L.mm_errors.argtypes = [_P, _P]
        L.mm_raise_flags.argtypes = [_P, _I]
        L.mm_rspace_field.argtypes = [_P, _P, _P, _P, _P]
        L.mm_kspace_field.argtypes = [_P, _P, _P]
        L.mm_nl_pairs.restype = _L
        L.mm_nl_pairs.argtypes = [_P, _I, _P, _L]
        L.mm_md_setup.argtypes = [_P, _D, _D, _D, _D]
        L.mm_md_run.argtypes = [_P, _I, _I]
        L.mm_stream.restype = _P
        L.mm_stream.argtypes = [_P]
        L.mm_sync.argtypes = [_P]
        L.mm_engine_sizes.argtypes = [_P, _P, _P]
        L.mm_remd_setup.argtypes = [_P, _I, _P, _P, C.c_ulonglong]
        L.mm_remd_run.argtypes = [_P, _I, _I, _I, _I]
        L.mm_remd_run_stochastic.argtypes = [_P, _I, _I, _I, _I, _P, _P]
        L.mm_remd_stats.argtypes = [_P, _P, _P]
        L.mm_tune_kspace.argtypes = [_D, _I, _D, _D, _I, _P]
        L.mm_param_gradients.argtypes = [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P]
        L.mm_param_gradients_basis.argtypes = [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P]
        L.mm_set_site_params.argtypes = [_P, _I, _P, _P, _P]
        L.mm_set_perturbation.argtypes = [_P, _I, _P, _P, _P]
        L.mm_set_lambda.argtypes = [_P, _I, _D]
        L.mm_scale_system.argtypes = [_P, _I, _D]
        L.mm_nl_builds.restype = C.c_longlong
        L.mm_nl_builds.argtypes = [_P]
        L.mm_desc_init.restype = None
        L.mm_desc_init.argtypes = [_P]
        L.mm_desc_resolve.argtypes = [_P, _D, _D]
        L.mm_derive_sites.argtypes = [_I, _P, _P, _P, _P, _P, _P, _L, _P, _P]
        L.mm_shape_default.restype = None
        L.mm_shape_default.argtypes = [_P]
        L.mm_engine_create_shaped.argtypes = [_I, _P, _P, C.POINTER(_P)]
        L.mm_tune_shape.argtypes = [_I, _P, _P, _P, _D, _D, _D, _D, _I, _P, _P, _I, _P]
        L.mm_free.restype = None
        L.mm_free.argtypes = [_P]
        L.mm_msim_preprocess.argtypes = [_I, _P, C.POINTER(_P)]
        L.mm_msim_eleclj_read.argtypes = [C.c_char_p, C.POINTER(_P)]
        L.mm_msim_eleclj_free.restype = None
        L.mm_msim_eleclj_free.argtypes = [_P]
        L.mm_msim_eleclj_get.argtypes = [_P, C.c_char_p, _P]
        L.mm_msim_eleclj_table.argtypes = [_P, C.c_char_p, C.c_char_p, _P]
        L.mm_msim_eleclj_desc.argtypes = [_P, _P]
        L.mm_msim_site_params.argtypes = [_P, _I, _P, _P, _P, _P, _P]
        L.mm_xyz_read.argtypes = [C.c_char_p, _I, _P, _P, _P, _P, _P, C.c_char_p, _I]
        L.mm_xyz_write.argtypes = [C.c_char_p, _I, _P, _P, _P, _P, C.c_char_p, C.c_char_p, _I]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


class Shape(C.Structure):
    _fields_ = [("cell_groups", _I), ("pair_slots", _I), ("spread_block", _I),
                ("graph_steps", _I), ("spread_halo", _I)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ShapeTrial(C.Structure):
    _fields_ = [("shape", Shape), ("ms_per_step", _D)]


def default_shape():
    s = Shape()
    lib().mm_shape_default(C.byref(s))
    return s.as_dict()


def _shape(d):
    s = Shape()
    lib().mm_shape_default(C.byref(s))
    for k, v in (d or {}).items():
        setattr(s, k, int(v))
    return s


def _descs(systems, params):
    if not isinstance(systems, (list, tuple)):
        systems, params = [systems], [params]
    elif not isinstance(params, (list, tuple)):
        params = [params] * len(systems)
    descs, keep = [], []
    for s, p in zip(systems, params):
        d, k = system_desc(s, p)
        descs.append(d)
        keep.append(k)
    return list(systems), (SystemDesc * len(descs))(*descs), keep


def tune_shape(systems, params, dt=0.001, bond=1.0, angle=109.47, tol=1e-8, steps=100,
               r=None, v=None):
    """The autotuner (mm_tune_shape): the fastest shape for this batch from
    timed MD blocks on the device; returns (best shape dict, trials list of
    (shape dict, ms per step))."""
    systems, arr, keep = _descs(systems, params)
    r = np.ascontiguousarray(np.concatenate([s.r for s in systems]) if r is None else r,
                             dtype=np.float64).reshape(-1)
    v = np.ascontiguousarray(np.concatenate([s.v for s in systems]) if v is None else v,
                             dtype=np.float64).reshape(-1)
    best = Shape()
    trials = (ShapeTrial * 32)()
    nt = C.c_int(0)
    _check(lib().mm_tune_shape(len(systems), C.cast(arr, _P), _ptr(r), _ptr(v), dt, bond, angle,
                               tol, int(steps), C.byref(best), C.cast(trials, _P), 32,
                               C.byref(nt)))
    return best.as_dict(), [(trials[i].shape.as_dict(), trials[i].ms_per_step)
                            for i in range(min(nt.value, 32))]


class MMError(RuntimeError):
    pass


class KspaceChoice(C.Structure):
    _fields_ = [("kmode", _I), ("eta", _D), ("kc", _D), ("mesh", _I * 3), ("order", _I),
                ("err_rspace", _D), ("err_kspace", _D), ("cost", _D)]


def tune_kspace(boxl, nsite, rc, tol=1e-4, allow_ewald=True):
    """The library's reciprocal-space tuner (host only, no GPU needed).
    Returns a parameter dict for a waterbox-style params entry plus the
    estimates: {eta, kmode, kc | mesh, order, err_rspace, err_kspace, cost}."""
    ch = KspaceChoice()
    _check(lib().mm_tune_kspace(float(boxl), int(nsite), float(rc), float(tol),
                                1 if allow_ewald else 0, C.byref(ch)))
    out = dict(eta=ch.eta, kmode=ch.kmode, err_rspace=ch.err_rspace,
               err_kspace=ch.err_kspace, cost=ch.cost)
    if ch.kmode == 0:
        out["kc"] = ch.kc
    else:
        out["mesh"] = tuple(ch.mesh)
        out["order"] = ch.order
    return out


def eleclj_defaults(boxl, p3m=False, p3m_grid_spacing=0.5, p3m_grid_density=-1.0, **given):
    """ElecLJ's parameter defaults for a cubic box (mm_desc_init, then
    mm_desc_resolve): pass any of rc, skin, smoothing_width, eta, kc, order,
    nalias, ik, mesh to fix them; returns the resolved parameters as a dict
    of the waterbox params keys.  Host only."""
    d = SystemDesc()
    lib().mm_desc_init(C.byref(d))
    d.boxl = float(boxl)
    d.kmode = 1 if p3m else 0
    for k in ("rc", "skin", "smoothing_width", "eta", "kc"):
        if k in given:
            setattr(d, k, float(given[k]))
    for k in ("order", "nalias", "ik"):
        if k in given:
            setattr(d, k, int(given[k]))
    if "mesh" in given:
        d.mesh = (_I * 3)(*given["mesh"])
    _check(lib().mm_desc_resolve(C.byref(d), float(p3m_grid_spacing), float(p3m_grid_density)))
    out = dict(rc=d.rc, skin=d.skin, smoothing_width=d.smoothing_width, eta=d.eta,
               kmode=d.kmode, kc=d.kc, order=d.order, nalias=d.nalias, ik=d.ik,
               include_lj_correction=d.include_lj_correction, geometric_sigma=d.geometric_sigma)
    if p3m:
        out["mesh"] = tuple(d.mesh)
    return out


def derive_sites(natom, bonds_of):
    """ElecLJ::init_sites for atom sites (mm_derive_sites).  bonds_of[a] is
    the neighbour list of atom a in Atom::neighbors order.  Returns
    (atom_of_site, groups [ngroup, 2], excluded pairs [nexcl, 2], ncross)."""
    start = np.zeros(natom + 1, dtype=np.int32)
    for a in range(natom):
        start[a + 1] = start[a] + len(bonds_of[a])
    nbr = np.array([b for a in range(natom) for b in bonds_of[a]] or [0], dtype=np.int32)
    atom_of_site = np.zeros(natom, dtype=np.int32)
    group = np.zeros(2 * natom, dtype=np.int32)
    ngroup, nexcl, ncross = C.c_int(0), C.c_long(0), C.c_long(0)
    args = lambda ex, cap: (natom, _ptr(start), _ptr(nbr), _ptr(atom_of_site), _ptr(group),
                            C.byref(ngroup), _ptr(ex), cap, C.byref(nexcl), C.byref(ncross))
    _check(lib().mm_derive_sites(*args(None, 0)))
    excl = np.zeros(2 * max(nexcl.value, 1), dtype=np.int32)
    _check(lib().mm_derive_sites(*args(excl, nexcl.value)))
    return (atom_of_site, group[:2 * ngroup.value].reshape(-1, 2),
            excl[:2 * nexcl.value].reshape(-1, 2), ncross.value)


def msim_preprocess(args):
    """msim's input preprocessor over the files (and -name value pairs) of
    an msim command line (mm_msim_preprocess); returns the processed text."""
    arr = (C.c_char_p * len(args))(*[a.encode() for a in args])
    out = _P()
    _check(lib().mm_msim_preprocess(len(args), C.cast(arr, _P), C.byref(out)))
    try:
        return C.string_at(out.value).decode()
    finally:
        lib().mm_free(out)


class MsimElecLJ:
    """ElecLJ's io parameters read from msim text (mm_msim_eleclj_read)."""

    def __init__(self, deck):
        self.h = _P()
        _check(lib().mm_msim_eleclj_read(deck.encode(), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().mm_msim_eleclj_free(self.h)
            self.h = None

    def __getitem__(self, field):
        v = C.c_double()
        _check(lib().mm_msim_eleclj_get(self.h, field.encode(), C.byref(v)))
        return v.value

    def table(self, table, key):
        v = C.c_double()
        rc = _check(lib().mm_msim_eleclj_table(self.h, table.encode(), key.encode(), C.byref(v)))
        return None if rc == 1 else v.value

    def desc(self, boxl):
        """(rc, smoothing_width, skin, eta, kmode, kc, mesh, order, nalias, ik,
        include_lj_correction, geometric_sigma) after mm_msim_eleclj_desc."""
        d = SystemDesc()
        d.boxl = boxl
        _check(lib().mm_msim_eleclj_desc(self.h, C.byref(d)))
        return dict(rc=d.rc, smoothing_width=d.smoothing_width, skin=d.skin, eta=d.eta,
                    kmode=d.kmode, kc=d.kc, mesh=tuple(d.mesh), order=d.order, nalias=d.nalias,
                    ik=d.ik, include_lj_correction=d.include_lj_correction,
                    geometric_sigma=d.geometric_sigma)

    def site_params(self, types):
        n = len(types)
        arr = (C.c_char_p * max(n, 1))(*[t.encode() for t in types])
        q, sig, eps = (np.zeros(n) for _ in range(3))
        mask = np.zeros(n, dtype=np.int32)
        _check(lib().mm_msim_site_params(self.h, n, C.cast(arr, _P), _ptr(q), _ptr(sig),
                                         _ptr(eps), _ptr(mask)))
        return q, sig, eps, mask


def xyz_read(fname):
    """An xyz coordinate / restart file (mm_xyz_read): dict of symbols,
    types, r [n, 3], v [n, 3] or None, bc (str)."""
    n = _check(lib().mm_xyz_read(fname.encode(), 0, None, None, None, None, None, None, 0))
    sym = C.create_string_buffer(16 * n)
    typ = C.create_string_buffer(16 * n)
    r, v = np.zeros((n, 3)), np.zeros((n, 3))
    hv = C.c_int(0)
    bc = C.create_string_buffer(256)
    _check(lib().mm_xyz_read(fname.encode(), n, sym, typ, _ptr(r), _ptr(v), C.byref(hv), bc, 256))
    cut = lambda b, i: b.raw[16 * i:16 * i + 16].split(b"\0")[0].decode()
    return dict(symbols=[cut(sym, i) for i in range(n)], types=[cut(typ, i) for i in range(n)],
                r=r, v=v if hv.value else None, bc=bc.value.decode())


def xyz_write(fname, symbols, r, bc, types=None, v=None, comment="", precision=0):
    """Write an xyz coordinate / restart file (mm_xyz_write)."""
    n = len(symbols)
    slot = lambda names: b"".join(x.encode()[:15].ljust(16, b"\0") for x in names)
    sym = C.create_string_buffer(slot(symbols), 16 * max(n, 1))
    typ = C.create_string_buffer(slot(types), 16 * max(n, 1)) if types is not None else None
    r = np.ascontiguousarray(r, dtype=np.float64)
    vv = np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
    _check(lib().mm_xyz_write(fname.encode(), n, sym, typ, _ptr(r), _ptr(vv) if vv is not None else None,
                              bc.encode(), comment.encode(), precision))


def _check(rc):
    if rc < 0:
        raise MMError(f"molmodel_amd error {rc}: {lib().mm_last_error().decode()}")
    return rc


def system_desc(system, params):
    """Build the C descriptor for a waterbox.System and method parameters;
    returns (desc, keepalive arrays)."""
    keep = dict(
        group=np.ascontiguousarray(system.group, dtype=np.int32).reshape(-1),
        q=np.ascontiguousarray(system.q, dtype=np.float64),
        sig=np.ascontiguousarray(system.sig, dtype=np.float64),
        eps=np.ascontiguousarray(system.eps, dtype=np.float64),
        mask=np.ascontiguousarray(system.site_mask, dtype=np.int32),
        mass=np.ascontiguousarray(system.mass, dtype=np.float64),
        excl=np.ascontiguousarray(system.excl, dtype=np.int32).reshape(-1),
    )
    mesh = params.get("mesh", (0, 0, 0))
    d = SystemDesc(system.nsite, system.ngroup, _ptr(keep["group"]), _ptr(keep["q"]),
                   _ptr(keep["sig"]), _ptr(keep["eps"]), _ptr(keep["mask"]), _ptr(keep["mass"]),
                   len(keep["excl"]) // 2, _ptr(keep["excl"]), system.boxl, params["rc"],
                   params["smoothing_width"], params["skin"], params["eta"], params["kmode"],
                   params.get("kc", 0.0), (_I * 3)(*mesh), params.get("order", 5),
                   params.get("nalias", 2), params.get("ik", 0),
                   params.get("include_lj_correction", 1), params.get("geometric_sigma", 0))
    return d, keep


class Engine:
    """A batch of systems resident on the GPU."""

    def __init__(self, systems, params, shape=None):
        """shape: a dict of mm_shape fields (the rest default), or None"""
        systems, arr, self._keep = _descs(systems, params)
        self.systems = systems
        h = _P()
        if shape is None:
            _check(lib().mm_engine_create(len(systems), C.cast(arr, _P), C.byref(h)))
        else:
            sh = _shape(shape)
            _check(lib().mm_engine_create_shaped(len(systems), C.cast(arr, _P), C.byref(sh),
                                                 C.byref(h)))
        self.h = h
        self.nsys = len(systems)
        self.nsite = sum(s.nsite for s in systems)
        self.offsets = np.cumsum([0] + [s.nsite for s in systems])

    def close(self):
        if getattr(self, "h", None):
            lib().mm_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state ---------------------------------------------------------------
    def set_positions(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
        assert r.size == 3 * self.nsite
        _check(lib().mm_set_positions(self.h, _ptr(r)))

    def positions(self):
        r = np.zeros((self.nsite, 3))
        _check(lib().mm_get_positions(self.h, _ptr(r)))
        return r

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        _check(lib().mm_set_velocities(self.h, _ptr(v)))

    def velocities(self):
        v = np.zeros((self.nsite, 3))
        _check(lib().mm_get_velocities(self.h, _ptr(v)))
        return v

    # evaluation ------------------------------------------------------------
    def compute(self, virial=False):
        _check(lib().mm_compute(self.h, VIRIAL if virial else 0))

    def forces(self):
        f = np.zeros((self.nsite, 3))
        _check(lib().mm_get_forces(self.h, _ptr(f)))
        return f

    def potential(self):
        p = np.zeros(self.nsite)
        _check(lib().mm_get_potential(self.h, _ptr(p)))
        return p

    def energies(self):
        e = np.zeros((self.nsys, NE))
        _check(lib().mm_get_energies(self.h, _ptr(e)))
        return e

    def virial(self):
        w = np.zeros((self.nsys, 9))
        _check(lib().mm_get_virial(self.h, _ptr(w)))
        return w.reshape(self.nsys, 3, 3)

    def errors(self):
        f = np.zeros(1, dtype=np.int32)
        _check(lib().mm_errors(self.h, _ptr(f)))
        return int(f[0])

    def raise_flags(self, flags):
        """Raise overflow flags on the device (tests of the recovery path)."""
        _check(lib().mm_raise_flags(self.h, int(flags)))

    def rspace_field(self):
        phi = np.zeros(self.nsite); ev = np.zeros((self.nsite, 3)); ljf = np.zeros((self.nsite, 3))
        lje = np.zeros(self.nsys)
        _check(lib().mm_rspace_field(self.h, _ptr(phi), _ptr(ev), _ptr(ljf), _ptr(lje)))
        return phi, ev, ljf, lje

    def kspace_field(self):
        phi = np.zeros(self.nsite); ev = np.zeros((self.nsite, 3))
        _check(lib().mm_kspace_field(self.h, _ptr(phi), _ptr(ev)))
        return phi, ev

    def nl_pairs(self, sys=0):
        n = _check(lib().mm_nl_pairs(self.h, sys, None, 0))
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        _check(lib().mm_nl_pairs(self.h, sys, _ptr(out), n))
        return out[:n]

    # dynamics --------------------------------------------------------------
    def md_setup(self, dt=0.001, bond=1.0, angle=109.47, tol=1e-8):
        _check(lib().mm_md_setup(self.h, dt, bond, angle, tol))

    def md_run(self, nsteps, graph=True):
        _check(lib().mm_md_run(self.h, int(nsteps), 1 if graph else 0))

    def sync(self):
        _check(lib().mm_sync(self.h))

    def nl_builds(self):
        return _check(lib().mm_nl_builds(self.h))

    # parameter gradients -----------------------------------------------------
    def param_gradients(self, sys=0, f_ref=None, maxtypes=16, basis=False):
        """dE/d(q, sigma, epsilon) and dL/d(...) per site type of system
        `sys` for L = sum |F - f_ref|^2: the adjoint kernels
        (mm_param_gradients; 'ms' = their device time), or with basis=True
        the basis of forward evaluations (mm_param_gradients_basis;
        'basis_err')."""
        n = self.systems[sys].nsite
        nt = np.zeros(1, dtype=np.int32)
        st = np.zeros(n, dtype=np.int32)
        tp = np.zeros(3 * maxtypes)
        dE = np.zeros(3 * maxtypes)
        dL = np.zeros(3 * maxtypes)
        loss = np.zeros(1)
        extra = np.zeros(1)
        fr = None if f_ref is None else np.ascontiguousarray(f_ref, dtype=np.float64).reshape(-1)
        fn = lib().mm_param_gradients_basis if basis else lib().mm_param_gradients
        _check(fn(self.h, sys, _ptr(fr), maxtypes, _ptr(nt), _ptr(st), _ptr(tp), _ptr(dE),
                  _ptr(dL), _ptr(loss), _ptr(extra)))
        k = int(nt[0])
        out = dict(ntype=k, site_type=st, params=tp[:3 * k].reshape(k, 3),
                   dE=dE[:3 * k].reshape(k, 3), dL=dL[:3 * k].reshape(k, 3),
                   loss=float(loss[0]))
        out["basis_err" if basis else "ms"] = float(extra[0])
        return out

    # site parameters and free-energy perturbation ----------------------------
    def set_site_params(self, sys, q, sigma, epsilon):
        """New per-site charges (internal units), sigma, epsilon of system sys."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (q, sigma, epsilon)]
        assert all(x.size == self.systems[sys].nsite for x in a)
        _check(lib().mm_set_site_params(self.h, sys, _ptr(a[0]), _ptr(a[1]), _ptr(a[2])))

    def set_perturbation(self, sys, qB, sigmaB, epsilonB):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (qB, sigmaB, epsilonB)]
        assert all(x.size == self.systems[sys].nsite for x in a)
        _check(lib().mm_set_perturbation(self.h, sys, _ptr(a[0]), _ptr(a[1]), _ptr(a[2])))

    def set_lambda(self, sys, lam):
        _check(lib().mm_set_lambda(self.h, sys, float(lam)))

    # replica exchange --------------------------------------------------------
    def remd_setup(self, ladders, kT, seed=1):
        """ladders: lists of system indices by ascending temperature;
        kT: per system, kcal/mol."""
        flat = []
        for i, lad in enumerate(ladders):
            if i:
                flat.append(-1)
            flat.extend(int(x) for x in lad)
        lad = np.asarray(flat, dtype=np.int32)
        kT = np.ascontiguousarray(kT, dtype=np.float64)
        assert kT.size == self.nsys
        _check(lib().mm_remd_setup(self.h, lad.size, _ptr(lad), _ptr(kT), int(seed)))

    def remd_run(self, nintervals, steps, reinit_velocities=True, graph=True):
        _check(lib().mm_remd_run(self.h, int(nintervals), int(steps),
                                 1 if reinit_velocities else 0, 1 if graph else 0))

    def remd_run_stochastic(self, nsteps, exchange_interval, reinit_interval=0, graph=True):
        """nsteps MD steps with the reference's per-step move draws
        (do_every, mdyn.cpp:49-52); returns (exchanges, velocity redraws)."""
        nx, nr = C.c_longlong(0), C.c_longlong(0)
        _check(lib().mm_remd_run_stochastic(self.h, int(nsteps), int(exchange_interval),
                                            int(reinit_interval), 1 if graph else 0,
                                            C.byref(nx), C.byref(nr)))
        return nx.value, nr.value

    def remd_stats(self):
        st = np.zeros((self.nsys, 2), dtype=np.int64)
        cfg = np.zeros(self.nsys, dtype=np.int32)
        _check(lib().mm_remd_stats(self.h, _ptr(st), _ptr(cfg)))
        return st, cfg
