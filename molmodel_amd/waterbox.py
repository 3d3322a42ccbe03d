"""Seeded synthetic systems for the nonbonded path: rigid 3-site water boxes,
optionally with Na+/Cl- ions.

Reference context: the reference builds systems through MSys (msys.cpp,
lattice tiling and solvation) and a parameter directory that is not shipped
(atom.cpp:41-45).  This module replaces that with an explicit generator:

* water: O-H 1.0 A, H-O-H 109.47 deg, q_O = -0.8476 e, q_H = +0.4238 e,
  LJ on O only (sigma 3.166 A, epsilon 0.1553 kcal/mol); masses O 15.9994,
  H 1.008 g/mol;
* molecules centred (centre of mass) on an n^3 simple-cubic lattice, uniform
  random orientation, +-0.1 A uniform jitter of the centre;
* Maxwell-Boltzmann velocities, projected onto the rigid-body constraints and
  with the total momentum removed;
* groups are molecules [start, end) with the heavy atom first, as
  ElecLJ::init_sites forms them (eleclj.cpp:653-677); the excluded site pairs
  are the 1-2 and 1-3 pairs (eleclj.cpp:712-724).

Random numbers come from a counter-based generator (splitmix64 of
seed, stream and index), so every array is reproducible on any host.
Charges are returned in the reference's internal unit sqrt(kcal A/mol)
(units.h:91-94) and masses in (kcal/mol) ps^2/A^2 (units.h:166-169).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# units.h:23-29
_BOHR, _AVOGADRO, _HARTREE, _CALORIE, _BOLTZMANN = (
    0.5291772108, 6.0221415, 4.35974417, 4.184, 1.3806505)
CHARGE_UNIT = math.sqrt(_BOHR * _HARTREE * _AVOGADRO * 100 / _CALORIE)   # e -> internal


def e_to_charge_unit(x):
    """units.h:91-94"""
    return x * CHARGE_UNIT


def g_mol_to_mass_unit(x):
    """units.h:166-169"""
    return x / (100 * _CALORIE)


def K_to_kcal_mol(t):
    """units.h:201-204"""
    return t * (_BOLTZMANN * _AVOGADRO * 1e-3 / _CALORIE)


SITE_HAS_Q = 1   # rspace.h:7
SITE_HAS_LJ = 2  # rspace.h:8

WATER = dict(bond=1.0, angle=109.47, qO=-0.8476, qH=0.4238, sigO=3.166, epsO=0.1553,
             mO=15.9994, mH=1.008)
IONS = {"Na": dict(q=1.0, sig=2.35, eps=0.13, m=22.98977),
        "Cl": dict(q=-1.0, sig=4.40, eps=0.10, m=35.453)}

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniforms in [0,1) from counter (seed, stream, index)."""
    idx = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        z = _splitmix64(idx ^ key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 7919, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2 * np.pi * u2)


def _random_rotations(seed: int, n: int) -> np.ndarray:
    """Uniform random rotation matrices (Shoemake's quaternion method)."""
    u = [uniform(seed, 11 + k, n) for k in range(3)]
    a, b = np.sqrt(1 - u[0]), np.sqrt(u[0])
    q = np.stack([a * np.sin(2 * np.pi * u[1]), a * np.cos(2 * np.pi * u[1]),
                  b * np.sin(2 * np.pi * u[2]), b * np.cos(2 * np.pi * u[2])], 1)
    w, x, y, z = q[:, 3], q[:, 0], q[:, 1], q[:, 2]
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


@dataclass
class System:
    """A periodic cubic system in the reference's internal units."""
    boxl: float
    r: np.ndarray            # (N,3) float64 positions, A
    v: np.ndarray            # (N,3) float64 velocities, A/ps
    q: np.ndarray            # (N,) charges, sqrt(kcal A/mol)
    sig: np.ndarray          # (N,) LJ sigma, A
    eps: np.ndarray          # (N,) LJ epsilon, kcal/mol
    site_mask: np.ndarray    # (N,) int32
    mass: np.ndarray         # (N,) internal mass units
    group: np.ndarray        # (G,2) int32 [start,end)
    excl: np.ndarray         # (E,2) int32 excluded site pairs
    nwater: int = 0
    nion: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def nsite(self):
        return int(self.r.shape[0])

    @property
    def ngroup(self):
        return int(self.group.shape[0])


def water_geometry(bond=WATER["bond"], angle=WATER["angle"]):
    """Body-frame O, H1, H2 with the centre of mass at the origin."""
    t = math.radians(angle) / 2
    o = np.zeros(3)
    h1 = np.array([bond * math.sin(t), 0.0, bond * math.cos(t)])
    h2 = np.array([-bond * math.sin(t), 0.0, bond * math.cos(t)])
    m = np.array([WATER["mO"], WATER["mH"], WATER["mH"]])
    x = np.stack([o, h1, h2])
    com = (m[:, None] * x).sum(0) / m.sum()
    return x - com


def water_box(nside: int, boxl: float, seed: int = 2024, temperature: float = 300.0,
              jitter: float = 0.1, ions: tuple[int, int] = (0, 0),
              ion_min_dist: float = 2.0) -> System:
    """nside^3 waters on a simple-cubic lattice in a cubic box of side boxl.

    ions = (n_Na, n_Cl) inserts ions by seeded rejection sampling at least
    ion_min_dist from every site already placed (waters first)."""
    nw = nside ** 3
    spacing = boxl / nside
    g = np.arange(nside)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    com = (lat + 0.5) * spacing
    com += (uniform(seed, 1, 3 * nw).reshape(nw, 3) * 2 - 1) * jitter
    R = _random_rotations(seed, nw)
    body = water_geometry()
    r = (com[:, None, :] + np.einsum("nij,kj->nki", R, body)).reshape(-1, 3)
    qO, qH = e_to_charge_unit(WATER["qO"]), e_to_charge_unit(WATER["qH"])
    q = np.tile([qO, qH, qH], nw)
    sig = np.tile([WATER["sigO"], 0.0, 0.0], nw)
    eps = np.tile([WATER["epsO"], 0.0, 0.0], nw)
    mask = np.tile([SITE_HAS_Q | SITE_HAS_LJ, SITE_HAS_Q, SITE_HAS_Q], nw).astype(np.int32)
    mass = np.tile([g_mol_to_mass_unit(WATER["mO"]), g_mol_to_mass_unit(WATER["mH"]),
                    g_mol_to_mass_unit(WATER["mH"])], nw)
    group = np.stack([np.arange(nw) * 3, np.arange(nw) * 3 + 3], 1).astype(np.int32)
    b = np.arange(nw)[:, None] * 3
    excl = np.concatenate([b + [0, 1], b + [0, 2], b + [1, 2]], 0).astype(np.int32)
    nna, ncl = ions
    if nna + ncl:
        rr, qq, ss, ee, mm, kk = [r], [q], [sig], [eps], [mass], [mask]
        placed = r.copy()
        u = uniform(seed, 101, 3 * 200000).reshape(-1, 3)
        t = 0
        for name, n in (("Na", nna), ("Cl", ncl)):
            p = IONS[name]
            for _ in range(n):
                while True:
                    if t >= len(u):
                        raise RuntimeError("ion placement failed")
                    cand = u[t] * boxl
                    t += 1
                    d = placed - cand
                    d -= boxl * np.floor(d / boxl + 0.5)
                    if (d * d).sum(1).min() >= ion_min_dist ** 2:
                        break
                placed = np.vstack([placed, cand])
                rr.append(cand[None]); qq.append([e_to_charge_unit(p["q"])])
                ss.append([p["sig"]]); ee.append([p["eps"]])
                mm.append([g_mol_to_mass_unit(p["m"])]); kk.append([SITE_HAS_Q | SITE_HAS_LJ])
        n0 = r.shape[0]
        r = np.concatenate(rr, 0); q = np.concatenate(qq); sig = np.concatenate(ss)
        eps = np.concatenate(ee); mass = np.concatenate(mm); mask = np.concatenate(kk).astype(np.int32)
        ig = np.arange(n0, r.shape[0])
        group = np.concatenate([group, np.stack([ig, ig + 1], 1)], 0).astype(np.int32)
    v = thermal_velocities(mass, group, r, nw, temperature, seed)
    return System(boxl=float(boxl), r=r, v=v, q=q, sig=sig, eps=eps, site_mask=mask,
                  mass=mass, group=group, excl=excl, nwater=nw, nion=nna + ncl,
                  meta=dict(seed=seed, nside=nside, temperature=temperature))


def subset(s: System, groups) -> System:
    """The system made of the given groups of s (in that order), in the same
    box; exclusions inside the kept groups are kept."""
    groups = np.asarray(groups, dtype=np.int64)
    sites, new_group, where = [], [], {}
    for g in groups:
        a, b = int(s.group[g, 0]), int(s.group[g, 1])
        new_group.append((len(sites), len(sites) + b - a))
        for i in range(a, b):
            where[i] = len(sites)
            sites.append(i)
    sites = np.asarray(sites, dtype=np.int64)
    excl = [(where[i], where[j]) for i, j in s.excl.tolist() if i in where and j in where]
    nw = int(sum(1 for g in groups if g < s.nwater))
    return System(boxl=s.boxl, r=s.r[sites].copy(), v=s.v[sites].copy(), q=s.q[sites].copy(),
                  sig=s.sig[sites].copy(), eps=s.eps[sites].copy(),
                  site_mask=s.site_mask[sites].copy(), mass=s.mass[sites].copy(),
                  group=np.asarray(new_group, dtype=np.int32).reshape(-1, 2),
                  excl=np.asarray(excl, dtype=np.int32).reshape(-1, 2), nwater=nw,
                  nion=len(groups) - nw, meta=dict(s.meta, subset=len(groups)))


def thermal_velocities(mass, group, r, nwater, temperature, seed):
    """Maxwell-Boltzmann velocities; each water's velocities are projected
    onto its rigid-body (constraint-preserving) subspace; total momentum 0."""
    n = mass.shape[0]
    kT = K_to_kcal_mol(temperature)
    v = normal(seed, 31, 3 * n).reshape(n, 3) * np.sqrt(kT / mass)[:, None]
    for w in range(nwater):
        s = slice(3 * w, 3 * w + 3)
        v[s] = _rigid_project(r[s], v[s], mass[s])
    p = (mass[:, None] * v).sum(0)
    v -= p / mass.sum()
    # the uniform shift keeps the constraints' time derivatives zero
    return v


def _rigid_project(x, v, m):
    """Mass-weighted projection of v onto {d/dt |r_ij|^2 = 0 for all pairs}."""
    pairs = ((0, 1), (0, 2), (1, 2))
    A = np.zeros((3, 9))
    for c, (i, j) in enumerate(pairs):
        d = x[i] - x[j]
        A[c, 3 * i:3 * i + 3] = d
        A[c, 3 * j:3 * j + 3] = -d
    Minv = np.repeat(1.0 / m, 3)
    vv = v.reshape(9)
    S = (A * Minv) @ A.T
    lam = np.linalg.solve(S, -A @ vv)
    return (vv + Minv * (A.T @ lam)).reshape(3, 3)


# Named workloads ---------------------------------------------------------

def config(name: str, seed: int = 2024) -> tuple[System, dict]:
    """Systems and method parameters of the named workloads.

    water648_ewald     216 waters, L 18.64, Ewald eta 0.35, kc 7, rc 9
    water12k_p3m       4096 waters, L 49.7, P3M 48^3 order 5, eta 0.32, rc 10
    water98k_p3m       32768 waters, L 99.4, P3M 64^3 order 5, eta 0.32, rc 10
    water98k_p3m_ik    the same with the P3M field formed as -i k phi on the mesh
    water1000_ions     1000 waters + 20 Na + 20 Cl, L 31.4, Ewald (eta 0.32, kc 7), rc 10
    """
    if name == "water648_ewald":
        s = water_box(6, 18.64, seed)
        p = dict(rc=9.0, smoothing_width=1.0, eta=0.35, kmode=0, kc=7.0)
    elif name == "water12k_p3m":
        s = water_box(16, 49.7, seed)
        p = dict(rc=10.0, smoothing_width=1.0, eta=0.32, kmode=1, mesh=(48, 48, 48), order=5)
    elif name in ("water98k_p3m", "water98k_p3m_ik"):
        s = water_box(32, 99.4, seed)
        p = dict(rc=10.0, smoothing_width=1.0, eta=0.32, kmode=1, mesh=(64, 64, 64), order=5,
                 ik=int(name.endswith("_ik")))
    elif name == "water1000_ions":
        s = water_box(10, 31.4, seed, ions=(20, 20))
        p = dict(rc=10.0, smoothing_width=1.0, eta=0.32, kmode=0, kc=7.0)
    elif name.startswith("water") and name.endswith("_small"):
        # small test boxes, e.g. "water64_small": n^3 waters at water density
        n = int(round(int(name[5:-6]) ** (1 / 3)))
        L = n * 3.1067
        s = water_box(n, L, seed)
        rc = min(9.0, 0.45 * L)
        p = dict(rc=rc, smoothing_width=1.0, eta=3.2 / rc, kmode=0, kc=6.0)
    else:
        raise KeyError(name)
    p.setdefault("skin", default_skin(p["rc"], s.boxl))
    p.setdefault("nalias", 2)
    p.setdefault("ik", 0)
    p.setdefault("include_lj_correction", 1)
    return s, p


def default_skin(rc, boxl):
    """ElecLJ::init (eleclj.cpp:172-177): 0.2 rc, capped below half the box."""
    nlw = 0.2 * rc
    if rc + nlw >= 0.5 * boxl:
        nlw = 0.499 * boxl - rc
    return nlw


def smooth_mesh_size(n: int) -> int:
    """Smallest size >= n (and >= 8) whose prime factors are 2, 3, 5, 7."""
    n = max(int(n), 8)
    while True:
        m = n
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 1


def remd_batch(nrep: int = 16, t_lo: float = 300.0, t_hi: float = 450.0, seed: int = 7,
               sizes=((6, 18.64), (8, 24.86), (10, 31.07)), kspace=None):
    """Replica-exchange batch: for each (n, L) in sizes, nrep replicas of an
    n^3-water box at temperatures geometric from t_lo to t_hi.  Replicas of
    one size share the box and topology and start from different seeded
    configurations.  kspace(boxl, nsite, rc) -> dict of reciprocal-space
    parameters (eta, kmode, kc or mesh/order) chooses them per size; the
    default is Ewald (eta 0.35, kc 7) below 20 A and P3M (eta 0.32, order 5,
    mesh spacing about 1 A) above.  Returns (systems, params, kT, ladders)
    with kT in kcal/mol per system and one ladder (ascending T) per size."""
    systems, params, kT, ladders = [], [], [], []
    temps = [t_lo * (t_hi / t_lo) ** (k / max(nrep - 1, 1)) for k in range(nrep)]
    for si, (n, L) in enumerate(sizes):
        lad = []
        rc = 9.0 if L < 20 else min(10.0, 0.49 * L)
        if kspace is not None:
            kp = dict(kspace(L, 3 * n ** 3, rc))
        elif L < 20:
            kp = dict(eta=0.35, kmode=0, kc=7.0)
        else:
            m = smooth_mesh_size(round(L))
            kp = dict(eta=0.32, kmode=1, mesh=(m, m, m), order=5)
        for k, T in enumerate(temps):
            s = water_box(n, L, seed + 1000 * si + k, temperature=T)
            p = dict(rc=rc, smoothing_width=1.0, **kp)
            p.setdefault("skin", default_skin(p["rc"], L))
            p.update(nalias=2, ik=0, include_lj_correction=1)
            lad.append(len(systems))
            systems.append(s)
            params.append(p)
            kT.append(K_to_kcal_mol(T))
        ladders.append(lad)
    return systems, params, kT, ladders
