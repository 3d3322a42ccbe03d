"""Recovery paths of the engine (every kernel exits normally; conditions are
forced from the host or arise from valid inputs).

* A non-finite position in one system of a batch is reported, and every
  other system's next evaluation is that of a fresh engine (the pair search
  leaves all rows empty after such an error, so every list is rebuilt).
* Sites much denser than the box average (a two-molecule-thick liquid slab)
  overflow the P3M column bins: the engine grows them and evaluates again,
  matching the FP64 oracle; the p3m_field drop-in does the same.
* An overflow inside an MD block or an exchange interval (raised from the
  host) makes the block run again from its start with larger buffers; the
  trajectory equals an uninterrupted one.
* Positions set by the caller before mm_md_run are evaluated before the
  first half kick, as after mm_md_setup.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine, MMError, lib

pytestmark = pytest.mark.gpu

FLAG_NL_OVERFLOW, FLAG_BIN_OVERFLOW = 1, 16


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rigid_shift(s, size, seed):
    """Each molecule translated rigidly by a random vector of length size."""
    d = W.normal(seed, 3, 3 * s.ngroup).reshape(-1, 3)
    d *= size / np.linalg.norm(d, axis=1, keepdims=True)
    out = s.r.copy()
    for g, (a, b) in enumerate(s.group):
        out[a:b] += d[g]
    return out


def test_nan_in_one_system_leaves_the_others_correct(gpu_lib):
    s0, p = W.config("water216_small")
    s1, _ = W.config("water216_small", seed=5)
    e = Engine([s0, s1], [p, p])
    e.set_positions(np.concatenate([s0.r, s1.r]))
    e.compute()
    # system 0 gets a NaN while system 1 moves by more than the skin, so
    # system 1 rebuilds in the same pass as the failing system
    r0 = s0.r.copy()
    r0[4, 0] = np.nan
    r1 = rigid_shift(s1, 0.6 * p["skin"], 9)
    e.set_positions(np.concatenate([r0, r1]))
    print("[recovery] NaN in system 0 of 2", flush=True)
    with pytest.raises(MMError, match="non-finite"):
        e.compute()
    r_ok = np.concatenate([s0.r, r1])
    e.set_positions(r_ok)
    e.compute()
    fresh = Engine([s0, s1], [p, p])
    fresh.set_positions(r_ok)
    fresh.compute()
    n0 = s0.nsite
    f, g = e.forces(), fresh.forces()
    print(f"[recovery] system 1 rel diff {rel(f[n0:], g[n0:]):.2e}, system 0 {rel(f[:n0], g[:n0]):.2e}")
    assert rel(f[n0:], g[n0:]) < 1e-12
    assert rel(f[:n0], g[:n0]) < 1e-12
    assert np.allclose(e.energies(), fresh.energies(), rtol=1e-12, atol=1e-9)


def slab_system():
    """A liquid slab two molecules thick (x < 6.2 A) of the 98k-atom box's
    lattice: 2048 waters in a 99.4 A box, so that the mesh column bins over
    the slab hold about eight times the average."""
    big, _ = W.config("water98k_p3m")
    first = big.r[big.group[:, 0]]
    x = first[:, 0] - big.boxl * np.floor(first[:, 0] / big.boxl)
    keep = np.nonzero(x < 6.2)[0]
    return W.subset(big, keep)


def test_p3m_bin_overflow_grows_and_matches_oracle(gpu_lib):
    s = slab_system()
    p = dict(rc=10.0, smoothing_width=1.0, eta=0.32, kmode=1, mesh=(64, 64, 64), order=5,
             skin=2.0, nalias=2, ik=0, include_lj_correction=1)
    nbins = 32 * 32
    cap = (4 * ((s.nsite + nbins - 1) // nbins) + 64 + 31) // 32 * 32
    print(f"[recovery] slab of {s.ngroup} waters, bin capacity {cap}", flush=True)
    e = Engine(s, p)
    e.set_positions(s.r)
    e.compute(virial=True)
    f_ref, en_ref, w_ref = O.evaluate(s, p, virial=True)
    err = rel(e.forces(), f_ref)
    en = e.energies()[0]
    print(f"[recovery] slab force rel err {err:.2e}, U {en[3]:.6f} vs {en_ref[3]:.6f}")
    assert err <= 1e-5
    # the slab's elec (+177) and LJ (-224) terms nearly cancel
    assert abs(en[3] - en_ref[3]) <= 1e-5 * (abs(en_ref[0]) + abs(en_ref[1]))
    assert e.errors() == 0


def test_p3m_dropin_bin_overflow_grows(gpu_lib):
    s = slab_system()
    L = lib()
    L.p3m_new.restype = C.c_void_p
    L.p3m_new.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_void_p]
    L.p3m_init.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    L.p3m_field.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.p3m_delete.argtypes = [C.c_void_p]
    ptv = np.array([s.boxl, 0, 0, 0, s.boxl, 0, 0, 0, s.boxl])
    ng = np.zeros(3, dtype=np.int32)
    h = L.p3m_new(s.nsite, ptv.ctypes.data, s.boxl / 64, 2, 0, 5, ng.ctypes.data)
    assert list(ng) == [64, 64, 64]
    chi = C.c_double(0)
    L.p3m_init(h, 0.32, C.byref(chi))
    r = np.ascontiguousarray(s.r)
    phi = np.zeros(s.nsite)
    ev = np.zeros((s.nsite, 3))
    print("[recovery] p3m_field drop-in on the slab", flush=True)
    L.p3m_field(h, r.ctypes.data, s.q.ctypes.data, phi.ctypes.data, ev.ctypes.data)
    L.p3m_delete(h)
    pp = dict(mesh=(64, 64, 64), eta=0.32, order=5, nalias=2, ik=0)
    phi_o, ev_o, _, _ = O.p3m_parts(s, pp)
    print(f"[recovery] drop-in phi {rel(phi, phi_o):.2e} evec {rel(ev, ev_o):.2e}")
    # FP32 mesh: the slab's potentials are smaller than a full box's
    assert rel(phi, phi_o) <= 5e-6
    assert rel(ev, ev_o) <= 5e-6


@pytest.mark.parametrize("name,flag", [("water648_ewald", FLAG_NL_OVERFLOW),
                                       ("water12k_p3m", FLAG_BIN_OVERFLOW)])
def test_overflow_inside_md_block_runs_it_again(gpu_lib, name, flag):
    s, p = W.config(name)

    def start():
        e = Engine(s, p)
        e.set_positions(s.r)
        e.set_velocities(s.v)
        e.md_setup(dt=0.001)
        return e

    a, b = start(), start()
    a.md_run(30, graph=True)
    b.raise_flags(flag)
    print(f"[recovery] {name}: overflow flag {flag} raised before a 30-step block", flush=True)
    b.md_run(30, graph=True)
    assert b.errors() == 0
    d = np.abs(a.positions() - b.positions()).max()
    dv = np.abs(a.velocities() - b.velocities()).max()
    print(f"[recovery] {name}: replayed block max |dr| {d:.2e} A, |dv| {dv:.2e}")
    assert d <= 1e-9 and dv <= 1e-7
    assert np.allclose(a.energies()[0][:5], b.energies()[0][:5], rtol=1e-9, atol=1e-7)


def test_overflow_inside_exchange_interval_runs_it_again(gpu_lib):
    systems, params, kT, ladders = W.remd_batch(nrep=3, sizes=((6, 18.64),))
    r0 = np.concatenate([s.r for s in systems])
    v0 = np.concatenate([s.v for s in systems])

    def start():
        e = Engine(systems, params)
        e.set_positions(r0)
        e.set_velocities(v0)
        e.md_setup(dt=0.001)
        e.remd_setup(ladders, kT, seed=3)
        return e

    a, b = start(), start()
    a.remd_run(3, 20)
    b.remd_run(1, 20)
    b.raise_flags(FLAG_NL_OVERFLOW)
    print("[recovery] overflow raised before the second of three exchange intervals", flush=True)
    b.remd_run(2, 20)
    sa, ca = a.remd_stats()
    sb, cb = b.remd_stats()
    assert np.array_equal(sa, sb) and np.array_equal(ca, cb)
    d = np.abs(a.positions() - b.positions()).max()
    print(f"[recovery] exchange run replayed, max |dr| {d:.2e} A")
    assert d <= 1e-9


def test_positions_set_before_md_run_are_evaluated_first(gpu_lib):
    s, p = W.config("water648_ewald")
    a = Engine(s, p)
    a.set_positions(s.r)
    a.set_velocities(s.v)
    a.md_setup(dt=0.001)
    r1, v1 = a.positions(), a.velocities()   # constrained start
    a.md_run(15, graph=True)                 # somewhere else
    a.set_positions(r1)
    a.set_velocities(v1)
    a.md_run(20, graph=True)
    b = Engine(s, p)
    b.set_positions(r1)
    b.set_velocities(v1)
    b.md_setup(dt=0.001)
    b.md_run(20, graph=True)
    d = np.abs(a.positions() - b.positions()).max()
    print(f"[recovery] set_positions + md_run vs md_setup + md_run: max |dr| {d:.2e} A")
    assert d <= 1e-9
