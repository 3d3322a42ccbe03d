"""GPU parity at the full benchmark size: 98 304-atom water (32 768 groups),
P3M 64^3 order 5, eta 0.32, rc 10 Å, skin 2 Å.

Bars (as for the smaller boxes):
  * neighbour-list pair set bit-exact against the oracle's build
    (nbrlist.cpp:222-251 restated) on the same positions;
  * forces ||F_gpu - F_oracle|| / ||F_oracle|| <= 1e-5, energies within 1e-5
    of |E_total|, LJ tail exact (FP32 pair and mesh arithmetic, FP64
    accumulation);
  * size-independent properties: the net force vanishes (Newton's third law
    in real space, a neutral mesh), a second evaluation and a second engine
    give bit-identical forces, and 200 MD steps keep the constraints and the
    total energy (within 2% of the kinetic energy) while the neighbour list
    is rebuilt along the way.
"""
import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine

pytestmark = pytest.mark.gpu

NAME = "water98k_p3m"


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def full(gpu_lib):
    s, p = W.config(NAME)
    e = Engine(s, p)
    e.set_positions(s.r)
    print(f"[full] {NAME} N={s.nsite}", flush=True)
    e.compute()
    return s, p, e


def test_full_size_pair_set_bit_exact(full):
    s, p, e = full
    got = O.pair_set(e.nl_pairs(0))
    ref = O.pair_set(O.nl_pairs(O.first_sites(s), s.boxl, p["rc"], p["skin"]))
    print(f"[full] pairs {got.size} vs {ref.size}", flush=True)
    assert got.size == ref.size
    assert np.array_equal(got, ref)


def test_full_size_forces_and_energies(full):
    s, p, e = full
    f = e.forces()
    en = e.energies()[0]
    of, oen = O.evaluate(s, p)
    ferr = rel(f, of)
    print(f"[full] force rel {ferr:.3e} elec {en[0]:.6f} vs {oen[0]:.6f} "
          f"lj {en[1]:.6f} vs {oen[1]:.6f}", flush=True)
    assert ferr <= 1e-5
    assert abs(en[0] - oen[0]) <= 1e-5 * abs(oen[3])
    assert abs(en[1] - oen[1]) <= 1e-5 * abs(oen[3])
    assert abs(en[2] - oen[2]) <= 1e-12 * abs(oen[2])


def test_full_size_net_force_and_determinism(full):
    s, p, e = full
    f1 = e.forces()
    # the total force is zero up to the mesh's aliasing and FP32 rounding
    net = np.abs(f1.sum(0)).max() / np.abs(f1).sum(0).max()
    print(f"[full] net force / sum |F| = {net:.2e}", flush=True)
    assert net < 1e-5
    e.compute()
    assert np.array_equal(e.forces(), f1)
    other = Engine(s, p)
    other.set_positions(s.r)
    other.compute()
    assert np.array_equal(other.forces(), f1)


def test_full_size_md_keeps_constraints_and_energy(gpu_lib):
    s, p = W.config(NAME)
    e = Engine(s, p)
    e.set_positions(s.r)
    e.set_velocities(s.v)
    e.md_setup(dt=0.001)
    e.md_run(1, graph=False)
    en0 = e.energies()[0]           # [3] potential total, [4] kinetic
    b0 = e.nl_builds()
    e.md_run(200)
    en1 = e.energies()[0]
    r = e.positions().reshape(-1, 3, 3)
    a, b = r[:, 1] - r[:, 0], r[:, 2] - r[:, 0]
    bond = max(np.abs(0.5 * ((a * a).sum(1) - 1)).max(), np.abs(0.5 * ((b * b).sum(1) - 1)).max())
    tot0, tot1 = en0[3] + en0[4], en1[3] + en1[4]
    print(f"[full] 200 steps: builds {e.nl_builds() - b0}, bond err {bond:.2e}, "
          f"E {tot0:.3f} -> {tot1:.3f} (kinetic {en1[4]:.3f})", flush=True)
    assert e.errors() == 0
    assert e.nl_builds() - b0 >= 1
    assert bond <= 1e-8
    assert abs(tot1 - tot0) <= 0.02 * abs(en1[4])
