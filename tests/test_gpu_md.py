"""GPU rigid-water dynamics against the FP64 oracle's velocity Verlet with the
reference's constraint iteration (mdyn.cpp:142-160, constraint.cpp:295-358).

Bars:
  * 10 steps of the 648-atom Ewald box: max |r_gpu - r_oracle| <= 1e-6 A and
    per-step potential energy within 1e-5 of |U| (FP32 pair arithmetic seeds a
    ~1e-7 relative force difference that the short trajectory keeps small);
  * constraints hold to the iteration tolerance: |0.5 (r_OH^2 - 1)| <= 1e-8,
    |theta - theta0| <= 1e-8;
  * a captured hipGraph replay gives bit-identical states to stream launches;
  * 1000 steps conserve the total energy to 2% of the kinetic energy.
"""
import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine

pytestmark = pytest.mark.gpu


def setup(name, seed=2024):
    s, p = W.config(name, seed)
    e = Engine(s, p)
    e.set_positions(s.r)
    e.set_velocities(s.v)
    e.md_setup(dt=0.001)
    return s, p, e


def water_geometry_errors(r):
    x = r.reshape(-1, 3, 3)
    a, b = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
    la, lb = (a * a).sum(1), (b * b).sum(1)
    cos = (a * b).sum(1) / np.sqrt(la * lb)
    th = np.arccos(np.clip(cos, -1, 1))
    return (np.abs(0.5 * (la - 1)).max(), np.abs(0.5 * (lb - 1)).max(),
            np.abs(th - np.radians(109.47)).max())


def test_md_short_trajectory_matches_oracle(gpu_lib):
    s, p, e = setup("water648_ewald")
    print("[md] water648 10 steps", flush=True)
    pot = []
    for _ in range(10):
        e.md_run(1, graph=False)
        pot.append(e.energies()[0][3])
    r_gpu, v_gpu = e.positions(), e.velocities()
    r_ref, v_ref, en_ref = O.md_run(s, p, 10)
    dr = np.abs(r_gpu - r_ref).max()
    du = np.abs(np.array(pot) - en_ref[:, 0]).max()
    print(f"max |dr| {dr:.3e} A, max |dU| {du:.3e}, |U| {abs(en_ref[-1, 0]):.3f}")
    assert dr <= 1e-6
    assert du <= 1e-5 * abs(en_ref[-1, 0])
    assert np.abs(v_gpu - v_ref).max() <= 1e-3
    errs = water_geometry_errors(r_gpu)
    print("constraint residuals", errs)
    assert max(errs) <= 1e-8


def test_graph_replay_bit_identical(gpu_lib):
    s, p, e1 = setup("water648_ewald")
    _, _, e2 = setup("water648_ewald")
    e1.md_run(40, graph=True)
    e2.md_run(40, graph=False)
    assert np.array_equal(e1.positions(), e2.positions())
    assert np.array_equal(e1.velocities(), e2.velocities())
    # replay again after host work in between
    e1.md_run(25, graph=True)
    e2.md_run(25, graph=False)
    assert np.array_equal(e1.positions(), e2.positions())


def test_energy_conservation_1000_steps(gpu_lib):
    s, p, e = setup("water648_ewald")
    e0 = e.energies()[0]
    e.md_run(1, graph=False)
    first = e.energies()[0]
    etot0 = first[3] + first[4]
    e.md_run(999, graph=True)
    last = e.energies()[0]
    etot1 = last[3] + last[4]
    print(f"E0 {etot0:.5f} E1000 {etot1:.5f} KE {last[4]:.3f} drift {etot1 - etot0:.5f}")
    assert e.errors() == 0
    assert abs(etot1 - etot0) <= 0.02 * last[4]
    assert max(water_geometry_errors(e.positions())) <= 1e-8


def test_p3m_md_runs(gpu_lib):
    s, p, e = setup("water12k_p3m")
    print("[md] water12k p3m 50 steps", flush=True)
    e.md_run(50, graph=True)
    en = e.energies()[0]
    print("energies", en)
    assert np.all(np.isfinite(en))
    assert max(water_geometry_errors(e.positions())) <= 1e-8
