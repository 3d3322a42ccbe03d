"""The 48-system replica-exchange batch at its real shape: 16 replicas each of
216, 512 and 1000 waters (L = 18.64, 24.86, 31.07 A), temperatures geometric
from 300 to 450 K within each size, reciprocal space chosen per size by the
library's tuner at 1e-4 (Ewald for 216 waters, P3M order 6 on 30^3 and 36^3
meshes for the others; the tuner keeps to orders the reference runs), all
in one engine, two 100-step intervals through the captured step graph, each
ending in an exchange attempt and a velocity redraw.

Bars, per system:
  * forces, energies and the virial against the FP64 oracle at the start and
    at the end of the run: force relative RMS <= 1e-5, energies <= 1e-5 of
    |E_elec| + |E_LJ| (a replica's two terms can nearly cancel), virial
    <= 1e-5 relative (Frobenius);
  * the exchange decisions of both attempts equal the host restatement of
    the reference rule (mdyn.cpp:317-351) driven by the same counter-based
    random numbers and the energies the device held at the attempt;
  * the rigid-water constraints hold to the iteration tolerance (1e-8).
"""
import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine, E_ELEC, E_LJ, tune_kspace
from test_gpu_remd import predict

pytestmark = pytest.mark.gpu

SEED = 11


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def kspace(boxl, nsite, rc):
    c = tune_kspace(boxl, nsite, rc, 1e-4)
    return {k: v for k, v in c.items() if k in ("eta", "kmode", "kc", "mesh", "order")}


@pytest.fixture(scope="module")
def batch():
    systems, params, kT, ladders = W.remd_batch(nrep=16, kspace=kspace)
    assert [s.nwater for s in systems] == [216] * 16 + [512] * 16 + [1000] * 16
    e = Engine(systems, params)
    e.set_positions(np.concatenate([s.r for s in systems]))
    e.set_velocities(np.concatenate([s.v for s in systems]))
    e.md_setup(dt=0.001)
    e.remd_setup(ladders, kT, seed=SEED)
    return systems, params, np.asarray(kT), ladders, e


def check_against_oracle(systems, params, e, tag):
    e.compute(virial=True)
    f = e.forces()
    en = e.energies()
    w = e.virial()
    r = e.positions()
    off = e.offsets
    worst = [0.0, 0.0, 0.0]
    for k, (s, p) in enumerate(zip(systems, params)):
        a, b = off[k], off[k + 1]
        of, oen, ow = O.evaluate(s, p, r=r[a:b], virial=True)
        ferr = rel(f[a:b], of)
        # relative to the terms' size: a replica's elec and LJ terms can
        # nearly cancel (one 216-water start sums to -1.0 from +82 and -83)
        eerr = max(abs(en[k, 0] - oen[0]), abs(en[k, 1] - oen[1])) / (abs(oen[0]) + abs(oen[1]))
        werr = rel(w[k], ow.reshape(3, 3))
        worst = [max(worst[0], ferr), max(worst[1], eerr), max(worst[2], werr)]
        assert ferr <= 1e-5, (tag, k, ferr)
        assert eerr <= 1e-5, (tag, k, eerr)
        assert werr <= 1e-5, (tag, k, werr)
    print(f"[remd48] {tag}: worst force {worst[0]:.2e} energy {worst[1]:.2e} virial {worst[2]:.2e}",
          flush=True)


def constraint_residual(r):
    x = r.reshape(-1, 3, 3)
    a, b = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
    la, lb = (a * a).sum(1), (b * b).sum(1)
    th = np.arccos(np.clip((a * b).sum(1) / np.sqrt(la * lb), -1, 1))
    return max(np.abs(0.5 * (la - 1)).max(), np.abs(0.5 * (lb - 1)).max(),
               np.abs(th - np.radians(109.47)).max())


def test_remd48_two_intervals(gpu_lib, batch):
    systems, params, kT, ladders, e = batch
    print("[remd48] kspace per size:",
          [{k: params[i][k] for k in ("kmode", "mesh", "order", "kc") if k in params[i]}
           for i in (0, 16, 32)], flush=True)
    assert params[0]["kmode"] == 0
    assert params[16]["kmode"] == 1 and params[32]["kmode"] == 1
    assert params[16]["order"] <= 6 and params[32]["order"] <= 6
    check_against_oracle(systems, params, e, "start")
    nsys = len(systems)
    cfg = list(range(nsys))
    total = np.zeros(nsys, dtype=np.int64)
    for attempt in range(2):
        e.md_run(100, graph=True)
        en = e.energies()
        u = en[:, E_ELEC] + en[:, E_LJ]
        perm, acc = predict(ladders, kT, u, SEED, attempt, nsys)
        e.remd_run(1, 0, reinit_velocities=True, graph=True)
        st, got = e.remd_stats()
        cfg = [cfg[perm[s]] for s in range(nsys)]
        total += acc
        print(f"[remd48] attempt {attempt}: accepted {int(acc.sum())} of {3 * 15}", flush=True)
        assert list(got) == cfg, (attempt, list(got), cfg)
        assert st[:, 1].tolist() == total.tolist()
        tries = np.array([0 if k in (15, 31, 47) else attempt + 1 for k in range(nsys)])
        assert st[:, 0].tolist() == tries.tolist()
    assert e.errors() == 0
    r = e.positions()
    off = e.offsets
    res = max(constraint_residual(r[off[k]:off[k + 1]]) for k in range(nsys))
    print(f"[remd48] constraint residual {res:.2e}", flush=True)
    assert res <= 1e-8
    check_against_oracle(systems, params, e, "end")
