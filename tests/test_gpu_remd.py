"""Replica exchange over the systems of one engine (mdyn.cpp:317-351, 476-497).

Checks:
  * the exchange decisions equal a host restatement of the reference rule
    du = -U1/kT0 - U0/kT1 + U0/kT0 + U1/kT1, accept if du < 0 or u < exp(-du),
    resolved from the hottest pair down with each replica's last energy,
    driven by the same counter-based random numbers; positions move
    bit-exactly with the permutation, and equal temperatures always swap;
  * forces recomputed after a swap equal the forces the configuration had
    in its previous system (same Hamiltonian) to FP32 summation order;
  * redrawn velocities satisfy the rigid-water constraints, carry no net
    momentum, and have a kinetic temperature near the replica's;
  * a mixed batch (two box sizes, Ewald and P3M, several replicas each)
    runs exchange intervals through the captured step graph without error.
"""
import numpy as np
import pytest

from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine, E_ELEC, E_LJ, E_KINETIC

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def _splitmix(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform01(seed, stream, idx):
    k = _splitmix((seed * 0x100000001B3 + stream) & M64)
    return (_splitmix((idx ^ k) & M64) >> 11) * (1.0 / 9007199254740992.0)


def predict(ladders, kT, u, seed, attempt, nsys):
    perm = list(range(nsys))
    acc = np.zeros(nsys, dtype=int)
    for lad in ladders:
        for k in range(len(lad) - 2, -1, -1):
            a, b = lad[k], lad[k + 1]
            du = -u[b] / kT[a] - u[a] / kT[b] + u[a] / kT[a] + u[b] / kT[b]
            if du < 0 or uniform01(seed, attempt, a) < np.exp(-du):
                perm[a], perm[b] = perm[b], perm[a]
                acc[a] += 1
    return perm, acc


def replicas(n, L, temps, seed=11):
    systems, params = [], []
    for k, T in enumerate(temps):
        s = W.water_box(n, L, seed + k, temperature=T)
        rc = min(9.0, 0.45 * L)
        p = dict(rc=rc, smoothing_width=1.0, eta=3.2 / rc, kmode=0, kc=6.0)
        p["skin"] = W.default_skin(rc, L)
        p.update(nalias=2, ik=0, include_lj_correction=1)
        systems.append(s)
        params.append(p)
    return systems, params


def make_engine(systems, params):
    e = Engine(systems, params)
    e.set_positions(np.concatenate([s.r for s in systems]))
    e.set_velocities(np.concatenate([s.v for s in systems]))
    e.md_setup(dt=0.001)
    return e


def test_exchange_matches_reference_rule(gpu_lib):
    temps = [300.0, 300.0, 340.0, 390.0, 450.0]   # first pair: equal kT always swaps
    systems, params = replicas(4, 4 * 3.1067, temps)
    e = make_engine(systems, params)
    kT = np.array([W.K_to_kcal_mol(T) for T in temps])
    ladders = [[0, 1, 2, 3, 4]]
    seed = 12345
    e.remd_setup(ladders, kT, seed=seed)
    e.md_run(20)
    en = e.energies()
    u = en[:, E_ELEC] + en[:, E_LJ]
    r_before = e.positions()
    f_before = e.forces()
    perm, acc = predict(ladders, kT, u, seed, 0, len(systems))
    print("energies", u, "predicted perm", perm)
    e.remd_run(1, 0, reinit_velocities=False)
    st, cfg = e.remd_stats()
    assert list(cfg) == perm
    assert st[:4, 0].tolist() == [1, 1, 1, 1] and st[4, 0] == 0
    assert st[:, 1].tolist() == acc.tolist()
    assert perm[0] != 0, "equal temperatures must swap"
    r_after = e.positions()
    f_after = e.forces()
    off = e.offsets
    for s_, src in enumerate(perm):
        a, b = off[s_], off[s_ + 1]
        c, d = off[src], off[src + 1]
        assert np.array_equal(r_after[a:b], r_before[c:d]), (s_, src)
        fr = np.linalg.norm(f_after[a:b] - f_before[c:d]) / np.linalg.norm(f_before[c:d])
        print("system", s_, "from", src, "force rel diff", fr)
        assert fr < 1e-5
    en2 = e.energies()
    u2 = en2[:, E_ELEC] + en2[:, E_LJ]
    for s_, src in enumerate(perm):
        assert abs(u2[s_] - u[src]) <= 1e-5 * abs(u[src])


def test_velocity_redraw(gpu_lib):
    temps = [300.0, 360.0, 450.0]
    systems, params = replicas(6, 18.64, temps, seed=3)
    e = make_engine(systems, params)
    kT = np.array([W.K_to_kcal_mol(T) for T in temps])
    e.remd_setup([[0, 1, 2]], kT, seed=99)
    e.remd_run(1, 0, reinit_velocities=True)
    v = e.velocities()
    r = e.positions()
    m = np.concatenate([s.mass for s in systems])
    off = e.offsets
    for s_, T in enumerate(temps):
        a, b = off[s_], off[s_ + 1]
        vs, rs, ms = v[a:b], r[a:b], m[a:b]
        p = (ms[:, None] * vs).sum(0)
        assert np.abs(p).max() < 1e-9 * np.abs(ms[:, None] * vs).sum()
        x = rs.reshape(-1, 3, 3)
        w = vs.reshape(-1, 3, 3)
        # d/dt of the O-H lengths and of the H-H distance (rigid water)
        for i, j in ((0, 1), (0, 2), (1, 2)):
            rate = ((x[:, i] - x[:, j]) * (w[:, i] - w[:, j])).sum(1)
            assert np.abs(rate).max() < 1e-6
        nw = x.shape[0]
        ndof = 6 * nw - 3
        ke = 0.5 * (ms[:, None] * vs ** 2).sum()
        T_kin = 2 * ke / (ndof * W.K_to_kcal_mol(1.0))
        print("system", s_, "target", T, "kinetic T", T_kin)
        assert abs(T_kin / T - 1) < 0.25
    en = e.energies()
    assert np.all(np.isfinite(en))


def test_mixed_batch_intervals(gpu_lib):
    systems, params, kT, ladders = W.remd_batch(nrep=3, sizes=((6, 18.64), (8, 24.86)))
    e = make_engine(systems, params)
    e.remd_setup(ladders, kT, seed=5)
    b0 = e.nl_builds()
    e.remd_run(3, 10, reinit_velocities=True, graph=True)
    st, cfg = e.remd_stats()
    en = e.energies()
    print("tries/accepts", st.tolist(), "config", cfg.tolist(), "nl builds", e.nl_builds() - b0)
    assert sorted(cfg[:3]) == [0, 1, 2] and sorted(cfg[3:]) == [3, 4, 5]
    assert st[0, 0] == 3 and st[2, 0] == 0
    assert np.all(np.isfinite(en[:, :5]))
    assert e.nl_builds() > b0


STREAM_REINIT = 0x5343484544000001
STREAM_EXCHANGE = 0x5343484544000002


def schedule(seed, first, nsteps, xint, rint):
    """do_every (mdyn.cpp:49-52) per step, restated: (exchanges, redraws)."""
    nx = nr = 0
    for k in range(first, first + nsteps):
        nr += rint > 0 and rint * uniform01(seed, STREAM_REINIT, k) < 1
        nx += xint > 0 and xint * uniform01(seed, STREAM_EXCHANGE, k) < 1
    return nx, nr


def test_stochastic_schedule(gpu_lib):
    """The reference's per-step move draws: the engine makes exactly the
    exchanges and velocity redraws a host restatement of the schedule
    predicts, continues the step count across calls, and every exchange
    attempt is counted in the ladders' tries."""
    systems, params, kT, ladders = W.remd_batch(nrep=3, sizes=((6, 18.64), (8, 24.86)))
    e = make_engine(systems, params)
    seed = 9
    e.remd_setup(ladders, kT, seed=seed)
    got = [e.remd_run_stochastic(40, 7, 13), e.remd_run_stochastic(25, 5, 0)]
    want = [schedule(seed, 0, 40, 7, 13), schedule(seed, 40, 25, 5, 0)]
    print("[remd] stochastic moves", got, "host schedule", want, flush=True)
    assert got == want and want[0][0] > 0 and want[0][1] > 0
    st, cfg = e.remd_stats()
    assert st[0, 0] == got[0][0] + got[1][0] and st[3, 0] == st[0, 0]
    assert sorted(cfg[:3]) == [0, 1, 2] and sorted(cfg[3:]) == [3, 4, 5]
    assert e.errors() == 0 and np.all(np.isfinite(e.energies()[:, :5]))
    # interval 1: an attempt after every step
    assert e.remd_run_stochastic(3, 1, 0) == (3, 0)
