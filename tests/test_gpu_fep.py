"""Free-energy perturbation windows (ElecLJ::types_changed(pert) + set_lambda,
eleclj.cpp:267-308; msimprog.cpp:80-114) on the water + ions box.

State B takes the sodium ions' charge to zero and changes their sigma and
epsilon (2.35 -> 3.0 A, 0.13 -> 0.2 kcal/mol); the box then carries a net
charge for lambda > 0, so the neutralising background term is exercised.
Three windows (lambda 0, 0.5, 1) are three systems of one engine, evaluated
in one launch.  Each window's forces, energies and virial are checked
against the FP64 oracle evaluated with the interpolated site parameters
q = (1 - l) qA + l qB, sigma and epsilon likewise on sites with LJ (the tail
correction over the "%12.6g" keys of those parameters): force relative RMS
<= 1e-5, energies <= 1e-5 of |U|, virial <= 1e-5.  The FEP energy
difference U(1) - U(0) matches the oracle's, and lambda = 0 gives back the
unperturbed system's forces exactly.
"""
import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine

pytestmark = pytest.mark.gpu

LAMBDAS = (0.0, 0.5, 1.0)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def state_b(s):
    qB, sB, eB = s.q.copy(), s.sig.copy(), s.eps.copy()
    na = np.nonzero(np.isclose(s.q, W.e_to_charge_unit(1.0)))[0]
    assert len(na) == 20
    qB[na] = 0.0
    sB[na] = 3.0
    eB[na] = 0.2
    return qB, sB, eB


def interpolated(s, lam, qB, sB, eB):
    t = W.System(**{k: getattr(s, k) for k in s.__dataclass_fields__})
    t.q = (1 - lam) * s.q + lam * qB
    lj = (s.site_mask & W.SITE_HAS_LJ) != 0
    t.sig = np.where(lj, (1 - lam) * s.sig + lam * sB, s.sig)
    t.eps = np.where(lj, (1 - lam) * s.eps + lam * eB, s.eps)
    return t


@pytest.mark.parametrize("kspace", ["config", "p3m_ik"])
def test_fep_windows_match_oracle(gpu_lib, kspace):
    s, p = W.config("water1000_ions")
    if kspace == "p3m_ik":   # the windows' net charges differ; ik field on a 32^3 mesh
        p = dict(p, kmode=1, mesh=(32, 32, 32), order=5, ik=1)
    qB, sB, eB = state_b(s)
    n = len(LAMBDAS)
    e = Engine([s] * n, [p] * n)
    e.set_positions(np.concatenate([s.r] * n))
    for k, lam in enumerate(LAMBDAS):
        e.set_perturbation(k, qB, sB, eB)
        e.set_lambda(k, lam)
    print(f"[fep] three windows in one engine ({kspace})", flush=True)
    e.compute(virial=True)
    f, en, w = e.forces(), e.energies(), e.virial()
    off = e.offsets
    u = []
    for k, lam in enumerate(LAMBDAS):
        t = interpolated(s, lam, qB, sB, eB)
        of, oen, ow = O.evaluate(t, p, virial=True)
        a, b = off[k], off[k + 1]
        ferr = rel(f[a:b], of)
        eerr = max(abs(en[k, 0] - oen[0]), abs(en[k, 1] - oen[1])) / abs(oen[3])
        terr = abs(en[k, 2] - oen[2]) / abs(oen[2])
        werr = rel(w[k], ow.reshape(3, 3))
        print(f"[fep] lambda {lam}: force {ferr:.2e} energy {eerr:.2e} tail {terr:.2e} "
              f"virial {werr:.2e} U {en[k, 3]:.6f} vs {oen[3]:.6f}")
        assert ferr <= 1e-5 and eerr <= 1e-5 and werr <= 1e-5
        assert terr <= 1e-12
        u.append((en[k, 3], oen[3]))
    du, dou = u[-1][0] - u[0][0], u[-1][1] - u[0][1]
    print(f"[fep] U(1) - U(0): {du:.6f} vs oracle {dou:.6f}")
    assert abs(du - dou) <= 1e-5 * (abs(u[0][1]) + abs(u[-1][1]))
    # lambda 0 is the unperturbed system
    plain = Engine(s, p)
    plain.set_positions(s.r)
    plain.compute()
    assert np.array_equal(plain.forces(), f[off[0]:off[1]])


def test_lambda_moves_back_and_forth(gpu_lib):
    """Setting lambda 1 then 0 on one system restores its forces, and the
    state-B window matches a fresh engine built with state B's parameters."""
    s, p = W.config("water1000_ions")
    qB, sB, eB = state_b(s)
    e = Engine(s, p)
    e.set_positions(s.r)
    e.compute()
    f0 = e.forces()
    e.set_perturbation(0, qB, sB, eB)
    e.set_lambda(0, 1.0)
    e.compute()
    fb = e.forces()
    fresh = Engine(interpolated(s, 1.0, qB, sB, eB), p)
    fresh.set_positions(s.r)
    fresh.compute()
    assert rel(fb, fresh.forces()) <= 1e-12
    e.set_lambda(0, 0.0)
    e.compute()
    assert np.array_equal(e.forces(), f0)
