"""Synthetic system generator: geometry, units, charge neutrality,
rigid-body velocities and zero momentum."""
import math

import numpy as np

from molmodel_amd import waterbox as W


def test_water_geometry_and_charges():
    s, p = W.config("water648_ewald")
    x = s.r.reshape(-1, 3, 3)
    a, b = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
    assert np.allclose(np.linalg.norm(a, axis=1), 1.0, atol=1e-12)
    ang = np.degrees(np.arccos((a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))))
    assert np.allclose(ang, 109.47, atol=1e-9)
    assert abs(s.q.sum()) < 1e-12
    assert abs(s.q[0] / W.CHARGE_UNIT + 0.8476) < 1e-15
    assert s.nsite == 648 and s.ngroup == 216 and len(s.excl) == 648


def test_velocities_rigid_and_momentum_free():
    s, p = W.config("water648_ewald")
    P = (s.mass[:, None] * s.v).sum(0)
    assert np.abs(P).max() < 1e-10
    x = s.r.reshape(-1, 3, 3); v = s.v.reshape(-1, 3, 3)
    for i, j in ((0, 1), (0, 2), (1, 2)):
        d = (x[:, i] - x[:, j]) * (v[:, i] - v[:, j])
        assert np.abs(d.sum(1)).max() < 1e-10
    # temperature near 300 K over 6N - 3N_constraints - 3 degrees of freedom
    ke = 0.5 * (s.mass[:, None] * s.v ** 2).sum()
    dof = 6 * s.nwater - 3
    T = 2 * ke / (dof * W.K_to_kcal_mol(1.0))
    assert 250 < T < 350


def test_counter_rng_reproducible():
    assert np.array_equal(W.uniform(5, 3, 100), W.uniform(5, 3, 100))
    assert not np.array_equal(W.uniform(5, 3, 100), W.uniform(6, 3, 100))
    u = W.uniform(1, 1, 200000)
    assert abs(u.mean() - 0.5) < 5e-3 and u.min() >= 0 and u.max() < 1


def test_ion_box():
    s, p = W.config("water1000_ions")
    assert s.nsite == 3040 and s.ngroup == 1040 and s.nion == 40
    assert abs(s.q.sum()) < 1e-9
    ions = s.r[3000:]
    d = s.r[None, :, :] - ions[:, None, :]
    d -= s.boxl * np.floor(d / s.boxl + 0.5)
    r = np.sqrt((d * d).sum(-1))
    r[np.arange(40), 3000 + np.arange(40)] = 99
    assert r.min() >= 2.0
