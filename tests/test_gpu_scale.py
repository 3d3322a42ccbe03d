"""Box scaling of one system (mm_scale_system, ElecLJ::scale,
eleclj.cpp:1268-1280 with p3m_scale, pppm.c:643-664): after scaling by f
and moving each molecule rigidly with its first site (the volume move's
centroid scaling), forces and energies equal those of a fresh engine built
for the box f L with eta / f, and the FP64 oracle's for the same inputs."""
import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import waterbox as W
from molmodel_amd.engine import Engine, lib

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def scaled_positions(s, f):
    r = s.r.copy()
    for g0, g1 in s.group:
        r[g0:g1] += (f - 1) * s.r[g0]
    return r


@pytest.mark.parametrize("name,ik", [("water648_ewald", 0), ("water12k_p3m", 0),
                                     ("water12k_p3m", 1)])
def test_scaled_box_matches_fresh_engine_and_oracle(gpu_lib, name, ik):
    s, p = W.config(name)
    p = dict(p, ik=ik)
    f = 1.015
    e = Engine(s, p)
    e.set_positions(s.r)
    e.compute()
    f0 = e.forces()
    assert lib().mm_scale_system(e.h, 0, f) == 0, lib().mm_last_error()
    r1 = scaled_positions(s, f)
    e.set_positions(r1)
    e.compute()
    fs, es = e.forces(), e.energies()[0]
    s2 = W.System(**{k: getattr(s, k) for k in s.__dataclass_fields__})
    s2.boxl = s.boxl * f
    s2.r = r1
    p2 = dict(p, eta=p["eta"] / f)
    e2 = Engine(s2, p2)
    e2.set_positions(r1)
    e2.compute()
    ff, ef = e2.forces(), e2.energies()[0]
    fo, eo = O.evaluate(s2, p2, r=r1)
    print(name, "scaled vs fresh", rel(fs, ff), "vs oracle", rel(fs, fo),
          "energy", es[3], ef[3], eo[3])
    assert rel(fs, ff) < 1e-6
    assert abs(es[3] - ef[3]) <= 1e-6 * abs(ef[3])
    assert rel(fs, fo) < 1e-5
    assert abs(es[3] - eo[3]) <= 1e-5 * abs(eo[3])
    # and back
    assert lib().mm_scale_system(e.h, 0, 1 / f) == 0
    e.set_positions(s.r)
    e.compute()
    assert rel(e.forces(), f0) < 1e-6
