"""Engine shapes (mm_shape) and the autotuner (mm_tune_shape).

A shape changes how the batch is laid onto the device, not what is
computed:
  * pair workgroups per CU, P3M spreading blocks and spreading by owner
    into halos instead of by pulling: forces bit-identical (a pair tile's
    sums do not depend on the workgroup that takes it; the mesh is
    accumulated in exact fixed point, the halos summed as integers);
  * steps per captured graph: trajectories bit-identical;
  * neighbour-list cells of 1 or 4 groups on average instead of the
    reference's 2: the same unordered pair set as the oracle's build (which
    row holds a pair follows the cells), forces equal to FP64 summation
    order (1e-12 relative) and within the parity bar of the oracle.
The tuner must return one of the shapes it timed.
"""
import numpy as np
import pytest

import oracle_lib as O
from molmodel_amd import engine as E
from molmodel_amd import waterbox as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def unordered(pairs):
    p = np.sort(np.asarray(pairs, dtype=np.int64), axis=1)
    return np.unique(p[:, 0] * (1 << 32) + p[:, 1])


def forces_of(s, p, shape):
    e = E.Engine(s, p, shape=shape)
    e.set_positions(s.r)
    e.compute()
    return e, e.forces(), e.energies()[0]


@pytest.fixture(scope="module")
def p3m_box():
    s, p = W.config("water12k_p3m")
    e, f, en = forces_of(s, p, None)
    return s, p, f, en, e.nl_pairs(0)


@pytest.mark.parametrize("shape", [dict(pair_slots=2), dict(pair_slots=4), dict(pair_slots=1),
                                   dict(spread_block=4), dict(spread_block=1),
                                   dict(spread_halo=1), dict(spread_halo=1, spread_block=4),
                                   dict(spread_halo=1, spread_block=1)])
def test_tile_and_spread_shapes_are_bit_identical(gpu_lib, p3m_box, shape):
    s, p, f0, en0, _ = p3m_box
    print(f"[shape] {shape}", flush=True)
    _, f, en = forces_of(s, p, shape)
    assert np.array_equal(f, f0) and np.array_equal(en, en0)


@pytest.mark.parametrize("cells", [1, 4])
def test_cell_sizes_keep_the_pair_set(gpu_lib, p3m_box, cells):
    s, p, f0, en0, pairs0 = p3m_box
    print(f"[shape] cell_groups {cells}", flush=True)
    e, f, en = forces_of(s, p, dict(cell_groups=cells))
    ref = unordered(O.nl_pairs(O.first_sites(s), s.boxl, p["rc"], p["skin"]))
    assert np.array_equal(unordered(e.nl_pairs(0)), ref)
    assert np.array_equal(unordered(pairs0), ref)
    assert rel(f, f0) < 1e-12
    assert np.allclose(en, en0, rtol=1e-12, atol=0)
    of, _ = O.evaluate(s, p)
    assert rel(f, of) <= 1e-5


def test_graph_steps_give_the_same_trajectory(gpu_lib):
    s, p = W.config("water648_ewald")
    runs = []
    for k in (1, 7, 20):
        print(f"[shape] graph_steps {k}", flush=True)
        e = E.Engine(s, p, shape=dict(graph_steps=k))
        e.set_positions(s.r)
        e.set_velocities(s.v)
        e.md_setup(dt=0.001)
        e.md_run(45)
        runs.append((e.positions(), e.velocities()))
    for r, v in runs[1:]:
        assert np.array_equal(r, runs[0][0]) and np.array_equal(v, runs[0][1])


@pytest.mark.parametrize("name", ["water648_ewald", "water12k_p3m"])
def test_tuner_picks_a_timed_shape(gpu_lib, name):
    s, p = W.config(name)
    print(f"[shape] tuning {name}", flush=True)
    best, trials = E.tune_shape(s, p, dt=0.001, steps=30)
    print(f"[shape] {name}: best {best}; " +
          ", ".join(f"{t}: {ms:.4f}" for t, ms in trials), flush=True)
    shapes = [t for t, _ in trials]
    assert best in shapes and all(ms > 0 for _, ms in trials)
    assert trials[0][0] == E.default_shape()
    # one choice at a time: pair slots (3), spreading blocks and the way of
    # spreading for P3M (3 + 2), cells (3), graph steps (3), each shape
    # timed once
    assert len(shapes) == len({tuple(t.values()) for t in shapes})
    assert len(shapes) >= (7 if p["kmode"] == 0 else 10)
    ms = dict((tuple(t.values()), m) for t, m in trials)
    assert ms[tuple(best.values())] == min(ms.values())
