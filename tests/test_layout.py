"""Host check of the engine's device-buffer layout (mm_check_layout): for
mixed batches of Ewald and P3M systems of different sizes, every system's
slice of each concatenated buffer (sites, groups, cells, stencils, erfc and
LJ tables, wave vectors, meshes, influence functions, twiddles, spreading
bins) lies inside the buffer and the slices are disjoint.  Runs without a
GPU: the layout is the host half of mm_engine_create."""
import ctypes as C

import pytest

from molmodel_amd import engine as E
from molmodel_amd import waterbox as W


def check(systems, params):
    descs, keep = [], []
    for s, p in zip(systems, params):
        d, k = E.system_desc(s, p)
        descs.append(d)
        keep.append(k)
    arr = (E.SystemDesc * len(descs))(*descs)
    L = E.lib()
    L.mm_check_layout.argtypes = [C.c_int, C.c_void_p]
    rc = L.mm_check_layout(len(descs), C.cast(arr, C.c_void_p))
    return rc, L.mm_last_error().decode()


def tuned(boxl, nsite, rc):
    c = E.tune_kspace(boxl, nsite, rc, 1e-4)
    return {k: v for k, v in c.items() if k in ("eta", "kmode", "kc", "mesh", "order")}


def test_replica_batch_of_three_sizes():
    systems, params, _, _ = W.remd_batch(nrep=4, kspace=tuned)
    assert len(systems) == 12 and len({s.nsite for s in systems}) == 3
    rc, msg = check(systems, params)
    assert rc == 0, msg


@pytest.mark.parametrize("order", [
    ("water648_ewald", "water12k_p3m", "water1000_ions"),
    ("water12k_p3m", "water648_ewald", "water12k_p3m", "water1000_ions"),
    ("water1000_ions", "water1000_ions", "water648_ewald"),
])
def test_mixed_ewald_p3m_batches(order):
    cfgs = [W.config(n) for n in order]
    # a second P3M mesh size in the same batch
    s, p = W.config("water12k_p3m")
    p2 = dict(p, mesh=(40, 40, 40), order=4)
    systems = [c[0] for c in cfgs] + [s]
    params = [c[1] for c in cfgs] + [p2]
    rc, msg = check(systems, params)
    assert rc == 0, msg


def test_large_and_small_together():
    a, pa = W.config("water98k_p3m")
    b, pb = W.config("water648_ewald")
    rc, msg = check([b, a, b], [pb, pa, pb])
    assert rc == 0, msg


def test_invalid_input_is_reported():
    s, p = W.config("water648_ewald")
    rc, msg = check([s], [dict(p, rc=0.6 * s.boxl)])
    assert rc < 0 and "0.49" in msg


def test_default_shape_and_shape_checks():
    assert E.default_shape() == dict(cell_groups=2, pair_slots=3, spread_block=2, graph_steps=10,
                                     spread_halo=0)
    s, p = W.config("water648_ewald")
    d, keep = E.system_desc(s, p)
    h = C.c_void_p()
    L = E.lib()
    for bad in (dict(pair_slots=0), dict(pair_slots=5), dict(cell_groups=0),
                dict(spread_block=9), dict(graph_steps=65), dict(spread_halo=2)):
        sh = E._shape(bad)
        rc = L.mm_engine_create_shaped(1, C.byref(d), C.byref(sh), C.byref(h))
        assert rc == -1 and "mm_shape" in L.mm_last_error().decode()
