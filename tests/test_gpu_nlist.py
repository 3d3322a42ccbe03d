"""The standalone device neighbour list (mm_nlist_*, the NeighborList
surface of nbrlist.hpp:8-40) against the reference's own output held in
tests/golden/ref_nl_small.npz: random point sets (some outside the central
box) in small boxes, and a two-step call that exercises the skin test.
Bar: the pair set and each pair's PairBinary row (pbin.c:6 owner) equal
the reference's exactly; the rebuild decision equals the reference's."""
import ctypes as C
import os

import numpy as np
import pytest

from molmodel_amd import engine as E

pytestmark = pytest.mark.gpu
P = C.c_void_p

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_nl_small.npz")


def lib():
    L = E.lib()
    L.mm_nlist_create.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(P)]
    L.mm_nlist_update.argtypes = [P, P]
    L.mm_nlist_rows.restype = C.c_long
    L.mm_nlist_rows.argtypes = [P, P, P, C.c_long]
    L.mm_nlist_destroy.argtypes = [P]
    return L


def rows_as_pairs(L, h, n):
    npairs = L.mm_nlist_rows(h, None, None, 0)
    assert npairs >= 0, L.mm_last_error()
    start = np.zeros(n + 1, dtype=np.int32)
    cols = np.zeros(max(npairs, 1), dtype=np.int32)
    L.mm_nlist_rows(h, start.ctypes.data_as(P), cols.ctypes.data_as(P), npairs)
    out = [(i, int(c)) for i in range(n) for c in cols[start[i]:start[i + 1]]]
    return sorted(out)


def new_list(L, n, rc, skin, boxl):
    h = P()
    assert L.mm_nlist_create(n, rc, skin, boxl, C.byref(h)) == 0, L.mm_last_error()
    return h


def test_nlist_matches_reference_cases(gpu_lib):
    L = lib()
    g = np.load(GOLD)
    i = 0
    while f"c{i}_n" in g:
        n = int(g[f"c{i}_n"])
        x = np.ascontiguousarray(g[f"c{i}_x"], dtype=np.float64)
        boxl, rc, skin = float(g[f"c{i}_boxl"]), float(g[f"c{i}_rc"]), float(g[f"c{i}_skin"])
        h = new_list(L, n, rc, skin, boxl)
        assert L.mm_nlist_update(h, x.ctypes.data_as(P)) == 1
        mine = rows_as_pairs(L, h, n)
        ref = sorted(map(tuple, g[f"c{i}_pairs"].tolist()))
        print("case", i, "n", n, "pairs", len(mine), "reference", len(ref))
        assert mine == ref, i
        L.mm_nlist_destroy(h)
        i += 1
    assert i >= 5


def test_nlist_skin_test_two_step(gpu_lib):
    L = lib()
    g = np.load(GOLD)
    n, boxl, rc, skin = g["two_params"]
    n = int(n)
    x0 = np.ascontiguousarray(g["two_x0"])
    for tag, expect_rebuild in (("small", 0), ("large", 1)):
        h = new_list(L, n, rc, skin, boxl)
        assert L.mm_nlist_update(h, x0.ctypes.data_as(P)) == 1
        x1 = np.ascontiguousarray(g[f"two_{tag}_x1"])
        assert L.mm_nlist_update(h, x1.ctypes.data_as(P)) == expect_rebuild, tag
        mine = rows_as_pairs(L, h, n)
        assert mine == sorted(map(tuple, g[f"two_{tag}_pairs"].tolist())), tag
        L.mm_nlist_destroy(h)
