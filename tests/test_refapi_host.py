"""Host-side reference-named entry points of libmolmodel_amd.so (no GPU):
PairBinary (pbin.c) against the reference's known-answer example, and
RSpaceChebyshev_new (rspace.c:25-37) against the reference's table."""
import ctypes as C
import json
import os

import numpy as np

from molmodel_amd import engine as E

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def lib():
    L = E.lib()
    L.PairBinary_new.restype = C.c_void_p
    L.PairBinary_new.argtypes = [C.c_int]
    L.PairBinary_add.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.PairBinary_init.argtypes = [C.c_void_p]
    L.PairBinary_delete.argtypes = [C.c_void_p]
    L.PairBinary_exists.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.PairBinary_number_of_elements.restype = C.POINTER(C.c_int)
    L.PairBinary_number_of_elements.argtypes = [C.c_void_p]
    L.PairBinary_elements.restype = C.POINTER(C.POINTER(C.c_int))
    L.PairBinary_elements.argtypes = [C.c_void_p]
    L.RSpaceChebyshev_new.restype = C.c_void_p
    L.RSpaceChebyshev_new.argtypes = [C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    return L


def test_pairbinary_kat():
    d = json.load(open(os.path.join(GOLD, "pbin_kat.json")))
    L = lib()
    p = L.PairBinary_new(10)
    for i, j in d["adds"]:
        L.PairBinary_add(p, i, j)
    L.PairBinary_init(p)
    nn = L.PairBinary_number_of_elements(p)
    el = L.PairBinary_elements(p)
    rows = [[el[i][k] for k in range(nn[i])] for i in range(10)]
    assert rows == d["expected_rows_from_comment"]
    ex = [[L.PairBinary_exists(p, i, j) for j in range(10)] for i in range(10)]
    assert ex == d["exists"]
    L.PairBinary_delete(p)


def test_chebyshev_table_matches_reference():
    g = np.load(os.path.join(GOLD, "ref_water648.npz"), allow_pickle=False)
    L = lib()
    t = L.RSpaceChebyshev_new(9.0, 0.35, None, None)
    arr = np.ctypeslib.as_array((C.c_double * (3 + 512 * 7)).from_address(t))
    assert np.array_equal(arr, g["cheby"])
