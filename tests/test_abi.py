"""The C ABI library loads on a host without a GPU and exports every
function that include/*.h declares; creating an engine without a device
fails with an error instead of falling back to anything."""
import glob
import os
import re

import numpy as np
import pytest

from molmodel_amd import engine as E
from molmodel_amd import waterbox as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#.*", "", text)
        for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mm_engine_create" in names and "mm_compute" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = E.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_engine_create_without_device_raises():
    if E.lib().mm_device_count() > 0:
        pytest.skip("a GPU is visible")
    s, p = W.config("water27_small")
    with pytest.raises(E.MMError):
        E.Engine(s, p)
