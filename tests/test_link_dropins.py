"""The drop-in boundary builds and links on the reference's side.

* tests/link/dropin_calls.c includes the reference's rspace.h, kspace.h,
  pppm.h, cheby.h and pbin.h and takes the address of every function of
  theirs that eleclj.cpp, nbrlist.cpp, contact.cpp, ppot.cpp, msys.cpp and
  bcifit.cpp call; it must link with -Wl,--no-undefined against
  libmolmodel_amd.so alone.
* The NeighborList of molmodel_amd/compat compiles with -Wall -Werror
  against the reference's cvec.hpp/bcond.hpp, the reference's own callers
  of NeighborList (ppot.cpp, contact.cpp, eleclj.cpp) compile against it,
  and a driver that uses it as ppot.cpp does links against the library
  (plus the reference's boundary-condition objects that oracle/Makefile
  builds into oracle/_ref).  Without a GPU the driver must stop with the
  library's error message: the class has no host path.

These need the reference tree and gcc; they are skipped where either is
absent.  Nothing here runs on a GPU.
"""
import os
import shutil
import subprocess

import pytest

from molmodel_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
REF_INC = os.path.join(REF, "include")
COMPAT = os.path.join(ROOT, "molmodel_amd", "compat")
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "o2")
BCOND_OBJS = ["bcond.o", "bcondtype.o", "out.o", "util.o", "fns.o", "geograd.o"]


def need_reference():
    if not os.path.isdir(REF_INC):
        pytest.skip("reference headers not present")
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not present")
    if not os.path.exists(E.LIB_PATH):
        pytest.skip("libmolmodel_amd.so not built")


def run(cmd, **kw):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300, **kw)
    return p.returncode, p.stdout + p.stderr


def link_flags():
    return [E.LIB_PATH, "-Wl,-rpath," + os.path.dirname(E.LIB_PATH),
            "-Wl,-rpath-link=/opt/rocm/lib", "-Wl,--no-undefined", "-lm"]


def test_c_dropins_link_without_undefined_symbols(tmp_path):
    need_reference()
    exe = str(tmp_path / "dropin_calls")
    # no -Werror: cheby.h defines a static function the TU does not use
    rc, out = run(["gcc", "-std=gnu99", "-Wall", "-I", REF_INC,
                   os.path.join(ROOT, "tests", "link", "dropin_calls.c"), "-o", exe]
                  + link_flags())
    assert rc == 0, out
    rc, out = run([exe])
    assert rc == 0, out
    assert int(out.split()[0]) == 34


@pytest.mark.parametrize("caller", ["ppot.cpp", "contact.cpp", "eleclj.cpp"])
def test_reference_callers_compile_against_neighborlist(caller):
    need_reference()
    rc, out = run(["g++", "-std=gnu++11", "-fsyntax-only", "-Wall", "-I", COMPAT,
                   "-I", os.path.join(ROOT, "include"), "-I", REF_INC,
                   os.path.join(REF, "src", caller)])
    assert rc == 0, out


def test_eleclj_gpu_adapter_compiles_against_the_public_surface(tmp_path):
    """ElecLJGpu compiles with -Wall -Werror against the reference's
    eleclj.hpp/msys.hpp; every ElecLJ member it touches is public (a
    private one would not compile)."""
    need_reference()
    rc, out = run(["g++", "-std=gnu++11", "-O1", "-Wall", "-Werror", "-I", COMPAT,
                   "-I", os.path.join(ROOT, "include"), "-I", REF_INC, "-c",
                   os.path.join(COMPAT, "eleclj_gpu.cpp"), "-o", str(tmp_path / "e.o")])
    assert rc == 0, out
    # its undefined symbols are the library's mm_* calls and reference code
    rc, out = run(["nm", "-u", "-C", str(tmp_path / "e.o")])
    assert rc == 0, out
    mm = {l.split()[-1] for l in out.splitlines() if "mm_" in l}
    assert {"mm_derive_sites", "mm_desc_resolve", "mm_engine_create", "mm_compute",
            "mm_get_forces", "mm_set_perturbation", "mm_set_lambda"} <= mm
    lib_syms = set(dir(E.lib())) | {n for n in mm if hasattr(E.lib(), n)}
    assert mm <= lib_syms


def test_neighborlist_links_and_has_no_host_path(tmp_path):
    need_reference()
    objs = [os.path.join(REF_OBJ, o) for o in BCOND_OBJS]
    if not all(os.path.exists(o) for o in objs):
        pytest.skip("oracle/_ref objects not built (make -C oracle ref)")
    inc = ["-I", COMPAT, "-I", os.path.join(ROOT, "include"), "-I", REF_INC]
    o1, o2, exe = (str(tmp_path / x) for x in ("nbrlist.o", "driver.o", "driver"))
    for src, obj in ((os.path.join(COMPAT, "nbrlist.cpp"), o1),
                     (os.path.join(ROOT, "tests", "link", "nbrlist_driver.cpp"), o2)):
        rc, out = run(["g++", "-std=gnu++11", "-O1", "-Wall", "-Werror"] + inc
                      + ["-c", src, "-o", obj])
        assert rc == 0, out
    rc, out = run(["g++", "-o", exe, o2, o1] + objs + link_flags())
    assert rc == 0, out
    if E.lib().mm_device_count() > 0:
        pytest.skip("a GPU is visible: the no-device path is not reachable")
    rc, out = run([exe])
    assert rc == 1, out
    assert "NeighborList: mm_nlist_create" in out and "no HIP device" in out
