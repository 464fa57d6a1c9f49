"""CPU checks of the C ABI: the library loads, exports every function that
include/bh_potential.h declares, and rejects bad descriptors before it
touches a GPU."""
import ctypes
import os
import re

import pytest

from bipoly_hydro_amd import potential as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bh_potential.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("bh_plan_create", "bh_plan_destroy", "bh_solve", "bh_solve_device", "bh_set_potential",
              "bh_last_error"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = P.load_library()
    for n in declared_functions():
        assert hasattr(lib, n), n
    assert set(declared_functions()) <= set(P.EXPORTS)


@pytest.mark.parametrize("desc,msg", [
    (dict(numr=66, numz=66, numphi=48), "power of two"),
    (dict(numr=66, numz=65, numphi=64), "numz must be even"),
    (dict(numr=3, numz=66, numphi=64), ">= 5"),
    (dict(numr=66, numz=66, numphi=64, isym=4), "isym must be 1, 2 or 3"),
    (dict(numr=66, numz=66, numphi=64, isym=2, precision=1), "need precision BH_PREC_FP64"),
    (dict(numr=66, numz=66, numphi=64, isym=3, precision=2), "need precision BH_PREC_FP64"),
    (dict(numr=66, numz=66, numphi=64, mmax=-1), "mmax"),
    (dict(numr=600, numz=66, numphi=64), "<= 512"),
    (dict(numr=66, numz=66, numphi=64, passes=5), "refine_passes"),
    (dict(numr=66, numz=66, numphi=64, passes=-1), "refine_passes"),
    (dict(numr=66, numz=66, numphi=64, isym=2, precision=3), "need precision BH_PREC_FP64"),
    (dict(numr=66, numz=66, numphi=64, precision=4), "precision"),
    (dict(numr=66, numz=66, numphi=64, precision=-1), "precision"),
])
def test_invalid_descriptors_are_rejected_without_a_gpu(desc, msg):
    lib = P.load_library()
    d = P.GridDesc(desc["numr"], desc["numz"], desc["numphi"], desc.get("isym", 1), 1.0 / 127.0,
                   desc.get("mmax", 30), 20, 5, desc.get("passes", 0))
    h = ctypes.c_void_p()
    rc = lib.bh_plan_create(ctypes.byref(d), desc.get("precision", 0), 0, ctypes.byref(h))
    assert rc == -1
    assert msg in lib.bh_last_error().decode()
    assert not h.value


def test_null_arguments_are_errors():
    lib = P.load_library()
    assert lib.bh_solve(None, None, None, 0) == -1
    assert lib.bh_plan_destroy(None) == 0


def test_grid_desc_layout_matches_the_header():
    """The ctypes mirror has the header's fields in the header's order."""
    text = open(os.path.join(ROOT, "include", "bh_potential.h")).read()
    body = re.search(r"typedef struct bh_grid_desc \{(.*?)\} bh_grid_desc;", text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\b(?:int|double)\s+([a-z_]+)\s*;", body)
    assert fields == [f for f, _ in P.GridDesc._fields_]
