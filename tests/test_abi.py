"""CPU checks of the C-ABI library: it is built, loads, and exports every
function that include/*.h declares.  No compute call is made here."""
import ctypes
import glob
import os
import re

import pytest

import cuhre

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(cuhre_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert "cuhre_integrate" in names and "cuhre_eval_regions" in names
    lib = ctypes.CDLL(cuhre.LIBPATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(cuhre.EXPORTS) <= set(names)


def test_points_per_region_without_device():
    assert [cuhre.points_per_region(d) for d in (3, 5, 6, 7, 8)] == [33, 93, 149, 241, 401]
    assert cuhre.points_per_region(1) == -1
    # key 9: the reference's degree-9 rule, 1 + 8d + 6d(d-1) + 4d(d-1)(d-2)/3 + 2^d points
    assert [cuhre.points_per_region(d, key=9) for d in (2, 3, 5, 7, 8)] == [33, 77, 273, 717, 1105]
    assert cuhre.points_per_region(8, key=5) == -1


def test_context_without_device_is_an_error_code():
    import subprocess
    import sys
    # probe in a child so this process never initialises a GPU runtime
    code = ("import sys; sys.path.insert(0, %r); import cuhre\n"
            "try:\n    cuhre.Context(0)\n    print('HAVE_GPU')\n"
            "except cuhre.CuhreError as e:\n    print('ERR', e.code)\n") % os.path.join(ROOT, "old-gpuintegration_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    if "HAVE_GPU" in out:
        pytest.skip("a HIP device is present")
    assert "ERR -4" in out
