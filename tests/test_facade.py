"""The C++ facade (include/gpucuhre.hpp) builds against the C-ABI library and,
on a GPU, reproduces the reference driver's call pattern (cudaCuhre.cu)."""
import os
import subprocess

import pytest

import cuhre
import oracle as O
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_driver(tmp_path):
    exe = str(tmp_path / "facade_driver")
    libdir = os.path.dirname(cuhre.LIBPATH)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "facade_driver.cpp"), "-o", exe,
                           "-L", libdir, "-lcuhre_amd", f"-Wl,-rpath,{libdir}"])
    return exe


def test_facade_compiles_and_links(tmp_path):
    exe = build_driver(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_facade_driver_matches_oracle(tmp_path):
    exe = build_driver(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, check=True).stdout
    lines = [l.split("\t") for l in out.strip().splitlines()]
    f = W.cosine_sum(3)
    import numpy as np
    # the driver's Volume holds float bounds, widened exactly to double
    lows = [1.0, float(np.float32(0.1)), 0.0]
    highs = [2.0, float(np.float32(0.3)), 1.0]
    ex, _ = O.exact(f, lows, highs)
    for row, eps in zip(lines[:2], (1e-7, 1e-10)):
        integral, error, nreg, neval, flag = float(row[2]), float(row[3]), int(row[4]), int(row[5]), int(row[6])
        o = O.integrate_bfs(W.Problem("drv", f, lows, highs, eps, 1e-12, 10**9), nthreads=8)
        # at 1e-10 every region passes its local test before the global one
        # does (mixed-sign estimates), so both stop with flag 1
        assert flag == o["rc"] == (0 if eps == 1e-7 else 1)
        assert nreg == o["nregions"] and neval == o["neval"]
        assert abs(integral - o["integral"]) <= 1e-12 * abs(o["integral"])
        assert abs(integral - ex) <= max(eps * abs(ex), error)
    acc, acc_r = float(lines[2][1]), int(lines[2][2])
    assert abs(acc - 2 * float(lines[0][2])) <= 1e-12 * abs(acc)
    assert acc_r == 2 * int(lines[0][4])


@pytest.mark.gpu
def test_facade_driver_two_phase_matches_oracle(tmp_path):
    """The driver with --two-phase runs the reference's own method through
    the facade (key 9, Phase 1 then Phase 2) and agrees with the oracle's
    two-phase restatement."""
    exe = build_driver(tmp_path)
    out = subprocess.run([exe, "--two-phase"], capture_output=True, text=True, timeout=120, check=True).stdout
    lines = [l.split("\t") for l in out.strip().splitlines()]
    f = W.cosine_sum(3)
    import numpy as np
    lows = [1.0, float(np.float32(0.1)), 0.0]
    highs = [2.0, float(np.float32(0.3)), 1.0]
    ex, _ = O.exact(f, lows, highs)
    for row, eps in zip(lines[:2], (1e-7, 1e-10)):
        integral, error, nreg, neval, flag = float(row[2]), float(row[3]), int(row[4]), int(row[5]), int(row[6])
        o = O.integrate_two_phase(W.Problem("drv", f, lows, highs, eps, 1e-12, 10**9), key=9)
        assert (flag == 0) == (o["fail"] == 0), (flag, o)
        assert abs(nreg - o["nregions"]) <= 5e-3 * o["nregions"], (nreg, o["nregions"])
        assert abs(integral - o["integral"]) <= max(error, o["error"]) + 1e-12 * abs(o["integral"])
        assert abs(integral - ex) <= max(eps * abs(ex), 10 * error)


@pytest.mark.gpu
@pytest.mark.parametrize("two_phase", [False, True])
def test_facade_driver_lclt_matches_oracle(tmp_path, two_phase):
    """The reference driver's actual run: LC_LT_t over its Volume at epsrel
    1e-12 and 1e-7 (cudaCuhre.cu:43-79), through the facade, against the
    oracle (breadth-first GM7/5, or the reference's two-phase degree-9
    method with --two-phase)."""
    import json
    import numpy as np
    with open(os.path.join(ROOT, "tests", "golden", "lc_lt_t_tables.json")) as fh:
        t = json.load(fh)
    path = str(tmp_path / "lclt.bin")
    with open(path, "wb") as fh:
        np.array([len(t["lt"]), len(t["zt"])], dtype=np.int32).tofile(fh)
        np.concatenate([np.asarray(t[k], dtype=np.float32) for k in ("lt", "zt", "tau", "mu", "sigma", "fmsk",
                                                                      "fprj")]).tofile(fh)
    exe = build_driver(tmp_path)
    args = [exe, "--lclt", path] + (["--two-phase"] if two_phase else [])
    out = subprocess.run(args, capture_output=True, text=True, timeout=120, check=True).stdout
    lines = [l.split("\t") for l in out.strip().splitlines()]
    assert len(lines) == 2
    for row, eps in zip(lines, (1e-12, 1e-7)):
        integral, error, nreg, flag = float(row[2]), float(row[3]), int(row[4]), int(row[6])
        p = W.lclt_driver(t, eps)
        o = O.integrate_two_phase(p, key=9) if two_phase else O.integrate_bfs(p, nthreads=8)
        assert (flag == 0) == (o["rc"] == 0), (flag, o)
        if two_phase:
            assert abs(nreg - o["nregions"]) <= 5e-3 * o["nregions"]
            assert abs(integral - o["integral"]) <= max(error, o["error"])
        else:
            assert nreg == o["nregions"]
            assert abs(integral - o["integral"]) <= 1e-12 * abs(o["integral"])
