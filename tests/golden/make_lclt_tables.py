"""Writes lc_lt_t_tables.json: the data tables of the reference's LC_LT_t
integrand (lc_lt_t.h:91-253, the five 22 x 5 float tables tau, mu, sigma,
fmsk, fprj and the node arrays cInterpR (lt, 22 rows) and cInterpC (zt, 5
columns)), read from the reference's header as text.  Each value is stored
as the double of its float.  The same tables exist in FP64 in the
reference's CPU twin (integrands/lc_lt_t.cc:21-170); the script checks that
every float value equals the rounded double one, so the two transcriptions
in the reference pin each other.

Run once, in the container that holds the reference:
    python3 tests/golden/make_lclt_tables.py /root/reference
"""
import json
import os
import re
import sys

import numpy as np


def arrays(text, kind):
    pat = re.compile(r"%s\s+(\w+)\s*\[\s*\]\s*=\s*\{([^}]*)\}" % kind) if kind == "float" else \
        re.compile(r"std::array<double,\s*\d+>\s+constexpr\s+(\w+)\s*=\s*\{([^}]*)\}")
    out = {}
    for name, body in pat.findall(text):
        vals = [v.strip() for v in body.replace("\n", " ").split(",") if v.strip()]
        out[name] = vals
    return out


def main(ref):
    with open(os.path.join(ref, "lc_lt_t.h")) as fh:
        hdr = fh.read()
    with open(os.path.join(ref, "integrands", "lc_lt_t.cc")) as fh:
        twin = fh.read()
    fa = arrays(hdr, "float")
    da = arrays(twin, "double")
    out = {"source": "lc_lt_t.h:91-253 (float tables), cross-checked against integrands/lc_lt_t.cc:21-170",
           "layout": "field[j*nx + i] = value at (lt_i, zt_j); nx = len(lt), ny = len(zt)"}
    out["lt"] = [float(np.float32(float(v))) for v in fa["cInterpR"]]
    out["zt"] = [float(np.float32(float(v))) for v in fa["cInterpC"]]
    nx, ny = len(out["lt"]), len(out["zt"])
    for name in ("tau", "mu", "sigma", "fmsk", "fprj"):
        f32 = np.array([float(v) for v in fa[name + "_arr"]], dtype=np.float32)
        f64 = np.array([float(v) for v in da[name + "_arr"]], dtype=np.float64)
        assert f32.size == nx * ny == f64.size, (name, f32.size, f64.size)
        assert np.array_equal(f32, f64.astype(np.float32)), name
        out[name] = [float(v) for v in f32]
    # the twin's bin arrays end with a symbolic upper edge (lt_max, zt_max)
    twin_lt = [float(v) for v in da["lt_bins"][:nx]]
    twin_zt = [float(v) for v in da["zt_bins"][:ny]]
    assert np.array_equal(np.float32(twin_lt), np.float32(out["lt"]))
    assert np.array_equal(np.float32(twin_zt), np.float32(out["zt"]))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lc_lt_t_tables.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", path, nx, ny)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
