"""Extract the golden vectors of the reference's test_001 into a JSON fixture.

Reads, as data, ``tests/test_001_pec_cubic_cavity.py`` (the WP[::50] and
Z[::20] regression arrays, lines 18-67, and the two cumsum pins, lines 155-173)
and ``tests/stl/001_cubic_cavity.stl`` (vertex coordinates, to recover the
float32 domain bounds the test uses and the box decomposition of the solid)
from the reference tree. Nothing from the reference is imported or executed.

Usage: python tests/golden/make_test001_golden.py /root/reference
"""

import ast
import json
import re
import sys
from pathlib import Path

import numpy as np


def _array_literal(src, name):
    m = re.search(name + r"\s*=\s*np\.array\(\s*(\[.*?\])\s*,?\s*\)", src, re.S)
    body = m.group(1).replace("\n", " ")
    body = re.sub(r"([0-9.eE+-]+)\s*([+-])\s*([0-9.eE+-]+)j", r"complex(\1,\2\3)", body)
    body = body.replace("+0.j", "")
    vals = []
    for tok in re.findall(r"complex\([^)]*\)|[-+]?[0-9.]+e[-+][0-9]+", body):
        if tok.startswith("complex"):
            a, b = tok[len("complex("):-1].split(",")
            vals.append([float(a), float(b)])
        else:
            vals.append([float(tok), 0.0])
    return vals


def main(ref_root):
    ref = Path(ref_root)
    src = (ref / "tests" / "test_001_pec_cubic_cavity.py").read_text()
    wp = [v[0] for v in _array_literal(src, "WP")]
    z = _array_literal(src, "Z")
    cs_wp = float(re.search(r"np\.cumsum\(np\.abs\(wake\.WP\)\)\[-1\] == pytest\.approx\(\s*"
                            r"([0-9.]+)", src).group(1))
    cs_z = float(re.search(r"np\.cumsum\(np\.abs\(wake\.Z\)\)\[-1\] == pytest\.approx\(\s*"
                           r"([0-9.]+)", src).group(1))
    stl = (ref / "tests" / "stl" / "001_cubic_cavity.stl").read_text().split()
    verts = np.array([[float(stl[i + 1]), float(stl[i + 2]), float(stl[i + 3])]
                      for i, t in enumerate(stl) if t == "vertex"], dtype=np.float32)
    bounds = [float(verts[:, 0].min()), float(verts[:, 0].max()),
              float(verts[:, 1].min()), float(verts[:, 1].max()),
              float(verts[:, 2].min()), float(verts[:, 2].max())]
    levels = [sorted({float(v) for v in verts[:, a]}) for a in range(3)]
    out = {
        "source": "reference tests/test_001_pec_cubic_cavity.py:18-67,151-173 and "
                  "tests/stl/001_cubic_cavity.stl",
        "grid": {"Nx": 50, "Ny": 50, "Nz": 150, "bounds_float32": bounds},
        "solid_levels_float32": levels,
        "solid": "union of box |x|,|y|<=0.025, |z|<=0.015 and box |x|,|y|<=0.0075, "
                 "|z|<=0.05 (vacuum) in a PEC background",
        "beam": {"q": 1e-9, "sigmaz": 18.5e-3, "beta": 1.0, "xs": 0.0, "ys": 0.0,
                 "xt": 0.0, "yt": 0.0, "skip_cells": 12, "wakelength": 1.0},
        "bc": {"low": ["pec", "pec", "pec"], "high": ["pec", "pec", "pec"]},
        "dtype": "float32",
        "WP_every50": wp,
        "Z_every20_re_im": z,
        "tol_WP": {"rtol": 5e-5, "atol": 5e-5},
        "tol_Z": {"rtol": 5e-4, "atol": 5e-4},
        "cumsum_abs_WP": cs_wp,
        "cumsum_abs_Z": cs_z,
        "cumsum_rel": 0.1,
    }
    dst = Path(__file__).with_name("test001_golden.json")
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst, len(wp), "WP samples,", len(z), "Z samples")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
