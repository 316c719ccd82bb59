"""Build libwakis_fit.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m wakis_amd.build   (or wakis_amd.build.build())
"""

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwakis_fit.so"
SOURCES = ["libwakis_fit.hip"]
HEADERS = ["fit_common.h", "step_general.h", "step_fused.h", "step_fused4.h", "wake.h",
           "reduce.h", "adjoint.h", "adjoint_fused.h", "monitor.h"]

# -ffp-contract=off keeps every multiply and add separately rounded, as numpy
# and scipy's csr_matvec do on the host; -packed-fp32-ops keeps the compiler
# from emitting packed FP32 arithmetic.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def build(verbose=False):
    """Compile the shared library; always recompiles."""
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", str(PKG.parent / "include"), "-o", str(tmp)] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libwakis_fit.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
