"""Build libwakis_fit.so (HIP, gfx950) in-tree with hipcc.

The library is split into translation units that compile in parallel (the
fused-step template instantiations dominate the compile time), then linked.

Usage: python -m wakis_amd.build   (or wakis_amd.build.build())
"""

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwakis_fit.so"
SOURCES = ["lib_fused_f32_a.hip", "lib_fused_f32_b.hip", "lib_fused_f32_c.hip",
           "lib_fused_f32_d.hip", "lib_fused_f64_a.hip", "lib_fused_f64_b.hip",
           "lib_fused_f32.hip", "lib_fused_f64.hip", "lib_fused_t2.hip", "lib_batch.hip",
           "lib_adjoint.hip", "lib_core.hip"]
HEADERS = ["lib_internal.h", "lib_fused.h", "fit_common.h", "step_general.h", "step_fused.h",
           "step_fused4.h", "step_fused_t2.h", "wake.h", "reduce.h", "adjoint.h",
           "adjoint_fused.h", "monitor.h"]

# -ffp-contract=off keeps every multiply and add separately rounded, as numpy
# and scipy's csr_matvec do on the host; -packed-fp32-ops keeps the compiler
# from emitting packed FP32 arithmetic.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    return res.returncode, res.stdout + res.stderr


def build(verbose=False, jobs=None):
    """Compile the shared library; always recompiles every translation unit."""
    cc = hipcc()
    inc = ["-I", str(PKG.parent / "include")]
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory(prefix="wakis_fit_build_") as tmpd:
        objs = [str(Path(tmpd) / (Path(s).stem + ".o")) for s in SOURCES]
        cmds = [[cc] + HIPCC_FLAGS + inc + ["-c", str(CSRC / s), "-o", o]
                for s, o in zip(SOURCES, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(lambda c: _run(c, verbose), cmds))
        bad = [(s, out) for s, (rc, out) in zip(SOURCES, results) if rc != 0]
        if bad:
            for s, out in bad:
                sys.stderr.write(f"--- {s}\n{out}")
            raise RuntimeError("hipcc failed building libwakis_fit.so: "
                               + ", ".join(s for s, _ in bad))
        tmp = LIB.with_suffix(".so.tmp")
        rc, out = _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp)] + objs,
                       verbose)
        if rc != 0:
            sys.stderr.write(out)
            raise RuntimeError("linking libwakis_fit.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
