"""FP64 oracle outputs for the 48x48x128 cavity/pipe wakefield run.

The run: PEC cube 50 mm with a beam pipe r = 5 mm in a 60x60x160 mm domain,
48x48x128 cells, vacuum in the cavity and pipe, PEC background, PEC on x/y
and ABC on z; Gaussian beam q = 1e-9 C, sigma_z = 10 mm, beta = 1 on axis,
wakelength 0.5 m (2062 steps), skip_cells = 10, impedance up to 5 GHz.
Fields and all wake arithmetic in FP64 through the oracle (oracle/), the
CPU restatement of the reference's scipy path. Writes
tests/golden/cavity_48x48x128_fp64_oracle.npz: s, WP_3d, WPx, WPy, f, Z, Zx, Zy,
lambdas, and a stride-97 sample plus exactly rounded seeded weighted sums
of the final E, H, J.

Usage: python tests/golden/make_cavity_48x48x128_golden.py   (about two minutes)
"""

import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import cases  # noqa: E402
from oracle import fit_oracle as fo  # noqa: E402
from oracle import wake_oracle as wo  # noqa: E402

OUT = Path(__file__).resolve().with_name("cavity_48x48x128_fp64_oracle.npz")
RUN = dict(dims=(48, 48, 128), q=1e-9, sigmaz=10e-3, beta=1.0, wakelength=0.5,
           skip_cells=10, fmax=5e9)
STRIDE = 97


def field_digest(F, seed):
    """Stride-97 sample and the exactly rounded sum of F * w (math.fsum), a
    weighted checksum independent of BLAS threading."""
    w = np.random.default_rng(seed).standard_normal(F.size)
    return F[::STRIDE].copy(), math.fsum((F * w).tolist())


def main():
    t0 = time.time()
    so = cases.cavity_pipe_solver(*RUN["dims"], dtype=np.float64)
    ti = wo.wake_ti(RUN["sigmaz"], RUN["beta"])
    cap = fo.wakesolve_fields(so, RUN["wakelength"], RUN["q"], RUN["sigmaz"], RUN["beta"],
                              0.0, 0.0, 0.0, 0.0, ti)
    out = wo.wake_solve(cap, RUN["q"], RUN["sigmaz"], RUN["beta"], RUN["wakelength"],
                        skip_cells=RUN["skip_cells"], fmax=RUN["fmax"])
    arrays = {k: out[k] for k in ("s", "WP_3d", "WPx", "WPy", "f", "Z", "Zx", "Zy", "lambdas")}
    for name, F, seed in (("E", so.E, 1), ("H", so.H, 2), ("J", so.J, 3)):
        sample, dot = field_digest(F, seed)
        arrays[name + "_sample"] = sample
        arrays[name + "_dot"] = np.array(dot)
    arrays["Nt"] = np.array(cap["Nt"])
    arrays["dt"] = np.array(so.dt)
    with open(OUT, "wb") as fh:
        np.savez(fh, **arrays)
    print(f"wrote {OUT} (Nt={cap['Nt']}, {time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
