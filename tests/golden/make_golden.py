"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

Each fixture holds seeded inputs and the oracle's FP32 compat and FP64 outputs
for one reference driver at a tiny size, so a regression in the oracle (or a
GPU run compared against the fixture) shows up without re-deriving anything.
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "ct-xafs_analysis_amd")]

import oracle as O  # noqa: E402
from ctrecon import phantom as ph  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def main():
    n, pa, ss = 32, 16, 4
    img, sino, ang = ph.make_slice(n, pa, 6, seed=77, i0=1e4)
    y = sino[None]
    g = O.geom(n, n, pa, ss=ss, zp=64)
    x0 = np.ones((1, n, n), np.float32)
    L = O.power_iteration(g, ang, 30)
    cases = {
        "osem_it3": lambda d: O.osem_execution(g, ang, x0, y, 3, dtype=d),
        "aart_it2": lambda d: O.aart_execution(g, ang, x0, y, 2, 1.0, dtype=d),
        "fbp": lambda d: O.fbp_execution(g, ang, y, dtype=d),
        "fista_osem_it2": lambda d: O.fista_osem_execution(g, ang, x0, y, 2, 1e-3, dtype=d),
        "fista_it2": lambda d: O.fista_execution(g, ang, L, x0, y, 2, 1.0, 1e-3, dtype=d),
        "power_L": lambda d: O.power_iteration(g, ang, 30, dtype=d),
    }
    res = {k: (f(np.float32), f(np.float64)) for k, f in cases.items()}
    np.savez_compressed(os.path.join(OUT, "golden_32x16.npz"), angles=ang, sino=y, L=L,
                        geom=np.array([n, n, pa, ss, 64], np.int32),
                        **{f"{k}_f32": v[0] for k, v in res.items()},
                        **{f"{k}_f64": v[1] for k, v in res.items()})
    print("wrote", os.path.join(OUT, "golden_32x16.npz"))


if __name__ == "__main__":
    main()
