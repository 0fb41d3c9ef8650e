"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

Each fixture holds seeded inputs and the oracle's FP32 compat and FP64 outputs
for one reference driver at a tiny size, so a regression in the oracle (or a
GPU run compared against the fixture) shows up without re-deriving anything.
Run from the repository root:  python tests/golden/make_golden.py [drivers|around|all]
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


def main_around():
    """The steps around the reconstruction: mode 8 (hybrid), sinogram
    formation (mu-t conversion, reslice with shifts, spike removal,
    x-projection, z-correction) and the rotation-centre search."""
    rng = np.random.default_rng(91)
    n, pa, ss = 32, 16, 4
    img, sino, ang = ph.make_slice(n, pa, 6, seed=78, i0=1e4)
    y = sino[None]
    g = O.geom(n, n, pa, ss=ss, zp=64)
    out = {"angles": ang, "sino": y, "geom": np.array([n, n, pa, ss, 64], np.int32)}
    out["hybrid_it2_f32"] = O.hybrid_execution(g, ang, y, 2, baseup_order=4)
    # frames: dark / I0 (3 scans) and It (pa angles) of a 6-row, 32-column camera
    W, H = 32, 6
    stride = W * H + 32
    dark = rng.integers(90, 110, 3 * stride).astype(np.uint16)
    i0 = rng.integers(20000, 30000, 3 * stride).astype(np.uint16)
    it = rng.integers(3000, 20000, pa * stride).astype(np.uint16)
    dm = O.merge_frames(dark, 3, W, H, stride)
    im = O.merge_frames(i0, 3, W, H, stride)
    mt = O.mt_conversion(dm, im, it, pa, W, H, stride)
    out.update(dark=dark, i0=i0, it=it, frame_stride=np.array([stride], np.int64), mt_f32=mt)
    xs = rng.uniform(-2.0, 2.0, pa).astype(np.float32)
    ys = rng.uniform(-1.5, 1.5, pa).astype(np.float32)
    out.update(xshift=xs, yshift=ys)
    out["reslice_f32"] = O.reslice(mt, xs, ys, baseup=0.25, y0=1, n_layers=4)
    spiky = out["reslice_f32"].copy()
    spiky[1, 5, 7] += 3.0
    out["spiky"] = spiky
    out["denoise_f32"] = O.denoise_sinogram(spiky, 0.5)
    Yv = np.arange(256, dtype=np.float64)
    drift = rng.uniform(-2.0, 2.0, 10)
    xp = np.stack([0.5 * (Yv + d) + 3.0 + 0.2 * np.sin((Yv + d) / 9.0) for d in drift]).astype(np.float32)
    out["xproj_in"] = xp
    out["zcorr_f32"] = O.zcorrection(xp, 20, 230)
    out["xprojection_f32"] = O.xprojection(mt, 2, 29)
    rec, f = O.rotation_center_search(g, ang, sino, 7, -1.5, 0.5, it=1, init=1.0, correction_mode=1)
    out.update(rcs_recon_f32=rec, rcs_findex_f32=f)
    np.savez_compressed(os.path.join(OUT, "golden_around_32x16.npz"), **out)
    print("wrote", os.path.join(OUT, "golden_around_32x16.npz"))


if __name__ == "__main__":
    if len(sys.argv) < 2 or sys.argv[1] in ("all", "drivers"):
        main()
    if len(sys.argv) < 2 or sys.argv[1] in ("all", "around"):
        main_around()
