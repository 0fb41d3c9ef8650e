"""Golden fixtures (tests/golden/golden_32x16.npz and golden_around_32x16.npz, written by make_golden.py):
the oracle reproduces them bit for bit; the GPU drivers match the FP32 values."""
import os

import numpy as np
import pytest

import oracle as O
from _util import dev, host, relerr, say

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_32x16.npz")


def _load():
    z = np.load(FIX)
    n, _, pa, ss, zp = (int(v) for v in z["geom"])
    return z, n, pa, ss, zp


def test_oracle_reproduces_fixtures():
    z, n, pa, ss, zp = _load()
    g = O.geom(n, n, pa, ss=ss, zp=zp)
    ang, y = z["angles"], z["sino"]
    x0 = np.ones((1, n, n), np.float32)
    assert np.array_equal(O.osem_execution(g, ang, x0, y, 3), z["osem_it3_f32"])
    assert np.array_equal(O.fbp_execution(g, ang, y, dtype=np.float64), z["fbp_f64"])
    assert np.array_equal(O.power_iteration(g, ang, 30), z["power_L_f32"])
    assert np.array_equal(O.fista_execution(g, ang, z["L"], x0, y, 2, 1.0, 1e-3), z["fista_it2_f32"])


@pytest.mark.gpu
def test_gpu_drivers_match_fixtures(ctx):
    z, n, pa, ss, zp = _load()
    ang, y = z["angles"], z["sino"]
    ctx.set_geometry(n, n, pa, ss, zp=zp, angles_deg=ang, cs_lambda=1e-3, wt1=1.0)
    x0 = np.ones((1, n, n), np.float32)
    runs = {
        "osem_it3": lambda x: ctx.osem_execution(x, dev(y), 3),
        "aart_it2": lambda x: ctx.aart_execution(x, dev(y), 2),
        "fbp": lambda x: ctx.fbp_execution(x, dev(y)),
        "fista_osem_it2": lambda x: ctx.fista_osem_execution(x, dev(y), 2),
        "fista_it2": lambda x: ctx.fista_execution(dev(z["L"]), x, dev(y), 2),
    }
    for name, run in runs.items():
        say("golden", name)
        x = dev(x0)
        run(x)
        e = relerr(host(x), z[f"{name}_f32"])
        print(f"[parity] golden {name}: relerr vs FP32 fixture {e:.3e}, vs FP64 {relerr(host(x), z[name + '_f64']):.3e}",
              flush=True)
        assert e < 1e-4, (name, e)
    say("golden power")
    L = dev(np.zeros(ss))
    ctx.power_iteration(L, 30)
    assert relerr(host(L), z["power_L_f32"]) < 1e-4


FIX2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_around_32x16.npz")


def test_oracle_reproduces_around_fixtures():
    z = np.load(FIX2)
    n, _, pa, ss, zp = (int(v) for v in z["geom"])
    g = O.geom(n, n, pa, ss=ss, zp=zp)
    W, H, stride = 32, 6, int(z["frame_stride"][0])
    assert np.array_equal(O.hybrid_execution(g, z["angles"], z["sino"], 2, baseup_order=4), z["hybrid_it2_f32"])
    dm = O.merge_frames(z["dark"], 3, W, H, stride)
    im = O.merge_frames(z["i0"], 3, W, H, stride)
    assert np.array_equal(O.mt_conversion(dm, im, z["it"], pa, W, H, stride), z["mt_f32"])
    assert np.array_equal(O.reslice(z["mt_f32"], z["xshift"], z["yshift"], baseup=0.25, y0=1, n_layers=4),
                          z["reslice_f32"])
    assert np.array_equal(O.denoise_sinogram(z["spiky"], 0.5), z["denoise_f32"])
    assert np.array_equal(O.zcorrection(z["xproj_in"], 20, 230), z["zcorr_f32"])
    assert np.array_equal(O.xprojection(z["mt_f32"], 2, 29), z["xprojection_f32"])
    rec, f = O.rotation_center_search(g, z["angles"], z["sino"][0], 7, -1.5, 0.5, it=1, init=1.0, correction_mode=1)
    assert np.array_equal(rec, z["rcs_recon_f32"]) and np.array_equal(f, z["rcs_findex_f32"])


@pytest.mark.gpu
def test_gpu_around_fixtures(ctx):
    import torch
    z = np.load(FIX2)
    n, _, pa, ss, zp = (int(v) for v in z["geom"])
    ang = z["angles"]
    W, H, stride = 32, 6, int(z["frame_stride"][0])
    u16 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.uint16).view(np.int16)).cuda()
    say("golden around: hybrid")
    ctx.set_geometry(n, n, pa, ss, zp=zp, angles_deg=ang)
    x = dev(np.zeros((1, n, n)))
    ctx.hybrid_execution(x, dev(z["sino"]), 2)
    assert relerr(host(x), z["hybrid_it2_f32"]) < 1e-4
    say("golden around: frames, mu-t, reslice, spikes, x-projection, z-correction")
    dm = ctx.merge_frames(u16(z["dark"]), 3, W, H, stride)
    im = ctx.merge_frames(u16(z["i0"]), 3, W, H, stride)
    mt = ctx.mt_conversion(dm, im, u16(z["it"]), pa, stride)
    assert np.allclose(host(mt), z["mt_f32"], rtol=2e-6, atol=2e-6)
    r = ctx.reslice(dev(z["mt_f32"]), dev(z["xshift"]), dev(z["yshift"]), baseup=0.25, y0=1, n_layers=4)
    assert np.array_equal(host(r), z["reslice_f32"])
    assert np.array_equal(host(ctx.denoise_sinogram(dev(z["spiky"]), 0.5)), z["denoise_f32"])
    assert np.array_equal(host(ctx.zcorrection(dev(z["xproj_in"]), 20, 230)), z["zcorr_f32"])
    assert relerr(host(ctx.xprojection(dev(z["mt_f32"]), 2, 29)), z["xprojection_f32"]) < 1e-6
    say("golden around: centre search")
    ctx.set_geometry(n, n, pa, ss, correction_mode=1, angles_deg=ang)
    rec, f, _ = ctx.rotation_center_search(dev(z["sino"][0]), 7, -1.5, 0.5, it=1, init=1.0)
    assert relerr(host(rec), z["rcs_recon_f32"]) < 1e-5 and relerr(host(f), z["rcs_findex_f32"]) < 1e-4
