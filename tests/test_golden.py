"""Golden fixtures (tests/golden/golden_32x16.npz, written by make_golden.py):
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
