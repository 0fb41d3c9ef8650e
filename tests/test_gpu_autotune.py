"""ctr_autotune and kernel timing.  The tuner's pick is one of its candidates,
no slower than the default schedule on its own timing, reported by
ctr_get_schedule, and OS-EM under it matches the oracle.  Kernel timing
records one projection and one backprojection launch per subset and
iteration, with positive times, and leaves the images unchanged."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, phantom_stack, relerr, say

pytestmark = pytest.mark.gpu


def _geom(ctx, n, pa, ss):
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    ctx.set_geometry(n, n, pa, ss, angles_deg=ang)
    return ang, O.geom(n, n, pa, ss=ss)


def test_autotune_picks_a_schedule_that_runs(ctx):
    n, pa, ss = 128, 60, 6
    ang, g = _geom(ctx, n, pa, ss)
    say("autotune", n)
    try:
        t = ctx.autotune(n)
        print(f"[autotune] side {n}: {t}", flush=True)
        assert t["fp_rays"] in (0, 8, 16, 32, 64) and t["fp_waves"] in (2, 4)
        assert t["fp_rays"] != 0 or t["fp_tile"] in (32, 44, 60)
        assert t["bp_rows"] in (2, 4, 8, 16)
        assert 0.0 < t["fp_ms"] <= t["fp_ms_default"] and 0.0 < t["bp_ms"] <= t["bp_ms_default"]
        assert ctx.get_schedule(n) == {k: t[k] for k in ("fp_rays", "fp_waves", "fp_tile", "bp_rows")}
        _, sino, _ = phantom_stack(2, n, pa, 6, seed=21, i0=1e4)
        x = dev(np.ones((2, n, n)))
        ctx.osem_execution(x, dev(sino), 2)
        e = relerr(host(x), O.osem_execution(g, ang, np.ones((2, n, n), np.float32), sino, 2))
        print(f"[parity] OS-EM under the tuned schedule: relerr {e:.3e}", flush=True)
        assert e < 1e-4
    finally:
        ctx.set_schedule(n, 0, 4, 8)          # the defaults, for the tests after this one


def test_kernel_timing_counts_each_launch(ctx):
    n, pa, ss, it = 96, 60, 6, 2
    _geom(ctx, n, pa, ss)
    _, sino, _ = phantom_stack(1, n, pa, 6, seed=22, i0=1e4)
    y = dev(sino)
    x0 = dev(np.ones((1, n, n)))
    ctx.osem_execution(x0, y.clone(), it)
    say("kernel timing")
    ctx.set_kernel_timing(True)
    try:
        x = dev(np.ones((1, n, n)))
        ctx.osem_execution(x, y.clone(), it)
        ctx.synchronize()
        fp_ms, bp_ms, nfp, nbp = ctx.kernel_timing()
    finally:
        ctx.set_kernel_timing(False)
    print(f"[timing] FP {fp_ms:.4f} ms x {nfp}, BP {bp_ms:.4f} ms x {nbp}", flush=True)
    assert nfp == ss * it and nbp == ss * it
    assert fp_ms > 0.0 and bp_ms > 0.0
    assert np.array_equal(host(x), host(x0))
