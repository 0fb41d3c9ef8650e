"""ctr_osem_execution_noalloc (the entry point that allocates nothing: its
workspace is the library's per-device scratch array) against
ctr_osem_execution on two calls in a row, the second with a smaller input,
both through the C ABI; results must be bit-identical.  A stack beyond the
scratch array is refused before anything runs."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, phantom_stack, relerr, say

pytestmark = pytest.mark.gpu


def _case(ctx, n, pa, ss, nz, seed, it, correction_mode=0):
    _, sinos, ang = phantom_stack(nz, n, pa, 6, seed=seed, i0=1e4)
    ctx.set_geometry(n, n, pa, ss, angles_deg=ang, correction_mode=correction_mode)
    x0 = np.random.default_rng(seed).uniform(0.5, 1.5, (nz, n, n)).astype(np.float32)
    say("noalloc", n, pa, ss, nz, it, correction_mode)
    a, ya = dev(x0), dev(sinos)
    ctx.osem_execution_noalloc(a, ya, it)
    b, yb = dev(x0), dev(sinos)
    ctx.osem_execution(b, yb, it)
    ga, gb = host(a), host(b)
    assert np.array_equal(ga, gb)
    assert np.array_equal(host(ya), host(yb))          # the in-place correction too
    return x0, sinos, ang, ga


def test_noalloc_equals_osem_execution_twice_second_smaller(ctx):
    x0, sinos, ang, got = _case(ctx, 256, 360, 10, 3, seed=71, it=2, correction_mode=2)
    _case(ctx, 160, 180, 6, 2, seed=72, it=3)
    # and against the oracle on the first case
    want = O.osem_execution(O.geom(256, 256, 360, ss=10), ang, x0, sinos.copy(), 2, correction_mode=2)
    assert relerr(got, want) < 1e-5


def test_noalloc_refuses_a_stack_beyond_its_scratch(ctx):
    from ctrecon import CtrError
    n, pa = 1024, 120
    ctx.set_geometry(n, n, pa, 10, angles_deg=np.arange(pa, dtype=np.float32) * 1.5)
    x = dev(np.ones((6, n, n)))
    y = dev(np.ones((6, pa, n)))
    with pytest.raises(CtrError, match="scratch"):
        ctx.osem_execution_noalloc(x, y, 1)
    assert float(host(x).min()) == 1.0 and float(host(x).max()) == 1.0     # nothing ran
    ctx.osem_execution_noalloc(x[:5].contiguous(), y[:5].contiguous(), 1)   # 5 layers of 1024^2 fit
