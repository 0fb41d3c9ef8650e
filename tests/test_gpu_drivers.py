"""Driver-level parity: the GPU.cpp execution drivers restated on the GPU,
against the oracle's FP32 compat drivers on seeded phantoms.  The bar is
relative L2 <= 1e-4 against the compat oracle; the error against the FP64
oracle is printed for every case but not bounded: exact ties (a pixel on a bin
boundary at 30/60/120/150 degrees, the correction-mode-4 mask edge at 90
degrees) resolve differently in float and double arithmetic, in the
reference's own FP32 kernels as much as here."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, phantom_stack, relerr, say

pytestmark = pytest.mark.gpu


def _geom(ctx, n, pa, ss, **kw):
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    ctx.set_geometry(n, n, pa, ss, angles_deg=ang, **kw)
    return ang, O.geom(n, n, pa, ss=ss, zp=ctx.geom.zp)


def _report(name, got, c32, c64):
    e32, e64 = relerr(got, c32), relerr(got, c64)
    print(f"[parity] {name}: relerr vs FP32 compat {e32:.3e}, vs FP64 {e64:.3e}", flush=True)
    return e32, e64


@pytest.mark.parametrize("n,pa,ss,it,nz,corr", [(64, 90, 10, 3, 2, 0), (96, 120, 8, 2, 1, 3), (64, 60, 6, 2, 3, 4)])
def test_osem_execution(ctx, n, pa, ss, it, nz, corr):
    ang, g = _geom(ctx, n, pa, ss, correction_mode=corr)
    img0, sino, _ = phantom_stack(nz, n, pa, 10, seed=11, i0=1e4)
    x0 = np.ones_like(img0)
    say("osem_execution", n, pa, ss, it, nz, corr)
    x = dev(x0)
    ctx.osem_execution(x, dev(sino), it)
    got = host(x)
    c32 = O.osem_execution(g, ang, x0, sino, it, correction_mode=corr)
    c64 = O.osem_execution(g, ang, x0, sino, it, correction_mode=corr, dtype=np.float64)
    e32, e64 = _report("osem", got, c32, c64)
    assert e32 < 1e-4


def test_osem_512sq_900proj_20it(ctx):
    """512^2, 900 angles, 10 subsets, 20 iterations, 30 ellipses, Poisson I0=1e4."""
    n, pa, ss, it = 512, 900, 10, 20
    ang, g = _geom(ctx, n, pa, ss)
    img0, sino, _ = phantom_stack(1, n, pa, 30, seed=2, i0=1e4)
    x0 = np.ones_like(img0)
    say("osem_execution 512^2 x 900", n, pa, ss, it)
    x = dev(x0)
    ctx.osem_execution(x, dev(sino), it)
    got = host(x)
    c32 = O.osem_execution(g, ang, x0, sino, it)
    c64 = O.osem_execution(g, ang, x0, sino, it, dtype=np.float64)
    e32, e64 = relerr(got, c32), relerr(got, c64)
    print(f"[parity] OS-EM 512^2/900/S10/it20: relerr vs FP32 compat {e32:.3e}, vs FP64 {e64:.3e} "
          f"(FP32 compat vs FP64 {relerr(c32, c64):.3e})", flush=True)
    assert e32 < 1e-4


@pytest.mark.parametrize("mode", [1, 3])
def test_aart_execution(ctx, mode):
    n, pa, ss, it = 64, 90, 10, 2
    ang, g = _geom(ctx, n, pa, ss, mode=mode, wt1=1.0, wt2=0.01)
    img0, sino, _ = phantom_stack(2, n, pa, 10, seed=5)
    x0 = np.full_like(img0, 7.0)   # replaced by the backprojection init
    say("aart_execution", mode)
    x = dev(x0)
    ctx.aart_execution(x, dev(sino), it)
    got = host(x)
    alpha = 0.01 if mode == 3 else 1.0
    c32 = O.aart_execution(g, ang, x0, sino, it, alpha)
    c64 = O.aart_execution(g, ang, x0, sino, it, alpha, dtype=np.float64)
    e32, _ = _report(f"aart mode {mode}", got, c32, c64)
    assert e32 < 1e-4


def test_fbp_256sq_360proj(ctx):
    """256^2, 360 angles, 256 bins, Ram-Lak, ZP=512, 10 ellipses, analytic sinogram."""
    n, pa = 256, 360
    ang, g = _geom(ctx, n, pa, 1, zp=512)
    img0, sino, _ = phantom_stack(1, n, pa, 10, seed=1)
    say("fbp_execution 256^2 x 360")
    x = dev(np.zeros_like(img0))
    ctx.fbp_execution(x, dev(sino))
    got = host(x)
    c32 = O.fbp_execution(g, ang, sino)
    c64 = O.fbp_execution(g, ang, sino, dtype=np.float64)
    e32, _ = _report("fbp 256^2 x 360", got, c32, c64)
    assert e32 < 1e-4


def test_power_iteration(ctx):
    n, pa, ss = 64, 90, 10
    ang, g = _geom(ctx, n, pa, ss)
    say("power_iteration")
    L = dev(np.zeros(ss))
    ctx.power_iteration(L, 30)
    got = host(L)
    c32 = O.power_iteration(g, ang, 30)
    c64 = O.power_iteration(g, ang, 30, dtype=np.float64)
    e32, _ = _report("power iteration", got, c32, c64)
    assert e32 < 1e-4


@pytest.mark.parametrize("it", [1, 3])
def test_fista_osem_execution(ctx, it):
    n, pa, ss = 64, 90, 10
    ang, g = _geom(ctx, n, pa, ss, cs_lambda=1e-3)
    img0, sino, _ = phantom_stack(3, n, pa, 10, seed=8, i0=1e4)
    x0 = np.ones_like(img0)
    say("fista_osem_execution", it)
    x = dev(x0)
    ctx.fista_osem_execution(x, dev(sino), it)
    got = host(x)
    c32 = O.fista_osem_execution(g, ang, x0, sino, it, 1e-3)
    c64 = O.fista_osem_execution(g, ang, x0, sino, it, 1e-3, dtype=np.float64)
    e32, _ = _report(f"fista-osem it={it}", got, c32, c64)
    assert e32 < 1e-4


@pytest.mark.parametrize("it", [0, 1, 2])
def test_fista_execution(ctx, it):
    n, pa, ss = 64, 90, 10
    ang, g = _geom(ctx, n, pa, ss, cs_lambda=1e-3, wt1=1.0)
    img0, sino, _ = phantom_stack(2, n, pa, 10, seed=9)
    L = O.power_iteration(g, ang, 30)
    x0 = np.ones_like(img0)
    say("fista_execution", it)
    x = dev(x0)
    ctx.fista_execution(dev(L), x, dev(sino), it)
    got = host(x)
    c32 = O.fista_execution(g, ang, L, x0, sino, it, 1.0, 1e-3)
    e32 = relerr(got, c32)
    print(f"[parity] fista it={it}: relerr vs FP32 compat {e32:.3e}", flush=True)
    assert e32 < 1e-4


def test_batched_matches_per_slice(ctx):
    """Uneven batched OS-EM (graph and eager) equals per-slice OSEM_execution bit for bit."""
    import torch
    pa, ss, it = 60, 6, 2
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    ctx.set_geometry(32, 32, pa, ss, angles_deg=ang)
    sides = [48, 32, 64, 48]
    sl = []
    for i, n in enumerate(sides):
        _, sino, _ = phantom_stack(1, n, pa, 6, seed=20 + i, i0=1e4)
        sl.append((n, sino[0]))
    ref = []
    for n, sino in sl:
        ctx.set_geometry(n, n, pa, ss, angles_deg=ang)
        x = dev(np.ones((1, n, n)))
        say("per-slice", n)
        ctx.osem_execution(x, dev(sino[None]), it)
        ref.append(host(x)[0])
    ctx.set_geometry(sides[0], sides[0], pa, ss, angles_deg=ang)
    for use_graph in (False, True):
        imgs = [dev(np.ones((n, n))) for n, _ in sl]
        prjs = [dev(s) for _, s in sl]
        say("batched graph" if use_graph else "batched eager")
        ctx.osem_batched(imgs, prjs, it, use_graph=use_graph)
        ctx.synchronize()
        for r, t in zip(ref, imgs):
            assert np.array_equal(host(t), r)
    torch.cuda.synchronize()


def test_batched_graph_replays(ctx):
    """Replays of the captured iteration, with eager work and synchronisation in
    between, each equal the eager result.  The graph context keeps one plan
    (same slice list every call), so calls 2 and 3 replay the graph captured in
    call 1; a second context runs the same schedule eagerly."""
    import torch
    from ctrecon import Context
    pa, ss = 60, 6
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    ctx.set_geometry(40, 40, pa, ss, angles_deg=ang)
    ectx = Context(0)
    ectx.set_geometry(40, 40, pa, ss, angles_deg=ang)
    sides = [40, 56]
    sinos = [phantom_stack(1, n, pa, 6, seed=40 + i, i0=1e4)[1][0] for i, n in enumerate(sides)]
    eager = [dev(np.ones((n, n))) for n in sides]
    graph = [dev(np.ones((n, n))) for n in sides]
    pe = [dev(s) for s in sinos]
    pg = [dev(s) for s in sinos]
    for rep in range(3):
        say("graph replay", rep)
        ectx.osem_batched(eager, pe, 2, use_graph=False)
        ectx.synchronize()
        scratch = torch.randn(1000, device="cuda").sum()   # unrelated eager work
        torch.cuda.synchronize()
        ctx.osem_batched(graph, pg, 2, use_graph=True)
        ctx.synchronize()
        for a, b in zip(eager, graph):
            assert np.array_equal(host(a), host(b)), rep
        del scratch
    ectx.close()


@pytest.mark.parametrize("cs,corr", [(False, 0), (False, 2), (True, 0)])
def test_hybrid_execution(ctx, cs, corr):
    """Mode 8 (hybrid_thread, GPU.cpp:1459-1580): FBP, base-up, OS-EM."""
    n, pa, ss, it = 64, 90, 6, 2
    ang, g = _geom(ctx, n, pa, ss, correction_mode=corr, cs_lambda=2e-3)
    _, sino, _ = phantom_stack(2, n, pa, 6, seed=31, i0=1e4)
    x = dev(np.zeros((2, n, n)))
    say("hybrid", cs, corr)
    ctx.hybrid_execution(x, dev(sino), it, cs=cs)
    got = host(x)
    c32 = O.hybrid_execution(g, ang, sino, it, cs=cs, cs_lambda=2e-3, correction_mode=corr)
    c64 = O.hybrid_execution(g, ang, sino, it, cs=cs, cs_lambda=2e-3, correction_mode=corr, dtype=np.float64)
    e32, _ = _report(f"hybrid cs={cs} corr={corr}", got, c32, c64)
    assert e32 < 1e-4
