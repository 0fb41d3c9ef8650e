"""Edge cases of the HIP path against the oracle: every projector and
backprojector schedule the autotuner can pick, odd / rectangular / detector !=
grid geometries, one subset and one subset per angle, the largest side the
LDS-tiled projector takes and the gather projector past it, all-zero and NaN
sinograms, and the error statuses of bad calls (the context stays usable)."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, phantom_stack, relerr, say

pytestmark = pytest.mark.gpu

SCHEDULES = [(0, 4, 60), (0, 4, 44), (0, 4, 32), (64, 2, 60), (64, 4, 60), (32, 2, 60), (32, 4, 60), (16, 2, 60),
             (16, 4, 60), (8, 2, 60), (8, 4, 60)]


def _geom(ctx, nx, px, pa, ss, ny=None):
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    ctx.set_geometry(nx, px, pa, ss, ny=ny, angles_deg=ang)
    return ang, O.geom(nx, px, pa, ss=ss, ny=ny, zp=ctx.geom.zp)


@pytest.mark.parametrize("rays,waves,tile", SCHEDULES)
def test_every_projector_schedule(ctx, rays, waves, tile):
    n, pa, ss, nz = 96, 60, 6, 2
    ang, g = _geom(ctx, n, n, pa, ss)
    rng = np.random.default_rng(5)
    img = rng.uniform(0.0, 1.0, (nz, n, n)).astype(np.float32)
    ctx.set_adaptive_rows(False)          # run each row count as set, small as the stack is
    try:
        for bp_rows in (1, 2, 4, 8, 16):
            ctx.set_schedule(n, rays, waves, bp_rows, fp_tile=tile)
            say("schedule", rays, waves, tile, bp_rows)
            out = dev(np.zeros((nz, pa, n)))
            ctx.op("aart1", dev(img), dev(np.zeros((nz, pa, n))), out, nz, 1)
            assert relerr(host(out), O.aart1(g, ang, img, np.zeros((nz, pa, n), np.float32), 1)) < 1e-5
            _, sino, _ = phantom_stack(nz, n, pa, 6, seed=8, i0=1e4)
            x = dev(np.ones((nz, n, n)))
            ctx.osem_execution(x, dev(sino), 2)
            e = relerr(host(x), O.osem_execution(g, ang, np.ones((nz, n, n), np.float32), sino, 2))
            print(f"[parity] schedule rays={rays} waves={waves} tile={tile} rows={bp_rows}: OS-EM relerr {e:.3e}",
                  flush=True)
            assert e < 1e-4
    finally:
        ctx.set_schedule(n, 0, 4, 8)
        ctx.set_adaptive_rows(True)


@pytest.mark.parametrize("n,nz", [(512, 1), (96, 3), (1024, 1)])
def test_adaptive_rows_match_the_set_row_count(ctx, n, nz):
    """A stack too small for 4 backprojection blocks per compute unit at the
    side's 8 rows runs with fewer rows per thread (one 512^2 slice: 1 row).
    The images equal those of the 8-row launch bit for bit, for OS-EM, AART
    and the batched driver."""
    pa, ss = 60, 6
    ang, g = _geom(ctx, n, n, pa, ss)
    _, sino, _ = phantom_stack(nz, n, pa, 6, seed=11, i0=1e4)
    y = dev(sino)
    res = {}
    try:
        for adaptive in (True, False):
            ctx.set_adaptive_rows(adaptive)
            say("adaptive rows", adaptive, n, nz)
            x = dev(np.ones((nz, n, n)))
            ctx.osem_execution(x, y.clone(), 2)
            a = dev(np.zeros((nz, n, n)))
            ctx.aart_execution(a, y.clone(), 1)
            imgs = [dev(np.ones((n, n))) for _ in range(nz)]
            ctx.osem_batched(imgs, [y[z] for z in range(nz)], 2, use_graph=False)
            res[adaptive] = (host(x), host(a), np.stack([host(i) for i in imgs]))
    finally:
        ctx.set_adaptive_rows(True)
    for got, want, what in zip(res[True], res[False], ("OS-EM", "AART", "batched OS-EM")):
        assert np.array_equal(got, want), what
    # and the 8-row launch is the oracle's OS-EM
    e = relerr(res[False][0], O.osem_execution(g, ang, np.ones((nz, n, n), np.float32), sino, 2))
    assert e < 1e-4


@pytest.mark.parametrize("nx,ny,px,pa,ss", [
    (37, 37, 41, 30, 3),      # odd sizes, detector wider than the grid
    (80, 48, 64, 36, 4),      # wide grid
    (48, 80, 40, 36, 6),      # tall grid, narrow detector
    (64, 64, 96, 40, 5),
    (50, 50, 50, 24, 24),     # one angle per subset
    (64, 64, 64, 20, 1),      # one subset
])
def test_odd_and_rectangular_geometries(ctx, nx, ny, px, pa, ss):
    ang, g = _geom(ctx, nx, px, pa, ss, ny=ny)
    rng = np.random.default_rng(nx + ny)
    y = rng.uniform(0.0, 4.0, (2, pa, px)).astype(np.float32)
    say("osem", nx, ny, px, pa, ss)
    x = dev(np.ones((2, ny, nx)))
    ctx.osem_execution(x, dev(y), 2)
    e = relerr(host(x), O.osem_execution(g, ang, np.ones((2, ny, nx), np.float32), y, 2))
    say("aart", nx, ny, px, pa, ss)
    xa = dev(np.ones((2, ny, nx)))
    ctx.aart_execution(xa, dev(y), 2)
    ea = relerr(host(xa), O.aart_execution(g, ang, np.ones((2, ny, nx), np.float32), y, 2, 1.0))
    print(f"[parity] {nx}x{ny} px {px} pa {pa} ss {ss}: OS-EM {e:.3e}, AART {ea:.3e}", flush=True)
    assert e < 1e-4 and ea < 1e-4


def test_lds_projector_limit_and_gather_beyond(ctx):
    """3840 (64 tiles of 60 samples) is the largest side of the LDS-tiled
    projector: it and the gather projector both match the oracle there; 3856
    is refused for the LDS-tiled projector and runs on the gather projector."""
    from ctrecon import CtrError
    rng = np.random.default_rng(1)
    for n in (3840, 3856):
        pa, ss = 40, 10
        ang, g = _geom(ctx, n, n, pa, ss)
        img = rng.uniform(0.0, 1.0, (1, n, n)).astype(np.float32)
        zeros = np.zeros((1, pa, n), np.float32)
        outs = {}
        for rays in (0, 64):
            if n > 3840 and rays == 0:
                with pytest.raises(CtrError):
                    ctx.set_schedule(n, 0, 4, 8)
                continue
            ctx.set_schedule(n, rays, 4 if rays == 0 else 2, 8)
            say("large side", n, "rays", rays)
            out = dev(np.zeros((1, pa, n)))
            ctx.op("aart1", dev(img), dev(zeros), out, 1, 3)
            outs[rays] = host(out)
        want = O.aart1(g, ang, img, zeros, 3)
        for rays, got in outs.items():
            e = relerr(got, want)
            print(f"[parity] side {n} rays {rays}: relerr vs FP32 compat {e:.3e}", flush=True)
            assert e < 1e-5
        if n == 3840:
            assert relerr(outs[0], outs[64]) < 1e-5   # two summation orders over 3840 samples
    with pytest.raises(CtrError):
        ctx.set_schedule(4096, 0, 4, 8)


def test_zero_and_nan_sinograms(ctx):
    n, pa, ss = 64, 36, 6
    ang, g = _geom(ctx, n, n, pa, ss)
    rng = np.random.default_rng(2)
    cases = {"zero": np.zeros((1, pa, n), np.float32)}
    y = rng.uniform(0.0, 3.0, (1, pa, n)).astype(np.float32)
    y[0, rng.integers(0, pa, 40), rng.integers(0, n, 40)] = np.nan
    cases["nan"] = y
    for name, yy in cases.items():
        say("sinogram", name)
        x = dev(np.ones((1, n, n)))
        ctx.osem_execution(x, dev(yy), 2)
        got = host(x)
        want = O.osem_execution(g, ang, np.ones((1, n, n), np.float32), yy, 2)
        assert np.isfinite(got).all() and relerr(got, want) < 1e-4, name
        xa = dev(np.ones((1, n, n)))
        ctx.aart_execution(xa, dev(yy), 1)
        assert relerr(host(xa), O.aart_execution(g, ang, np.ones((1, n, n), np.float32), yy, 1, 1.0)) < 1e-4, name


def test_fbp_with_output_region(ctx):
    """An output region below the grid size changes nothing in the FBP
    image: the whole grid is written and its top-left ox x oy block is the
    reference's output (GPU.cpp:704, 848-870)."""
    n, pa = 64, 90
    _, sinos, ang = phantom_stack(2, n, pa, 5, seed=3)
    ctx.set_geometry(n, n, pa, 1, zp=128, angles_deg=ang)
    full = dev(np.zeros((2, n, n)))
    ctx.fbp_execution(full, dev(sinos))
    ctx.set_geometry(n, n, pa, 1, zp=128, angles_deg=ang, ox=40, oy=24)
    roi = dev(np.zeros((2, n, n)))
    say("FBP with a 40 x 24 output region")
    ctx.fbp_execution(roi, dev(sinos))
    assert np.array_equal(host(roi), host(full))
    want = O.fbp_execution(O.geom(n, n, pa, zp=128), ang, sinos)
    assert relerr(host(roi)[:, :24, :40], want[:, :24, :40]) < 1e-5


def test_error_statuses_leave_the_context_usable(ctx):
    import torch
    from ctrecon import CtrError
    n, pa, ss = 32, 12, 4
    ang, g = _geom(ctx, n, n, pa, ss)
    x = torch.ones((1, n, n), device="cuda")
    y = torch.ones((1, pa, n), device="cuda")
    bad = [
        lambda: ctx.osem_execution(x[:0], y[:0], 1),                       # dN = 0
        lambda: ctx.op("osem1", x, y, torch.zeros_like(y), 1, ss),        # subset out of range
        lambda: ctx.op("osem1", x, y, torch.zeros_like(y), 1, -1),
        lambda: ctx.set_geometry(n, n, 10, 4, angles_deg=ang[:10]),       # pa % ss != 0
        lambda: (ctx.set_geometry(n, n, pa, ss, zp=48, angles_deg=ang),  # FBP with zp not a power of two
                 ctx.fbp_execution(x, y.clone())),
        lambda: ctx.set_geometry(0, n, pa, ss, angles_deg=ang),           # empty grid
        lambda: ctx.set_schedule(n, 12, 4, 8),                             # no such schedule
        lambda: ctx.set_schedule(3000, 0, 4, 8, fp_tile=44),               # 64 tiles of 44 samples < 3000
        lambda: ctx.osem_batched([], [], 1),                               # empty batch
        lambda: ctx.set_geometry(n, n, pa, ss, angles_deg=ang, ox=n + 2),  # output region beyond the grid
        lambda: (ctx.set_geometry(n, n, pa, ss, angles_deg=ang, ox=n - 8),  # output region: FBP only
                 ctx.osem_execution(x, y, 1)),
    ]
    for i, f in enumerate(bad):
        say("bad call", i)
        with pytest.raises((CtrError, ValueError)):
            f()
    _geom(ctx, n, n, pa, ss)
    ctx.osem_execution(x, y, 1)
    ctx.synchronize()
    assert torch.isfinite(x).all()
