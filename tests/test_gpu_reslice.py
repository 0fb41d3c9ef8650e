"""Sinogram formation on the GPU (Reslice/, the mt conversion of ImageReg
share source/): frame merge, mu-t conversion, reslice with shifts, spike
removal, x-projection and the Levenberg-Marquardt z-correction against the
oracle; the whole reslice_thread order; resliced sinograms feeding OS-EM
directly; and a 2048 x 2048 camera with 1600 angles resliced exactly."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, relerr, say

pytestmark = pytest.mark.gpu


def _u16(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint16).view(np.int16)).cuda()


def test_merge_frames_and_mt_conversion(ctx):
    rng = np.random.default_rng(1)
    W, H, n = 96, 40, 5
    stride = W * H + 32
    raw = rng.integers(0, 65535, stride * n).astype(np.uint16)
    say("merge_rawhisdata")
    got = host(ctx.merge_frames(_u16(raw), n, W, H, stride))
    assert np.array_equal(got, O.merge_frames(raw, n, W, H, stride))
    dark = rng.uniform(90.0, 110.0, (H, W)).astype(np.float32)
    I0 = rng.uniform(3000.0, 4000.0, (H, W)).astype(np.float32)
    It = rng.integers(0, 5000, stride * n).astype(np.uint16)
    It[:50] = 100                                   # It = dark region: clamps and NaN rules
    say("mt_conversion")
    mt = host(ctx.mt_conversion(dev(dark), dev(I0), _u16(It), n, stride))
    want = O.mt_conversion(dark, I0, It, n, W, H, stride)
    d = np.abs(mt - want)
    print(f"[parity] mt_conversion: max |diff| {d.max():.3e}", flush=True)
    assert np.isfinite(mt).all() and np.allclose(mt, want, rtol=2e-6, atol=2e-6)


def test_reslice_shifts_baseup_correction(ctx):
    rng = np.random.default_rng(2)
    A, H, W = 24, 70, 90
    mt = rng.uniform(-1.0, 3.0, (A, H, W)).astype(np.float32)
    say("reslice transpose")
    assert np.array_equal(host(ctx.reslice(dev(mt))), mt.transpose(1, 0, 2))
    xs = rng.uniform(-4.0, 4.0, A).astype(np.float32)
    ys = rng.uniform(-4.0, 4.0, A).astype(np.float32)
    say("reslice shifts")
    got = host(ctx.reslice(dev(mt), dev(xs), dev(ys), baseup=0.3, y0=10, n_layers=50))
    want = O.reslice(mt, xs, ys, baseup=0.3, y0=10, n_layers=50)
    assert np.array_equal(got, want), np.abs(got - want).max()
    say("reslice correction")
    got = host(ctx.reslice(dev(mt), None, dev(ys), correction=1))
    want = O.reslice(mt, None, ys, correction=1)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6)
    # shifts that move whole rows and columns past the image and its border
    # band (the row-by-row path and the all-border path of the kernel)
    far_y = np.where(np.arange(A) % 2 == 0, -(H // 2 + 7.25), H // 2 + 6.5).astype(np.float32)
    far_x = np.where(np.arange(A) % 3 == 0, W + 9.5, 0.75).astype(np.float32)
    say("reslice far shifts")
    got = host(ctx.reslice(dev(mt), dev(far_x), dev(far_y), baseup=-0.5, y0=3, n_layers=H - 3))
    want = O.reslice(mt, far_x, far_y, baseup=-0.5, y0=3, n_layers=H - 3)
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_denoise_and_xprojection(ctx):
    rng = np.random.default_rng(3)
    nz, A, W = 6, 50, 64
    src = rng.uniform(0.0, 1.0, (nz, A, W)).astype(np.float32)
    src[rng.integers(0, nz, 40), rng.integers(0, A, 40), rng.integers(0, W, 40)] += 4.0
    say("denoiseSinogram")
    got = host(ctx.denoise_sinogram(dev(src), 0.8))
    assert np.array_equal(got, O.denoise_sinogram(src, 0.8))
    for W, sx, ex in ((120, 3, 110), (121, 0, 120), (64, 5, 5), (96, -3, 99)):
        mt = rng.uniform(0.0, 1.0, (7, 300, W)).astype(np.float32)
        say("xprojection", W, sx, ex)
        xp = host(ctx.xprojection(dev(mt), sx, ex))
        e = relerr(xp, O.xprojection(mt, sx, ex))
        print(f"[parity] xprojection W {W} columns {sx}..{ex}: relerr {e:.3e}", flush=True)
        assert e < 1e-6


def _drifted_ramps(A, H, W, seed):
    rng = np.random.default_rng(seed)
    Y = np.arange(H, dtype=np.float64)
    drift = rng.uniform(-2.0, 2.0, A)
    prof = [0.5 * (Y + drift[a]) + 3.0 + 0.2 * np.sin((Y + drift[a]) / 9.0) for a in range(A)]
    return np.stack([np.tile(p[:, None], (1, W)) for p in prof]).astype(np.float32)


@pytest.mark.parametrize("H,sz,ez", [(300, 10, 290), (64, 0, 63), (1000, 200, 700)])
def test_zcorrection_row_counts(ctx, H, sz, ez):
    """Profiles whose length is not a multiple of the 256-item group, and
    evaluation ranges touching the ends."""
    A = 12
    rng = np.random.default_rng(H)
    Y = np.arange(H, dtype=np.float64)
    xp = np.stack([0.5 * (Y + d) + 2.0 + 0.3 * np.sin((Y + d) / 7.0) for d in rng.uniform(-1.5, 1.5, A)])
    xp = xp.astype(np.float32)
    say("zcorrection", H, sz, ez)
    got = host(ctx.zcorrection(dev(xp), sz, ez))
    want = O.zcorrection(xp, sz, ez)
    print(f"[parity] zcorrection H {H}: max |diff| {np.abs(got - want).max():.3e} px", flush=True)
    assert np.abs(got - want).max() < 1e-4


def test_zcorrection_and_reslice_execution(ctx):
    A, H, W = 40, 512, 48
    mt = _drifted_ramps(A, H, W, 4)
    mt[5, 200, 17] += 7.0
    say("xprojection + zcorrection")
    xp = ctx.xprojection(dev(mt), 1, W - 2)
    ys = host(ctx.zcorrection(xp, 30, H - 30))
    want = O.zcorrection(host(xp), 30, H - 30)
    print(f"[parity] zcorrection on the same profiles: max |diff| {np.abs(ys - want).max():.3e} px, "
          f"bit-exact {np.array_equal(ys, want)}", flush=True)
    assert np.abs(ys - want).max() < 1e-4
    say("reslice_execution")
    prj, ys2 = ctx.reslice_execution(dev(mt), baseup=0.5, z_correction=True, start_x=1, end_x=W - 2, start_z=30,
                                     end_z=H - 30, denoise=True, threshold=2.0, y0=100, n_layers=300)
    wprj, wys = O.reslice_execution(mt, baseup=0.5, z_correction=True, start_x=1, end_x=W - 2, start_z=30,
                                    end_z=H - 30, denoise=True, threshold=2.0, y0=100, n_layers=300)
    e = relerr(host(prj), wprj)
    print(f"[parity] reslice_execution: shifts max |diff| {np.abs(host(ys2) - wys).max():.3e}, sinograms {e:.3e}",
          flush=True)
    assert np.abs(host(ys2) - wys).max() < 1e-3 and e < 1e-5


def test_resliced_sinograms_feed_osem(ctx):
    """Projections of a phantom volume, drifted by whole rows per angle, are
    resliced with those shifts straight into the OS-EM stack: the images are
    the ones OS-EM makes from the undrifted sinograms."""
    import torch
    from _util import phantom_stack
    n, pa, ss, nz = 64, 60, 6, 8
    _, sinos, ang = phantom_stack(nz, n, pa, 5, seed=9)
    mt = np.ascontiguousarray(sinos.transpose(1, 0, 2))             # [A][H = nz][W]
    drift = np.random.default_rng(0).integers(-2, 3, pa)
    H = nz + 8
    drifted = np.zeros((pa, H, n), np.float32)
    for a in range(pa):
        drifted[a, 4 - drift[a]:4 - drift[a] + nz] = mt[a]
    ctx.set_geometry(n, n, pa, ss, angles_deg=ang)
    say("reslice -> osem")
    y = ctx.reslice(dev(drifted), None, dev((4 - drift).astype(np.float32)), y0=0, n_layers=nz)
    x = torch.ones((nz, n, n), device="cuda")
    ctx.osem_execution(x, y, 2)
    x0 = dev(np.ones((nz, n, n)))
    ctx.osem_execution(x0, dev(sinos), 2)
    assert np.array_equal(host(y), sinos)
    assert np.array_equal(host(x), host(x0))


def test_reslice_2048_camera_1600_angles(ctx):
    """The reference's default camera (2048 x 2048) and angle count (1600):
    26.8 GB of projections resliced into 26.8 GB of sinograms.  Zero shifts
    give the exact transpose; whole-row shifts move every layer by that row."""
    import torch
    A, H, W = 1600, 2048, 2048
    say("reslice 2048 x 2048 x 1600: fill")
    g = torch.Generator(device="cuda").manual_seed(0)
    mt = torch.rand((A, H, W), device="cuda", generator=g)
    prj = torch.empty((H, A, W), device="cuda")
    say("reslice 2048 x 2048 x 1600: transpose")
    ctx.reslice(mt, out=prj)
    ctx.synchronize()
    for y in range(0, H, 128):
        assert torch.equal(prj[y:y + 128], mt[:, y:y + 128].transpose(0, 1)), y
    ys = torch.full((A,), 3.0, device="cuda")
    say("reslice 2048 x 2048 x 1600: shifted rows")
    ctx.reslice(mt, None, ys, y0=0, n_layers=H, out=prj)
    ctx.synchronize()
    for y in range(0, H - 3, 128):
        m = min(128, H - 3 - y)
        assert torch.equal(prj[y:y + m], mt[:, y + 3:y + 3 + m].transpose(0, 1)), y
    assert torch.count_nonzero(prj[H - 3:]) == 0
