"""Rotation-centre search on the GPU (RotationCenterSearch/): the shifted
candidates (rotCenterShift) bit-exact with the oracle, the masked focus index
(setMask + imgFocusIndex) and the whole search (candidates, OS-EM with the
projection correction, focus indices) against the FP32 restatement, and at
1024^2 / 900 angles the focus index over one common mask smallest at the
shift the data were made with."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, relerr, say
from test_oracle_pins import _filled_phantom, _zero_fill_shift

pytestmark = pytest.mark.gpu


def _angles(pa):
    return (np.arange(pa) * (180.0 / pa)).astype(np.float32)


def test_rot_center_shift_bit_exact(ctx):
    n, px, pa, ss = 64, 72, 30, 5
    ang = _angles(pa)
    ctx.set_geometry(n, px, pa, ss, angles_deg=ang)
    g = O.geom(n, px, pa, ss=ss)
    prj = np.random.default_rng(6).uniform(-1.0, 2.0, (pa, px)).astype(np.float32)
    for start, step, cnt in ((-3.0, 1.0, 7), (-2.35, 0.45, 13), (0.5, 0.0, 1), (-80.0, 40.0, 5)):
        say("rotCenterShift", start, step, cnt)
        out = dev(np.zeros((cnt, pa, px)))
        ctx.op("rot_center_shift", dev(prj), out, cnt, start, step)
        assert np.array_equal(host(out), O.rot_center_shift(g, prj, cnt, start, step)), (start, step)


@pytest.mark.parametrize("nx,ny,rng_deg", [(96, 96, (0.0, 180.0)), (96, 96, (0.0, 90.0)), (80, 64, (12.5, 171.0)),
                                            (33, 33, (0.0, 179.0))])
def test_focus_index(ctx, nx, ny, rng_deg):
    pa, ss = 12, 4
    ctx.set_geometry(nx, nx, pa, ss, ny=ny, angles_deg=_angles(pa))
    g = O.geom(nx, nx, pa, ss=ss, ny=ny)
    rng = np.random.default_rng(nx + ny)
    img = rng.uniform(0.5, 1.5, (9, ny, nx)).astype(np.float32)
    say("focus index", nx, ny, rng_deg)
    f = dev(np.zeros(9))
    ctx.op("focus_index", dev(img), 9, -4.5, 1.25, rng_deg[0], rng_deg[1], f)
    want = O.focus_index(g, img, -4.5, 1.25, rng_deg[0], rng_deg[1])
    e = relerr(host(f), want)
    print(f"[parity] focus index {nx}x{ny} {rng_deg}: relerr {e:.3e}", flush=True)
    assert np.isfinite(want).all() and e < 1e-5


@pytest.mark.parametrize("n,px,pa,ss,s0", [(64, 100, 90, 10, 2), (128, 128, 180, 10, -3), (96, 144, 120, 8, 0)])
def test_rotation_center_search_parity(ctx, n, px, pa, ss, s0):
    """px >= 1.5 n keeps the whole filled grid on the detector at every angle;
    px = n truncates the projections, a parity-only case."""
    ang = _angles(pa)
    g = O.geom(n, px, pa, ss=ss)
    sino = O.align_forward(g, ang, _filled_phantom(n, 2)[None])[0]
    y = _zero_fill_shift(sino, -s0).astype(np.float32)
    ctx.set_geometry(n, px, pa, ss, correction_mode=1, angles_deg=ang)
    say("rotation centre search", n, px, pa, ss, s0)
    recon, findex, best = ctx.rotation_center_search(dev(y), 11, -5.0, 1.0, it=1, init=1.0)
    want_rec, want_f = O.rotation_center_search(g, ang, y, 11, -5.0, 1.0, it=1, init=1.0, correction_mode=1)
    er, ef = relerr(host(recon), want_rec), relerr(host(findex), want_f)
    print(f"[parity] centre search n {n} s0 {s0}: recon {er:.3e}, focus index {ef:.3e}, best {best}", flush=True)
    assert er < 1e-5 and ef < 1e-4
    assert best == np.float32(-5.0 + np.argmin(want_f))
    if px >= 1.5 * n:
        assert abs(best - s0) <= 1


def test_rotation_center_search_1024(ctx):
    """1024^2 grid, 1536 bins (the filled grid stays on the detector), 900
    angles, 9 candidates around an axis 3 bins right of the detector centre;
    the data come from the library's own projector (xi = X - px/2 - center).
    Over one common mask the focus index is smallest at the true shift."""
    import torch
    n, px, pa, ss, s0 = 1024, 1536, 900, 10, 3
    ang = _angles(pa)
    ctx.set_geometry(n, px, pa, ss, correction_mode=1, angles_deg=ang)
    img = dev(_filled_phantom(n, 3)[None])
    y = torch.zeros((1, pa, px), device="cuda")
    say("centre search 1024: data")
    ctx.align_forward(img, None, float(s0), y)
    say("centre search 1024: search")
    recon, findex, best = ctx.rotation_center_search(y[0], 9, -1.0, 1.0, it=1, init=1.0)
    f = host(findex)
    print(f"[property] centre search 1024: focus index {np.round(f, 4)}, best {best}", flush=True)
    assert np.isfinite(f).all() and abs(best - s0) <= 2
    common = dev(np.zeros(9))
    ctx.op("focus_index", recon, 9, 0.0, 0.0, float(ang[0]), float(ang[-1]), common)
    assert int(np.argmin(host(common))) - 1 == s0
