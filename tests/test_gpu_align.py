"""Differentiable alignment model: forward and shift gradient against the FP64
oracle, the adjoint by the dot-product test (<= 1e-6) and against the oracle's
scatter adjoint, and the shift/centre gradient against FP64 central finite
differences on a 64^2 x 90-projection subproblem."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, relerr, say

pytestmark = pytest.mark.gpu


def _setup(ctx, n=64, pa=90, seed=3):
    rng = np.random.default_rng(seed)
    ang = (np.arange(pa) * (180.0 / pa) + 0.37).astype(np.float32)
    ctx.set_geometry(n, n, pa, 1, angles_deg=ang)
    return rng, ang, O.geom(n, n, pa, ss=1)


def test_forward_matches_oracle_and_osem_projector(ctx):
    rng, ang, g = _setup(ctx)
    img = rng.uniform(0, 1, (2, 64, 64)).astype(np.float32)
    shifts = rng.normal(0, 1.5, 90).astype(np.float32)
    say("align_forward")
    out = dev(np.zeros((2, 90, 64)))
    ctx.align_forward(dev(img), dev(shifts), 1.3, out)
    exp = O.align_forward(g, ang, img, shifts, 1.3, dtype=np.float64)
    assert relerr(host(out), exp) < 1e-5
    # zero shift/centre: the OSEM1 projector itself (AART1 with y = 0 gives -Ax)
    out0 = dev(np.zeros((2, 90, 64)))
    ctx.align_forward(dev(img), None, 0.0, out0)
    minus = O.aart1(g, ang, img, np.zeros((2, 90, 64), np.float32), 0)
    assert relerr(host(out0), -minus) < 1e-5


def test_adjoint_dot_product(ctx):
    """<Ax, y> = <x, A^T y>.  Positive random vectors: relative difference of the
    two inner products <= 1e-6.  Signed vectors (the inner product can cancel to
    a small number): difference scaled by ||Ax|| ||y|| <= 1e-6."""
    rng, ang, g = _setup(ctx)
    shifts = rng.normal(0, 1.5, 90).astype(np.float32)
    for trial in range(6):
        signed = trial >= 3
        if signed:
            x = rng.normal(0, 1, (1, 64, 64)).astype(np.float32)
            y = rng.normal(0, 1, (1, 90, 64)).astype(np.float32)
        else:
            x = rng.uniform(0, 1, (1, 64, 64)).astype(np.float32)
            y = rng.uniform(0, 1, (1, 90, 64)).astype(np.float32)
        say("dot test", trial)
        ax, aty = dev(np.zeros_like(y)), dev(np.zeros_like(x))
        ctx.align_forward(dev(x), dev(shifts), -0.8, ax)
        ctx.align_adjoint(dev(y), dev(shifts), -0.8, aty)
        axh = host(ax).astype(np.float64)
        lhs = float(np.dot(axh.ravel(), y.ravel().astype(np.float64)))
        rhs = float(np.dot(x.ravel().astype(np.float64), host(aty).ravel().astype(np.float64)))
        scale = np.linalg.norm(axh) * np.linalg.norm(y.astype(np.float64)) if signed else max(abs(lhs), abs(rhs))
        rel = abs(lhs - rhs) / scale
        print(f"[dot] {'signed' if signed else 'positive'} <Ax,y>={lhs:.9e} <x,ATy>={rhs:.9e} rel={rel:.2e}",
              flush=True)
        assert rel < 1e-6


def test_adjoint_matches_oracle_scatter(ctx):
    rng, ang, g = _setup(ctx)
    shifts = rng.normal(0, 1.5, 90).astype(np.float32)
    y = rng.normal(0, 1, (2, 90, 64)).astype(np.float32)
    say("align_adjoint")
    out = dev(np.zeros((2, 64, 64)))
    ctx.align_adjoint(dev(y), dev(shifts), 0.6, out)
    exp = O.align_adjoint(g, ang, y, shifts, 0.6, dtype=np.float64)
    assert relerr(host(out), exp) < 1e-5


def test_shift_gradient_vs_fp64_and_finite_differences(ctx):
    rng, ang, g = _setup(ctx, seed=7)
    from ctrecon import phantom as ph
    img = ph.rasterize(ph.random_ellipses(8, rng), 64, supersample=2)[None].astype(np.float32)
    shifts = rng.normal(0, 1.5, 90).astype(np.float32)
    center = 0.9
    gp = rng.normal(0, 1, (1, 90, 64)).astype(np.float32)
    say("align_grad_shift")
    gs = dev(np.zeros(90))
    ctx.align_grad_shift(dev(img), dev(gp), dev(shifts), center, gs)
    got = host(gs)
    exp = O.align_grad(g, ang, img, gp, shifts, center, dtype=np.float64)
    assert relerr(got, exp) < 1e-4
    # FP64 central differences of <gp, A(img; shifts, center)> on the oracle
    img64, s64, gp64 = img.astype(np.float64), shifts.astype(np.float64), gp.astype(np.float64)
    h = 1e-6   # small enough that few samples straddle a bilinear kink
    fd = np.zeros(90)
    for a in range(90):
        sp, sm = s64.copy(), s64.copy()
        sp[a] += h
        sm[a] -= h
        fp_ = np.sum(gp64 * O.align_forward(g, ang, img64, sp, center, dtype=np.float64))
        fm_ = np.sum(gp64 * O.align_forward(g, ang, img64, sm, center, dtype=np.float64))
        fd[a] = (fp_ - fm_) / (2 * h)
    e_fd = relerr(got, fd)
    fdc = (np.sum(gp64 * O.align_forward(g, ang, img64, s64, center + h, dtype=np.float64))
           - np.sum(gp64 * O.align_forward(g, ang, img64, s64, center - h, dtype=np.float64))) / (2 * h)
    print(f"[fd] shift grad relerr vs FD {e_fd:.2e}; centre grad {got.sum():.6e} vs FD {fdc:.6e}", flush=True)
    assert e_fd < 1e-3
    assert abs(got.sum() - fdc) <= 1e-3 * max(1.0, abs(fdc))


@pytest.mark.parametrize("nx,ny,px,pa", [(96, 64, 128, 60), (3900, 40, 3900, 6)])
def test_nonsquare_and_gather_fallback(ctx, nx, ny, px, pa):
    """nx != ny with a detector wider than the image (the LDS-tiled forward
    and gradient), and a side past 60 x 64 = 3840 samples per ray (the direct
    gather kernels): forward, adjoint and shift gradient against FP64."""
    rng = np.random.default_rng(nx + ny)
    ang = (np.arange(pa) * (180.0 / pa) + 0.37).astype(np.float32)
    ctx.set_geometry(nx, px, pa, 1, ny=ny, angles_deg=ang)
    g = O.geom(nx, px, pa, ss=1, ny=ny)
    # a smooth image that falls to 0 at its edges: the shift gradient of
    # bilinear sampling jumps where a sample crosses a texel edge, by the
    # image's second difference there (the zero border included), and FP32 and
    # FP64 coordinates put a few samples on either side of an edge
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    img = (np.sin(np.pi * (xx + 0.5) / nx) ** 2 * np.sin(np.pi * (yy + 0.5) / ny) ** 2
           * (1.0 + 0.3 * np.sin(xx / 7.0 + 0.3) * np.cos(yy / 5.0)))[None].astype(np.float32)
    shifts = rng.normal(0, 1.5, pa).astype(np.float32)
    gp = rng.normal(0, 1, (1, pa, px)).astype(np.float32)
    say("align ops", nx, ny, px, pa)
    # FP32 coordinates near 1800 carry 1.2e-4 px of rounding, so at large
    # sides the bound is the FP32 oracle's own distance from FP64
    f64 = O.align_forward(g, ang, img, shifts, 0.4)
    a64 = O.align_adjoint(g, ang, gp, shifts, 0.4)
    g64 = O.align_grad(g, ang, img, gp, shifts, 0.4)
    b_f = max(1e-5, 2.0 * relerr(O.align_forward(g, ang, img, shifts, 0.4, dtype=np.float32), f64))
    b_a = max(1e-5, 2.0 * relerr(O.align_adjoint(g, ang, gp, shifts, 0.4, dtype=np.float32), a64))
    b_g = max(1e-4, 2.0 * relerr(O.align_grad(g, ang, img, gp, shifts, 0.4, dtype=np.float32), g64))
    out = dev(np.zeros((1, pa, px)))
    ctx.align_forward(dev(img), dev(shifts), 0.4, out)
    e_f = relerr(host(out), f64)
    adj = dev(np.zeros((1, ny, nx)))
    ctx.align_adjoint(dev(gp), dev(shifts), 0.4, adj)
    e_a = relerr(host(adj), a64)
    gs = dev(np.zeros(pa))
    ctx.align_grad_shift(dev(img), dev(gp), dev(shifts), 0.4, gs)
    e_g = relerr(host(gs), g64)
    print(f"[parity] align {nx}x{ny}, {px} bins, {pa} angles: forward {e_f:.2e} (bound {b_f:.1e}), adjoint "
          f"{e_a:.2e} ({b_a:.1e}), shift gradient {e_g:.2e} ({b_g:.1e})", flush=True)
    assert e_f < b_f and e_a < b_a and e_g < b_g
