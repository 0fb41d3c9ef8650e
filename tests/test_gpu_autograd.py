"""torch.autograd front end of the alignment model (ctrecon.autograd) and the
joint image + shift + centre FISTA-TV (ctrecon.joint).

Gradient checks on a 64^2 x 90-projection subproblem: image, shift and centre
gradients of a weighted sum of the projection against the FP64 oracle's
adjoint and shift gradient, the centre gradient against FP64 central finite
differences, the dot-product test through torch.autograd.grad, and chunked
stacks against unchunked ones.  Joint recovery on a 256^2 x 360-projection
case with shifts N(0, 1.5) px and a centre offset U(-3, 3) px."""
import numpy as np
import pytest

import oracle as O
from _util import relerr, say

pytestmark = pytest.mark.gpu


def _setup(ctx, n=64, pa=90, seed=11):
    rng = np.random.default_rng(seed)
    ang = (np.arange(pa) * (180.0 / pa) + 0.37).astype(np.float32)
    ctx.set_geometry(n, n, pa, 1, angles_deg=ang)
    return rng, ang, O.geom(n, n, pa, ss=1)


def test_gradients_vs_fp64_oracle_and_finite_differences(ctx):
    import torch
    from ctrecon import phantom as ph
    from ctrecon.autograd import project
    rng, ang, g = _setup(ctx)
    img = np.stack([ph.rasterize(ph.random_ellipses(8, rng), 64, supersample=2) for _ in range(2)])
    img = img.astype(np.float32)
    sh = rng.normal(0, 1.5, 90).astype(np.float32)
    w = rng.normal(0, 1, (2, 90, 64)).astype(np.float32)
    center = 0.7
    x = torch.from_numpy(img).cuda().requires_grad_(True)
    s = torch.from_numpy(sh).cuda().requires_grad_(True)
    c = torch.tensor(center, device="cuda", requires_grad=True)
    wt = torch.from_numpy(w).cuda()
    say("autograd forward/backward")
    loss = (wt * project(ctx, x, s, c)).sum()
    loss.backward()
    gx = O.align_adjoint(g, ang, w, sh, center, dtype=np.float64)
    gs = O.align_grad(g, ang, img, w, sh, center, dtype=np.float64)
    ex, es = relerr(x.grad.cpu().numpy(), gx), relerr(s.grad.cpu().numpy(), gs)
    print(f"[grad] image {ex:.2e} shifts {es:.2e}", flush=True)
    assert ex < 1e-5 and es < 1e-4
    # centre: FP64 central differences of the oracle forward
    h = 1e-6
    f = lambda cc: float((w.astype(np.float64) * O.align_forward(g, ang, img, sh, cc, dtype=np.float64)).sum())  # noqa: E731
    fd = (f(center + h) - f(center - h)) / (2 * h)
    ec = abs(float(c.grad) - fd) / abs(fd)
    print(f"[grad] centre {float(c.grad):.6e} fd {fd:.6e} rel {ec:.2e}", flush=True)
    assert ec < 1e-3
    assert abs(float(c.grad) - float(s.grad.sum())) <= 1e-6 * abs(float(c.grad)) + 1e-6


def test_dot_product_through_autograd(ctx):
    import torch
    from ctrecon.autograd import project
    rng, ang, g = _setup(ctx)
    sh = torch.from_numpy(rng.normal(0, 1.5, 90).astype(np.float32)).cuda()
    c = torch.tensor(-0.4, device="cuda")
    for trial in range(3):
        v = torch.from_numpy(rng.uniform(0, 1, (1, 64, 64)).astype(np.float32)).cuda().requires_grad_(True)
        u = torch.from_numpy(rng.uniform(0, 1, (1, 90, 64)).astype(np.float32)).cuda()
        say("autograd dot test", trial)
        av = project(ctx, v, sh, c)
        (atu,) = torch.autograd.grad(av, v, grad_outputs=u)
        lhs = float((av.double() * u.double()).sum())
        rhs = float((v.detach().double() * atu.double()).sum())
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        print(f"[dot] autograd <Av,u>={lhs:.9e} <v,A^T u>={rhs:.9e} rel={rel:.2e}", flush=True)
        assert rel < 1e-6


@pytest.mark.parametrize("nx,ny,px,pa", [(64, 64, 64, 90), (96, 40, 110, 36), (150, 150, 100, 40),
                                         (3900, 24, 72, 12)])
def test_derivative_rows_match_recomputation(ctx, nx, ny, px, pa):
    """ctr_align_forward_deriv (the forward writing each ray's offset
    derivative) against ctr_align_forward and ctr_align_grad_shift (sampling
    the image again): the projections are identical, sum_X g * dprj equals the
    recomputed shift gradient, both are as close to the FP64 oracle as the
    FP32 oracle is (or within 1e-4), and
    autograd gives the same gradients either way.  The LDS-tiled kernel, and
    the direct-gather one past 3840 columns."""
    import torch
    from ctrecon.autograd import project
    rng = np.random.default_rng(nx + pa)
    ang = (np.arange(pa) * (180.0 / pa) + 0.37).astype(np.float32)
    ctx.set_geometry(nx, px, pa, 1, ny=ny, angles_deg=ang)
    g = O.geom(nx, px, pa, ss=1, ny=ny)
    img = rng.uniform(0, 1, (3, ny, nx)).astype(np.float32)
    sh = rng.normal(0, 1.5, pa).astype(np.float32)
    w = rng.normal(0, 1, (3, pa, px)).astype(np.float32)
    center = -0.6
    x, s_, wt = (torch.from_numpy(a).cuda() for a in (img, sh, w))
    say("forward with derivative rows", nx, ny, px, pa)
    p0 = torch.empty((3, pa, px), device="cuda")
    p1, d1 = torch.empty_like(p0), torch.empty_like(p0)
    ctx.align_forward(x, s_, center, p0)
    ctx.align_forward_deriv(x, s_, center, p1, d1)
    gs_re = torch.empty(pa, device="cuda")
    ctx.align_grad_shift(x, wt, s_, center, gs_re)
    torch.cuda.synchronize()
    assert torch.equal(p0, p1)
    gs_d = (wt * d1).sum(dim=(0, 2))
    gs64 = O.align_grad(g, ang, img, w, sh, center, dtype=np.float64)
    # FP32 coordinates near column 3900 carry ~2e-4 px of rounding, and a
    # sample that changes texel cells changes its derivative by a step: the
    # bound is twice the FP32 oracle's own distance from FP64 (or 1e-4)
    bound = max(1e-4, 2.0 * relerr(O.align_grad(g, ang, img, w, sh, center, dtype=np.float32), gs64))
    e_d, e_re = relerr(gs_d.cpu().numpy(), gs64), relerr(gs_re.cpu().numpy(), gs64)
    e_dr = relerr(gs_d.cpu().numpy(), gs_re.cpu().numpy().astype(np.float64))
    print(f"[grad] derivative rows vs FP64 {e_d:.2e}, recomputed vs FP64 {e_re:.2e} (bound {bound:.1e}), rows vs "
          f"recomputed {e_dr:.2e}", flush=True)
    assert e_d < bound and e_re < bound and e_dr < 1e-5
    grads = []
    for derivatives in (True, False):
        xv = x.clone().requires_grad_(True)
        sv = s_.clone().requires_grad_(True)
        cv = torch.tensor(center, device="cuda", requires_grad=True)
        (wt * project(ctx, xv, sv, cv, derivatives=derivatives)).sum().backward()
        grads.append((xv.grad, sv.grad, cv.grad))
    assert torch.equal(grads[0][0], grads[1][0])
    assert relerr(grads[0][1].cpu().numpy(), grads[1][1].cpu().numpy().astype(np.float64)) < 1e-5
    # the centre gradient sums the shift gradients, which partly cancel: bound
    # its difference by the sum of their magnitudes
    assert abs(float(grads[0][2] - grads[1][2])) <= 1e-5 * float(grads[1][1].abs().sum()) + 1e-6


def test_padded_grid_adjoint_and_gradients_vs_fp64(ctx):
    """The joint solver's padded grids (image 1.45x the detector width): the
    adjoint against the FP64 oracle's and the dot test <A x, y> = <x, A^T y>,
    on a 146^2 grid with 100 bins and 48 angles."""
    import torch
    n, px, pa = 146, 100, 48
    rng = np.random.default_rng(77)
    ang = (np.arange(pa) * (180.0 / pa) + 0.37).astype(np.float32)
    ctx.set_geometry(n, px, pa, 1, ny=n, angles_deg=ang)
    g = O.geom(n, px, pa, ss=1, ny=n)
    x = rng.uniform(0, 1, (2, n, n)).astype(np.float32)
    y = rng.normal(0, 1, (2, pa, px)).astype(np.float32)
    sh = rng.normal(0, 1.5, pa).astype(np.float32)
    center = 0.9
    say("padded grid adjoint", n, px, pa)
    ax = torch.empty((2, pa, px), device="cuda")
    aty = torch.empty((2, n, n), device="cuda")
    xs, ys, shs = (torch.from_numpy(a).cuda() for a in (x, y, sh))
    ctx.align_forward(xs, shs, center, ax)
    ctx.align_adjoint(ys, shs, center, aty)
    torch.cuda.synchronize()
    a64 = O.align_adjoint(g, ang, y, sh, center, dtype=np.float64)
    f64 = O.align_forward(g, ang, x, sh, center, dtype=np.float64)
    e_a, e_f = relerr(aty.cpu().numpy(), a64), relerr(ax.cpu().numpy(), f64)
    lhs = float((ax.double() * ys.double()).sum())
    rhs = float((xs.double() * aty.double()).sum())
    dot = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"[grad] padded grid: forward {e_f:.2e}, adjoint {e_a:.2e} vs FP64, dot test {dot:.2e}", flush=True)
    assert e_f < 1e-5 and e_a < 1e-5 and dot < 1e-6


def test_chunked_stack_matches(ctx):
    import torch
    from ctrecon.autograd import project
    rng, ang, g = _setup(ctx)
    img = torch.from_numpy(rng.uniform(0, 1, (5, 64, 64)).astype(np.float32)).cuda()
    sh0 = torch.from_numpy(rng.normal(0, 1, 90).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.normal(0, 1, (5, 90, 64)).astype(np.float32)).cuda()
    res = []
    for chunk in (0, 2):
        x = img.clone().requires_grad_(True)
        s = sh0.clone().requires_grad_(True)
        c = torch.tensor(0.3, device="cuda", requires_grad=True)
        say("chunk", chunk)
        p = project(ctx, x, s, c, chunk=chunk)
        (w * p).sum().backward()
        res.append((p.detach(), x.grad, s.grad, c.grad))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])
    assert torch.allclose(res[0][2], res[1][2], rtol=1e-5, atol=1e-5 * float(res[0][2].abs().max()))
    assert abs(float(res[0][3] - res[1][3])) <= 1e-5 * abs(float(res[0][3]))


def test_joint_alignment_recovery(ctx):
    """Shifts + centre are identifiable only modulo the translation components
    {cos theta, sin theta} of the offsets (moving the image moves every
    projection by x cos theta - y sin theta), so the offset error is measured
    after removing its least-squares fit on them."""
    import torch
    from ctrecon import phantom as ph
    from ctrecon.joint import joint_fista_tv
    n, pa = 256, 360
    rng = np.random.default_rng(4)
    shifts = rng.normal(0, 1.5, pa)
    center = rng.uniform(-3, 3)
    _, sino, ang = ph.make_slice(n, pa, 10, 321, shifts=shifts, center=center)
    ctx.set_geometry(n, n, pa, 1, angles_deg=ang)
    y = torch.from_numpy(sino[None].copy()).cuda()
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    say("joint FISTA-TV 256^2 x 360")
    x, s, c, hist = joint_fista_tv(ctx, y, iters=100, tv=1e-3)
    _, _, _, h0 = joint_fista_tv(ctx, y, iters=100, tv=1e-3, align=False)
    e0 = off_err(np.zeros(pa))
    e1 = off_err((s + c).cpu().numpy().astype(np.float64))
    print(f"[joint] offset error mod translation {e0:.3f} -> {e1:.3f} px; data term {h0[-1]:.3e} (no alignment)"
          f" -> {hist[-1]:.3e}", flush=True)
    assert e1 < 0.3 and e1 < 0.2 * e0
    assert hist[-1] < 0.2 * h0[-1]
    assert float(x.min()) >= 0.0          # the per-slice TV prox keeps x >= 0


@pytest.mark.parametrize("pad", [1.0, 1.45])
def test_joint_alignment_c4_size_on_a_padded_grid(pad):
    """The C4 case (512^2, 720 projections, shifts N(0, 1.5) px, a centre
    offset U(-3, 3), TV 1e-3, 100 iterations) with a phantom whose ellipses
    reach beyond the detector's square.  On the 512^2 grid the offsets
    absorb the mass the grid cannot hold and the error grows; on a grid
    1.45x the detector width (joint_grid) it falls below 0.05 px and keeps
    falling after iteration 20."""
    import torch
    from ctrecon import Context
    from ctrecon import phantom as ph
    from ctrecon.joint import joint_fista_tv, joint_grid
    n, pa = 512, 720
    rng = np.random.default_rng(2027)
    shifts = rng.normal(0.0, 1.5, pa)
    center = rng.uniform(-3.0, 3.0)
    _, sino, ang = ph.make_slice(n, pa, 10, 4, shifts=shifts, center=center)
    m = joint_grid(n, pad)
    ctx = Context(0)
    ctx.set_geometry(m, n, pa, 1, ny=m, angles_deg=ang)
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    say("joint FISTA-TV 512^2 x 720, grid", m)
    errs = []
    y = torch.from_numpy(sino[None].copy()).cuda()
    x, s, c, hist = joint_fista_tv(ctx, y, iters=100, tv=1e-3,
                                   callback=lambda k, x_, off, f: errs.append(
                                       off_err(off.cpu().numpy().astype(np.float64))))
    e0 = off_err(np.zeros(pa))
    print(f"[joint] C4 size, grid {m}: offset error {e0:.3f} -> {errs[19]:.3f} (it 20) -> {errs[-1]:.3f} px; "
          f"data {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    if pad == 1.0:
        assert errs[-1] > e0            # the truncated object pulls the offsets away
    else:
        assert errs[-1] < 0.05 and errs[-1] <= errs[59] <= errs[19]
        assert hist[-1] < 1e-4 * hist[0]
    ctx.close()


def test_joint_alignment_shared_offsets_uneven_stack():
    """Slices of three sizes (two of each) from one dataset share the
    per-projection offsets (shifts N(0, 2) px and a centre offset, Poisson
    noise at I0 = 5e3): one offset estimate from all six slices."""
    import torch
    from ctrecon import Context, phantom as ph
    from ctrecon.joint import joint_fista_tv_groups
    pa = 360
    rng = np.random.default_rng(12)
    shifts = rng.normal(0, 2.0, pa)
    center = rng.uniform(-3, 3)
    groups, ctxs = [], []
    for gi, n in enumerate((128, 192, 256)):
        ys = []
        for k in range(2):
            _, sino, ang = ph.make_slice(n, pa, 10, 500 + gi, shifts=shifts, center=center, i0=5e3, slice_index=k)
            ys.append(sino)
        c = Context(0)
        c.set_geometry(n, n, pa, 1, angles_deg=ang)
        ctxs.append(c)
        groups.append((c, torch.from_numpy(np.stack(ys)).cuda()))
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    say("joint FISTA-TV, 3 groups x 2 slices, shared offsets")
    xs, s, c, hist = joint_fista_tv_groups(groups, iters=100, tv=1e-3)
    e0, e1 = off_err(np.zeros(pa)), off_err((s + c).cpu().numpy().astype(np.float64))
    print(f"[joint] uneven stack: offset error mod translation {e0:.3f} -> {e1:.3f} px, data {hist[0]:.3e} -> "
          f"{hist[-1]:.3e}", flush=True)
    # the 128^2 phantoms fill the detector, so shifts of several px push them
    # past its edge, which the model cannot represent: the error settles near
    # 0.45 px here even without noise (0.19 px on the 256^2 case above)
    assert e1 < 0.6 and e1 < 0.3 * e0
    assert [tuple(x.shape) for x in xs] == [(2, 128, 128), (2, 192, 192), (2, 256, 256)]
    for cx in ctxs:
        cx.close()
