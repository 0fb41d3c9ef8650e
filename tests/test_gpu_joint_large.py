"""The batched joint-alignment workload at its real sizes (96 slices of
512^2..2048^2 with 1800 projections, shifts N(0, 2) px, Poisson noise at
I0 = 5e3): the alignment kernels on a 2048^2 slice with 1800 projections
against the FP64 oracle, the per-slice TV prox against the oracle, and the
joint solver on groups of the workload's slice sizes."""
import numpy as np
import pytest

import oracle as O
from _util import dev, host, relerr, say

pytestmark = pytest.mark.gpu


def test_tv_prox_2d_matches_oracle_and_keeps_slices_apart(ctx):
    rng = np.random.default_rng(41)
    nx, ny, pa = 200, 136, 8
    ctx.set_geometry(nx, nx, pa, 1, ny=ny, angles_deg=np.arange(pa, dtype=np.float32))
    g = O.geom(nx, nx, pa, ss=1, ny=ny)
    v = rng.uniform(-0.2, 1.0, (3, ny, nx)).astype(np.float32)
    say("tv_prox_2d", nx, ny)
    out = dev(np.zeros_like(v))
    ctx.tv_prox_2d(dev(v), out, 0.03, 0.0)
    got = host(out)
    want = O.tv_prox2d(g, v, 0.03, 0.0)
    assert np.array_equal(got, want)          # the median is one of the nine values, computed alike
    # slice 1 alone gives the same result: no slice is coupled to another
    one = dev(np.zeros((1, ny, nx)))
    ctx.tv_prox_2d(dev(v[1:2]), one, 0.03, 0.0)
    assert np.array_equal(host(one)[0], got[1])


def _smooth_image(n, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    a, b = rng.uniform(3.0, 9.0, 2)
    return (np.sin(np.pi * (xx + 0.5) / n) ** 2 * np.sin(np.pi * (yy + 0.5) / n) ** 2
            * (1.0 + 0.3 * np.sin(xx / (a * n / 256.0)) * np.cos(yy / (b * n / 256.0))))[None].astype(np.float32)


def test_align_kernels_2048sq_1800proj_vs_fp64(ctx):
    """Forward, adjoint, shift gradient and the forward with derivative rows on
    one 2048^2 slice with all 1800 projections on the device (the LDS-tiled forward / gradient and the
    pixel-driven adjoint at the workload's largest side), compared with the
    FP64 oracle on 12 of the angles: their forward rows and shift gradients
    depend on those angles alone, and an incoming gradient that is zero on
    every other angle makes the adjoint the oracle's adjoint over the 12."""
    n, pa = 2048, 1800
    rng = np.random.default_rng(42)
    ang = (np.arange(pa) * (180.0 / pa)).astype(np.float32)
    shifts = rng.normal(0, 2.0, pa).astype(np.float32)
    center = np.float32(rng.uniform(-3, 3))
    ctx.set_geometry(n, n, pa, 1, angles_deg=ang)
    pick = np.sort(rng.choice(pa, 12, replace=False))
    gs_ = O.geom(n, n, len(pick), ss=1)
    img = _smooth_image(n, 43)
    gp = np.zeros((1, pa, n), np.float32)
    gp[0, pick] = rng.normal(0, 1, (len(pick), n)).astype(np.float32)
    say("align kernels 2048^2 x 1800 projections")
    out = dev(np.zeros((1, pa, n)))
    ctx.align_forward(dev(img), dev(shifts), float(center), out)
    fwd = host(out)[0, pick]
    adj_t = dev(np.zeros((1, n, n)))
    ctx.align_adjoint(dev(gp), dev(shifts), float(center), adj_t)
    adj = host(adj_t)
    gsh = dev(np.zeros(pa))
    ctx.align_grad_shift(dev(img), dev(gp), dev(shifts), float(center), gsh)
    grad = host(gsh)
    assert np.all(grad[np.setdiff1d(np.arange(pa), pick)] == 0.0)
    # the forward with derivative rows: the same projections, and the shift
    # gradient as a reduction of the rows
    out2, drows = dev(np.zeros((1, pa, n))), dev(np.zeros((1, pa, n)))
    ctx.align_forward_deriv(dev(img), dev(shifts), float(center), out2, drows)
    assert np.array_equal(host(out2)[0, pick], fwd)
    grad_d = (gp[0, pick].astype(np.float64) * host(drows)[0, pick]).sum(axis=1)
    a_s, s_s, gp_s = ang[pick], shifts[pick], gp[:, pick]
    f64 = O.align_forward(gs_, a_s, img, s_s, float(center))
    a64 = O.align_adjoint(gs_, a_s, gp_s, s_s, float(center))
    g64 = O.align_grad(gs_, a_s, img, gp_s, s_s, float(center))
    # FP32 coordinates near 2048 carry ~1e-4 px of rounding: the bound is twice
    # the FP32 oracle's own distance from FP64 (or the op tolerance, if larger)
    b_f = max(1e-5, 2.0 * relerr(O.align_forward(gs_, a_s, img, s_s, float(center), dtype=np.float32), f64))
    b_a = max(1e-5, 2.0 * relerr(O.align_adjoint(gs_, a_s, gp_s, s_s, float(center), dtype=np.float32), a64))
    b_g = max(1e-4, 2.0 * relerr(O.align_grad(gs_, a_s, img, gp_s, s_s, float(center), dtype=np.float32), g64))
    e_f, e_a, e_g = relerr(fwd, f64[0]), relerr(adj, a64), relerr(grad[pick], g64)
    e_d = relerr(grad_d, g64)
    print(f"[parity] align 2048^2 x 1800 (12 angles vs FP64): forward {e_f:.2e} (bound {b_f:.1e}), adjoint "
          f"{e_a:.2e} ({b_a:.1e}), shift gradient {e_g:.2e} ({b_g:.1e}), from derivative rows {e_d:.2e}",
          flush=True)
    assert e_f < b_f and e_a < b_a and e_g < b_g and e_d < b_g


def test_joint_alignment_on_c5_sized_groups():
    """Two slices each of 512^2, 1024^2 and 2048^2 with 1800 projections,
    shifts N(0, 2) px, a centre offset U(-3, 3) and Poisson noise at
    I0 = 5e3, each reconstructed on a grid 1.45x its detector width (the
    phantoms reach beyond the detector's square): 30 joint iterations
    (offsets updated from iteration 11) cut the offset error modulo image
    translation below 0.25 of its start, the data term falls 100-fold, and
    every slice stays finite and >= 0."""
    import torch
    from ctrecon import Context, phantom as ph
    from ctrecon.joint import joint_fista_tv_groups, joint_grid
    pa = 1800
    rng = np.random.default_rng(44)
    shifts = rng.normal(0, 2.0, pa)
    center = rng.uniform(-3, 3)
    groups, ctxs = [], []
    for gi, n in enumerate((512, 1024, 2048)):
        ys = []
        for k in range(2):
            _, sino, ang = ph.make_slice(n, pa, 10, 700 + gi, shifts=shifts, center=center, i0=5e3, slice_index=k)
            ys.append(sino)
        c = Context(0)
        m = joint_grid(n)
        c.set_geometry(m, n, pa, 1, ny=m, angles_deg=ang)
        ctxs.append(c)
        groups.append((c, torch.from_numpy(np.stack(ys)).cuda()))
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    say("joint FISTA-TV, 512 / 1024 / 2048 x 2 slices, 1800 projections")
    errs = []
    xs, s, c, hist = joint_fista_tv_groups(groups, iters=30, tv=1e-3, warmup=10,
                                           callback=lambda k, xs_, off, f: errs.append(
                                               off_err(off.cpu().numpy().astype(np.float64))))
    e0, e1 = off_err(np.zeros(pa)), off_err((s + c).cpu().numpy().astype(np.float64))
    print(f"[joint] 1800 projections, sides 512-2048 on padded grids: offset error {e0:.3f} -> {e1:.3f} px "
          f"(iteration 20: {errs[19]:.3f}), data {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert e1 < 0.25 * e0
    assert hist[-1] < 0.01 * hist[0]
    for x in xs:
        assert torch.isfinite(x).all() and float(x.min()) >= 0.0
    for cx in ctxs:
        cx.close()


def test_joint_offsets_converge_when_the_object_fits_the_field_of_view():
    """The offset error of the workload's phantoms stops falling near
    iteration 20 and then creeps up: their ellipses reach radius 1.4 of the
    half-width, beyond the square grid, so rays through mass the image cannot
    hold still reach the detector and the solver absorbs the mismatch in the
    offsets.  With the same phantoms scaled by 0.6 (radius <= 0.84) the model
    fits, and the error keeps falling: 8 x 512^2 + 4 x 1024^2, 1800
    projections, shifts N(0, 2), centre U(-3, 3), I0 = 5e3, 60 iterations."""
    import torch
    from ctrecon import Context, phantom as ph
    from ctrecon.joint import joint_fista_tv_groups
    pa = 1800
    rng = np.random.default_rng(2026)
    shifts = rng.normal(0.0, 2.0, pa)
    center = rng.uniform(-3.0, 3.0)
    ang = ph.uniform_angles(pa)
    th = np.deg2rad(ang.astype(np.float64))
    B = np.stack([np.cos(th), np.sin(th)], 1)

    def off_err(off):
        e = off - (shifts + center)
        coef, *_ = np.linalg.lstsq(B, e, rcond=None)
        return float(np.sqrt(np.mean((e - B @ coef) ** 2)))

    groups, ctxs = [], []
    for n, cnt in ((512, 8), (1024, 4)):
        ys = []
        for k in range(cnt):
            r = np.random.default_rng(ph.stream_seed(900 + n, k))
            e = ph.random_ellipses(10, r)
            e = ph.Ellipses(e.cx * 0.6, e.cy * 0.6, e.a * 0.6, e.b * 0.6, e.phi, e.mu)
            ys.append(ph.transmission_noise(ph.line_integrals(e, n, ang, n, shifts=shifts, center=center), 5e3, r))
        c = Context(0)
        c.set_geometry(n, n, pa, 1, angles_deg=ang)
        ctxs.append(c)
        groups.append((c, torch.from_numpy(np.stack(ys).astype(np.float32)).cuda()))
    say("joint FISTA-TV, objects inside the field of view, 8 x 512^2 + 4 x 1024^2")
    errs = []
    xs, s, c, hist = joint_fista_tv_groups(groups, iters=60, tv=1e-3, warmup=10,
                                           callback=lambda k, xs_, off, f: errs.append(
                                               off_err(off.cpu().numpy().astype(np.float64))))
    print(f"[joint] objects inside the field of view: offset error {off_err(np.zeros(pa)):.3f} -> "
          f"{errs[19]:.3f} (it 20) -> {errs[29]:.3f} (30) -> {errs[59]:.3f} px (60)", flush=True)
    assert errs[59] < 0.08
    assert errs[59] <= errs[29] + 1e-3 and errs[29] <= errs[19] + 1e-3
    assert hist[-1] < 0.01 * hist[0]
    for cx in ctxs:
        cx.close()
