"""Joint image + projection-alignment reconstruction (FISTA-TV).

Minimises  1/2 || A(o) x - y ||^2 + tv * TV(x)  over the image stack x >= 0 and
the per-projection detector offsets o[a] = shifts[a] + center, where A(o) is
the alignment forward model of ctrecon.autograd (the OSEM1 projector with
moved detector coordinates).  Each iteration:

  * one FISTA step on x: the gradient A^T (A z - y) and the offset gradient
    come from one backward pass through AlignedProjection (the library's
    adjoint and shift-gradient kernels); the proximal step is the
    reference's median-form TV prox ISTA3D (CT_reconstruction/FISTA.cl:409-493,
    which also keeps x >= 1e-6), applied by the library with
    lambda = tv * step.  With precondition=True the data term is
    weighted by R = 1 / (A 1) (inverse ray lengths) and the step by
    C = 1 / (A^T 1) (the SIRT scaling), which makes a unit step stable and
    converges far faster than the 1 / ||A||^2 step; the TV weight is then
    applied with the mean of C as its step;
  * after `warmup` iterations, one damped Gauss-Newton step on every offset:
    o[a] -= g[a] / (h[a] + mu), with g[a] the offset gradient and
    h[a] = sum over layers and bins of (d p / d X)^2 the curvature of the
    projection row, each step clipped to +-max_step bins.

Shifts and centre enter the model only through their sum, so the split is a
convention: center = mean(o), shifts = o - center.
"""
import math

import torch

from .autograd import project


def _norm2_sq(lib, nz, ny, nx, off, device, iters):
    """||A(o)||^2 by power iteration on A^T A."""
    c = 0.0
    v = torch.ones((nz, ny, nx), dtype=torch.float32, device=device)
    v /= v.norm()
    g = lib.geom
    p = torch.empty((nz, g.pa, g.px), dtype=torch.float32, device=device)
    w = torch.empty_like(v)
    lam = 1.0
    for _ in range(iters):
        lib.align_forward(v, off, c, p)
        lib.align_adjoint(p, off, c, w)
        lam = float(w.norm())
        v = w / lam
    return lam


def joint_fista_tv(lib, y, *, iters=100, tv=1.0e-3, align=True, warmup=10, max_step=0.5, mu=1.0e-3,
                   precondition=False, power_iters=15, chunk=0, offsets0=None, x0=None, callback=None):
    """y: measured sinograms [nz][pa][px] (float32 GPU tensor) of the context's
    geometry; x0 an optional starting image (zeros by default).  Returns
    (x, shifts, center, history) with history a list of the data term
    1/2 ||A x - y||^2 per iteration."""
    g = lib.geom
    nz = y.shape[0]
    nx, ny, pa = g.nx, g.ny, g.pa
    dev = y.device
    off = (torch.zeros(pa, dtype=torch.float32, device=dev) if offsets0 is None
           else offsets0.to(dev, torch.float32).clone())
    zero = torch.zeros((), dtype=torch.float32, device=dev)
    if precondition:
        ones = torch.ones((nz, ny, nx), dtype=torch.float32, device=dev)
        a1 = torch.empty((nz, g.pa, g.px), dtype=torch.float32, device=dev)
        lib.align_forward(ones, off, 0.0, a1)
        rw = torch.where(a1 > 1.0e-3, 1.0 / a1.clamp(min=1.0e-3), torch.zeros_like(a1))
        at1 = torch.empty_like(ones)
        lib.align_adjoint(torch.ones_like(a1), off, 0.0, at1)
        cw = torch.where(at1 > 1.0e-3, 1.0 / at1.clamp(min=1.0e-3), torch.zeros_like(at1))
        step = float(cw.mean())
    else:
        rw = cw = None
        step = 1.0 / _norm2_sq(lib, nz, ny, nx, off, dev, power_iters)
    cs = float(g.cs_lambda)
    l2 = torch.tensor([0.5 * cs / (tv * step)], dtype=torch.float32, device=dev)   # ISTA3D: lambda = cs / l2 * 0.5
    x = (torch.zeros((nz, ny, nx), dtype=torch.float32, device=dev) if x0 is None
         else x0.to(dev, torch.float32).contiguous().clone())
    z = x.clone()
    t = 1.0
    hist = []
    for k in range(iters):
        zr = z.detach().requires_grad_(True)
        o = off.detach().requires_grad_(align and k >= warmup)
        p = project(lib, zr, o, zero, chunk=chunk)
        r = p - y
        loss = 0.5 * (r * r).sum() if rw is None else 0.5 * (rw * r * r).sum()
        loss.backward()
        hist.append(float(0.5 * (r.detach() * r.detach()).sum()))
        with torch.no_grad():
            v = (z - (step * zr.grad if cw is None else cw * zr.grad)).contiguous()
            x_new = torch.empty_like(v)
            lib.op("ista3d", v, x_new, nz, 0, l2)
            if o.requires_grad:
                pd = p.detach()
                dp = torch.zeros_like(pd)
                dp[..., 1:-1] = 0.5 * (pd[..., 2:] - pd[..., :-2])
                h = (dp * dp).sum(dim=(0, 2)) if rw is None else (rw * dp * dp).sum(dim=(0, 2))
                d = -o.grad / (h + mu * h.mean())
                off += d.clamp(-max_step, max_step)
            t_next = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
            z = x_new + ((t - 1.0) / t_next) * (x_new - x)
            x, t = x_new, t_next
        if callback is not None:
            callback(k, x, off, hist[-1])
    center = off.mean()
    return x, off - center, center, hist
