"""Joint image + projection-alignment reconstruction (FISTA-TV).

Minimises  1/2 || A(o) x - y ||^2 + tv * TV(x)  over the image stack x >= 0 and
the per-projection detector offsets o[a] = shifts[a] + center, where A(o) is
the alignment forward model of ctrecon.autograd (the OSEM1 projector with
moved detector coordinates).  Each iteration:

  * one FISTA step on x: the gradient A^T (A z - y) and the offset gradient
    come from one backward pass through AlignedProjection (the library's
    adjoint and shift-gradient kernels); the proximal step is the per-slice
    median-form TV prox with four in-plane neighbours (ctr_tv_prox_2d, the
    2-D form of CT_reconstruction/FISTA.cl:409-493), with lambda = tv * step
    and x >= 0.  The slices of a stack are unrelated, so no slice is coupled
    to another (the reference's ISTA3D would tie each slice to the next one
    and to layer 0).  With precondition=True the data term is
    weighted by R = 1 / (A 1) (inverse ray lengths) and the step by
    C = 1 / (A^T 1) (the SIRT scaling), with which a unit step is stable;
    the TV weight is then applied with the mean of C as its step;
  * after `warmup` iterations, one damped Gauss-Newton step on every offset:
    o[a] -= g[a] / (h[a] + mu), with g[a] the offset gradient and
    h[a] = sum over layers and bins of (d p / d X)^2 the curvature of the
    projection row, each step clipped to +-max_step bins.

Shifts and centre enter the model only through their sum, so the split is a
convention: center = mean(o), shifts = o - center.

joint_fista_tv_groups runs several stacks of different slice sizes (one
context each) against one set of offsets, as the slices of one CT-XAFS
dataset share the per-projection detector motion: the offset gradient and
curvature are summed over all slices of all groups.

The image grid of a context may be larger than its detector (nx, ny > px),
and should be whenever the object can reach beyond the square the detector
width spans: every point of the object projects onto the detector at some
angles, so mass the grid cannot hold makes the data inconsistent with the
model, and the solver absorbs the mismatch in the offsets.  joint_grid(px)
gives a grid side that holds an object of radius 1.45 px / 2.
"""
import math

import torch

from .autograd import project


def joint_grid(px: int, pad: float = 1.45) -> int:
    """An even image side of about pad * px for a detector of px bins."""
    return int(round(pad * px / 2.0)) * 2


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


class _Group:
    """One stack of equal-size slices (one library context / geometry)."""

    def __init__(self, lib, y, off, tv, precondition, power_iters, x0):
        g = lib.geom
        self.lib, self.y = lib, y
        nz, dev = y.shape[0], y.device
        self.nz, self.shape = nz, (nz, g.ny, g.nx)
        if precondition:
            ones = torch.ones(self.shape, dtype=torch.float32, device=dev)
            a1 = torch.empty((nz, g.pa, g.px), dtype=torch.float32, device=dev)
            lib.align_forward(ones, off, 0.0, a1)
            self.rw = torch.where(a1 > 1.0e-3, 1.0 / a1.clamp(min=1.0e-3), torch.zeros_like(a1))
            at1 = torch.empty_like(ones)
            lib.align_adjoint(torch.ones_like(a1), off, 0.0, at1)
            self.cw = torch.where(at1 > 1.0e-3, 1.0 / at1.clamp(min=1.0e-3), torch.zeros_like(at1))
            self.step = float(self.cw.mean())
        else:
            self.rw = self.cw = None
            self.step = 1.0 / _norm2_sq(lib, nz, g.ny, g.nx, off, dev, power_iters)
        self.lam = tv * self.step
        self.x = (torch.zeros(self.shape, dtype=torch.float32, device=dev) if x0 is None
                  else x0.to(dev, torch.float32).contiguous().clone())
        self.z = self.x.clone()


def joint_fista_tv_groups(groups, *, iters=100, tv=1.0e-3, align=True, warmup=10, max_step=0.5, mu=1.0e-3,
                          precondition=False, power_iters=15, chunk=0, offsets0=None, x0s=None, callback=None):
    """Joint reconstruction of several slice stacks that share the detector
    offsets: groups is a list of (context, y) with y [nz][pa][px] in that
    context's geometry (sides may differ between groups; pa must not).  The
    offset gradient and the Gauss-Newton curvature are summed over every
    slice of every group.  Returns (xs, shifts, center, history).
    callback(k, xs, offsets, data), if given, runs after iteration k with the
    data term as a 0-d device tensor, so that a callback that does not read
    it makes the host wait for nothing."""
    pa = groups[0][0].geom.pa
    if any(lib.geom.pa != pa for lib, _ in groups):
        raise ValueError("all groups need the same projection count")
    dev = groups[0][1].device
    off = (torch.zeros(pa, dtype=torch.float32, device=dev) if offsets0 is None
           else offsets0.to(dev, torch.float32).clone())
    zero = torch.zeros((), dtype=torch.float32, device=dev)
    st = [_Group(lib, y, off, tv, precondition, power_iters, None if x0s is None else x0s[i])
          for i, (lib, y) in enumerate(groups)]
    t = 1.0
    hist = []
    for k in range(iters):
        o = off.detach().requires_grad_(align and k >= warmup)
        h = torch.zeros(pa, dtype=torch.float32, device=dev)
        data = torch.zeros((), dtype=torch.float64, device=dev)
        new_x = []
        for G in st:
            zr = G.z.detach().requires_grad_(True)
            p = project(G.lib, zr, o, zero, chunk=chunk)
            r = p - G.y
            loss = 0.5 * (r * r).sum() if G.rw is None else 0.5 * (G.rw * r * r).sum()
            loss.backward()        # the offset gradient accumulates over the groups
            with torch.no_grad():
                data += 0.5 * (r * r).sum(dtype=torch.float64)
                v = (G.z - (G.step * zr.grad if G.cw is None else G.cw * zr.grad)).contiguous()
                x_new = torch.empty_like(v)
                G.lib.tv_prox_2d(v, x_new, G.lam, 0.0)
                new_x.append(x_new)
                if o.requires_grad:
                    pd = p.detach()
                    dp = torch.zeros_like(pd)
                    dp[..., 1:-1] = 0.5 * (pd[..., 2:] - pd[..., :-2])
                    h += (dp * dp).sum(dim=(0, 2)) if G.rw is None else (G.rw * dp * dp).sum(dim=(0, 2))
        hist.append(data)          # stays on the device: no host wait inside an iteration
        with torch.no_grad():
            if o.requires_grad:
                d = -o.grad / (h + mu * h.mean())
                off += d.clamp(-max_step, max_step)
            t_next = 0.5 * (1.0 + math.sqrt(1.0 + 4.0 * t * t))
            for G, x_new in zip(st, new_x):
                G.z = x_new + ((t - 1.0) / t_next) * (x_new - G.x)
                G.x = x_new
            t = t_next
        if callback is not None:
            callback(k, [G.x for G in st], off, hist[-1])
    center = off.mean()
    return [G.x for G in st], off - center, center, torch.stack(hist).tolist() if hist else []


def joint_fista_tv(lib, y, *, iters=100, tv=1.0e-3, align=True, warmup=10, max_step=0.5, mu=1.0e-3,
                   precondition=False, power_iters=15, chunk=0, offsets0=None, x0=None, callback=None):
    """y: measured sinograms [nz][pa][px] (float32 GPU tensor) of the context's
    geometry; x0 an optional starting image (zeros by default).  Returns
    (x, shifts, center, history) with history a list of the data term
    1/2 ||A x - y||^2 per iteration."""
    cb = None if callback is None else (lambda k, xs, off, f: callback(k, xs[0], off, f))
    xs, shifts, center, hist = joint_fista_tv_groups(
        [(lib, y)], iters=iters, tv=tv, align=align, warmup=warmup, max_step=max_step, mu=mu,
        precondition=precondition, power_iters=power_iters, chunk=chunk, offsets0=offsets0,
        x0s=None if x0 is None else [x0], callback=cb)
    return xs[0], shifts, center, hist
