"""torch.autograd front end of the differentiable alignment model.

The forward model is the OSEM1 projector (CT_reconstruction/OSEM.cl:56-62)
over all angles of the context's geometry, with the detector coordinate of
projection a moved by shifts[a] + center:

    prj[z][a][X] = sum_k bilinear(img[z], u(X, k), v(X, k)),
    xi = (X - px/2) - shifts[a] - center

Its gradients come from the library's hand-written kernels: the image
gradient is the exact adjoint (ctr_align_adjoint, the transpose of the
forward's sampling weights); the shift gradient is each ray's derivative
with respect to its offset, weighted by the incoming gradient and summed
over the stack's layers and the rays of a projection; the centre gradient
is the sum of the shift gradients.  When the offsets need a gradient, the
forward writes those derivatives from the same samples
(ctr_align_forward_deriv, one more projection-sized array kept for the
backward), so the backward only reduces them.  With derivatives=False the
backward instead samples the image again (ctr_align_grad_shift:
checkpointed recomputation), and a call holds only its input image.
project() can split a large stack into layer chunks to bound the library's
workspace.
"""
import torch

from ._lib import Context


class AlignedProjection(torch.autograd.Function):
    """prj = A(shifts, center) img for an image stack [nz][ny][nx]."""

    @staticmethod
    def forward(fctx, img, shifts, center, lib: Context, pa: int, px: int, derivatives: bool):
        img = img.detach().contiguous()
        sh = shifts.detach().contiguous()
        c = float(center.detach())
        prj = torch.empty((img.shape[0], pa, px), dtype=torch.float32, device=img.device)
        dprj = None
        if derivatives and (fctx.needs_input_grad[1] or fctx.needs_input_grad[2]):
            dprj = torch.empty_like(prj)
            lib.align_forward_deriv(img, sh, c, prj, dprj)
        else:
            lib.align_forward(img, sh, c, prj)
        fctx.save_for_backward(img, sh, dprj)
        fctx.center = c
        fctx.lib = lib
        return prj

    @staticmethod
    def backward(fctx, gprj):
        img, sh, dprj = fctx.saved_tensors
        lib, c = fctx.lib, fctx.center
        gprj = gprj.contiguous()
        g_img = g_sh = g_c = None
        if fctx.needs_input_grad[0]:
            g_img = torch.empty_like(img)
            lib.align_adjoint(gprj, sh, c, g_img)
        if fctx.needs_input_grad[1] or fctx.needs_input_grad[2]:
            if dprj is not None:
                gs = (gprj * dprj).sum(dim=(0, 2))
            else:
                gs = torch.empty_like(sh)
                lib.align_grad_shift(img, gprj, sh, c, gs)
            if fctx.needs_input_grad[1]:
                g_sh = gs
            if fctx.needs_input_grad[2]:
                g_c = gs.sum().reshape(())
        return g_img, g_sh, g_c, None, None, None, None


def project(lib: Context, img, shifts, center, chunk: int = 0, derivatives: bool = True):
    """A(shifts, center) img with autograd; `chunk` > 0 runs the stack in
    groups of that many layers (each group recomputed in its own backward);
    derivatives=False recomputes the offset derivatives in the backward
    instead of keeping them from the forward."""
    g = lib.geom
    if g is None:
        raise RuntimeError("set_geometry() first")
    pa, px = g.pa, g.px
    if shifts is None:
        shifts = torch.zeros(pa, dtype=torch.float32, device=img.device)
    if not torch.is_tensor(center):
        center = torch.tensor(float(center), dtype=torch.float32, device=img.device)
    if chunk <= 0 or chunk >= img.shape[0]:
        return AlignedProjection.apply(img, shifts, center, lib, pa, px, derivatives)
    parts = [AlignedProjection.apply(img[z:z + chunk], shifts, center, lib, pa, px, derivatives)
             for z in range(0, img.shape[0], chunk)]
    return torch.cat(parts, 0)
