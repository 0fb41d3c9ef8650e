"""The Python binding checks every tensor's shape against the geometry before
a device pointer reaches the library: a tensor of the wrong extent would make
a kernel read or write past its buffer.  CPU only: the context is built
without a device, its library is a stub that records calls, and every
mis-shaped call must raise ValueError with no call recorded."""
import numpy as np
import pytest
import torch

from ctrecon._lib import Context, Geometry


class _StubLib:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*a):
            self.calls.append(name)
            return 0
        return f


def _ctx(nx=16, ny=12, px=20, pa=12, ss=4):
    c = object.__new__(Context)
    c._torch = torch
    c.lib = _StubLib()
    c.h = None
    c.device = 0
    c._ext = None
    c.geom = Geometry(nx, ny, nx, ny, px, pa, ss, 64, 6, 0, 1.0, 0.01, 1e-4)
    c.angles = np.zeros(pa, np.float32)
    return c


def z(*shape):
    return torch.zeros(shape, dtype=torch.float32)


BAD_CALLS = {
    # drivers: image [dN][ny][nx], sinogram [dN][pa][px]
    "osem image transposed": lambda c: c.osem_execution(z(2, 16, 12), z(2, 12, 20), 1),
    "osem sinogram short": lambda c: c.osem_execution(z(2, 12, 16), z(2, 12, 19), 1),
    "osem layer counts differ": lambda c: c.osem_execution(z(2, 12, 16), z(3, 12, 20), 1),
    "osem 2-D image": lambda c: c.osem_execution(z(12, 16), z(1, 12, 20), 1),
    "fista_osem": lambda c: c.fista_osem_execution(z(1, 12, 16), z(1, 11, 20), 1),
    "aart": lambda c: c.aart_execution(z(1, 13, 16), z(1, 12, 20), 1),
    "fbp": lambda c: c.fbp_execution(z(1, 12, 16), z(1, 12, 21)),
    "hybrid": lambda c: c.hybrid_execution(z(1, 12, 17), z(1, 12, 20), 1),
    "fista L too short": lambda c: c.fista_execution(z(3), z(1, 12, 16), z(1, 12, 20), 1),
    "power L too short": lambda c: c.power_iteration(z(2)),
    # operators
    "osem1 out": lambda c: c.op("osem1", z(1, 12, 16), z(1, 12, 20), z(1, 12, 16), 1, 0),
    "osem1 dN beyond tensors": lambda c: c.op("osem1", z(1, 12, 16), z(1, 12, 20), z(1, 12, 20), 2, 0),
    "osem2 image": lambda c: c.op("osem2", z(1, 12, 15), z(1, 12, 20), z(1, 12, 16), 1, 0),
    "aart2 dst": lambda c: c.op("aart2", z(1, 12, 16), z(1, 12, 20), z(1, 12, 20), 1, 0, 0.5),
    "backprojection prj": lambda c: c.op("backprojection", z(1, 6, 20), z(1, 12, 16), 1, 0),
    "fbp_filter": lambda c: c.op("fbp_filter", z(2, 12, 20), z(1, 12, 20), 2),
    "fista_backprojection L": lambda c: c.op("fista_backprojection", z(1, 12, 16), z(1, 12, 20), z(1),
                                             z(1, 12, 16), 1, 0, 1.0),
    "ista3d": lambda c: c.op("ista3d", z(2, 12, 16), z(1, 12, 16), 2, 0, z(4)),
    "fista3d": lambda c: c.op("fista3d", z(1, 12, 16), z(1, 12, 16), z(1, 12, 8), z(1, 12, 16), None, None,
                              1, 0, z(4)),
    "power_iter1": lambda c: c.op("power_iter1", z(3, 12, 16), z(4, 12, 20)),
    "power_iter2": lambda c: c.op("power_iter2", z(4, 12, 20), z(4, 16, 12)),
    "rot_center_shift out": lambda c: c.op("rot_center_shift", z(12, 20), z(2, 12, 20), 3, -1.0, 1.0),
    "focus_index findex": lambda c: c.op("focus_index", z(3, 12, 16), 3, 0.0, 1.0, 0.0, 180.0, z(2)),
    # batched stacks: square slices, [pa][n] sinograms
    "batched rectangle": lambda c: c.osem_batched([z(16, 12)], [z(12, 16)], 1),
    "batched sinogram": lambda c: c.osem_batched([z(16, 16)], [z(12, 15)], 1),
    "batched counts": lambda c: c.osem_batched([z(16, 16)], [], 1),
    # alignment model
    "align_forward prj": lambda c: c.align_forward(z(1, 12, 16), None, 0.0, z(1, 12, 16)),
    "align_forward shifts": lambda c: c.align_forward(z(1, 12, 16), z(11), 0.0, z(1, 12, 20)),
    "align_adjoint img": lambda c: c.align_adjoint(z(2, 12, 20), None, 0.0, z(1, 12, 16)),
    "align_grad gshift": lambda c: c.align_grad_shift(z(1, 12, 16), z(1, 12, 20), z(12), 0.0, z(13)),
    "align_forward_deriv dprj": lambda c: c.align_forward_deriv(z(1, 12, 16), None, 0.0, z(1, 12, 20), z(1, 12, 19)),
    # sinogram formation
    "reslice shifts": lambda c: c.reslice(z(5, 4, 6), xshift=z(4)),
    "reslice out": lambda c: c.reslice(z(5, 4, 6), out=z(4, 5, 5)),
    "zcorrection yshift": lambda c: c.zcorrection(z(5, 4), 0, 3, yshift=z(4)),
    "rotation centre prj": lambda c: c.rotation_center_search(z(12, 19), 3, -1.0, 1.0),
}


@pytest.mark.parametrize("name", sorted(BAD_CALLS))
def test_misshaped_tensor_raises_before_any_launch(name):
    c = _ctx()
    with pytest.raises(ValueError):
        BAD_CALLS[name](c)
    assert c.lib.calls == [], c.lib.calls


def test_well_shaped_calls_pass_the_checks():
    """The same calls with matching shapes get past the checks (they then stop
    at the pointer conversion, which wants GPU tensors)."""
    c = _ctx()
    good = [
        lambda: c.osem_execution(z(2, 12, 16), z(2, 12, 20), 1),
        lambda: c.op("osem1", z(1, 12, 16), z(1, 12, 20), z(1, 12, 20), 1, 0),
        lambda: c.op("fista3d", z(1, 12, 16), z(1, 12, 16), z(1, 12, 16), z(1, 12, 16), None, None, 1, 0, z(4)),
        lambda: c.op("power_iter1", z(4, 12, 16), z(4, 12, 20)),
        lambda: c.align_grad_shift(z(1, 12, 16), z(1, 12, 20), z(12), 0.0, z(12)),
        lambda: c.osem_batched([z(16, 16)], [z(12, 16)], 1),
    ]
    for f in good:
        with pytest.raises(ValueError, match="GPU tensor"):
            f()
    assert c.lib.calls == []
