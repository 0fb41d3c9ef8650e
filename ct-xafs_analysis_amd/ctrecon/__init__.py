"""ctrecon -- Python front end of the MI355X CT projector hot path.

The compute path is libctr_hip.so (hand-written gfx950 kernels behind the C ABI
of include/ctr.h); this package binds it with ctypes and adds the
torch.autograd wrapper of the differentiable alignment model.
"""
from ._lib import CtrError, Context, Geometry, LIB_PATH, EXPORTED, load, subset_order  # noqa: F401

__all__ = ["CtrError", "Context", "Geometry", "LIB_PATH", "EXPORTED", "load", "subset_order"]
