"""The C ABI library loads and exports every entry point include/ctr.h
declares; host-only entry points behave as documented without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ctr.h")
LIB = os.path.join(ROOT, "ct-xafs_analysis_amd", "libctr_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ctr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("ctr_create", "ctr_osem_execution", "ctr_fbp_execution", "ctr_aart_execution",
                 "ctr_fista_execution", "ctr_fista_osem_execution", "ctr_power_iteration",
                 "ctr_osem_batched", "ctr_align_forward", "ctr_align_adjoint", "ctr_align_grad_shift"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    import ctrecon
    assert set(declared()) <= set(ctrecon.EXPORTED)


def test_host_only_entry_points():
    lib = C.CDLL(LIB)
    lib.ctr_status_string.restype = C.c_char_p
    assert lib.ctr_status_string(0) == b"ok"
    assert lib.ctr_status_string(2) == b"invalid geometry"
    sub = np.zeros(16, np.int32)
    assert lib.ctr_subset_order(16, sub.ctypes.data_as(C.c_void_p)) == 0
    import oracle as O
    assert sub.tolist() == O.subset_order(16).tolist()
    assert lib.ctr_subset_order(0, sub.ctypes.data_as(C.c_void_p)) == 1
    # null context -> CTR_INVALID_ARGUMENT, never a crash
    assert lib.ctr_synchronize(None) == 1
    assert lib.ctr_set_geometry(None, None, None, None) == 1
    for setter in ("ctr_set_sync_each_kernel", "ctr_set_kernel_timing", "ctr_set_adaptive_rows"):
        assert getattr(lib, setter)(None, 1) == 1, setter
    lib.ctr_last_error.restype = C.c_char_p
    assert lib.ctr_last_error(None) == b"null context"


def test_create_without_gpu_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = C.CDLL(LIB)
    h = C.c_void_p()
    assert lib.ctr_create(0, C.byref(h)) == 3   # CTR_NO_DEVICE


def test_subset_order_matches_the_oracle_for_every_count():
    """The library forms the order as a bit reversal, the oracle by doubling
    the list: they agree for every subset count up to 256."""
    import oracle as O
    from ctrecon import subset_order
    for ss in range(1, 257):
        assert subset_order(ss).tolist() == O.subset_order(ss).tolist(), ss


def test_query_layout_host_only():
    from ctrecon._lib import query_layout
    l = query_layout(1024, 1024, 1024)
    assert l.pad == 100 and l.ipitch % 16 == 0 and l.ipitch >= 1224 and l.tpitch >= 1224
    assert l.roff >= 1 and l.rpitch >= 1024 + 2 * l.roff and l.rpitch % 16 == 0
    lib = C.CDLL(LIB)
    assert lib.ctr_query_layout(0, 4, 4, None) == 1
