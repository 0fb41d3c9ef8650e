"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle_ct.so (built by oracle/Makefile) and exposes numpy
wrappers.  dtype selects the instantiation: np.float32 -> the FP32 compat
restatement of the reference kernels, np.float64 -> the FP64 oracle.
Only tests, the smoke check and the benchmark's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_ct.so")
_lib = None


class Geom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("px", C.c_int), ("pa", C.c_int),
                ("ss", C.c_int), ("zp", C.c_int)]


def build():
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def geom(nx, px, pa, ss=1, zp=None, ny=None):
    ny = nx if ny is None else ny
    if zp is None:
        zp = 1
        while zp < 2 * px:
            zp *= 2
    return Geom(nx, ny, px, pa, ss, zp)


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ang(ang):
    return np.ascontiguousarray(ang, dtype=np.float32)


def subset_order(ss):
    sub = np.zeros(ss, dtype=np.int32)
    lib().ctor_subset_order(C.c_int(ss), _p(sub))
    return sub


def num_threads():
    return lib().ctor_num_threads()


def _call(name, dtype, *args):
    f = getattr(lib(), f"ctor_{name}_{_sfx(dtype)}")
    f.restype = None
    return f(*args)


# ---- single operators -----------------------------------------------------

def osem1(g, ang, img, prj, sub, dtype=np.float32):
    img, prj = _arr(img, dtype), _arr(prj, dtype)
    nz = img.shape[0]
    out = np.zeros_like(prj)
    _call("osem1", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(prj), _p(out), C.c_int(nz), C.c_int(sub))
    return out


def osem2(g, ang, img, rprj, sub, dtype=np.float32):
    img, rprj = _arr(img, dtype), _arr(rprj, dtype)
    out = np.zeros_like(img)
    _call("osem2", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(rprj), _p(out), C.c_int(img.shape[0]), C.c_int(sub))
    return out


def aart1(g, ang, img, prj, sub, dtype=np.float32):
    img, prj = _arr(img, dtype), _arr(prj, dtype)
    out = np.zeros_like(prj)
    _call("aart1", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(prj), _p(out), C.c_int(img.shape[0]), C.c_int(sub))
    return out


def aart2(g, ang, img, dprj, sub, alpha, dtype=np.float32):
    img, dprj = _arr(img, dtype), _arr(dprj, dtype)
    out = np.zeros_like(img)
    _call("aart2", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(dprj), _p(out), C.c_int(img.shape[0]),
          C.c_int(sub), C.c_float(alpha))
    return out


def backprojection(g, ang, prj, sub, dtype=np.float32):
    prj = _arr(prj, dtype)
    nz = prj.shape[0]
    out = np.zeros((nz, g.ny, g.nx), dtype=dtype)
    _call("backprojection", dtype, C.byref(g), _p(_ang(ang)), _p(prj), _p(out), C.c_int(nz), C.c_int(sub))
    return out


def sinogram_correction(g, ang, prj, mode, a=0.5, b=0.5, dtype=np.float32):
    prj = _arr(prj, dtype)
    out = np.zeros_like(prj)
    _call("sinogram_correction", dtype, C.byref(g), _p(_ang(ang)), _p(prj), _p(out), C.c_int(prj.shape[0]),
          C.c_int(mode), C.c_float(a), C.c_float(b))
    return out


def fbp_filter(g, prj, dtype=np.float32):
    prj = _arr(prj, dtype)
    out = np.zeros_like(prj)
    _call("fbp_filter", dtype, C.byref(g), _p(prj), _p(out), C.c_int(prj.shape[0]))
    return out


def fbp_backprojection(g, ang, fprj, dtype=np.float32):
    fprj = _arr(fprj, dtype)
    nz = fprj.shape[0]
    out = np.zeros((nz, g.ny, g.nx), dtype=dtype)
    _call("fbp_backprojection", dtype, C.byref(g), _p(_ang(ang)), _p(fprj), _p(out), C.c_int(nz))
    return out


def power_iter1(g, ang, img, dtype=np.float32):
    img = _arr(img, dtype)
    out = np.zeros((g.ss, g.pa, g.px), dtype=dtype)
    _call("power_iter1", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(out))
    return out


def power_iter2(g, ang, prjc, dtype=np.float32):
    prjc = _arr(prjc, dtype)
    out = np.zeros((g.ss, g.ny, g.nx), dtype=dtype)
    _call("power_iter2", dtype, C.byref(g), _p(_ang(ang)), _p(prjc), _p(out))
    return out


def fista_backprojection(g, ang, img, dprj, L, sub, alpha, dtype=np.float32):
    img, dprj, L = _arr(img, dtype), _arr(dprj, dtype), _arr(L, dtype)
    out = np.zeros_like(img)
    _call("fista_backprojection", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(dprj), _p(L), _p(out),
          C.c_int(img.shape[0]), C.c_int(sub), C.c_float(alpha))
    return out


def ista3d(g, v, L, sub, lam, dtype=np.float32):
    v, L = _arr(v, dtype), _arr(L, dtype)
    out = np.zeros_like(v)
    _call("ista3d", dtype, C.byref(g), _p(v), _p(out), C.c_int(v.shape[0]), C.c_int(sub), _p(L), C.c_double(lam))
    return out


def fista3d(g, x, v, b, L, sub, lam, dtype=np.float32):
    x, v, b, L = _arr(x, dtype), _arr(v, dtype), _arr(b, dtype), _arr(L, dtype)
    w, xn, bn = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    _call("fista3d", dtype, C.byref(g), _p(x), _p(v), _p(b), _p(w), _p(xn), _p(bn), C.c_int(x.shape[0]),
          C.c_int(sub), _p(L), C.c_double(lam))
    return w, xn, bn


def tv_prox2d(g, v, lam, lower=0.0, dtype=np.float32):
    """Per-slice TV prox (four in-plane neighbours) of the joint alignment model."""
    v = _arr(v, dtype)
    out = np.zeros_like(v)
    R = C.c_float if np.dtype(dtype) == np.float32 else C.c_double
    _call("tv_prox2d", dtype, C.byref(g), _p(v), _p(out), C.c_int(v.shape[0]), R(lam), R(lower))
    return out


# ---- drivers ---------------------------------------------------------------

def osem_execution(g, ang, img, prj, it, sub=None, prj_correction=True, correction_mode=0, dtype=np.float32):
    img, prj = _arr(img, dtype).copy(), _arr(prj, dtype).copy()
    sub = subset_order(g.ss) if sub is None else np.ascontiguousarray(sub, dtype=np.int32)
    _call("osem_execution", dtype, C.byref(g), _p(_ang(ang)), _p(sub), _p(img), _p(prj), C.c_int(img.shape[0]),
          C.c_int(it), C.c_int(int(prj_correction)), C.c_int(correction_mode))
    return img


def fista_osem_execution(g, ang, img, prj, it, cs_lambda, sub=None, prj_correction=True, correction_mode=0,
                         dtype=np.float32):
    img, prj = _arr(img, dtype).copy(), _arr(prj, dtype).copy()
    sub = subset_order(g.ss) if sub is None else np.ascontiguousarray(sub, dtype=np.int32)
    _call("fista_osem_execution", dtype, C.byref(g), _p(_ang(ang)), _p(sub), _p(img), _p(prj),
          C.c_int(img.shape[0]), C.c_int(it), C.c_int(int(prj_correction)), C.c_int(correction_mode),
          C.c_double(cs_lambda))
    return img


def aart_execution(g, ang, img, prj, it, alpha, sub=None, correction_mode=0, dtype=np.float32):
    img, prj = _arr(img, dtype).copy(), _arr(prj, dtype).copy()
    sub = subset_order(g.ss) if sub is None else np.ascontiguousarray(sub, dtype=np.int32)
    _call("aart_execution", dtype, C.byref(g), _p(_ang(ang)), _p(sub), _p(img), _p(prj), C.c_int(img.shape[0]),
          C.c_int(it), C.c_int(correction_mode), C.c_float(alpha))
    return img


def fbp_execution(g, ang, prj, prj_correction=True, correction_mode=0, dtype=np.float32):
    prj = _arr(prj, dtype).copy()
    nz = prj.shape[0]
    img = np.zeros((nz, g.ny, g.nx), dtype=dtype)
    _call("fbp_execution", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(prj), C.c_int(nz),
          C.c_int(int(prj_correction)), C.c_int(correction_mode))
    return img


def layer_min(g, img, dtype=np.float32):
    img = _arr(img, dtype)
    mins = np.zeros(img.shape[0], dtype=dtype)
    _call("layer_min", dtype, C.byref(g), _p(img), _p(mins), C.c_int(img.shape[0]))
    return mins


def baseup(g, img, mins, order, dtype=np.float32):
    img = _arr(img, dtype).copy()
    mins = _arr(mins, dtype)
    _call("baseup", dtype, C.byref(g), _p(img), _p(mins), C.c_int(img.shape[0]), C.c_int(order))
    return img


def hybrid_execution(g, ang, prj, it, baseup_order=4, cs=False, cs_lambda=1e-4, sub=None, prj_correction=True,
                     correction_mode=0, dtype=np.float32):
    """Mode 8 (hybrid_thread, GPU.cpp:1459-1580); returns the image stack."""
    prj = _arr(prj, dtype).copy()
    nz = prj.shape[0]
    img = np.zeros((nz, g.ny, g.nx), dtype=dtype)
    sub = subset_order(g.ss) if sub is None else np.ascontiguousarray(sub, dtype=np.int32)
    _call("hybrid_execution", dtype, C.byref(g), _p(_ang(ang)), _p(sub), _p(img), _p(prj), C.c_int(nz),
          C.c_int(it), C.c_int(int(prj_correction)), C.c_int(correction_mode), C.c_int(baseup_order),
          C.c_int(int(cs)), C.c_double(cs_lambda))
    return img


def power_iteration(g, ang, trials=30, dtype=np.float32):
    L = np.zeros(g.ss, dtype=dtype)
    _call("power_iteration", dtype, C.byref(g), _p(_ang(ang)), _p(L), C.c_int(trials))
    return L


def fista_execution(g, ang, L, img, prj, it, alpha, cs_lambda, sub=None, prj_correction=True,
                    correction_mode=0, dtype=np.float32):
    img, prj, L = _arr(img, dtype).copy(), _arr(prj, dtype).copy(), _arr(L, dtype)
    sub = subset_order(g.ss) if sub is None else np.ascontiguousarray(sub, dtype=np.int32)
    _call("fista_execution", dtype, C.byref(g), _p(_ang(ang)), _p(L), _p(sub), _p(img), _p(prj),
          C.c_int(img.shape[0]), C.c_int(it), C.c_int(int(prj_correction)), C.c_int(correction_mode),
          C.c_float(alpha), C.c_double(cs_lambda))
    return img


# ---- differentiable alignment model --------------------------------------

def align_forward(g, ang, img, shift=None, center=0.0, dtype=np.float64):
    img = _arr(img, dtype)
    nz = img.shape[0]
    out = np.zeros((nz, g.pa, g.px), dtype=dtype)
    sh = None if shift is None else _arr(shift, dtype)
    R = C.c_float if np.dtype(dtype) == np.float32 else C.c_double
    _call("align_forward", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(sh), R(center), _p(out), C.c_int(nz))
    return out


def align_adjoint(g, ang, prj, shift=None, center=0.0, dtype=np.float64):
    prj = _arr(prj, dtype)
    nz = prj.shape[0]
    out = np.zeros((nz, g.ny, g.nx), dtype=dtype)
    sh = None if shift is None else _arr(shift, dtype)
    R = C.c_float if np.dtype(dtype) == np.float32 else C.c_double
    _call("align_adjoint", dtype, C.byref(g), _p(_ang(ang)), _p(prj), _p(sh), R(center), _p(out), C.c_int(nz))
    return out


def align_grad(g, ang, img, gprj, shift=None, center=0.0, dtype=np.float64):
    img, gprj = _arr(img, dtype), _arr(gprj, dtype)
    out = np.zeros(g.pa, dtype=dtype)
    sh = None if shift is None else _arr(shift, dtype)
    R = C.c_float if np.dtype(dtype) == np.float32 else C.c_double
    _call("align_grad", dtype, C.byref(g), _p(_ang(ang)), _p(img), _p(gprj), _p(sh), R(center), _p(out),
          C.c_int(img.shape[0]))
    return out


# ---- rotation-centre search (RotationCenterSearch/) -------------------------

def rcs_mask(nx, ny, n, start, step, ang_min, ang_max):
    """setMask (rotCenterSearch.cl:50-92): uint8 mask [n, ny, nx]."""
    out = np.zeros((n, ny, nx), dtype=np.uint8)
    lib().ctor_rcs_mask(C.c_int(nx), C.c_int(ny), C.c_int(n), C.c_float(start), C.c_float(step),
                        C.c_float(ang_min), C.c_float(ang_max), _p(out))
    return out


def rot_center_shift(g, prj, n, start, step, dtype=np.float32):
    """rotCenterShift (rotCenterSearch.cl:35-48) of one layer [pa, px] for n candidates."""
    prj = _arr(prj, dtype).reshape(g.pa, g.px)
    out = np.zeros((n, g.pa, g.px), dtype=dtype)
    _call("rot_center_shift", dtype, C.byref(g), _p(prj), _p(out), C.c_int(n), C.c_float(start), C.c_float(step))
    return out


def focus_index(g, img, start, step, ang_min, ang_max, dtype=np.float32):
    """imgFocusIndex (rotCenterSearch.cl:154-227) of each layer of img [n, ny, nx]."""
    img = _arr(img, dtype)
    n = img.shape[0]
    out = np.zeros(n, dtype=dtype)
    _call("focus_index", dtype, C.byref(g), _p(img), _p(out), C.c_int(n), C.c_float(start), C.c_float(step),
          C.c_float(ang_min), C.c_float(ang_max))
    return out


def rotation_center_search(g, ang, prj, n, start, step, it=1, init=1.0, ang_range=None, sub=None,
                           prj_correction=True, correction_mode=0, dtype=np.float32):
    """rotationCenterSearch from the resliced layer on
    (rotationCenterSearch_ocl.cpp:489-551); returns (recon, findex)."""
    prj = _arr(prj, dtype).reshape(g.pa, g.px)
    ang = _ang(ang)
    if ang_range is None:
        ang_range = (float(ang[0]), float(ang[-1]))
    sub = subset_order(g.ss) if sub is None else np.ascontiguousarray(sub, dtype=np.int32)
    recon = np.zeros((n, g.ny, g.nx), dtype=dtype)
    findex = np.zeros(n, dtype=dtype)
    _call("rotation_center_search", dtype, C.byref(g), _p(ang), _p(sub), _p(prj), C.c_int(n), C.c_float(start),
          C.c_float(step), C.c_float(ang_range[0]), C.c_float(ang_range[1]), C.c_int(it), C.c_float(init),
          C.c_int(int(prj_correction)), C.c_int(correction_mode), _p(recon), _p(findex))
    return recon, findex


# ---- sinogram formation (Reslice/, ImageReg share source/) ------------------

def merge_frames(raw, n, W, H, stride=None, dtype=np.float32):
    """merge_rawhisdata (imageregistration_kernel_src.cl:868-881)."""
    raw = np.ascontiguousarray(raw, dtype=np.uint16)
    stride = W * H if stride is None else stride
    out = np.zeros((H, W), dtype=dtype)
    _call("merge_frames", dtype, _p(raw), C.c_int(n), C.c_int(W), C.c_int(H), C.c_longlong(stride), _p(out))
    return out


def mt_conversion(dark, I0, It, n, W, H, stride=None, dtype=np.float32):
    """mt_conversion (imageregistration_kernel_src.cl:161-223), the mu-t value."""
    dark, I0 = _arr(dark, dtype), _arr(I0, dtype)
    It = np.ascontiguousarray(It, dtype=np.uint16)
    stride = W * H if stride is None else stride
    out = np.zeros((n, H, W), dtype=dtype)
    _call("mt_conversion", dtype, _p(dark), _p(I0), _p(It), C.c_int(n), C.c_int(W), C.c_int(H),
          C.c_longlong(stride), _p(out))
    return out


def xprojection(mt, start_x, end_x, dtype=np.float32):
    """xprojection (reslice.cl:72-88): [A][H] from projections [A][H][W]."""
    mt = _arr(mt, dtype)
    A, H, W = mt.shape
    out = np.zeros((A, H), dtype=dtype)
    _call("xprojection", dtype, _p(mt), C.c_int(A), C.c_int(H), C.c_int(W), C.c_int(start_x), C.c_int(end_x),
          _p(out))
    return out


def zcorrection(xproj, start_z, end_z, yshift=None, dtype=np.float32):
    """zcorrection (reslice.cl:90-219): per-angle vertical drift."""
    xproj = _arr(xproj, dtype)
    A, H = xproj.shape
    ys = np.zeros(A, dtype=dtype) if yshift is None else _arr(yshift, dtype).copy()
    _call("zcorrection", dtype, _p(xproj), C.c_int(A), C.c_int(H), C.c_int(start_z), C.c_int(end_z), _p(ys))
    return ys


def reslice(mt, xshift=None, yshift=None, baseup=0.0, correction=0, y0=0, n_layers=None, dtype=np.float32):
    """reslice (reslice.cl:52-70): sinograms [n_layers][A][W] from projections [A][H][W]."""
    mt = _arr(mt, dtype)
    A, H, W = mt.shape
    n_layers = H - y0 if n_layers is None else n_layers
    xs = None if xshift is None else _arr(xshift, dtype)
    ys = None if yshift is None else _arr(yshift, dtype)
    out = np.zeros((n_layers, A, W), dtype=dtype)
    _call("reslice", dtype, _p(mt), C.c_int(A), C.c_int(H), C.c_int(W), _p(xs), _p(ys), C.c_float(baseup),
          C.c_int(correction), C.c_int(y0), C.c_int(n_layers), _p(out))
    return out


def denoise_sinogram(src, threshold, dtype=np.float32):
    """denoiseSinogram (reslice.cl:222-248) of sinograms [nz][A][W]."""
    src = _arr(src, dtype)
    nz, A, W = src.shape
    out = np.zeros_like(src)
    _call("denoise_sinogram", dtype, _p(src), _p(out), C.c_int(nz), C.c_int(A), C.c_int(W), C.c_float(threshold))
    return out


def reslice_execution(mt, baseup=0.0, z_correction=False, start_x=0, end_x=0, start_z=0, end_z=0, denoise=False,
                      threshold=0.0, y0=0, n_layers=None, dtype=np.float32):
    """reslice_thread (reslice_ocl.cpp:258-350); returns (sinograms, yshift)."""
    mt = _arr(mt, dtype)
    A, H, W = mt.shape
    n_layers = H - y0 if n_layers is None else n_layers
    ys = np.zeros(A, dtype=dtype)
    out = np.zeros((n_layers, A, W), dtype=dtype)
    _call("reslice_execution", dtype, _p(mt), C.c_int(A), C.c_int(H), C.c_int(W), C.c_float(baseup),
          C.c_int(int(z_correction)), C.c_int(start_x), C.c_int(end_x), C.c_int(start_z), C.c_int(end_z),
          C.c_int(int(denoise)), C.c_float(threshold), C.c_int(y0), C.c_int(n_layers), _p(ys), _p(out))
    return out, ys
