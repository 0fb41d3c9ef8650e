"""ctypes binding of the C ABI in include/ctr.h (libctr_hip.so).

torch is only plumbing here: device tensors provide device memory and the
stream; every computation runs in the hand-written gfx950 kernels of the
library.  There is no fallback path: if the library or a GPU is missing the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libctr_hip.so")

CTR_OK = 0
_STATUS = {0: "ok", 1: "invalid argument", 2: "invalid geometry", 3: "no device", 4: "out of memory",
           5: "HIP error", 6: "I/O error", 7: "not supported"}


class CtrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class Geometry(C.Structure):
    """ctr_geometry (include/ctr.h)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ox", C.c_int), ("oy", C.c_int), ("px", C.c_int),
                ("pa", C.c_int), ("ss", C.c_int), ("zp", C.c_int), ("mode", C.c_int),
                ("correction_mode", C.c_int), ("wt1", C.c_float), ("wt2", C.c_float), ("cs_lambda", C.c_float)]


class Slice(C.Structure):
    """ctr_slice (include/ctr.h)."""
    _fields_ = [("n", C.c_int), ("img", C.c_void_p), ("prj", C.c_void_p)]


class Schedule(C.Structure):
    """ctr_schedule (include/ctr.h)."""
    _fields_ = [("fp_rays", C.c_int), ("fp_waves", C.c_int), ("fp_tile", C.c_int), ("bp_rows", C.c_int),
                ("fp_ms", C.c_float), ("bp_ms", C.c_float), ("fp_ms_default", C.c_float),
                ("bp_ms_default", C.c_float)]

    def as_dict(self):
        return {f: (round(getattr(self, f), 4) if t is C.c_float else getattr(self, f)) for f, t in self._fields_}


class Layout(C.Structure):
    """ctr_layout (include/ctr.h)."""
    _fields_ = [("pad", C.c_int), ("ipitch", C.c_int), ("tpitch", C.c_int), ("rpitch", C.c_int), ("roff", C.c_int)]


class ResliceParams(C.Structure):
    """ctr_reslice_params (include/ctr.h)."""
    _fields_ = [("baseup", C.c_float), ("z_correction", C.c_int), ("start_x", C.c_int), ("end_x", C.c_int),
                ("start_z", C.c_int), ("end_z", C.c_int), ("denoise", C.c_int), ("denoise_threshold", C.c_float),
                ("y0", C.c_int), ("n_layers", C.c_int)]


_lib = None

_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_SIGS = {
    "ctr_create": [_I, C.POINTER(_P)],
    "ctr_destroy": [_P],
    "ctr_set_stream": [_P, _P],
    "ctr_hybrid_execution": [_P, _P, _P, _I, _I, _I, _I, _I],
    "ctr_get_stream": [_P, _P],
    "ctr_synchronize": [_P],
    "ctr_set_sync_each_kernel": [_P, _I],
    "ctr_set_geometry": [_P, C.POINTER(Geometry), _P, _P],
    "ctr_subset_order": [_I, _P],
    "ctr_query_layout": [_I, _I, _I, C.POINTER(Layout)],
    "ctr_osem_execution": [_P, _P, _P, _I, _I, _I],
    "ctr_osem_execution_noalloc": [_P, _P, _P, _I, _I, _I],
    "ctr_fista_osem_execution": [_P, _P, _P, _I, _I, _I],
    "ctr_aart_execution": [_P, _P, _P, _I, _I],
    "ctr_fbp_execution": [_P, _P, _P, _I, _I],
    "ctr_power_iteration": [_P, _P, _I],
    "ctr_fista_execution": [_P, _P, _P, _P, _I, _I, _I],
    "ctr_op_osem1": [_P, _P, _P, _P, _I, _I],
    "ctr_op_osem2": [_P, _P, _P, _P, _I, _I],
    "ctr_op_aart1": [_P, _P, _P, _P, _I, _I],
    "ctr_op_aart2": [_P, _P, _P, _P, _I, _I, _F],
    "ctr_op_backprojection": [_P, _P, _P, _I, _I],
    "ctr_op_sinogram_correction": [_P, _P, _P, _I, _I],
    "ctr_op_fbp_filter": [_P, _P, _P, _I],
    "ctr_op_fbp_backprojection": [_P, _P, _P, _I],
    "ctr_op_power_iter1": [_P, _P, _P],
    "ctr_op_power_iter2": [_P, _P, _P],
    "ctr_op_fista_backprojection": [_P, _P, _P, _P, _P, _I, _I, _F],
    "ctr_op_ista3d": [_P, _P, _P, _I, _I, _P],
    "ctr_op_fista3d": [_P, _P, _P, _P, _P, _P, _P, _I, _I, _P],
    "ctr_tv_prox_2d": [_P, _P, _P, _I, _F, _F],
    "ctr_osem_batched": [_P, C.POINTER(Slice), _I, _I, _I],
    "ctr_autotune": [_P, _I, C.POINTER(Schedule)],
    "ctr_set_schedule": [_P, _I, C.POINTER(Schedule)],
    "ctr_get_schedule": [_P, _I, C.POINTER(Schedule)],
    "ctr_set_kernel_timing": [_P, _I],
    "ctr_set_adaptive_rows": [_P, _I],
    "ctr_kernel_timing": [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_I), C.POINTER(_I)],
    "ctr_align_forward": [_P, _P, _P, _F, _P, _I],
    "ctr_align_forward_deriv": [_P, _P, _P, _F, _P, _P, _I],
    "ctr_align_adjoint": [_P, _P, _P, _F, _P, _I],
    "ctr_align_grad_shift": [_P, _P, _P, _P, _F, _P, _I],
    "ctr_merge_frames": [_P, _P, _I, _I, _I, C.c_longlong, _P],
    "ctr_mt_conversion": [_P, _P, _P, _P, _I, _I, _I, C.c_longlong, _P],
    "ctr_xprojection": [_P, _P, _I, _I, _I, _I, _I, _P],
    "ctr_zcorrection": [_P, _P, _I, _I, _I, _I, _P],
    "ctr_reslice": [_P, _P, _I, _I, _I, _P, _P, _F, _I, _I, _I, _P],
    "ctr_denoise_sinogram": [_P, _P, _P, _I, _I, _I, _F],
    "ctr_reslice_execution": [_P, _P, _I, _I, _I, C.POINTER(ResliceParams), _P, _P],
    "ctr_op_rot_center_shift": [_P, _P, _P, _I, _F, _F],
    "ctr_op_focus_index": [_P, _P, _I, _F, _F, _F, _F, _P],
    "ctr_rotation_center_search": [_P, _P, _I, _F, _F, _F, _F, _I, _F, _I, _P, _P],
}

EXPORTED = tuple(list(_SIGS) + ["ctr_last_error", "ctr_status_string"])


def load():
    """Loads libctr_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or `make -C ct-xafs_analysis_amd`")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.ctr_last_error.argtypes = [_P]
    lib.ctr_last_error.restype = C.c_char_p
    lib.ctr_status_string.argtypes = [C.c_int]
    lib.ctr_status_string.restype = C.c_char_p
    _lib = lib
    return lib


def subset_order(ss: int) -> np.ndarray:
    sub = np.zeros(ss, dtype=np.int32)
    st = load().ctr_subset_order(ss, sub.ctypes.data_as(_P))
    if st != CTR_OK:
        raise CtrError(st, "ctr_subset_order")
    return sub


def query_layout(nx: int, ny: int, px: int) -> Layout:
    """The device layout of one slice (host-only; works without a GPU)."""
    out = Layout()
    st = load().ctr_query_layout(nx, ny, px, C.byref(out))
    if st != CTR_OK:
        raise CtrError(st, f"ctr_query_layout({nx}, {ny}, {px})")
    return out


def _ptr(t):
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor on the GPU")
    if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
        raise ValueError("expected a contiguous float32 GPU tensor")
    return C.c_void_p(t.data_ptr())


def _ptr_u16(t):
    """16-bit camera frames: a contiguous uint16 (or int16 view) GPU tensor."""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype not in (torch.uint16, torch.int16) \
            or not t.is_contiguous():
        raise ValueError("expected a contiguous 16-bit GPU tensor")
    return C.c_void_p(t.data_ptr())



def _shape_of(t):
    return tuple(int(v) for v in t.shape)


def _need(t, shape, what):
    """Raises ValueError unless tensor t has exactly `shape` (checked before
    any pointer reaches a kernel: a wrong extent would make it read or write
    past the buffer)."""
    if t is None:
        raise ValueError(f"{what}: missing tensor")
    if _shape_of(t) != tuple(shape):
        raise ValueError(f"{what}: shape {_shape_of(t)} does not match the expected {tuple(shape)}")


def _need_numel(t, n, what, at_least=False):
    if t is None:
        raise ValueError(f"{what}: missing tensor")
    m = int(t.numel())
    if (m < n) if at_least else (m != n):
        raise ValueError(f"{what}: {m} elements, expected {'at least ' if at_least else ''}{n}")


def _stack(t, what):
    """The layer count of an image or sinogram stack [dN][.][.]."""
    if t is None or t.dim() != 3 or t.shape[0] < 1:
        raise ValueError(f"{what}: expected a 3-D stack [dN][rows][columns], got "
                         f"{None if t is None else _shape_of(t)}")
    return int(t.shape[0])


class Context:
    """One library context (device + stream).  Methods take float32 GPU tensors
    and run on the context stream, ordered after the work queued on torch's
    current stream and before the work queued there afterwards (as a torch
    operation would be).  synchronize() waits for the context stream.

    use_torch_stream=True runs the library on torch's current stream instead
    of a stream of its own.  That stream must not be torch's default stream:
    its handle is null, which ctr_set_stream takes to mean the context's own
    stream (ValueError)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = False):
        import torch
        self._torch = torch
        if use_torch_stream and torch.cuda.current_stream(device).cuda_stream == 0:
            raise ValueError("use_torch_stream: torch's current stream is the default stream, whose null handle "
                             "ctr_set_stream reads as the context's own stream; run under a torch.cuda.Stream "
                             "(calls are ordered with torch's stream either way)")
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.ctr_create(device, C.byref(h))
        if st != CTR_OK:
            raise CtrError(st, f"ctr_create(device={device})")
        self.h = h
        self.device = device
        self.geom = None
        self.angles = None
        self._ext = None
        if use_torch_stream:
            self.set_stream(torch.cuda.current_stream(device).cuda_stream)
        else:
            self._bind_stream()

    def _bind_stream(self):
        """Wraps the library stream so calls can be ordered with torch's."""
        s = C.c_void_p()
        self._chk(self.lib.ctr_get_stream(self.h, C.byref(s)), "ctr_get_stream")
        cur = self._torch.cuda.current_stream(self.device)
        self._ext = None if s.value == cur.cuda_stream else \
            self._torch.cuda.ExternalStream(s.value, device=self.device)

    def library_stream(self):
        """The torch stream object of the stream the library launches on."""
        return self._ext if self._ext is not None else self._torch.cuda.current_stream(self.device)

    def _run(self, fn, args, what):
        """Calls a library entry point in stream order with torch: the library
        stream first waits for the work already queued on torch's current
        stream (inputs written there), and torch's current stream then waits
        for the library stream (outputs read there)."""
        cur = self._torch.cuda.current_stream(self.device)
        if self._ext is not None:
            self._ext.wait_stream(cur)
        st = fn(*args)
        if self._ext is not None:
            cur.wait_stream(self._ext)
        self._chk(st, what)

    def _g(self):
        if self.geom is None:
            raise ValueError("set_geometry() first")
        return self.geom

    def _img(self, t, dN, what):
        g = self._g()
        _need(t, (dN, g.ny, g.nx), what)

    def _sino(self, t, dN, what):
        g = self._g()
        _need(t, (dN, g.pa, g.px), what)

    def _chk(self, st, what):
        if st != CTR_OK:
            msg = self.lib.ctr_last_error(self.h).decode()
            raise CtrError(st, f"{what}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ctr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        self._chk(self.lib.ctr_set_stream(self.h, C.c_void_p(stream_handle)), "ctr_set_stream")
        self._bind_stream()

    def synchronize(self):
        self._run(self.lib.ctr_synchronize, (self.h,), "ctr_synchronize")

    def set_sync_each_kernel(self, on: bool):
        self._run(self.lib.ctr_set_sync_each_kernel, (self.h, int(bool(on)),), "ctr_set_sync_each_kernel")

    def set_geometry(self, nx, px, pa, ss=1, *, ny=None, zp=None, mode=6, correction_mode=0, wt1=1.0, wt2=0.01,
                     cs_lambda=1e-4, angles_deg=None, sub=None, ox=None, oy=None):
        """ox, oy: the output region (g_ox, g_oy), the grid size by default;
        a smaller one is accepted by fbp_execution only."""
        ny = nx if ny is None else ny
        ox = nx if ox is None else ox
        oy = ny if oy is None else oy
        if zp is None:
            zp = 1
            while zp < 2 * px:
                zp *= 2
        g = Geometry(nx, ny, ox, oy, px, pa, ss, zp, mode, correction_mode, wt1, wt2, cs_lambda)
        ang = np.ascontiguousarray(angles_deg, dtype=np.float32)
        if ang.shape != (pa,):
            raise ValueError("angles_deg must hold pa values")
        subp = None
        if sub is not None:
            sub = np.ascontiguousarray(sub, dtype=np.int32)
            subp = sub.ctypes.data_as(_P)
        self._run(self.lib.ctr_set_geometry, (self.h, C.byref(g), ang.ctypes.data_as(_P), subp,), "ctr_set_geometry")
        self.geom = g
        self.angles = ang
        return g

    # ---- drivers ----
    def _driver_args(self, img, prj, what):
        dN = _stack(img, what + " image")
        self._img(img, dN, what + " image")
        self._sino(prj, dN, what + " sinogram")
        return dN

    def osem_execution(self, img, prj, it, prj_correction=True):
        dN = self._driver_args(img, prj, "osem_execution")
        self._run(self.lib.ctr_osem_execution, (self.h, _ptr(img), _ptr(prj), dN, it, int(prj_correction),),
                  "ctr_osem_execution")

    def osem_execution_noalloc(self, img, prj, it, prj_correction=True):
        """ctr_osem_execution_noalloc: the same results with the workspace in
        the library's per-device scratch array (no allocation)."""
        dN = self._driver_args(img, prj, "osem_execution_noalloc")
        self._run(self.lib.ctr_osem_execution_noalloc, (self.h, _ptr(img), _ptr(prj), dN, it, int(prj_correction),),
                  "ctr_osem_execution_noalloc")

    def fista_osem_execution(self, img, prj, it, prj_correction=True):
        dN = self._driver_args(img, prj, "fista_osem_execution")
        self._run(self.lib.ctr_fista_osem_execution, (self.h, _ptr(img), _ptr(prj), dN, it,
                                                    int(prj_correction),), "ctr_fista_osem_execution")

    def aart_execution(self, img, prj, it):
        dN = self._driver_args(img, prj, "aart_execution")
        self._run(self.lib.ctr_aart_execution, (self.h, _ptr(img), _ptr(prj), dN, it,), "ctr_aart_execution")

    def hybrid_execution(self, img, prj, it, prj_correction=True, baseup_order=4, cs=False):
        dN = self._driver_args(img, prj, "hybrid_execution")
        self._run(self.lib.ctr_hybrid_execution, (self.h, _ptr(img), _ptr(prj), dN, it, int(prj_correction),
                                                  baseup_order, int(bool(cs))), "ctr_hybrid_execution")

    def fbp_execution(self, img, prj, prj_correction=True):
        """img: the whole grid [dN][ny][nx] (an output region is the caller's
        top-left ox x oy block of it)."""
        dN = self._driver_args(img, prj, "fbp_execution")
        self._run(self.lib.ctr_fbp_execution, (self.h, _ptr(img), _ptr(prj), dN, int(prj_correction),),
                  "ctr_fbp_execution")

    # ---- sinogram formation (Reslice/) ----
    def merge_frames(self, raw, n, W, H, frame_stride=None):
        """merge_rawhisdata: the mean [H, W] of n 16-bit frames."""
        t = self._torch
        out = t.empty((H, W), dtype=t.float32, device=raw.device)
        stride = W * H if frame_stride is None else frame_stride
        _need_numel(raw, (n - 1) * stride + W * H, "merge_frames raw", at_least=True)
        self._run(self.lib.ctr_merge_frames, (self.h, _ptr_u16(raw), n, W, H, stride, _ptr(out)), "ctr_merge_frames")
        return out

    def mt_conversion(self, dark, I0, It, n, frame_stride=None):
        """mt_conversion: mu-t [n, H, W] of n 16-bit frames against dark / I0 [H, W]."""
        t = self._torch
        if dark.dim() != 2:
            raise ValueError(f"mt_conversion dark: expected [H, W], got {_shape_of(dark)}")
        H, W = dark.shape
        _need(I0, (H, W), "mt_conversion I0")
        stride = W * H if frame_stride is None else frame_stride
        _need_numel(It, (n - 1) * stride + W * H, "mt_conversion It", at_least=True)
        out = t.empty((n, H, W), dtype=t.float32, device=dark.device)
        self._run(self.lib.ctr_mt_conversion, (self.h, _ptr(dark), _ptr(I0), _ptr_u16(It), n, W, H, stride, _ptr(out)),
                  "ctr_mt_conversion")
        return out

    def xprojection(self, mt, start_x, end_x):
        t = self._torch
        A, H, W = _shape_of(mt) if mt.dim() == 3 else (None, None, None)
        if A is None:
            raise ValueError(f"xprojection: expected mt [A, H, W], got {_shape_of(mt)}")
        out = t.empty((A, H), dtype=t.float32, device=mt.device)
        self._run(self.lib.ctr_xprojection, (self.h, _ptr(mt), A, H, W, start_x, end_x, _ptr(out)), "ctr_xprojection")
        return out

    def zcorrection(self, xproj, start_z, end_z, yshift=None):
        t = self._torch
        if xproj.dim() != 2:
            raise ValueError(f"zcorrection: expected xproj [A, H], got {_shape_of(xproj)}")
        A, H = xproj.shape
        ys = t.zeros((A,), dtype=t.float32, device=xproj.device) if yshift is None else yshift
        _need_numel(ys, A, "zcorrection yshift")
        self._run(self.lib.ctr_zcorrection, (self.h, _ptr(xproj), A, H, start_z, end_z, _ptr(ys)), "ctr_zcorrection")
        return ys

    def reslice(self, mt, xshift=None, yshift=None, baseup=0.0, correction=0, y0=0, n_layers=None, out=None):
        """Sinograms [n_layers, A, W] (camera rows y0 ..) from projections [A, H, W]."""
        t = self._torch
        if mt.dim() != 3:
            raise ValueError(f"reslice: expected mt [A, H, W], got {_shape_of(mt)}")
        A, H, W = mt.shape
        n_layers = H - y0 if n_layers is None else n_layers
        for sh, nm in ((xshift, "xshift"), (yshift, "yshift")):
            if sh is not None:
                _need_numel(sh, A, "reslice " + nm)
        if out is None:
            out = t.empty((n_layers, A, W), dtype=t.float32, device=mt.device)
        _need(out, (n_layers, A, W), "reslice out")
        self._run(self.lib.ctr_reslice, (self.h, _ptr(mt), A, H, W, _ptr(xshift), _ptr(yshift), float(baseup),
                                         int(correction), y0, n_layers, _ptr(out)), "ctr_reslice")
        return out

    def denoise_sinogram(self, src, threshold):
        if src.dim() != 3:
            raise ValueError(f"denoise_sinogram: expected [nz, A, W], got {_shape_of(src)}")
        out = self._torch.empty_like(src)
        nz, A, W = src.shape
        self._run(self.lib.ctr_denoise_sinogram, (self.h, _ptr(src), _ptr(out), nz, A, W, float(threshold)),
                  "ctr_denoise_sinogram")
        return out

    def reslice_execution(self, mt, baseup=0.0, z_correction=False, start_x=0, end_x=0, start_z=0, end_z=0,
                          denoise=False, threshold=0.0, y0=0, n_layers=None, out=None):
        """reslice_thread: [Z-correction] -> reslice -> [spike removal].
        Returns (sinograms [n_layers, A, W], yshift [A])."""
        t = self._torch
        if mt.dim() != 3:
            raise ValueError(f"reslice_execution: expected mt [A, H, W], got {_shape_of(mt)}")
        A, H, W = mt.shape
        n_layers = H - y0 if n_layers is None else n_layers
        p = ResliceParams(baseup, int(bool(z_correction)), start_x, end_x, start_z, end_z, int(bool(denoise)),
                          threshold, y0, n_layers)
        ys = t.empty((A,), dtype=t.float32, device=mt.device)
        if out is None:
            out = t.empty((n_layers, A, W), dtype=t.float32, device=mt.device)
        _need(out, (n_layers, A, W), "reslice_execution out")
        self._run(self.lib.ctr_reslice_execution, (self.h, _ptr(mt), A, H, W, C.byref(p), _ptr(ys), _ptr(out)),
                  "ctr_reslice_execution")
        return out, ys

    def rotation_center_search(self, prj, n_shift, shift_start, shift_step, it=1, init=1.0, ang_range=None,
                               prj_correction=True):
        """Rotation-centre search over one layer's sinogram prj [pa, px]
        (rotationCenterSearch_ocl.cpp:489-551).  ang_range defaults to the
        first and last angle of the geometry.  Returns (recon [n_shift, ny, nx],
        findex [n_shift], best shift)."""
        t = self._torch
        g = self._g()
        _need_numel(prj, g.pa * g.px, "rotation_center_search prj")
        if ang_range is None:
            ang_range = (float(self.angles[0]), float(self.angles[-1]))
        recon = t.empty((n_shift, g.ny, g.nx), dtype=t.float32, device=prj.device)
        findex = t.empty((n_shift,), dtype=t.float32, device=prj.device)
        self._run(self.lib.ctr_rotation_center_search,
                  (self.h, _ptr(prj), n_shift, shift_start, shift_step, ang_range[0], ang_range[1], it, init,
                   int(prj_correction), _ptr(recon), _ptr(findex)), "ctr_rotation_center_search")
        f = findex.cpu().numpy()
        best = int(np.argmin(f)) if np.isfinite(f).any() else 0
        return recon, findex, np.float32(shift_start) + np.float32(shift_step) * np.float32(best)

    def power_iteration(self, L, trials=30):
        _need_numel(L, self._g().ss, "power_iteration L2norm", at_least=True)
        self._run(self.lib.ctr_power_iteration, (self.h, _ptr(L), trials,), "ctr_power_iteration")

    def fista_execution(self, L, img, prj, it, prj_correction=True):
        dN = self._driver_args(img, prj, "fista_execution")
        _need_numel(L, self._g().ss, "fista_execution L2norm", at_least=True)
        self._run(self.lib.ctr_fista_execution, (self.h, _ptr(L), _ptr(img), _ptr(prj), dN, it,
                                               int(prj_correction),), "ctr_fista_execution")

    # ---- single operators ----
    def _op_shapes(self, name, args):
        """(tensor, expected shape or element count, label) for every tensor
        argument of ctr_op_<name> (include/ctr.h), from the geometry and the
        layer count argument."""
        g = self._g()
        im = lambda n: (n, g.ny, g.nx)            # noqa: E731
        sn = lambda n: (n, g.pa, g.px)            # noqa: E731
        L = ("L", g.ss)
        table = {
            "osem1": (3, [im, sn, sn]), "aart1": (3, [im, sn, sn]),
            "osem2": (3, [im, sn, im]), "aart2": (3, [im, sn, im]),
            "backprojection": (2, [sn, im]),
            "sinogram_correction": (2, [sn, sn]),
            "fbp_filter": (2, [sn, sn]), "fbp_backprojection": (2, [sn, im]),
            "fista_backprojection": (4, [im, sn, L, im]),
            "ista3d": (2, [im, im, None, None, L]),
            "fista3d": (6, [im, im, im, im, im, im, None, None, L]),
        }
        if name in ("power_iter1", "power_iter2"):
            specs = [im, sn] if name == "power_iter1" else [sn, im]
            return [(args[i], specs[i](g.ss), f"op {name} argument {i}") for i in range(2)]
        if name == "rot_center_shift":
            return [(args[0], g.pa * g.px, "op rot_center_shift prj"),
                    (args[1], int(args[2]) * g.pa * g.px, "op rot_center_shift out")]
        if name == "focus_index":
            n = int(args[1])
            return [(args[0], n * g.nx * g.ny, "op focus_index img"), (args[6], n, "op focus_index findex")]
        if name not in table:
            raise ValueError(f"unknown operator {name}")
        dn_at, specs = table[name]
        dN = int(args[dn_at])
        if dN < 1:
            raise ValueError(f"op {name}: dN must be >= 1")
        out = []
        for i, sp in enumerate(specs):
            if sp is None:
                continue
            if isinstance(sp, tuple):             # L2norm: at least ss entries
                out.append((args[i], ("min", sp[1]), f"op {name} L2norm"))
            elif args[i] is not None or name != "fista3d" or i < 3:
                out.append((args[i], sp(dN), f"op {name} argument {i}"))
        return out

    def op(self, name, *args):
        for t, want, what in self._op_shapes(name, args):
            if isinstance(want, tuple) and len(want) == 2 and want[0] == "min":
                _need_numel(t, want[1], what, at_least=True)
            elif isinstance(want, tuple):
                _need(t, want, what)
            else:
                _need_numel(t, want, what)
        f = getattr(self.lib, "ctr_op_" + name)
        conv = [(_ptr(a) if (a is None or isinstance(a, self._torch.Tensor)) else a) for a in args]
        self._run(f, (self.h, *conv), "ctr_op_" + name)

    def tv_prox_2d(self, v, x_new, lam, lower=0.0):
        """Per-slice TV prox with four in-plane neighbours (ctr_tv_prox_2d):
        x_new = median of the neighbours and v - 4 lam .. v + 4 lam, then
        max(lower, x).  v, x_new: distinct [dN][ny][nx] stacks."""
        dN = _stack(v, "tv_prox_2d v")
        self._img(v, dN, "tv_prox_2d v")
        self._img(x_new, dN, "tv_prox_2d x_new")
        self._run(self.lib.ctr_tv_prox_2d, (self.h, _ptr(v), _ptr(x_new), dN, float(lam), float(lower)),
                  "ctr_tv_prox_2d")

    # ---- batched ----
    def osem_batched(self, imgs, prjs, it, use_graph=True):
        g = self._g()
        if len(imgs) != len(prjs) or not imgs:
            raise ValueError("osem_batched: need as many images as sinograms, at least one")
        n = len(imgs)
        arr = (Slice * n)()
        for i, (im, pr) in enumerate(zip(imgs, prjs)):
            m = int(im.shape[-1])
            if im.dim() not in (2, 3) or _shape_of(im)[-2:] != (m, m) or int(im.numel()) != m * m:
                raise ValueError(f"osem_batched slice {i}: image shape {_shape_of(im)} is not one square slice")
            _need_numel(pr, g.pa * m, f"osem_batched slice {i} sinogram ([pa][{m}])")
            arr[i] = Slice(m, _ptr(im).value, _ptr(pr).value)
        self._keep = (imgs, prjs, arr)
        self._run(self.lib.ctr_osem_batched, (self.h, arr, n, it, int(use_graph),), "ctr_osem_batched")

    def autotune(self, n):
        """Times the schedules for side n and keeps the fastest.  Returns a dict:
        fp_rays (0 = LDS-tiled projector), fp_waves, fp_tile (samples per LDS
        tile), bp_rows, and the chosen and default candidates' launch times."""
        out = Schedule()
        self._run(self.lib.ctr_autotune, (self.h, n, C.byref(out)), "ctr_autotune")
        return out.as_dict()

    def set_schedule(self, n, fp_rays, fp_waves, bp_rows, fp_tile=60):
        sch = Schedule(fp_rays, fp_waves, fp_tile, bp_rows, 0.0, 0.0, 0.0, 0.0)
        self._run(self.lib.ctr_set_schedule, (self.h, n, C.byref(sch)), "ctr_set_schedule")

    def get_schedule(self, n):
        out = Schedule()
        self._run(self.lib.ctr_get_schedule, (self.h, n, C.byref(out)), "ctr_get_schedule")
        d = out.as_dict()
        return {k: d[k] for k in ("fp_rays", "fp_waves", "fp_tile", "bp_rows")}

    def set_adaptive_rows(self, on):
        """on (the default): stacks too small to give the device 4
        backprojection blocks per compute unit run with fewer rows per thread
        than their side's schedule; off: every stack runs its schedule as set.
        The images are the same either way."""
        self._run(self.lib.ctr_set_adaptive_rows, (self.h, int(bool(on)),), "ctr_set_adaptive_rows")

    def set_kernel_timing(self, on):
        self._run(self.lib.ctr_set_kernel_timing, (self.h, int(bool(on)),), "ctr_set_kernel_timing")

    def kernel_timing(self):
        a, b, na, nb = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self._run(self.lib.ctr_kernel_timing, (self.h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb),),
                  "ctr_kernel_timing")
        return a.value, b.value, na.value, nb.value

    # ---- differentiable model ----
    def _align_args(self, img, prj, shifts, what):
        dN = _stack(img, what + " image")
        self._img(img, dN, what + " image")
        self._sino(prj, dN, what + " projections")
        if shifts is not None:
            _need_numel(shifts, self._g().pa, what + " shifts")
        return dN

    def align_forward(self, img, shifts, center, prj):
        dN = self._align_args(img, prj, shifts, "align_forward")
        self._run(self.lib.ctr_align_forward, (self.h, _ptr(img), _ptr(shifts), float(center), _ptr(prj), dN,),
                  "ctr_align_forward")

    def align_forward_deriv(self, img, shifts, center, prj, dprj):
        """align_forward, also writing dprj = d prj / d shifts[a] per ray."""
        dN = self._align_args(img, prj, shifts, "align_forward_deriv")
        self._sino(dprj, dN, "align_forward_deriv derivative rows")
        self._run(self.lib.ctr_align_forward_deriv, (self.h, _ptr(img), _ptr(shifts), float(center), _ptr(prj),
                                                   _ptr(dprj), dN,), "ctr_align_forward_deriv")

    def align_adjoint(self, prj, shifts, center, img):
        dN = self._align_args(img, prj, shifts, "align_adjoint")
        self._run(self.lib.ctr_align_adjoint, (self.h, _ptr(prj), _ptr(shifts), float(center), _ptr(img), dN,),
                  "ctr_align_adjoint")

    def align_grad_shift(self, img, gprj, shifts, center, gshift):
        dN = self._align_args(img, gprj, shifts, "align_grad_shift")
        _need_numel(gshift, self._g().pa, "align_grad_shift gshift")
        self._run(self.lib.ctr_align_grad_shift, (self.h, _ptr(img), _ptr(gprj), _ptr(shifts), float(center),
                                                _ptr(gshift), dN,), "ctr_align_grad_shift")
