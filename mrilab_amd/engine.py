"""Python entry points over the HIP engine's C ABI.

The MATLAB scan functions take no arguments and mutate seven global
structs. Here the structs are passed explicitly, as mappings of numpy arrays
with MRiLab's field names and dtypes, and are mutated in place the same
way:

    DoScanAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig)

writes the final magnetization into VObj.Mx/My/Mz, adds the signal into
VSig.Sx/Sy and the k-space positions into VSig.Kx/Ky/Kz, and leaves the
decoder state (utsi, rfi, ..., TRCount, Kx/Ky/Kz) in VVar, as
Lib/src/engine/GPUEngine/DoScanAtGPU.cu:76 and
Lib/src/engine/CPUEngine/DoScanAtCPU.cpp:70 do. A failing call raises
MRSError carrying the library's message, where the MEX would call
mexErrMsgTxt.

The library is loaded from mrilab_amd/libmrilab_amd.so; there is no CPU
fallback: importing works without a GPU, calling raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from typing import Callable, Optional

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmrilab_amd.so")
_lib = None
_lock = threading.Lock()

STATUS = {0: "MRS_OK", 1: "MRS_ERR_ARG", 2: "MRS_ERR_DEVICE", 3: "MRS_ERR_CALLBACK",
          4: "MRS_ERR_UNSUPPORTED", 5: "MRS_ERR_NOMEM"}

# every symbol declared in include/mrilab_amd.h
EXPORTS = ("mrs_last_error", "mrs_abi_version", "mrs_device_count", "mrs_DoScanAtCPU",
           "mrs_DoScanAtGPU", "mrs_DoGMScanAtGPU", "mrs_DoMTScanAtGPU", "mrs_DoCESTScanAtGPU",
           "mrs_DoMEScanAtGPU", "mrs_plan_create", "mrs_plan_run", "mrs_plan_sync",
           "mrs_plan_fetch", "mrs_plan_get_info", "mrs_plan_destroy", "mrs_batch_create", "mrs_batch_run",
           "mrs_batch_sync", "mrs_batch_fetch", "mrs_batch_get_info", "mrs_batch_destroy",
           "mrs_DoScanBatchAtGPU", "mrs_adjoint_create", "mrs_adjoint_forward", "mrs_adjoint_backward",
           "mrs_adjoint_get_info", "mrs_adjoint_destroy")


class MRSError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def load() -> C.CDLL:
    """Loads libmrilab_amd.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mrilab_amd.build`")
        # PyTorch for ROCm ships its own libamdhip64 with the same SONAME as
        # ROCm's. Loaded first, it also serves this library (one HIP runtime
        # per process, so device pointers pass between the two); loaded after
        # ours, it would bring a second runtime that finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        P = C.POINTER
        scan_args = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC), P(abi.VSeqC),
                     P(abi.VVarC), P(abi.VSigC), P(abi.CallbacksC), P(abi.OptionsC)]
        for fn in ("mrs_DoScanAtCPU", "mrs_DoScanAtGPU", "mrs_DoGMScanAtGPU", "mrs_DoMTScanAtGPU",
                   "mrs_DoCESTScanAtGPU", "mrs_DoMEScanAtGPU"):
            getattr(lib, fn).argtypes = scan_args
            getattr(lib, fn).restype = C.c_int
        lib.mrs_last_error.restype = C.c_char_p
        lib.mrs_abi_version.restype = C.c_int32
        lib.mrs_device_count.restype = C.c_int32
        lib.mrs_plan_create.argtypes = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC),
                                        P(abi.VSeqC), P(abi.VVarC), P(abi.VSigC), P(abi.OptionsC),
                                        P(C.c_void_p)]
        lib.mrs_plan_create.restype = C.c_int
        for fn in ("mrs_plan_run", "mrs_plan_sync"):
            getattr(lib, fn).argtypes = [C.c_void_p]
            getattr(lib, fn).restype = C.c_int
        lib.mrs_plan_fetch.argtypes = [C.c_void_p, P(abi.VObjC), P(abi.VVarC), P(abi.VSigC)]
        lib.mrs_plan_fetch.restype = C.c_int
        lib.mrs_plan_get_info.argtypes = [C.c_void_p, P(abi.PlanInfoC)]
        lib.mrs_plan_get_info.restype = C.c_int
        lib.mrs_plan_destroy.argtypes = [C.c_void_p]
        lib.mrs_plan_destroy.restype = None
        PP = lambda t: P(P(t))  # noqa: E731  (array of struct pointers)
        batch_structs = [PP(abi.VObjC), PP(abi.VMagC), PP(abi.VCoiC), PP(abi.VCtlC), PP(abi.VSeqC),
                         PP(abi.VVarC), PP(abi.VSigC)]
        lib.mrs_batch_create.argtypes = [C.c_int32] + batch_structs + [P(abi.OptionsC), P(C.c_void_p)]
        lib.mrs_batch_create.restype = C.c_int
        lib.mrs_DoScanBatchAtGPU.argtypes = [C.c_int32] + batch_structs + [P(abi.OptionsC)]
        lib.mrs_DoScanBatchAtGPU.restype = C.c_int
        for fn in ("mrs_batch_run", "mrs_batch_sync"):
            getattr(lib, fn).argtypes = [C.c_void_p]
            getattr(lib, fn).restype = C.c_int
        lib.mrs_batch_fetch.argtypes = [C.c_void_p, PP(abi.VObjC), PP(abi.VVarC), PP(abi.VSigC)]
        lib.mrs_batch_fetch.restype = C.c_int
        lib.mrs_batch_get_info.argtypes = [C.c_void_p, P(abi.BatchInfoC)]
        lib.mrs_batch_get_info.restype = C.c_int
        lib.mrs_batch_destroy.argtypes = [C.c_void_p]
        lib.mrs_batch_destroy.restype = None
        lib.mrs_adjoint_create.argtypes = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.ExcitationC),
                                           P(abi.OptionsC), P(C.c_void_p)]
        lib.mrs_adjoint_create.restype = C.c_int
        lib.mrs_adjoint_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int32, P(C.c_double), P(C.c_double)]
        lib.mrs_adjoint_forward.restype = C.c_int
        lib.mrs_adjoint_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_int32]
        lib.mrs_adjoint_backward.restype = C.c_int
        lib.mrs_adjoint_get_info.argtypes = [C.c_void_p, P(abi.AdjointInfoC)]
        lib.mrs_adjoint_get_info.restype = C.c_int
        lib.mrs_adjoint_destroy.argtypes = [C.c_void_p]
        lib.mrs_adjoint_destroy.restype = None
        _lib = lib
        return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise MRSError(rc, load().mrs_last_error().decode(errors="replace"))


def make_options(block_size=0, spins_per_thread=0, max_adc_per_launch=0, multipool_scheme=0,
                 precision_mode=0, timing=0, model=0) -> abi.OptionsC:
    return abi.OptionsC(block_size=block_size, spins_per_thread=spins_per_thread,
                        max_adc_per_launch=max_adc_per_launch, multipool_scheme=multipool_scheme,
                        precision_mode=precision_mode, timing=timing, model=model)


def _callbacks(ext: Optional[Callable], gpu_fetch: Optional[Callable]):
    """Wraps Python callables as mrs_callbacks. ``ext(code, vvar_struct,
    vobj_struct) -> None`` may raise to signal an error; ``gpu_fetch(code)
    -> bool`` says whether the magnetization is synced to VObj first."""
    if ext is None and gpu_fetch is None:
        return None, ()

    def _ext(user, code, vvar, vobj):
        try:
            if ext is not None:
                ext(code, vvar.contents, vobj.contents)
            return 0
        except Exception:  # noqa: BLE001 - any error maps to MRS_ERR_CALLBACK
            return 1

    def _fetch(user, code, out):
        try:
            out[0] = int(bool(gpu_fetch(code))) if gpu_fetch is not None else 0
            return 0
        except Exception:  # noqa: BLE001
            return 1

    f_ext = abi.EXT_FN(_ext)
    f_fetch = abi.GPU_FETCH_FN(_fetch)
    cb = abi.CallbacksC(gpu_fetch=f_fetch, ext=f_ext, user=None)
    return cb, (f_ext, f_fetch)


# multi-pool engines: options.model and the planes of VObj.K in each layout
MODELS = {"GM": 1, "MT": 2, "CEST": 3, "ME": 4}


def k_planes(model: int, ty: int) -> int:
    return 2 * (ty - 1) if model == MODELS["CEST"] else ty * ty


def _scan(fn: str, VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None,
          options: Optional[abi.OptionsC] = None, multipool: bool = False, model: int = 1) -> None:
    lib = load()
    kp = k_planes(model, int(VObj["TypeNum"])) if multipool else None
    P = abi.pack_scan(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, multipool=multipool, k_planes=kp)
    cb, keep = _callbacks(ext, gpu_fetch)
    rc = getattr(lib, fn)(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi), C.byref(P.vctl),
                          C.byref(P.vseq), C.byref(P.vvar), C.byref(P.vsig),
                          C.byref(cb) if cb is not None else None,
                          C.byref(options) if options is not None else None)
    del keep
    _check(rc)
    abi.unpack_vvar(P, VVar)


def DoScanAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """Single-pool scan on the GPU (DoScanAtGPU.cu:76 contract)."""
    _scan("mrs_DoScanAtGPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options)


def DoGMScanAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """Multi-pool scan on the GPU (DoGMScanAtGPU.cu contract): VObj carries K
    [R,C,Z,Ty,Ty] (K(a,b) = rate a -> b) and TypeFlag (0 free, 1 bound)."""
    _scan("mrs_DoGMScanAtGPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options,
          multipool=True)


def DoMTScanAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """Two-pool MT scan (DoMTScanAtGPU.cu contract): VObj.K [R,C,Z,4] with
    K[...,1] water -> bound and K[...,2] bound -> water (0-based planes)."""
    _scan("mrs_DoMTScanAtGPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options,
          multipool=True, model=MODELS["MT"])


def DoCESTScanAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """CEST scan (DoCESTScanAtGPU.cu contract): VObj.K [R,C,Z,2*(Ty-1)]."""
    _scan("mrs_DoCESTScanAtGPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options,
          multipool=True, model=MODELS["CEST"])


def DoMEScanAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """Multi-exchange scan (DoMEScanAtGPU.cu contract): VObj.K [R,C,Z,Ty^2],
    plane a*Ty + b the rate from pool a into pool b."""
    _scan("mrs_DoMEScanAtGPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options,
          multipool=True, model=MODELS["ME"])


def DoScanAtCPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """The DoScanAtCPU entry point (DoScanAtCPU.cpp:70 contract), executed on the GPU."""
    _scan("mrs_DoScanAtCPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options)


class Plan:
    """A decoded scan resident on the GPU (mrs_plan_*): run it many times."""

    def __init__(self, VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, options=None):
        self._lib = load()
        mp = options is not None and options.model >= 1
        kp = k_planes(options.model, int(VObj["TypeNum"])) if mp else None
        self._packed = abi.pack_scan(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, multipool=mp, k_planes=kp)
        self._h = C.c_void_p()
        P = self._packed
        _check(self._lib.mrs_plan_create(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi),
                                         C.byref(P.vctl), C.byref(P.vseq), C.byref(P.vvar),
                                         C.byref(P.vsig),
                                         C.byref(options) if options is not None else None,
                                         C.byref(self._h)))

    def run(self) -> None:
        _check(self._lib.mrs_plan_run(self._h))

    def sync(self) -> None:
        _check(self._lib.mrs_plan_sync(self._h))

    def fetch(self, VObj=None, VVar=None, VSig=None) -> None:
        """Copies results into the (same-shaped) structs given."""
        P = self._packed
        if VObj is not None or VSig is not None or VVar is not None:
            Q = abi.pack_scan(VObj if VObj is not None else P.keep[0], P.keep[1], P.keep[2], P.keep[3],
                              P.keep[4], VVar if VVar is not None else P.keep[5],
                              VSig if VSig is not None else P.keep[6])
        else:
            Q = P
        _check(self._lib.mrs_plan_fetch(self._h, C.byref(Q.vobj) if VObj is not None else None,
                                        C.byref(Q.vvar) if VVar is not None else None,
                                        C.byref(Q.vsig) if VSig is not None else None))
        if VVar is not None:
            abi.unpack_vvar(Q, VVar)

    def info(self) -> dict:
        inf = abi.PlanInfoC()
        _check(self._lib.mrs_plan_get_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in abi.PlanInfoC._fields_}

    def close(self) -> None:
        if self._h:
            self._lib.mrs_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _ptr_array(ctype, items):
    arr = (C.POINTER(ctype) * len(items))()
    for i, it in enumerate(items):
        arr[i] = C.pointer(it) if it is not None else C.POINTER(ctype)()
    return arr


class Batch:
    """Many independent single-pool scans run as one persistent launch
    (mrs_batch_*): ``scans`` is a sequence of (VObj, VMag, VCoi, VCtl, VSeq,
    VVar, VSig) tuples. No reference counterpart; BatchSim.m:129-166 runs
    saved scans one after another through the engine."""

    def __init__(self, scans, options=None):
        self._lib = load()
        self._packed = [abi.pack_scan(*sc) for sc in scans]
        self.n = len(self._packed)
        P = self._packed
        self._args = [_ptr_array(abi.VObjC, [p.vobj for p in P]), _ptr_array(abi.VMagC, [p.vmag for p in P]),
                      _ptr_array(abi.VCoiC, [p.vcoi for p in P]), _ptr_array(abi.VCtlC, [p.vctl for p in P]),
                      _ptr_array(abi.VSeqC, [p.vseq for p in P]), _ptr_array(abi.VVarC, [p.vvar for p in P]),
                      _ptr_array(abi.VSigC, [p.vsig for p in P])]
        self._h = C.c_void_p()
        _check(self._lib.mrs_batch_create(C.c_int32(self.n), *self._args,
                                          C.byref(options) if options is not None else None,
                                          C.byref(self._h)))

    def run(self) -> None:
        _check(self._lib.mrs_batch_run(self._h))

    def sync(self) -> None:
        _check(self._lib.mrs_batch_sync(self._h))

    def fetch(self, scans=None, vobj=True, vvar=True, vsig=True) -> None:
        """Writes results into ``scans`` (default: the structs given at
        construction), each a (VObj, ..., VSig) tuple of the same shapes."""
        if scans is None:
            Q = self._packed
        else:
            Q = [abi.pack_scan(*sc) for sc in scans]
        a_obj = _ptr_array(abi.VObjC, [q.vobj for q in Q]) if vobj else None
        a_var = _ptr_array(abi.VVarC, [q.vvar for q in Q]) if vvar else None
        a_sig = _ptr_array(abi.VSigC, [q.vsig for q in Q]) if vsig else None
        _check(self._lib.mrs_batch_fetch(self._h, a_obj, a_var, a_sig))
        if vvar:
            src = scans if scans is not None else [p.keep for p in self._packed]
            for q, sc in zip(Q, src):
                abi.unpack_vvar(q, sc[5])

    def info(self) -> dict:
        inf = abi.BatchInfoC()
        _check(self._lib.mrs_batch_get_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in abi.BatchInfoC._fields_}

    def close(self) -> None:
        if self._h:
            self._lib.mrs_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def DoScanBatchAtGPU(scans, options=None) -> None:
    """Runs every scan of ``scans`` ((VObj, ..., VSig) tuples) in one launch
    and mutates each tuple's structs as DoScanAtGPU would."""
    with Batch(scans, options) as b:
        b.run()
        b.fetch()


def _vec_ptr(a):
    """(pointer, is_device, keep-alive) of a float32 numpy array or torch tensor."""
    if hasattr(a, "data_ptr"):  # torch tensor
        t = a.detach().contiguous()
        if t.dtype != __import__("torch").float32:
            raise TypeError("adjoint tensors must be float32")
        return C.c_void_p(t.data_ptr()), 1 if t.is_cuda else 0, t
    arr = np.ascontiguousarray(a, dtype=np.float32)
    return C.c_void_p(arr.ctypes.data), 0, arr


class Adjoint:
    """rf-design problem on the GPU (mrs_adjoint_*): forward loss and its
    gradients w.r.t. rf [n, 2], T1, T2 and dB0 [nvox]. No reference
    counterpart. ``exc`` is a workloads.Excitation (or anything with the same
    fields); parameters may be numpy arrays (host) or float32 torch tensors on
    the GPU (device pointers)."""

    def __init__(self, exc, checkpoint: int = 64, options=None, segment_mode: int = 0):
        """segment_mode 0 rebuilds checkpoint segments backwards where that
        is safe (no scratch memory), 1 always recomputes them into scratch."""
        self._lib = load()
        S = int(exc.VObj["SpinNum"])
        self._packed = abi.pack_scan(exc.VObj, exc.VMag, exc.VCoi, *exc.aux)
        self.n = int(exc.grad.shape[0])
        self.nvox = int(exc.target.shape[0])
        self.spins = S
        self._grad = np.ascontiguousarray(exc.grad, dtype=np.float32)
        self._target = np.ascontiguousarray(exc.target, dtype=np.float32)
        if self._grad.shape != (self.n, 3) or self._target.shape != (self.nvox, 2):
            raise ValueError("grad must be [n, 3] and target [nvox, 2]")
        ex = abi.ExcitationC(n_steps=self.n, checkpoint=int(checkpoint), segment_mode=int(segment_mode),
                             dt=float(exc.dt),
                             grad=self._grad.ctypes.data_as(C.POINTER(C.c_float)),
                             target=self._target.ctypes.data_as(C.POINTER(C.c_float)))
        self._h = C.c_void_p()
        P = self._packed
        _check(self._lib.mrs_adjoint_create(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi), C.byref(ex),
                                            C.byref(options) if options is not None else None,
                                            C.byref(self._h)))

    def forward(self, rf, T1, T2, dB0, want_mxy: bool = False):
        """Returns the loss (and the voxel sums [nvox, 2] if asked)."""
        ptrs = [_vec_ptr(a) for a in (rf, T1, T2, dB0)]
        dev = {p[1] for p in ptrs}
        if len(dev) != 1:
            raise ValueError("parameters must all be on the host or all on the GPU")
        if dev == {1}:
            import torch
            torch.cuda.synchronize()
        loss = C.c_double()
        mxy = np.zeros((self.nvox, 2)) if want_mxy else None
        _check(self._lib.mrs_adjoint_forward(self._h, ptrs[0][0], ptrs[1][0], ptrs[2][0], ptrs[3][0],
                                             C.c_int32(dev.pop()), C.byref(loss),
                                             mxy.ctypes.data_as(C.POINTER(C.c_double)) if want_mxy else None))
        del ptrs
        return (loss.value, mxy) if want_mxy else loss.value

    def backward(self, out=None):
        """Gradients (g_rf [n,2], g_T1, g_T2, g_dB0 [nvox]) of the last forward
        pass, as numpy arrays, or written into ``out`` (four float32 torch
        tensors on the GPU)."""
        if out is None:
            g = (np.zeros((self.n, 2), np.float32), np.zeros(self.nvox, np.float32),
                 np.zeros(self.nvox, np.float32), np.zeros(self.nvox, np.float32))
            _check(self._lib.mrs_adjoint_backward(self._h, *[C.c_void_p(a.ctypes.data) for a in g], C.c_int32(0)))
            return g
        _check(self._lib.mrs_adjoint_backward(self._h, *[C.c_void_p(t.data_ptr()) for t in out], C.c_int32(1)))
        return out

    def info(self) -> dict:
        inf = abi.AdjointInfoC()
        _check(self._lib.mrs_adjoint_get_info(self._h, C.byref(inf)))
        return {f: getattr(inf, f) for f, _ in abi.AdjointInfoC._fields_}

    def close(self) -> None:
        if self._h:
            self._lib.mrs_adjoint_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def device_count() -> int:
    return int(load().mrs_device_count())
