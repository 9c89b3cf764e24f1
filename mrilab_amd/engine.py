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
           "mrs_adjoint_get_info", "mrs_adjoint_destroy", "mrs_recon_dims", "mrs_DoCartRecon", "mrs_DoEchoTrainRecon",
           "mrs_plan_recon", "mrs_DoAddNoise", "mrs_noise_scale", "mrs_DoCalSARAtGPU",
           "mrs_DoSARAverageAtGPU", "mrs_seq_compile", "mrs_seq_counts", "mrs_seq_copy", "mrs_seq_destroy",
           "mrs_seq_check", "mrs_multipool_batch_create", "mrs_multipool_batch_fetch", "mrs_DoMultiPoolBatchAtGPU",
           "mrs_DoSARAverageAtGPU_noalloc")


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
        lib.mrs_recon_dims.argtypes = [P(abi.ReconDescC), P(C.c_int64)]
        lib.mrs_recon_dims.restype = C.c_int
        lib.mrs_DoCartRecon.argtypes = [P(abi.VSigC), P(abi.ReconDescC), P(abi.VImgC), C.c_int32]
        lib.mrs_DoCartRecon.restype = C.c_int
        lib.mrs_DoEchoTrainRecon.argtypes = [P(abi.VSigC), P(abi.ReconDescC), C.c_int32, C.c_int32, C.c_int32,
                                             P(abi.VImgC), C.c_int32]
        lib.mrs_DoEchoTrainRecon.restype = C.c_int
        lib.mrs_plan_recon.argtypes = [C.c_void_p, P(abi.ReconDescC), P(abi.VImgC), C.c_int32]
        lib.mrs_plan_recon.restype = C.c_int
        lib.mrs_DoAddNoise.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_uint64, C.c_int32]
        lib.mrs_DoAddNoise.restype = C.c_int
        lib.mrs_noise_scale.argtypes = [C.c_double] * 10
        lib.mrs_noise_scale.restype = C.c_double
        lib.mrs_DoCalSARAtGPU.argtypes = [P(abi.SarObjC), P(abi.VSeqC), P(abi.VVarC), P(abi.SarSigC)]
        lib.mrs_DoCalSARAtGPU.restype = C.c_int
        lib.mrs_DoSARAverageAtGPU.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(C.c_int64), C.c_double,
                                              C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                              C.c_int32]
        lib.mrs_DoSARAverageAtGPU.restype = C.c_int
        lib.mrs_DoSARAverageAtGPU_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, P(C.c_int64), C.c_double,
                                                      C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                                      C.c_void_p]
        lib.mrs_DoSARAverageAtGPU_noalloc.restype = C.c_int
        lib.mrs_seq_compile.argtypes = [C.c_char_p, P(abi.SeqParamC), C.c_int32, C.c_double, P(C.c_void_p)]
        lib.mrs_seq_compile.restype = C.c_int
        lib.mrs_seq_counts.argtypes = [C.c_void_p, P(C.c_int64)]
        lib.mrs_seq_counts.restype = C.c_int
        lib.mrs_seq_copy.argtypes = [C.c_void_p, P(abi.SeqLinesC)]
        lib.mrs_seq_copy.restype = C.c_int
        lib.mrs_seq_destroy.argtypes = [C.c_void_p]
        lib.mrs_seq_destroy.restype = None
        lib.mrs_seq_check.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32]
        lib.mrs_seq_check.restype = C.c_int
        one = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC)]
        lib.mrs_multipool_batch_create.argtypes = ([C.c_int32] + one + [PP(abi.VSeqC), PP(abi.VVarC), PP(abi.VSigC),
                                                    P(abi.OptionsC), P(C.c_void_p)])
        lib.mrs_multipool_batch_create.restype = C.c_int
        lib.mrs_multipool_batch_fetch.argtypes = [C.c_void_p, PP(abi.VObjC), PP(abi.VVarC), PP(abi.VSigC)]
        lib.mrs_multipool_batch_fetch.restype = C.c_int
        lib.mrs_DoMultiPoolBatchAtGPU.argtypes = ([C.c_int32] + one + [PP(abi.VSeqC), PP(abi.VVarC), PP(abi.VSigC),
                                                  PP(abi.VObjC), P(abi.OptionsC)])
        lib.mrs_DoMultiPoolBatchAtGPU.restype = C.c_int
        _lib = lib
        return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise MRSError(rc, load().mrs_last_error().decode(errors="replace"))


def make_options(block_size=0, spins_per_thread=0, max_adc_per_launch=0, multipool_scheme=0,
                 precision_mode=0, timing=0, model=0, coil_sum=0, free_steps=0) -> abi.OptionsC:
    return abi.OptionsC(block_size=block_size, spins_per_thread=spins_per_thread,
                        max_adc_per_launch=max_adc_per_launch, multipool_scheme=multipool_scheme,
                        precision_mode=precision_mode, timing=timing, model=model, coil_sum=coil_sum,
                        free_steps=free_steps)


def make_callbacks(ext: Optional[Callable] = None, gpu_fetch: Optional[Callable] = None):
    """Wraps Python callables as mrs_callbacks. ``ext(code, vvar, state)``
    gets the VVarC struct (edits steer the rest of the scan) and the
    ExtStateC struct (``state.fetched`` says whether VObj.Mx/My/Mz hold the
    current magnetization; VSig holds the samples acquired so far); it may
    edit the scan's numpy arrays in place, which the engine reads again when
    the state was fetched, and may raise to signal an error. ``gpu_fetch(code)
    -> bool`` decides the fetch on the GPU entry points (DoGPUFetch.m); without
    it codes 6 and 8 fetch."""
    if ext is None and gpu_fetch is None:
        return None, ()

    def _ext(user, code, vvar, state):
        try:
            if ext is not None:
                ext(code, vvar.contents, state.contents)
            return 0
        except Exception:  # noqa: BLE001 - any error maps to MRS_ERR_CALLBACK
            return 1

    def _fetch(user, code, out):
        try:
            out[0] = int(bool(gpu_fetch(code)))
            return 0
        except Exception:  # noqa: BLE001
            return 1

    f_ext = abi.EXT_FN(_ext)
    f_fetch = abi.GPU_FETCH_FN(_fetch) if gpu_fetch is not None else abi.GPU_FETCH_FN()
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
    cb, keep = make_callbacks(ext, gpu_fetch)
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

    def recon(self, desc, device: bool = False) -> dict:
        """Cartesian reconstruction of the k-space this plan's last run left
        on the GPU (mrs_plan_recon): numpy arrays, or float64 CUDA tensors
        when ``device``."""
        return _recon_call(lambda img, dev: self._lib.mrs_plan_recon(self._h, C.byref(desc), C.byref(img), dev),
                           desc, device)

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


class MultiPoolBatch:
    """One phantom under several independent sequences on the GPU
    (mrs_multipool_batch_*): ``scans`` share VObj/VMag/VCoi/VCtl (those of
    scans[0] are used) and each brings its VSeq/VVar/VSig, e.g. the offsets of
    workloads.mt_cest(per_offset=True). ``model`` is "GM", "MT", "CEST" or
    "ME" (VObj.K in that engine's layout). No reference counterpart."""

    def __init__(self, scans, model: str = "GM", options=None):
        self._lib = load()
        self.scans = list(scans)
        m = MODELS[model]
        opt = abi.OptionsC.from_buffer_copy(options if options is not None else make_options())
        opt.model = m
        self._kp = k_planes(m, int(self.scans[0].VObj["TypeNum"]))
        self._packs = [abi.pack_scan(*sc.structs, multipool=True, k_planes=self._kp) for sc in self.scans]
        P0 = self._packs[0]
        self._args = [_ptr_array(abi.VSeqC, [p.vseq for p in self._packs]),
                      _ptr_array(abi.VVarC, [p.vvar for p in self._packs]),
                      _ptr_array(abi.VSigC, [p.vsig for p in self._packs])]
        self._h = C.c_void_p()
        _check(self._lib.mrs_multipool_batch_create(C.c_int32(len(self._packs)), C.byref(P0.vobj), C.byref(P0.vmag),
                                                    C.byref(P0.vcoi), C.byref(P0.vctl), *self._args,
                                                    C.byref(opt), C.byref(self._h)))

    def run(self) -> None:
        _check(self._lib.mrs_plan_run(self._h))

    def sync(self) -> None:
        _check(self._lib.mrs_plan_sync(self._h))

    def fetch(self, out_scans) -> None:
        """Sequence q's final magnetization, signal and decoder state go into
        out_scans[q] (Scans of the same shapes, each with its own VObj
        magnetization arrays)."""
        packs = [abi.pack_scan(*sc.structs, multipool=True, k_planes=self._kp) for sc in out_scans]
        a_obj = _ptr_array(abi.VObjC, [p.vobj for p in packs])
        a_var = _ptr_array(abi.VVarC, [p.vvar for p in packs])
        a_sig = _ptr_array(abi.VSigC, [p.vsig for p in packs])
        _check(self._lib.mrs_multipool_batch_fetch(self._h, a_obj, a_var, a_sig))
        for p, sc in zip(packs, out_scans):
            abi.unpack_vvar(p, sc.VVar)

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
        self.generation = 0   # forward passes run; the backward differentiates the last
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
        self.generation += 1
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


# ---------------------------------------------------------------------------
# Signal to image: Cartesian reconstruction and k-space noise
# ---------------------------------------------------------------------------
def swap_xy(scan_plane: str, freq_dir: str) -> bool:
    """Whether the scan plane and frequency direction select the second
    permute of the reconstruction (DoCartRecon.m:45-73)."""
    table = {"Axial": freq_dir == "A/P", "Sagittal": freq_dir == "S/I", "Coronal": freq_dir != "L/R"}
    if scan_plane not in table:
        raise ValueError(f"unknown ScanPlane {scan_plane!r}")
    return table[scan_plane]


def recon_desc(res_freq: int, res_phase: int, slice_num: int = 1, te_per_tr: int = 1, rx: int = 1, types: int = 1,
               zf=None, scan_plane: str = "Axial", freq_dir: str = "R/L") -> abi.ReconDescC:
    """mrs_recon_desc from VCtl-style values; ``zf`` = (ZF_Kx, ZF_Ky, ZF_Kz
    factor) or None."""
    zx, zy, zz = zf if zf is not None else (0, 0, 0)
    return abi.ReconDescC(res_freq=res_freq, res_phase=res_phase, slice_num=slice_num, te_per_tr=te_per_tr, rx=rx,
                          types=types, zf_kx=zx, zf_ky=zy, zf_kz=zz,
                          swap_xy=1 if swap_xy(scan_plane, freq_dir) else 0)


def recon_dims(desc) -> tuple:
    dims = (C.c_int64 * 5)()
    _check(load().mrs_recon_dims(C.byref(desc), dims))
    return tuple(int(d) for d in dims)


_IMG_FIELDS = ("Real", "Imag", "Mag", "Phase", "Sx", "Sy")


def _recon_call(call, desc, device: bool) -> dict:
    dims = recon_dims(desc)
    n = int(np.prod(dims))
    if device:
        import torch
        out = {k: torch.empty(n, dtype=torch.float64, device="cuda") for k in _IMG_FIELDS}
        img = abi.VImgC(**{k: C.c_void_p(v.data_ptr()) for k, v in out.items()})
        torch.cuda.synchronize()
        _check(call(img, 1))
        # column-major [d0, d1, d2, Rx, TEPerTR]: reverse the dims for torch
        return {k: v.view(*dims[::-1]) for k, v in out.items()}
    out = {k: np.empty(n, dtype=np.float64) for k in _IMG_FIELDS}
    img = abi.VImgC(**{k: C.c_void_p(v.ctypes.data) for k, v in out.items()})
    _check(call(img, 0))
    return {k: v.reshape(dims, order="F") for k, v in out.items()}


def DoCartRecon(Sx, Sy, desc, device: bool = False, echo_train=None) -> dict:
    """VImg from VSig.Sx/Sy (numpy float64 arrays, or float64 CUDA tensors
    with ``device``); arrays [d0, d1, d2, Rx, TEPerTR] in MATLAB order
    (numpy order "F"; torch tensors with the dims reversed).
    ``echo_train=(etl, shots, epi)`` reconstructs an EPI (epi True) or FSE
    scan as DoEPIRecon/DoFSERecon do (mrs_DoEchoTrainRecon)."""
    lib = load()
    if device:
        import torch
        sx, sy = Sx.contiguous(), Sy.contiguous()
        if sx.dtype != torch.float64 or not sx.is_cuda:
            raise TypeError("device recon takes float64 CUDA tensors")
        n = sx.numel() // (desc.rx * desc.types)
        vs = abi.VSigC(Sx=sx.data_ptr(), Sy=sy.data_ptr(), SignalNum=n)
        keep = (sx, sy)
    else:
        sx = np.ascontiguousarray(np.asarray(Sx, dtype=np.float64).ravel(order="F"))
        sy = np.ascontiguousarray(np.asarray(Sy, dtype=np.float64).ravel(order="F"))
        n = sx.size // (desc.rx * desc.types)
        vs = abi.VSigC(Sx=sx.ctypes.data, Sy=sy.ctypes.data, SignalNum=n)
        keep = (sx, sy)
    if echo_train is None:
        call = lambda img, dev: lib.mrs_DoCartRecon(C.byref(vs), C.byref(desc), C.byref(img), dev)
    else:
        etl, shots, epi = echo_train
        call = lambda img, dev: lib.mrs_DoEchoTrainRecon(C.byref(vs), C.byref(desc), int(etl), int(shots),
                                                         1 if epi else 0, C.byref(img), dev)
    out = _recon_call(call, desc, device)
    del keep
    return out


def DoAddNoise(Sx, Sy, scale: float, seed: int) -> None:
    """Adds scale * N(0, 1) to Sx and Sy in place (float64 numpy arrays or
    CUDA tensors of equal size)."""
    lib = load()
    if hasattr(Sx, "data_ptr"):
        if not (Sx.is_contiguous() and Sy.is_contiguous()) or Sx.numel() != Sy.numel():
            raise ValueError("Sx and Sy must be contiguous tensors of equal size")
        _check(lib.mrs_DoAddNoise(C.c_void_p(Sx.data_ptr()), C.c_void_p(Sy.data_ptr()), Sx.numel(), float(scale),
                                  int(seed), 1 if Sx.is_cuda else 0))
        return
    for a in (Sx, Sy):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and (a.flags.c_contiguous or a.flags.f_contiguous)):
            raise TypeError("Sx and Sy must be contiguous float64 numpy arrays")
    if Sx.size != Sy.size:
        raise ValueError("Sx and Sy must have equal sizes")
    _check(lib.mrs_DoAddNoise(C.c_void_p(Sx.ctypes.data), C.c_void_p(Sy.ctypes.data), Sx.size, float(scale),
                              int(seed), 0))


def noise_scale(bandwidth, noise_level, nex, b0, rfreq, rphase, rslice, res_freq, res_phase, slice_num) -> float:
    """NoiseScale of DoAddNoise.m:8-16 (VCtl fields)."""
    return float(load().mrs_noise_scale(*[float(x) for x in (bandwidth, noise_level, nex, b0, rfreq, rphase,
                                                              rslice, res_freq, res_phase, slice_num)]))


# ---------------------------------------------------------------------------
# Local SAR
# ---------------------------------------------------------------------------
def DoCalSARAtGPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig) -> None:
    """Local SAR of the sequence (DoCalSARAtCPU.cpp semantics) on the GPU, in
    place: VVar.SAR/SARi, the decoder fields of VVar and VSig.SAR, Muts and
    tRealSample are written. The structs carry the SAR fields
    (VObj.MassDen/ECon, VCoi.TxE1x/y/z, VVar.SAR/SARi, VSig.SAR/tSample/
    tRealSample/Muts; see workloads.sar_scan)."""
    lib = load()
    P = abi.pack_scan(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig)
    obj, sig, keep = abi.pack_sar(VObj, VCoi, VCtl, VVar, VSig, int(P.vseq.n_uts))
    _check(lib.mrs_DoCalSARAtGPU(C.byref(obj), C.byref(P.vseq), C.byref(P.vvar), C.byref(sig)))
    abi.unpack_vvar(P, VVar)
    VVar["SARi"] = np.int32(sig.SARi)
    VSig["SAR_rf_steps"] = int(sig.n_rf_steps)  # steps with rf on in this call
    del keep


def sar_window(VObj, VSig, t_start, t_end):
    """Unaveraged local SAR and power of the windows [t_start[i], t_end[i]]
    from the sampled VSig.SAR (DoSARWatchExec.m:109-126): per window
    (VSig.SAR(:,:,:,idx2) - VSig.SAR(:,:,:,idx1)) / (tReal(idx2) - tReal(idx1))
    divided by 2*MassDen for SAR (W/kg) and scaled by dx*dy*dz/2 for power
    (W), with NaN set to 0; idx1/idx2 are the samples at t_start/t_end. The
    arithmetic is single precision like the reference's single VSig.SAR and
    MassDen; returns float64 [rows, cols, slices, windows] arrays."""
    ts = np.asarray(VSig["tSample"], dtype=np.float64).ravel()
    tr = np.asarray(VSig["tRealSample"], dtype=np.float64).ravel()
    S = np.asarray(VSig["SAR"], dtype=np.float32)
    S = S.reshape(S.shape[:3] + (-1,), order="F") if S.ndim >= 4 else S.reshape(S.shape + (1,) * (4 - S.ndim),
                                                                                      order="F")
    m = np.asarray(VObj["MassDen"], dtype=np.float32).reshape(S.shape[:3], order="F")
    vox = np.float32(float(VObj["XDimRes"]) * float(VObj["YDimRes"]) * float(VObj["ZDimRes"]) / 2)
    t_start = np.atleast_1d(np.asarray(t_start, dtype=np.float64))
    t_end = np.atleast_1d(np.asarray(t_end, dtype=np.float64))
    sar = np.zeros(S.shape[:3] + (t_end.size,), order="F")
    power = np.zeros_like(sar)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i, (a, b) in enumerate(zip(t_start, t_end)):
            i1, i2 = np.flatnonzero(ts == a), np.flatnonzero(ts == b)
            if i1.size != 1 or i2.size != 1:
                raise ValueError(f"window {i}: t_start/t_end must each match one VSig.tSample")
            rate = (S[..., i2[0]] - S[..., i1[0]]) / np.float32(tr[i2[0]] - tr[i1[0]])
            sar[..., i] = rate / (np.float32(2) * m)
            power[..., i] = rate * vox
    sar[np.isnan(sar)] = 0
    power[np.isnan(power)] = 0
    return sar, power


def DoSARAverageAtGPU(SAR, Power, MassDen, dx: float, dy: float, dz: float, grams: float):
    """N-gram local SAR and power averages ([aveSAR, avePower] of
    DoSARAverageAtCPU.cpp), bit for bit the reference's fp64 arithmetic.
    SAR and Power are [rows, cols, slices(, times)], MassDen [rows, cols,
    slices] in kg/m^3, dx/dy/dz in m. Numpy inputs (any float dtype, read
    in column-major order) give numpy float64 outputs; CUDA float64
    tensors, contiguous with the rows axis fastest (a permuted view of a
    [times, slices, cols, rows] tensor, or flat), give CUDA tensors shaped
    like SAR with the same strides."""
    lib = load()
    if hasattr(SAR, "data_ptr"):
        import torch

        ts = (SAR, Power, MassDen)
        if not all(t.is_cuda and t.dtype == torch.float64 for t in ts):
            raise TypeError("SAR, Power and MassDen must all be CUDA float64 tensors")
        if SAR.shape != Power.shape or SAR.stride() != Power.stride():
            raise ValueError("SAR and Power differ in shape or strides")
        shp = (tuple(SAR.shape) + (1, 1, 1, 1))[:4]
        if MassDen.numel() != shp[0] * shp[1] * shp[2]:
            raise ValueError("MassDen does not match SAR's first three dimensions")
        for t in ts:
            if t.dim() > 1 and not t.permute(*reversed(range(t.dim()))).is_contiguous():
                raise ValueError("arrays must be column-major contiguous (rows fastest)")
        a = torch.empty_like(SAR)
        b = torch.empty_like(SAR)
        dims = (C.c_int64 * 4)(*shp)
        _check(lib.mrs_DoSARAverageAtGPU(C.c_void_p(SAR.data_ptr()), C.c_void_p(Power.data_ptr()),
                                         C.c_void_p(MassDen.data_ptr()), dims, float(dx), float(dy), float(dz),
                                         float(grams), C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), 1))
        return a, b
    sar = np.asfortranarray(SAR, dtype=np.float64)
    pw = np.asfortranarray(Power, dtype=np.float64)
    m = np.asfortranarray(MassDen, dtype=np.float64)
    shp = (tuple(sar.shape) + (1, 1, 1, 1))[:4]
    if len(sar.shape) > 4 or pw.shape != sar.shape or m.size != shp[0] * shp[1] * shp[2]:
        raise ValueError("SAR, Power and MassDen shapes differ")
    a = np.zeros(sar.shape, order="F")
    b = np.zeros(sar.shape, order="F")
    dims = (C.c_int64 * 4)(*shp)
    _check(lib.mrs_DoSARAverageAtGPU(C.c_void_p(sar.ctypes.data), C.c_void_p(pw.ctypes.data),
                                     C.c_void_p(m.ctypes.data), dims, float(dx), float(dy), float(dz), float(grams),
                                     C.c_void_p(a.ctypes.data), C.c_void_p(b.ctypes.data), 0))
    return a, b


# ---------------------------------------------------------------------------
# Sequence compiler
# ---------------------------------------------------------------------------
SEQ_LINES = ("utsLine", "tsLine", "rfAmpLine", "rfPhaseLine", "rfFreqLine", "rfCoilLine", "GzAmpLine",
             "GyAmpLine", "GxAmpLine", "ADCLine", "ExtLine", "flagsLine")


def compile_psd(psd_xml: str, VCtl: dict, gyro: float = 267538030.3797068, check: bool = False,
                tx_coil_num: int = 0) -> dict:
    """VSeq lines of a PSD XML (its text) for the scan parameters ``VCtl``
    (numbers, numeric arrays, or strings for TEAnchor / MultiTransmit /
    DP_Flag), as DoPulseGen builds them (mrs_seq_compile). Returns a dict
    with the VSeq line names (flagsLine [6, n_ts] column-major) and
    ``TRNum``, the number of TRs. ``check`` makes DoPulseGen's checks
    (mrs_seq_check: multi-transmit against ``tx_coil_num`` = VCoi.TxCoilNum
    when it is > 0, then gradient and update rate with VCtl MaxGrad,
    MaxSlewRate, MinUpdRate) and raises MRSError naming the first failure."""
    lib = load()
    keep, params = [], []
    for name, v in VCtl.items():
        p = abi.SeqParamC()
        nb = name.encode()
        keep.append(nb)
        p.name = nb
        if isinstance(v, str):
            tb = v.encode()
            keep.append(tb)
            p.text = tb
        else:
            a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)).ravel())
            keep.append(a)
            p.value = a.ctypes.data_as(C.POINTER(C.c_double))
            p.n = a.size
        params.append(p)
    arr = (abi.SeqParamC * max(1, len(params)))(*params)
    h = C.c_void_p()
    _check(lib.mrs_seq_compile(psd_xml.encode(), arr, len(params), float(gyro), C.byref(h)))
    try:
        if check:
            num = lambda k: float(np.atleast_1d(VCtl[k])[0])
            _check(lib.mrs_seq_check(h, num("MaxGrad"), num("MaxSlewRate"), num("MinUpdRate"), int(tx_coil_num)))
        cnt = (C.c_int64 * 9)()
        _check(lib.mrs_seq_counts(h, cnt))
        n_ts, n_uts, n_rf, n_gz, n_gy, n_gx, n_adc, n_ext, n_tr = list(cnt)
        sizes = (n_uts, n_ts, n_rf, n_rf, n_rf, n_rf, n_gz, n_gy, n_gx, n_adc, n_ext, 6 * n_ts)
        out = {k: np.zeros(n) for k, n in zip(SEQ_LINES, sizes)}
        dst = abi.SeqLinesC(*[out[k].ctypes.data_as(C.POINTER(C.c_double)) for k in SEQ_LINES])
        _check(lib.mrs_seq_copy(h, C.byref(dst)))
    finally:
        lib.mrs_seq_destroy(h)
    out["flagsLine"] = np.asfortranarray(out["flagsLine"].reshape(6, n_ts, order="F"))
    out["TRNum"] = int(n_tr)
    return out
