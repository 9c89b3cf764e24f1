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
           "mrs_DoScanAtGPU", "mrs_DoGMScanAtGPU", "mrs_plan_create", "mrs_plan_run", "mrs_plan_sync",
           "mrs_plan_fetch", "mrs_plan_get_info", "mrs_plan_destroy")


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
        lib = C.CDLL(LIB_PATH)
        P = C.POINTER
        scan_args = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC), P(abi.VSeqC),
                     P(abi.VVarC), P(abi.VSigC), P(abi.CallbacksC), P(abi.OptionsC)]
        for fn in ("mrs_DoScanAtCPU", "mrs_DoScanAtGPU", "mrs_DoGMScanAtGPU"):
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


def _scan(fn: str, VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None,
          options: Optional[abi.OptionsC] = None, multipool: bool = False) -> None:
    lib = load()
    P = abi.pack_scan(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, multipool=multipool)
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


def DoScanAtCPU(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext=None, gpu_fetch=None, options=None):
    """The DoScanAtCPU entry point (DoScanAtCPU.cpp:70 contract), executed on the GPU."""
    _scan("mrs_DoScanAtCPU", VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, ext, gpu_fetch, options)


class Plan:
    """A decoded scan resident on the GPU (mrs_plan_*): run it many times."""

    def __init__(self, VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, options=None):
        self._lib = load()
        mp = options is not None and options.model == 1
        self._packed = abi.pack_scan(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, multipool=mp)
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


def device_count() -> int:
    return int(load().mrs_device_count())
