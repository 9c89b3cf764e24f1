"""CPU oracle of the Bloch scan hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
may import this package: as the checker of the HIP engine, or as the timed
CPU baseline. The product (mrilab_amd) never imports it.

``scan(scan_obj, precision)`` runs the C restatement of MRiLab's
DoScanAtCPU (oracle/bloch_oracle.c) on a copy of a workloads.Scan and
returns the copy with outputs filled in. ``precision`` is "f32" (the
reference's fp32 operation order) or "f64" (the same scheme in fp64).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

from mrilab_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_libs: dict[str, C.CDLL] = {}


def build() -> None:
    """Compile the oracle libraries (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(precision: str) -> C.CDLL:
    if precision not in ("f32", "f64"):
        raise ValueError("precision must be 'f32' or 'f64'")
    if precision not in _libs:
        path = os.path.join(_BUILD, f"liborc_{precision}.so")
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        P = C.POINTER
        lib.orc_DoScanAtCPU.argtypes = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC),
                                        P(abi.VSeqC), P(abi.VVarC), P(abi.VSigC), P(abi.CallbacksC)]
        lib.orc_DoScanAtCPU.restype = C.c_int
        lib.orc_DoGMScan.argtypes = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC),
                                     P(abi.VSeqC), P(abi.VVarC), P(abi.VSigC), P(abi.CallbacksC), C.c_int]
        lib.orc_DoGMScan.restype = C.c_int
        lib.orc_superlorentzian.argtypes = [C.c_double, C.c_double]
        lib.orc_superlorentzian.restype = C.c_double
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_set_relax_mode.argtypes = [C.c_int]
        _libs[precision] = lib
    return _libs[precision]


def set_threads(n: int, precision: str = "f32") -> None:
    """Thread count of the slice loop; 0 restores min(slices, MaxThreadNum)."""
    _lib(precision).orc_set_threads(int(n))


def set_relax_mode(mode: int, precision: str = "f32") -> None:
    """0: the reference's relaxation order; 1: expm1 form (error study)."""
    _lib(precision).orc_set_relax_mode(int(mode))


def scan(sc, precision: str = "f64", callbacks=None, inplace: bool = False):
    """Run the oracle scan; returns the Scan with VObj/VVar/VSig outputs."""
    out = sc if inplace else sc.copy()
    lib = _lib(precision)
    P = abi.pack_scan(*out.structs)
    cb = C.pointer(callbacks) if callbacks is not None else None
    rc = lib.orc_DoScanAtCPU(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi), C.byref(P.vctl),
                             C.byref(P.vseq), C.byref(P.vvar), C.byref(P.vsig), cb)
    if rc != 0:
        raise RuntimeError(f"oracle scan failed with status {rc}")
    abi.unpack_vvar(P, out.VVar)
    return out


def gm_scan(sc, precision: str = "f64", scheme: int = 0, inplace: bool = False):
    """Multi-pool scan (MRiLab's GM model): scheme 0 = operator split as
    BlochKernelGM.cuh, 1 = fp64 matrix exponential of the same generator."""
    out = sc if inplace else sc.copy()
    lib = _lib(precision)
    P = abi.pack_scan(*out.structs, multipool=True)
    rc = lib.orc_DoGMScan(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi), C.byref(P.vctl),
                          C.byref(P.vseq), C.byref(P.vvar), C.byref(P.vsig), None, int(scheme))
    if rc != 0:
        raise RuntimeError(f"oracle GM scan failed with status {rc}")
    abi.unpack_vvar(P, out.VVar)
    return out


def superlorentzian(dw: float, t2b: float, precision: str = "f64") -> float:
    """The bound-pool lineshape G(dw) as the GM kernel's Riemann sum computes it."""
    return float(_lib(precision).orc_superlorentzian(float(dw), float(t2b)))
