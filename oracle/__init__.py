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

import numpy as np

from mrilab_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
_libs: dict[str, C.CDLL] = {}


def build() -> None:
    """Compile the oracle libraries (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _configure(lib: C.CDLL) -> C.CDLL:
    P = C.POINTER
    lib.orc_DoScanAtCPU.argtypes = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC),
                                    P(abi.VSeqC), P(abi.VVarC), P(abi.VSigC), P(abi.CallbacksC)]
    lib.orc_DoScanAtCPU.restype = C.c_int
    lib.orc_DoGMScan.argtypes = [P(abi.VObjC), P(abi.VMagC), P(abi.VCoiC), P(abi.VCtlC),
                                 P(abi.VSeqC), P(abi.VVarC), P(abi.VSigC), P(abi.CallbacksC), C.c_int,
                                 C.c_int]
    lib.orc_DoGMScan.restype = C.c_int
    lib.orc_superlorentzian.argtypes = [C.c_double, C.c_double]
    lib.orc_superlorentzian.restype = C.c_double
    if hasattr(lib, "orc_DoCalSAR"):
        lib.orc_DoCalSAR.argtypes = [P(abi.SarObjC), P(abi.VSeqC), P(abi.VVarC), P(abi.SarSigC)]
        lib.orc_DoCalSAR.restype = C.c_int
    if hasattr(lib, "orc_DoSARAverage"):
        D = C.POINTER(C.c_double)
        lib.orc_DoSARAverage.argtypes = [D, D, D, C.POINTER(C.c_int64), C.c_double, C.c_double, C.c_double,
                                         C.c_double, D, D, C.c_int64, C.c_int64]
        lib.orc_DoSARAverage.restype = C.c_int
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_set_relax_mode.argtypes = [C.c_int]
    return lib


def _lib(precision: str) -> C.CDLL:
    if precision not in ("f32", "f64"):
        raise ValueError("precision must be 'f32' or 'f64'")
    if precision not in _libs:
        path = os.path.join(_BUILD, f"liborc_{precision}.so")
        if not os.path.exists(path):
            build()
        _libs[precision] = _configure(C.CDLL(path))
    return _libs[precision]


def build_mutant(mutation: int, out_dir: str, precision: str = "f64") -> C.CDLL:
    """Compiles bloch_oracle.c with planted mistake ``mutation`` (the
    ORC_MUTATION list in its header) into ``out_dir`` and loads it. Used
    only by the oracle's own tests, to show that a pin catches the mistake."""
    real = "double" if precision == "f64" else "float"
    path = os.path.join(out_dir, f"liborc_mut{mutation}_{precision}.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    f"-DORC_REAL={real}", f"-DORC_MUTATION={int(mutation)}", "-o", path,
                    os.path.join(_HERE, "bloch_oracle.c"), "-lm"], check=True)
    return _configure(C.CDLL(path))


def set_threads(n: int, precision: str = "f32") -> None:
    """Thread count of the slice loop; 0 restores min(slices, MaxThreadNum)."""
    _lib(precision).orc_set_threads(int(n))


def set_relax_mode(mode: int, precision: str = "f32") -> None:
    """0: the reference's relaxation order; 1: expm1 form (error study)."""
    _lib(precision).orc_set_relax_mode(int(mode))


def scan(sc, precision: str = "f64", callbacks=None, inplace: bool = False, lib=None):
    """Run the oracle scan; returns the Scan with VObj/VVar/VSig outputs.
    ``lib``: a library from build_mutant in place of the default one."""
    out = sc if inplace else sc.copy()
    lib = lib or _lib(precision)
    P = abi.pack_scan(*out.structs)
    cb = C.pointer(callbacks) if callbacks is not None else None
    rc = lib.orc_DoScanAtCPU(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi), C.byref(P.vctl),
                             C.byref(P.vseq), C.byref(P.vvar), C.byref(P.vsig), cb)
    if rc != 0:
        raise RuntimeError(f"oracle scan failed with status {rc}")
    abi.unpack_vvar(P, out.VVar)
    return out


VARIANTS = {"GM": 0, "MT": 1, "CEST": 2, "ME": 3}


def k_planes(variant: str, ty: int) -> int:
    """Planes of VObj.K for each multi-pool layout (CEST: 2*(Ty-1))."""
    return 2 * (ty - 1) if variant == "CEST" else ty * ty


def gm_scan(sc, precision: str = "f64", scheme: int = 0, inplace: bool = False, variant: str = "GM", lib=None):
    """Multi-pool scan: variant "GM" (BlochKernelGM.cuh), "MT"
    (BlochKernelMT.cuh), "CEST" (BlochKernelCEST.cuh) or "ME"
    (BlochKernelME.cuh) with VObj.K in that kernel's layout; scheme 0 =
    the kernel's operator split, 1 = fp64 matrix exponential of the same
    generator."""
    out = sc if inplace else sc.copy()
    lib = lib or _lib(precision)
    ty = int(out.VObj["TypeNum"])
    P = abi.pack_scan(*out.structs, multipool=True, k_planes=k_planes(variant, ty))
    rc = lib.orc_DoGMScan(C.byref(P.vobj), C.byref(P.vmag), C.byref(P.vcoi), C.byref(P.vctl),
                          C.byref(P.vseq), C.byref(P.vvar), C.byref(P.vsig), None, int(scheme),
                          VARIANTS[variant])
    if rc != 0:
        raise RuntimeError(f"oracle GM scan failed with status {rc}")
    abi.unpack_vvar(P, out.VVar)
    return out


def superlorentzian(dw: float, t2b: float, precision: str = "f64") -> float:
    """The bound-pool lineshape G(dw) as the GM kernel's Riemann sum computes it."""
    return float(_lib(precision).orc_superlorentzian(float(dw), float(t2b)))


class _ExcC(C.Structure):
    _F = C.POINTER(C.c_float)
    _fields_ = [("nvox", C.c_int32), ("S", C.c_int32), ("n", C.c_int32), ("dt", C.c_double),
                ("gyro", C.c_double), ("x", _F), ("y", _F), ("z", _F), ("mg", _F), ("pe", _F),
                ("dWRnd", _F), ("rho", _F), ("M0x", _F), ("M0y", _F), ("M0z", _F), ("grad", _F),
                ("target", _F)]


def _exc(exc):
    import numpy as np
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F"))  # noqa: E731
    arrs = {"x": f(exc.VMag["Gxgrid"]), "y": f(exc.VMag["Gygrid"]), "z": f(exc.VMag["Gzgrid"]),
            "mg": f(exc.VCoi["TxCoilmg"]), "pe": f(exc.VCoi["TxCoilpe"]), "dWRnd": f(exc.VMag["dWRnd"]),
            "rho": f(exc.VObj["Rho"]), "M0x": f(exc.VObj["Mx"]), "M0y": f(exc.VObj["My"]),
            "M0z": f(exc.VObj["Mz"]), "grad": np.ascontiguousarray(exc.grad, dtype=np.float32).ravel(),
            "target": np.ascontiguousarray(exc.target, dtype=np.float32).ravel()}
    P = C.POINTER(C.c_float)
    s = _ExcC(nvox=int(exc.target.shape[0]), S=int(exc.VObj["SpinNum"]), n=int(exc.grad.shape[0]),
              dt=float(exc.dt), gyro=float(exc.VObj["Gyro"]),
              **{k: v.ctypes.data_as(P) for k, v in arrs.items()})
    return s, arrs


def _dptr(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def exc_loss(exc, rf, T1, T2, dB0, precision: str = "f64"):
    """Loss and voxel sums [nvox, 2] of the excitation problem (no reference
    counterpart; forward model of oracle/adjoint_oracle.c)."""
    import numpy as np
    lib = _lib(precision)
    lib.orc_exc_loss.restype = C.c_double
    s, keep = _exc(exc)
    ps = [_dptr(a) for a in (rf, T1, T2, dB0)]
    mxy = np.zeros((s.nvox, 2))
    L = lib.orc_exc_loss(C.byref(s), *[p[1] for p in ps], mxy.ctypes.data_as(C.POINTER(C.c_double)))
    del keep, ps
    return L, mxy


def exc_grad(exc, rf, T1, T2, dB0, precision: str = "f64"):
    """Loss and reverse-mode gradients (g_rf [n,2], g_T1, g_T2, g_dB0)."""
    import numpy as np
    lib = _lib(precision)
    lib.orc_exc_grad.restype = C.c_double
    s, keep = _exc(exc)
    ps = [_dptr(a) for a in (rf, T1, T2, dB0)]
    g = [np.zeros(s.n * 2), np.zeros(s.nvox), np.zeros(s.nvox), np.zeros(s.nvox)]
    L = lib.orc_exc_grad(C.byref(s), *[p[1] for p in ps], *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in g])
    del keep, ps
    return L, g[0].reshape(s.n, 2), g[1], g[2], g[3]


def sar(sc, precision: str = "f32", inplace: bool = False):
    """Local SAR (orc_DoCalSAR, DoCalSARAtCPU.cpp restated) of a Scan whose
    structs carry the SAR fields (workloads.sar_scan); returns the Scan with
    VVar.SAR/SARi and VSig.SAR/Muts/tRealSample written."""
    out = sc if inplace else sc.copy()
    lib = _lib(precision)
    P = abi.pack_scan(*out.structs)
    obj, sig, keep = abi.pack_sar(out.VObj, out.VCoi, out.VCtl, out.VVar, out.VSig, int(P.vseq.n_uts))
    rc = lib.orc_DoCalSAR(C.byref(obj), C.byref(P.vseq), C.byref(P.vvar), C.byref(sig))
    if rc != 0:
        raise RuntimeError(f"oracle SAR failed with status {rc}")
    abi.unpack_vvar(P, out.VVar)
    out.VVar["SARi"] = np.int32(sig.SARi)
    out.VSig["SAR_rf_steps"] = int(sig.n_rf_steps)  # steps with rf on in this call
    del keep
    return out


def sar_average(SAR, Power, MassDen, dx: float, dy: float, dz: float, grams: float, slices=None):
    """N-gram local SAR average (orc_DoSARAverage, DoSARAverageAtCPU.cpp
    restated): SAR and Power [rows, cols, slices(, times)] and MassDen
    [rows, cols, slices], all read as fp64 in column-major order. Returns
    (aveSAR, avePower) shaped as SAR. ``slices = (s0, s1)`` computes only
    those slices (the others stay 0)."""
    sar = np.asfortranarray(SAR, dtype=np.float64)
    pw = np.asfortranarray(Power, dtype=np.float64)
    m = np.asfortranarray(MassDen, dtype=np.float64)
    shp = (tuple(sar.shape) + (1, 1, 1, 1))[:4]
    if pw.shape != sar.shape or (tuple(m.shape) + (1, 1, 1))[:3] != shp[:3]:
        raise ValueError("SAR, Power and MassDen shapes differ")
    s0, s1 = (0, shp[2]) if slices is None else slices
    a = np.zeros(sar.shape, order="F")
    b = np.zeros(sar.shape, order="F")
    D = C.POINTER(C.c_double)
    dims = (C.c_int64 * 4)(*shp)
    rc = _lib("f64").orc_DoSARAverage(sar.ctypes.data_as(D), pw.ctypes.data_as(D), m.ctypes.data_as(D), dims,
                                       float(dx), float(dy), float(dz), float(grams), a.ctypes.data_as(D),
                                       b.ctypes.data_as(D), int(s0), int(s1))
    if rc == -1:
        raise ValueError("the volume's mass does not exceed grams*1e-3 kg")
    if rc != 0:
        raise RuntimeError(f"oracle SAR average failed with status {rc}")
    return a, b
