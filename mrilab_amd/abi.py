"""ctypes mirror of include/mrilab_amd.h.

The structs carry the field names of MRiLab's global scan structs
(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig). ``pack_scan`` turns Python
mappings of numpy arrays (one per MATLAB struct) into these structs after
checking every dtype against the contract of Src/Main/DoDataTypeConv.m:19-138.
Where the reference MEX would silently reinterpret a wrongly typed array,
``pack_scan`` raises TypeError naming the field.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Mapping

import numpy as np

F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)
# Array fields of the input/output structs are declared as raw addresses
# (the same size as the C pointers): filling them from numpy is a plain
# integer store instead of a ctypes cast per field. VVarC keeps typed
# pointers, since Ext callbacks receive it.
AP = C.c_void_p


class VObjC(C.Structure):
    _fields_ = [("Gyro", C.c_double), ("SpinNum", C.c_int32), ("TypeNum", C.c_int32),
                ("dims", C.c_int32 * 3),
                ("Mx", AP), ("My", AP), ("Mz", AP),
                ("Rho", AP), ("T1", AP), ("T2", AP),
                ("K", AP), ("TypeFlag", AP)]


class VMagC(C.Structure):
    _fields_ = [("dB0", AP), ("dWRnd", AP), ("Gzgrid", AP), ("Gygrid", AP),
                ("Gxgrid", AP)]


class VCoiC(C.Structure):
    _fields_ = [("TxCoilmg", AP), ("TxCoilpe", AP), ("RxCoilx", AP), ("RxCoily", AP),
                ("TxCoilNum", C.c_int32), ("RxCoilNum", C.c_int32),
                ("TxCoilDefault", C.c_double), ("RxCoilDefault", C.c_double)]


class VCtlC(C.Structure):
    _fields_ = [("CS", AP), ("TRNum", C.c_int32), ("RunMode", C.c_int32),
                ("MaxThreadNum", C.c_int32), ("GPUIndex", C.c_int32),
                ("MultiTransmit", C.c_int32), ("MasterTxCoil", C.c_int32)]


class VSeqC(C.Structure):
    _fields_ = [(n, AP) for n in ("utsLine", "tsLine", "rfAmpLine", "rfPhaseLine",
                                    "rfFreqLine", "rfCoilLine", "GzAmpLine", "GyAmpLine",
                                    "GxAmpLine", "ADCLine", "ExtLine", "flagsLine")] + \
              [(n, C.c_int64) for n in ("n_ts", "n_uts", "n_rf", "n_Gz", "n_Gy", "n_Gx",
                                        "n_ADC", "n_Ext")]


class VVarC(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double),
                ("rfAmp", F64P), ("rfPhase", F64P), ("rfFreq", F64P),
                ("rfCoil", C.c_double), ("rfRef", C.c_double),
                ("GzAmp", C.c_double), ("GyAmp", C.c_double), ("GxAmp", C.c_double),
                ("ADC", C.c_double), ("Ext", C.c_double),
                ("Kz", C.c_double), ("Ky", C.c_double), ("Kx", C.c_double),
                ("KzKeeper", C.c_double), ("KyKeeper", C.c_double), ("KxKeeper", C.c_double),
                ("gpuFetch", C.c_double)] + \
              [(n, C.c_int32) for n in ("utsi", "rfi", "Gzi", "Gyi", "Gxi", "ADCi", "Exti",
                                        "TRCount")]


class VSigC(C.Structure):
    _fields_ = [("Sx", AP), ("Sy", AP), ("Kz", AP), ("Ky", AP), ("Kx", AP),
                ("SignalNum", C.c_int32),
                ("Mx", AP), ("My", AP), ("Mz", AP), ("Muts", AP)]


class ExtStateC(C.Structure):
    """mrs_ext_state: what an Ext callback may read and re-point."""
    _fields_ = [("vobj", C.POINTER(VObjC)), ("vmag", C.POINTER(VMagC)), ("vcoi", C.POINTER(VCoiC)),
                ("vsig", C.POINTER(VSigC)), ("fetched", C.c_int32), ("reserved", C.c_int32)]


GPU_FETCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(C.c_int))
EXT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.POINTER(VVarC), C.POINTER(ExtStateC))


class CallbacksC(C.Structure):
    _fields_ = [("gpu_fetch", GPU_FETCH_FN), ("ext", EXT_FN), ("user", C.c_void_p)]


class OptionsC(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("spins_per_thread", C.c_int32),
                ("max_adc_per_launch", C.c_int32), ("multipool_scheme", C.c_int32),
                ("precision_mode", C.c_int32), ("timing", C.c_int32), ("model", C.c_int32),
                ("coil_sum", C.c_int32), ("free_steps", C.c_int32)]


class PlanInfoC(C.Structure):
    _fields_ = [("n_steps", C.c_int64), ("n_adc", C.c_int64), ("n_iso", C.c_int64),
                ("n_launches", C.c_int64), ("n_rf_steps", C.c_int64),
                ("n_cached_steps", C.c_int64),
                ("block_size", C.c_int32), ("spins_per_thread", C.c_int32),
                ("grid_blocks", C.c_int32), ("last_run_ms", C.c_float),
                ("fwd_kernel_ms", C.c_double), ("reduce_kernel_ms", C.c_double),
                ("deferred_free_steps", C.c_int32)]


class BatchInfoC(C.Structure):
    _fields_ = [("n_scans", C.c_int64), ("n_tiles", C.c_int64), ("n_iso", C.c_int64),
                ("n_iso_steps", C.c_int64), ("n_adc", C.c_int64),
                ("spins_per_thread", C.c_int32), ("block_size", C.c_int32), ("grid_blocks", C.c_int32),
                ("tuned", C.c_int32), ("n_candidates", C.c_int32), ("last_run_ms", C.c_float),
                ("tune_best_ms", C.c_float), ("tune_worst_ms", C.c_float), ("decode_ms", C.c_double),
                ("tune_ms", C.c_double), ("create_ms", C.c_double), ("deferred_free_steps", C.c_int32)]


class ExcitationC(C.Structure):
    _fields_ = [("n_steps", C.c_int32), ("checkpoint", C.c_int32), ("segment_mode", C.c_int32),
                ("reserved", C.c_int32), ("dt", C.c_double),
                ("grad", C.POINTER(C.c_float)), ("target", C.POINTER(C.c_float))]


class ReconDescC(C.Structure):
    _fields_ = [("res_freq", C.c_int32), ("res_phase", C.c_int32), ("slice_num", C.c_int32),
                ("te_per_tr", C.c_int32), ("rx", C.c_int32), ("types", C.c_int32), ("zf_kx", C.c_int32),
                ("zf_ky", C.c_int32), ("zf_kz", C.c_int32), ("swap_xy", C.c_int32)]


class VImgC(C.Structure):
    _fields_ = [("Real", C.c_void_p), ("Imag", C.c_void_p), ("Mag", C.c_void_p), ("Phase", C.c_void_p),
                ("Sx", C.c_void_p), ("Sy", C.c_void_p)]


class SarObjC(C.Structure):
    _fields_ = [("dims", C.c_int32 * 3), ("ECon", AP), ("TxE1x", AP),
                ("TxE1y", AP), ("TxE1z", AP), ("TxCoilNum", C.c_int32),
                ("TRNum", C.c_int32)]


class SarSigC(C.Structure):
    _fields_ = [("SARBase", AP), ("SARi", C.c_int32), ("reserved", C.c_int32),
                ("SAR", AP), ("Muts", AP),
                ("tSample", AP), ("tRealSample", AP),
                ("n_samples", C.c_int64), ("n_rf_steps", C.c_int64)]


class AdjointInfoC(C.Structure):
    _fields_ = [("n_iso", C.c_int64), ("n_steps", C.c_int64), ("checkpoint", C.c_int32),
                ("n_checkpoints", C.c_int32), ("spins_per_thread", C.c_int32), ("block_size", C.c_int32),
                ("segments_rebuilt", C.c_int32), ("reserved", C.c_int32),
                ("forward_ms", C.c_float), ("backward_ms", C.c_float)]


# ---------------------------------------------------------------------------
# dtype contract, Src/Main/DoDataTypeConv.m
# ---------------------------------------------------------------------------
_VVAR_F64 = ("t", "dt", "rfCoil", "rfRef", "GzAmp", "GyAmp", "GxAmp", "ADC", "Ext",
             "Kz", "Ky", "Kx", "gpuFetch")
_VVAR_I32 = ("utsi", "rfi", "Gzi", "Gyi", "Gxi", "ADCi", "Exti", "TRCount")


def _req(struct: Mapping[str, Any], sname: str, field: str, dtype, size: int | None = None,
         writable: bool = False) -> np.ndarray:
    if field not in struct:
        raise KeyError(f"{sname}.{field} is missing")
    a = struct[field]
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{sname}.{field} must be a numpy array of {np.dtype(dtype).name}")
    if a.dtype != np.dtype(dtype):
        raise TypeError(f"{sname}.{field} has dtype {a.dtype}, expected {np.dtype(dtype).name}")
    if not (a.flags.f_contiguous or a.flags.c_contiguous and a.ndim <= 1):
        raise TypeError(f"{sname}.{field} must be contiguous in column-major (MATLAB) order")
    if size is not None and a.size != size:
        raise ValueError(f"{sname}.{field} has {a.size} elements, expected {size}")
    if writable and not a.flags.writeable:
        raise ValueError(f"{sname}.{field} must be writable")
    return a


def _scalar(struct, sname, field, dtype):
    if field not in struct:
        raise KeyError(f"{sname}.{field} is missing")
    v = struct[field]
    if isinstance(v, np.generic):  # the common case, without an array round trip
        if v.dtype != dtype:
            raise TypeError(f"{sname}.{field} has dtype {v.dtype}, expected {np.dtype(dtype).name}")
        return v.item()
    a = np.asarray(v)
    if isinstance(v, np.ndarray) or isinstance(v, np.generic):
        if a.dtype != np.dtype(dtype):
            raise TypeError(f"{sname}.{field} has dtype {a.dtype}, expected {np.dtype(dtype).name}")
    if a.size != 1:
        raise ValueError(f"{sname}.{field} must be a scalar")
    return a.reshape(-1)[0].item()


def _a(a: np.ndarray | None):
    """Address of a numpy buffer (None for NULL) for an AP field."""
    return None if a is None else a.ctypes.data


def _p(a: np.ndarray | None, typ):
    if a is None:
        return C.cast(None, typ)
    return a.ctypes.data_as(typ)


class Packed:
    """The ABI structs plus references that keep the numpy buffers alive."""

    def __init__(self):
        self.keep: list[Any] = []


def pack_scan(VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, multipool: bool = False,
              k_planes: int | None = None) -> Packed:
    """Checks dtypes and sizes of the seven structs and returns the ABI
    structs pointing at their buffers. multipool: pass VObj.K (with
    ``k_planes`` planes of R*C*Z, default TypeNum^2) and VObj.TypeFlag
    (optional for the MT, CEST and ME layouts)."""
    P = Packed()
    Mz = _req(VObj, "VObj", "Mz", np.float32, writable=True)
    shp = Mz.shape + (1,) * (5 - Mz.ndim)
    R, Cc, Z = int(shp[0]), int(shp[1]), int(shp[2])
    S = int(_scalar(VObj, "VObj", "SpinNum", np.int32))
    Ty = int(_scalar(VObj, "VObj", "TypeNum", np.int32))
    nvox = R * Cc * Z
    if Mz.size != nvox * S * Ty:
        raise ValueError(f"VObj.Mz has {Mz.size} elements, expected R*C*Z*SpinNum*TypeNum = {nvox * S * Ty}")
    Mx = _req(VObj, "VObj", "Mx", np.float32, Mz.size, writable=True)
    My = _req(VObj, "VObj", "My", np.float32, Mz.size, writable=True)
    Rho = _req(VObj, "VObj", "Rho", np.float32, nvox * Ty)
    T1 = _req(VObj, "VObj", "T1", np.float32, nvox * Ty)
    T2 = _req(VObj, "VObj", "T2", np.float32, nvox * Ty)
    K = TF = None
    if multipool:
        K = _req(VObj, "VObj", "K", np.float32, nvox * (Ty * Ty if k_planes is None else int(k_planes)))
        if "TypeFlag" in VObj or k_planes is None:
            TF = _req(VObj, "VObj", "TypeFlag", np.float64, Ty)
    vobj = VObjC(Gyro=float(_scalar(VObj, "VObj", "Gyro", np.float64)), SpinNum=S, TypeNum=Ty)
    vobj.dims[0], vobj.dims[1], vobj.dims[2] = R, Cc, Z
    vobj.Mx, vobj.My, vobj.Mz = _a(Mx), _a(My), _a(Mz)
    vobj.Rho, vobj.T1, vobj.T2 = _a(Rho), _a(T1), _a(T2)
    vobj.K, vobj.TypeFlag = _a(K), _a(TF)

    vmag = VMagC(dB0=_a(_req(VMag, "VMag", "dB0", np.float32, nvox)),
                 dWRnd=_a(_req(VMag, "VMag", "dWRnd", np.float32, nvox * S * Ty)),
                 Gzgrid=_a(_req(VMag, "VMag", "Gzgrid", np.float32, nvox)),
                 Gygrid=_a(_req(VMag, "VMag", "Gygrid", np.float32, nvox)),
                 Gxgrid=_a(_req(VMag, "VMag", "Gxgrid", np.float32, nvox)))

    Tx = int(_scalar(VCoi, "VCoi", "TxCoilNum", np.int32))
    Rx = int(_scalar(VCoi, "VCoi", "RxCoilNum", np.int32))
    vcoi = VCoiC(TxCoilmg=_a(_req(VCoi, "VCoi", "TxCoilmg", np.float32, nvox * Tx)),
                 TxCoilpe=_a(_req(VCoi, "VCoi", "TxCoilpe", np.float32, nvox * Tx)),
                 RxCoilx=_a(_req(VCoi, "VCoi", "RxCoilx", np.float32, nvox * Rx)),
                 RxCoily=_a(_req(VCoi, "VCoi", "RxCoily", np.float32, nvox * Rx)),
                 TxCoilNum=Tx, RxCoilNum=Rx,
                 TxCoilDefault=float(_scalar(VCoi, "VCoi", "TxCoilDefault", np.float64)),
                 RxCoilDefault=float(_scalar(VCoi, "VCoi", "RxCoilDefault", np.float64)))

    mt = VCtl.get("MultiTransmit", "off")
    vctl = VCtlC(CS=_a(_req(VCtl, "VCtl", "CS", np.float64, Ty)),
                 TRNum=int(_scalar(VCtl, "VCtl", "TRNum", np.int32)),
                 RunMode=int(_scalar(VCtl, "VCtl", "RunMode", np.int32)),
                 MaxThreadNum=int(_scalar(VCtl, "VCtl", "MaxThreadNum", np.int32)),
                 GPUIndex=int(VCtl.get("GPUIndex", np.int32(0))),
                 MultiTransmit=1 if mt == "on" else 0,
                 MasterTxCoil=int(VCtl.get("MasterTxCoil", 1)))

    seq_fields = {}
    for f in ("utsLine", "tsLine", "rfAmpLine", "rfPhaseLine", "rfFreqLine", "rfCoilLine",
              "GzAmpLine", "GyAmpLine", "GxAmpLine", "ADCLine", "ExtLine", "flagsLine"):
        seq_fields[f] = _req(VSeq, "VSeq", f, np.float64)
    n_ts = seq_fields["tsLine"].size
    if seq_fields["flagsLine"].size != 6 * n_ts:
        raise ValueError("VSeq.flagsLine must be 6 x numel(tsLine)")
    if seq_fields["rfCoilLine"].size != seq_fields["rfAmpLine"].size:
        raise ValueError("VSeq.rfCoilLine and rfAmpLine differ in length")
    vseq = VSeqC(**{f: _a(a) for f, a in seq_fields.items()},
                 n_ts=n_ts, n_uts=seq_fields["utsLine"].size, n_rf=seq_fields["rfAmpLine"].size,
                 n_Gz=seq_fields["GzAmpLine"].size, n_Gy=seq_fields["GyAmpLine"].size,
                 n_Gx=seq_fields["GxAmpLine"].size, n_ADC=seq_fields["ADCLine"].size,
                 n_Ext=seq_fields["ExtLine"].size)

    rfA = _req(VVar, "VVar", "rfAmp", np.float64, Tx, writable=True)
    rfP = _req(VVar, "VVar", "rfPhase", np.float64, Tx, writable=True)
    rfF = _req(VVar, "VVar", "rfFreq", np.float64, Tx, writable=True)
    vvar = VVarC(rfAmp=_p(rfA, F64P), rfPhase=_p(rfP, F64P), rfFreq=_p(rfF, F64P))
    for f in _VVAR_F64:
        if f == "gpuFetch":
            vvar.gpuFetch = float(VVar.get(f, 0.0))
        else:
            setattr(vvar, f, float(_scalar(VVar, "VVar", f, np.float64)))
    for f in ("KzKeeper", "KyKeeper", "KxKeeper"):
        setattr(vvar, f, float(VVar.get(f, 0.0)))
    for f in _VVAR_I32:
        setattr(vvar, f, int(_scalar(VVar, "VVar", f, np.int32)))

    N = int(_scalar(VSig, "VSig", "SignalNum", np.int32))
    Sx = _req(VSig, "VSig", "Sx", np.float64, N * Ty * Rx, writable=True)
    Sy = _req(VSig, "VSig", "Sy", np.float64, N * Ty * Rx, writable=True)
    vsig = VSigC(Sx=_a(Sx), Sy=_a(Sy),
                 Kz=_a(_req(VSig, "VSig", "Kz", np.float64, N, writable=True)),
                 Ky=_a(_req(VSig, "VSig", "Ky", np.float64, N, writable=True)),
                 Kx=_a(_req(VSig, "VSig", "Kx", np.float64, N, writable=True)),
                 SignalNum=N)
    if vctl.RunMode == 1:
        nuts = seq_fields["utsLine"].size
        vsig.Mx = _a(_req(VSig, "VSig", "Mx", np.float32, Mz.size * nuts, writable=True))
        vsig.My = _a(_req(VSig, "VSig", "My", np.float32, Mz.size * nuts, writable=True))
        vsig.Mz = _a(_req(VSig, "VSig", "Mz", np.float32, Mz.size * nuts, writable=True))
        vsig.Muts = _a(_req(VSig, "VSig", "Muts", np.float64, nuts, writable=True))
    P.vobj, P.vmag, P.vcoi, P.vctl, P.vseq, P.vvar, P.vsig = vobj, vmag, vcoi, vctl, vseq, vvar, vsig
    P.keep += [VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig, seq_fields, K, TF]
    return P


def unpack_vvar(P: Packed, VVar) -> None:
    """Write the scalar decoder state back into the VVar mapping, as the
    reference leaves it in the MATLAB global after the MEX returns."""
    for f in _VVAR_F64 + ("KzKeeper", "KyKeeper", "KxKeeper"):
        VVar[f] = np.float64(getattr(P.vvar, f))
    for f in _VVAR_I32:
        VVar[f] = np.int32(getattr(P.vvar, f))


def pack_sar(VObj, VCoi, VCtl, VVar, VSig, n_uts: int):
    """The SAR engine's structs (DoCalSARAtCPU.cpp:115-166): VObj.MassDen
    (dims) and ECon [R,C,Z,3]; VCoi.TxE1x/y/z [R,C,Z,Tx]; VCtl.TRNum;
    VVar.SAR [R,C,Z,3] and SARi; VSig.SAR [R,C,Z,n_samples], Muts [n_uts],
    tSample and tRealSample [n_samples]. Returns (obj, sig, keep-alive)."""
    md = np.asarray(VObj["MassDen"])
    shp = md.shape + (1,) * (3 - md.ndim)
    nvox = int(np.prod(shp[:3]))
    Tx = int(_scalar(VCoi, "VCoi", "TxCoilNum", np.int32))
    ECon = _req(VObj, "VObj", "ECon", np.float32, 3 * nvox)
    E = [_req(VCoi, "VCoi", f"TxE1{k}", np.float32, nvox * Tx) for k in "xyz"]
    base = _req(VVar, "VVar", "SAR", np.float32, 3 * nvox, writable=True)
    tS = _req(VSig, "VSig", "tSample", np.float64)
    ns = tS.size
    SAR = _req(VSig, "VSig", "SAR", np.float32, nvox * ns, writable=True)
    Muts = _req(VSig, "VSig", "Muts", np.float64, n_uts, writable=True)
    tR = _req(VSig, "VSig", "tRealSample", np.float64, ns, writable=True)
    obj = SarObjC(ECon=_a(ECon), TxE1x=_a(E[0]), TxE1y=_a(E[1]), TxE1z=_a(E[2]),
                  TxCoilNum=Tx, TRNum=int(_scalar(VCtl, "VCtl", "TRNum", np.int32)))
    obj.dims[0], obj.dims[1], obj.dims[2] = (int(x) for x in shp[:3])
    sig = SarSigC(SARBase=_a(base), SARi=int(VVar.get("SARi", 0)), SAR=_a(SAR),
                  Muts=_a(Muts), tSample=_a(tS), tRealSample=_a(tR), n_samples=ns)
    return obj, sig, (ECon, E, base, tS, SAR, Muts, tR)


class SeqParamC(C.Structure):
    _fields_ = [("name", C.c_char_p), ("value", C.POINTER(C.c_double)), ("n", C.c_int32),
                ("reserved", C.c_int32), ("text", C.c_char_p)]


class SeqLinesC(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in
                ("utsLine", "tsLine", "rfAmpLine", "rfPhaseLine", "rfFreqLine", "rfCoilLine", "GzAmpLine",
                 "GyAmpLine", "GxAmpLine", "ADCLine", "ExtLine", "flagsLine")]
