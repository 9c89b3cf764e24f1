"""CPU restatement of MRiLab's Cartesian reconstruction and k-space noise.

Test infrastructure only: tests/ and the bench's CPU leg call these to check
the HIP path (mrs_DoCartRecon, mrs_DoAddNoise); nothing in the product
imports them.

cart_recon follows Src/Recon/Default/DoCartRecon.m:12-84 in MATLAB's
column-major order (numpy arrays with order="F" reshapes):
  * VSig.Sx/Sy reshaped to [ResFreq*TEPerTR, ResPhase, SliceNum, Rx, Ty]
    (:12-13) and summed over the types (:15-16);
  * per echo e: rows e*ResFreq .. (e+1)*ResFreq-1 (:19-21), the last
    readout sample dropped (:25-27);
  * optional zero filling into [ZF_Kx, ZF_Ky, ZF_Kz*SliceNum] at the
    offsets of :29-40;
  * permute [2 1 3 4] (:41-42), and once more when the scan plane and
    frequency direction ask for it (:45-73);
  * per coil fftshift(ifftn(fftshift(S))) (:79-84) -> Real, Imag, Mag,
    Phase; the permuted k-space goes to VImg.Sx/Sy (:86-87).
noise_scale follows Src/Main/DoAddNoise.m:8-16.
"""
from __future__ import annotations

import numpy as np


def swap_xy(scan_plane: str, freq_dir: str) -> bool:
    """Whether DoCartRecon.m:45-73 applies the second permute [2 1 3 4]."""
    if scan_plane == "Axial":
        return freq_dir == "A/P"
    if scan_plane == "Sagittal":
        return freq_dir == "S/I"
    if scan_plane == "Coronal":
        return freq_dir != "L/R"
    raise ValueError(f"unknown scan plane {scan_plane!r}")


def zf_offsets(zf, res_freq_kept: int, res_phase: int, slices: int):
    """0-based offsets of DoCartRecon.m:31-37 (1-based there); MATLAB's
    indexing needs them to be integers."""
    zx, zy, zz = zf
    ox = zx / 2 - res_freq_kept / 2
    oy = zy / 2 - res_phase / 2
    oz = zz * (slices // 2) - slices // 2
    for o in (ox, oy):
        if o != int(o) or o < 0:
            raise ValueError("zero-filling offsets must be non-negative integers")
    return int(ox), int(oy), int(oz)


def cart_recon(Sx, Sy, res_freq: int, res_phase: int, slice_num: int, te_per_tr: int, rx: int, types: int,
               zf=None, swap: bool = False):
    """Returns dict(Real, Imag, Mag, Phase, Sx, Sy), each [d0, d1, d2, Rx,
    TEPerTR] float64 in column-major order. ``zf`` = (ZF_Kx, ZF_Ky,
    ZF_Kz factor) or None."""
    shp = (res_freq * te_per_tr, res_phase, slice_num, rx, types)
    SX = np.asarray(Sx, dtype=np.float64).reshape(shp, order="F").sum(axis=4)
    SY = np.asarray(Sy, dtype=np.float64).reshape(shp, order="F").sum(axis=4)
    out = None
    for e in range(te_per_tr):
        sx = SX[e * res_freq:(e + 1) * res_freq][:-1]
        sy = SY[e * res_freq:(e + 1) * res_freq][:-1]
        rf = res_freq - 1
        if zf is not None:
            zx, zy, zz = zf
            ox, oy, oz = zf_offsets(zf, rf, res_phase, slice_num)
            bx = np.zeros((zx, zy, zz * slice_num, rx))
            by = np.zeros_like(bx)
            bx[ox:ox + rf, oy:oy + res_phase, oz:oz + slice_num] = sx
            by[ox:ox + rf, oy:oy + res_phase, oz:oz + slice_num] = sy
            sx, sy = bx, by
        sx = sx.transpose(1, 0, 2, 3)
        sy = sy.transpose(1, 0, 2, 3)
        if swap:
            sx = sx.transpose(1, 0, 2, 3)
            sy = sy.transpose(1, 0, 2, 3)
        S = sx + 1j * sy
        img = np.empty(S.shape, dtype=np.complex128)
        for c in range(rx):
            img[..., c] = np.fft.fftshift(np.fft.ifftn(np.fft.fftshift(S[..., c])))
        if out is None:
            d = S.shape + (te_per_tr,)
            out = {k: np.zeros(d) for k in ("Real", "Imag", "Mag", "Phase", "Sx", "Sy")}
        out["Real"][..., e] = img.real
        out["Imag"][..., e] = img.imag
        out["Mag"][..., e] = np.abs(img)
        out["Phase"][..., e] = np.angle(img)
        out["Sx"][..., e] = sx
        out["Sy"][..., e] = sy
    return out


def echo_train_order(Sx, Sy, res_freq: int, etl: int, shots: int, slice_num: int, rx: int, types: int, epi: bool):
    """DoEPIRecon.m:8-26 / DoFSERecon.m:8-22: the signal of each shot (ETL
    echoes of res_freq samples) into Cartesian order, echo e of shot p on
    line (e-1)*ShotNum + p, EPI's even echoes reversed along the readout;
    per coil and type, so cart_recon then sums the types."""
    shape = (res_freq * etl, shots, slice_num, rx, types)
    out = []
    for S in (Sx, Sy):
        X = np.asarray(S, dtype=np.float64).reshape(shape, order="F")
        Y = np.zeros((res_freq, etl * shots, slice_num, rx, types))
        for e in range(1, etl + 1):
            blk = X[(e - 1) * res_freq:e * res_freq]
            if epi and (-1) ** (e - 1) != 1:
                blk = blk[::-1]
            Y[:, (e - 1) * shots:e * shots] = blk
        out.append(Y.ravel(order="F"))
    return out[0], out[1]


def noise_scale(bandwidth, noise_level, nex, b0, rfreq, rphase, rslice, res_freq, res_phase, slice_num) -> float:
    """Standard deviation of the k-space noise, DoAddNoise.m:8-16."""
    bw_ref, b0_ref, nex_ref, adc_ref, vol_ref, noise_ref = 1e3, 1.5, 1.0, 1e4, 1e-9, 1.0
    return ((np.sqrt(bandwidth / bw_ref) * (noise_level / noise_ref))
            / (np.sqrt(nex / nex_ref) * (b0 / b0_ref) * (rfreq * rphase * rslice / vol_ref)
               * np.sqrt(res_freq * res_phase * slice_num / adc_ref)))


# Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), vectorised over
# counters: the generator mrs_DoAddNoise uses, restated for the tests.
_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox4x32_10(ctr, key):
    """ctr: uint32 array [..., 4]; key: (k0, k1). Returns uint32 [..., 4]."""
    c = [np.asarray(ctr, dtype=np.uint64)[..., i] for i in range(4)]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(_M0) * c[0]
        p1 = np.uint64(_M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + np.uint64(_W0)) & mask
        k1 = (k1 + np.uint64(_W1)) & mask
    return np.stack(c, axis=-1).astype(np.uint32)


def noise_normals(n: int, seed: int):
    """The pair of N(0, 1) values mrs_DoAddNoise adds (times scale) to
    Sx[i] and Sy[i]: Philox of counter (i lo, i hi, 0, 0) and key (seed lo,
    seed hi), two 53-bit uniforms, Box-Muller in fp64."""
    i = np.arange(n, dtype=np.uint64)
    ctr = np.stack([i & np.uint64(0xFFFFFFFF), i >> np.uint64(32), np.zeros_like(i), np.zeros_like(i)], axis=-1)
    w = philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)).astype(np.float64)
    u1 = (np.floor(w[:, 0] / 32) * 67108864.0 + np.floor(w[:, 1] / 64) + 0.5) / 9007199254740992.0
    u2 = (np.floor(w[:, 2] / 32) * 67108864.0 + np.floor(w[:, 3] / 64) + 0.5) / 9007199254740992.0
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)
