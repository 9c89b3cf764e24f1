"""Seeded synthetic scans in MRiLab's struct layout.

Each builder returns a :class:`Scan` holding the seven structs the scan
engine reads (VObj, VMag, VCoi, VCtl, VSeq, VVar, VSig), laid out and typed
as Src/Main/DoPreScan.m and Src/Main/DoDataTypeConv.m leave them:

* arrays are Fortran-ordered [R, C, Z, SpinNum, TypeNum];
* Mz starts at Rho / SpinNum, Mx = My = 0 (DoPreScan.m:409-415);
* positions follow DoPreScan.m's meshgrid: Gxgrid varies along columns,
  Gygrid along rows, Gzgrid along slices, in metres;
* VVar starts at zero (DoPreScan.m:517-547) with TRCount = 1
  (DoDataTypeConv.m:125);
* VSig.Sx/Sy have SignalNum*TypeNum*RxCoilNum entries
  (DoDataTypeConv.m:20-30).

Workloads (sizes are parameters so tests can shrink them):

* :func:`gre2d` -- spoiled 2-D gradient echo, sinc excitation, uniform 10 us
  grid (the single-slice parity case).
* :func:`bssfp3d` -- 3-D balanced SSFP, hard pulse with alternating phase,
  rfRef demodulation, several receive coils.
* :func:`mt_cest` -- three-pool exchange (free water, myelin water, MT bound
  pool with a super-Lorentzian line) under Gaussian saturation trains at a
  series of off-resonance offsets (a Z-spectrum), for the multi-pool engine.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass

import numpy as np

from . import psd
from .psd import GYRO

F = "F"


@dataclass
class Scan:
    VObj: dict
    VMag: dict
    VCoi: dict
    VCtl: dict
    VSeq: dict
    VVar: dict
    VSig: dict
    name: str = ""

    def copy(self) -> "Scan":
        return copy.deepcopy(self)

    @property
    def structs(self):
        return (self.VObj, self.VMag, self.VCoi, self.VCtl, self.VSeq, self.VVar, self.VSig)

    @property
    def n_iso(self) -> int:
        return int(self.VObj["Mz"].size)

    def n_steps(self) -> int:
        """Decoded steps with dt > 0 (one per unique time point that is
        followed by a later point of the same TR)."""
        uts = self.VSeq["utsLine"]
        return int(np.count_nonzero(np.diff(uts) > 0))


def tissue_maps(shape, rng, n_types=1):
    """T1 ~ U[0.3, 2.0] s, T2 ~ U[0.02, 0.3] s clipped below T1,
    PD ~ U[0, 1], dB0 = 2*pi*N(0, 20 Hz)/gamma T."""
    full = tuple(shape) + (n_types,)
    T1 = rng.uniform(0.3, 2.0, full)
    T2 = np.minimum(rng.uniform(0.02, 0.3, full), 0.99 * T1)
    Rho = rng.uniform(0.0, 1.0, full)
    dB0 = 2 * np.pi * rng.normal(0.0, 20.0, tuple(shape)) / GYRO
    return Rho, T1, T2, dB0


def make_structs(Rho, T1, T2, dB0, res, spin_num=1, dWRnd=None, cs=None,
                 rx=None, tx=None, seq=None, tr_num=1, run_mode=0, max_threads=64,
                 multitransmit=False, master_tx=1):
    """Assemble the seven structs from maps (all [R,C,Z(,Ty)] arrays)."""
    R, C_, Z = dB0.shape
    Ty = Rho.shape[3]
    S = int(spin_num)
    f32 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float32))
    xv = (np.arange(C_) - C_ // 2) * res[0]
    yv = (np.arange(R) - R // 2) * res[1]
    zv = (np.arange(Z) - Z // 2) * res[2]
    Gx, Gy, Gz = np.meshgrid(xv, yv, zv, indexing="xy")  # MATLAB meshgrid: x along columns
    Mz = np.zeros((R, C_, Z, S, Ty))
    for t in range(Ty):
        Mz[..., t] = np.repeat((Rho[..., t].astype(np.float32).astype(np.float64) / S)[..., None], S, axis=3)
    VObj = {"Gyro": np.float64(GYRO), "SpinNum": np.int32(S), "TypeNum": np.int32(Ty),
            "Mx": f32(np.zeros_like(Mz)), "My": f32(np.zeros_like(Mz)), "Mz": f32(Mz),
            "Rho": f32(Rho), "T1": f32(T1), "T2": f32(T2),
            "ChemShift": np.float64(0.0), "XDimRes": np.float64(res[0]), "YDimRes": np.float64(res[1]),
            "ZDimRes": np.float64(res[2])}
    if dWRnd is None:
        dWRnd = np.zeros((R, C_, Z, S, Ty))
    VMag = {"dB0": f32(dB0), "dWRnd": f32(dWRnd), "Gzgrid": f32(Gz), "Gygrid": f32(Gy),
            "Gxgrid": f32(Gx), "FRange": np.ones((R, C_, Z))}
    if tx is None:
        VCoi = {"TxCoilmg": f32(np.ones((R, C_, Z))), "TxCoilpe": f32(np.zeros((R, C_, Z))),
                "TxCoilNum": np.int32(1), "TxCoilDefault": np.float64(1.0)}
    else:
        mg, pe = tx
        VCoi = {"TxCoilmg": f32(mg), "TxCoilpe": f32(pe), "TxCoilNum": np.int32(mg.shape[3] if mg.ndim == 4 else 1),
                "TxCoilDefault": np.float64(0.0)}
    if rx is None:
        VCoi.update({"RxCoilx": f32(np.ones((R, C_, Z))), "RxCoily": f32(np.zeros((R, C_, Z))),
                     "RxCoilNum": np.int32(1), "RxCoilDefault": np.float64(1.0)})
    else:
        cx, cy = rx
        VCoi.update({"RxCoilx": f32(cx), "RxCoily": f32(cy), "RxCoilNum": np.int32(cx.shape[3]),
                     "RxCoilDefault": np.float64(0.0)})
    if cs is None:
        cs = np.zeros(Ty)
    VCtl = {"CS": np.asarray(cs, dtype=np.float64), "TRNum": np.int32(tr_num),
            "RunMode": np.int32(run_mode), "MaxThreadNum": np.int32(max_threads),
            "GPUIndex": np.int32(0), "MultiTransmit": "on" if multitransmit else "off",
            "MasterTxCoil": int(master_tx)}
    Tx = int(VCoi["TxCoilNum"])
    VVar = {"rfAmp": np.zeros(Tx), "rfPhase": np.zeros(Tx), "rfFreq": np.zeros(Tx),
            "rfCoil": np.float64(0), "rfRef": np.float64(0), "GzAmp": np.float64(0),
            "GyAmp": np.float64(0), "GxAmp": np.float64(0), "ADC": np.float64(0),
            "Ext": np.float64(0), "t": np.float64(0), "dt": np.float64(0),
            "Kz": np.float64(0), "Ky": np.float64(0), "Kx": np.float64(0),
            "KzKeeper": np.float64(0), "KyKeeper": np.float64(0), "KxKeeper": np.float64(0),
            "gpuFetch": np.float64(0)}
    for k in ("rfi", "Gzi", "Gyi", "Gxi", "ADCi", "Exti", "utsi"):
        VVar[k] = np.int32(0)
    VVar["TRCount"] = np.int32(1)
    N = int(np.count_nonzero(seq["ADCLine"] == 1))
    Rx = int(VCoi["RxCoilNum"])
    VSig = {"Sx": np.zeros(N * Ty * Rx), "Sy": np.zeros(N * Ty * Rx), "Kz": np.zeros(N),
            "Ky": np.zeros(N), "Kx": np.zeros(N), "SignalNum": np.int32(N)}
    if run_mode == 1:
        nuts = seq["utsLine"].size
        shape6 = (R, C_, Z, S, Ty, nuts)
        VSig["Mx"] = np.zeros(shape6, dtype=np.float32, order=F)
        VSig["My"] = np.zeros(shape6, dtype=np.float32, order=F)
        VSig["Mz"] = np.zeros(shape6, dtype=np.float32, order=F)
        VSig["Muts"] = np.zeros(nuts)
    else:
        VSig["Mx"] = VSig["My"] = VSig["Mz"] = np.zeros(1, dtype=np.float32)
        VSig["Muts"] = np.zeros(1)
    return Scan(VObj, VMag, VCoi, VCtl, seq, VVar, VSig)


def psd_scan(psd_xml: str, vctl: dict, shape=(16, 16, 4), spins=1, seed=3) -> Scan:
    """A scan whose VSeq comes from compiling the PSD XML ``psd_xml`` for the
    scan parameters ``vctl`` (engine.compile_psd): the phantom is ``shape``
    voxels of tissue_maps over the field of view vctl FOVFreq x FOVPhase x
    FOVSlice (columns x rows x slices), one uniform receive coil."""
    from mrilab_amd import engine  # the compiler lives in the engine library

    seq = engine.compile_psd(psd_xml, vctl)
    n_tr = seq.pop("TRNum")
    rng = np.random.default_rng(seed)
    Rho, T1, T2, dB0 = tissue_maps(shape, rng)
    R, C_, Z = shape
    res = (float(vctl["FOVFreq"]) / C_, float(vctl["FOVPhase"]) / R, float(vctl["FOVSlice"]) / Z)
    sc = make_structs(Rho, T1, T2, dB0, res, spin_num=spins, seq=seq, tr_num=n_tr)
    sc.name = f"psd_{shape[0]}x{shape[1]}x{shape[2]}_TR{n_tr}"
    return sc


def with_run_mode1(sc: Scan) -> Scan:
    """A copy of ``sc`` with VCtl.RunMode = 1 and the per-step capture arrays
    VSig.Mx/My/Mz [R,C,Z,S,Ty,n_uts] and VSig.Muts [n_uts] (DoPreScan.m:417-420)."""
    out = sc.copy()
    out.VCtl["RunMode"] = np.int32(1)
    nuts = out.VSeq["utsLine"].size
    shape6 = tuple(out.VObj["Mz"].shape) + (1,) * (5 - out.VObj["Mz"].ndim) + (nuts,)
    for k in ("Mx", "My", "Mz"):
        a = np.zeros(shape6, dtype=np.float32, order=F)
        a[..., 0] = out.VObj[k].reshape(shape6[:5], order="F")
        out.VSig[k] = a
    out.VSig["Muts"] = np.zeros(nuts)
    return out


# --------------------------------------------------------------------------
# Sequences on an exact time grid
# --------------------------------------------------------------------------

def _trap_breaks(k0, k1, ramp, amp):
    """Breakpoints (grid indices) of a trapezoid starting at k0, ending at k1."""
    return [k0, k0 + ramp, k1 - ramp, k1], [0.0, amp, amp, 0.0]


def _pwl_on_grid(segments, n):
    """Sum of trapezoids, each (k0, k1, ramp, amp), sampled at grid 0..n."""
    g = np.zeros(n + 1)
    kk = np.arange(n + 1)
    for k0, k1, ramp, amp in segments:
        bt, bg = _trap_breaks(k0, k1, ramp, amp)
        g += np.interp(kk, bt, bg, left=0.0, right=0.0)
    return g


def _held_area_grid(g, dt):
    a = np.where(np.abs(g[:-1]) <= np.abs(g[1:]), g[:-1], g[1:])
    return np.sum(a) * dt


def gre2d(n_pe=64, n_ro=64, fov=0.256, tr_steps=1000, te_steps=500, dt=10e-6, fa=30.0,
          rf_steps=100, tbp=4.0, seed=0, spoil=True, spins=1) -> Scan:
    """Spoiled 2-D gradient echo on a uniform grid of ``dt``.

    64x64x1 phantom, 1 isochromat per voxel, TR = tr_steps*dt (10 ms),
    TE = te_steps*dt (5 ms), sinc excitation (TBW 4, 1 ms), n_pe phase
    encodes of n_ro samples, one uniform receive coil. Every grid point
    carries a gradient sample, so utsLine is the uniform grid.
    """
    rng = np.random.default_rng(seed)
    shape = (n_pe, n_ro, 1)
    Rho, T1, T2, dB0 = tissue_maps(shape, rng)
    res = (fov / n_ro, fov / n_pe, 5e-3)
    gbar = GYRO / (2 * np.pi)
    dk_x, dk_y = 1.0 / fov, 1.0 / fov
    g_ro = dk_x / (gbar * dt)
    n = tr_steps
    k_ro0 = te_steps - n_ro // 2            # first ADC grid point
    k_ro1 = k_ro0 + n_ro - 1                # last ADC grid point
    ramp = 10
    ro_seg = (k_ro0 - 4 - ramp, k_ro1 + 4 + ramp, ramp, g_ro)
    pre_k0, pre_k1 = rf_steps + 10, rf_steps + 110
    # prephaser amplitude such that the held k-space is 0 at TE
    g_ro_w = _pwl_on_grid([ro_seg], n)
    area_ro_to_te = _held_area_grid(g_ro_w[: te_steps + 1], dt)
    g_unit = _pwl_on_grid([(pre_k0, pre_k1, ramp, 1.0)], n)
    area_unit = _held_area_grid(g_unit, dt)
    g_pre = -area_ro_to_te / area_unit
    gx = _pwl_on_grid([(pre_k0, pre_k1, ramp, g_pre), ro_seg], n)
    gz_ss = _pwl_on_grid([(0, rf_steps, 5, 2e-3), (rf_steps + 5, rf_steps + 55, 5, -2e-3 * 0.5 * (rf_steps - 5) / 45)], n)
    times = np.arange(n + 1) * dt
    trs = []
    for j in range(n_pe):
        ky = (j - n_pe // 2) * dk_y
        gy = _pwl_on_grid([(pre_k0, pre_k1, ramp, ky / (gbar * area_unit))], n)
        tr = psd.TR(duration=n * dt)
        tr.add_rf(psd.rf_sinc(times[: rf_steps + 1], fa, tbp=tbp))
        tr.add_gz(times, gz_ss)
        tr.add_gy(times, gy)
        tr.add_gx(times, gx)
        adc_k = np.arange(k_ro0 - 1, k_ro1 + 2)
        adc_v = np.ones(adc_k.size); adc_v[0] = adc_v[-1] = 0
        tr.add_adc(times[adc_k], adc_v)
        e, et = psd.ext_bit(0, dt, 0)
        tr.add_ext(times[[rf_steps, rf_steps + 1, rf_steps + 2]], [0, 1, 0])     # reset k (Ext 1)
        if spoil:
            k = n - 50
            tr.add_ext(times[[k, k + 1, k + 2]], [0, 6, 0])                     # ideal spoiler (Ext 6)
        trs.append(tr.finalize())
    seq = psd.concat(trs)
    dWRnd = None
    if spins > 1:
        q = np.linspace(0.01, 0.99, spins)
        dWRnd = np.zeros(shape + (spins, 1))
        for s in range(spins):
            dWRnd[..., s, 0] = 50.0 * np.tan(np.pi * (q[s] - 0.5))   # DoPreScan.m:381-390 with 1/T2*-1/T2 = 50/s
    sc = make_structs(Rho, T1, T2, dB0, res, spin_num=spins, dWRnd=dWRnd, seq=seq, tr_num=n_pe)
    sc.name = f"gre2d_{n_pe}x{n_ro}_S{spins}"
    return sc


def gaussian_coils(shape, res, n_coils, rng, sigma_frac=0.4):
    """Receive coils: Gaussian blobs (sigma = sigma_frac * FOV) centred on a
    ring of radius FOV/2 in the central slice, each with a random phase."""
    R, C_, Z = shape
    xv = (np.arange(C_) - C_ // 2) * res[0]
    yv = (np.arange(R) - R // 2) * res[1]
    zv = (np.arange(Z) - Z // 2) * res[2]
    X, Y, Zg = np.meshgrid(xv, yv, zv, indexing="xy")
    fov = max(C_ * res[0], R * res[1])
    sig = sigma_frac * fov
    cx = np.zeros(shape + (n_coils,))
    cy = np.zeros(shape + (n_coils,))
    for c in range(n_coils):
        ang = 2 * np.pi * c / n_coils
        x0, y0 = 0.5 * fov * np.cos(ang), 0.5 * fov * np.sin(ang)
        mag = np.exp(-((X - x0) ** 2 + (Y - y0) ** 2 + Zg ** 2) / (2 * sig ** 2))
        ph = rng.uniform(-np.pi, np.pi)
        cx[..., c] = mag * np.cos(ph)
        cy[..., c] = mag * np.sin(ph)
    return cx, cy


def bssfp3d(shape=(128, 128, 64), n_ro=128, n_pe_y=None, n_pe_z=None, spins=4, n_coils=8,
            tr_steps=500, rf_steps=10, dt=10e-6, fa=45.0, seed=1, n_tr=None,
            fov=(0.256, 0.256, 0.128)) -> Scan:
    """3-D balanced SSFP on a uniform grid of ``dt``.

    Phantom [R, C, Z] with ``spins`` isochromats per voxel spread uniformly
    over +-pi of intravoxel dephasing per TR; hard ``fa`` pulse of
    ``rf_steps`` grid steps whose phase alternates 0/pi; Ext 7 at the pulse
    centre sets rfRef so the signal is demodulated; n_pe_y x n_pe_z phase
    encodes (default: rows x slices) of ``n_ro`` samples at TE = TR/2;
    ``n_coils`` Gaussian receive coils. ``n_tr`` truncates the scan to its
    first TRs (for tests and bounded CPU baselines).
    """
    rng = np.random.default_rng(seed)
    R, C_, Z = shape
    n_pe_y = R if n_pe_y is None else n_pe_y
    n_pe_z = Z if n_pe_z is None else n_pe_z
    Rho, T1, T2, dB0 = tissue_maps(shape, rng)
    res = (fov[0] / C_, fov[1] / R, fov[2] / Z)
    gbar = GYRO / (2 * np.pi)
    n = tr_steps
    te = n // 2
    TR = n * dt
    ramp = 10
    k_ro0 = te - n_ro // 2
    k_ro1 = k_ro0 + n_ro - 1
    g_ro = (1.0 / fov[0]) / (gbar * dt)
    ro_seg = (k_ro0 - 2 - ramp, k_ro1 + 2 + ramp, ramp, g_ro)
    pre = (rf_steps + 5, k_ro0 - 2 - ramp - 2)          # prephase window
    post = (k_ro1 + 2 + ramp + 2, n - 5)                 # rewind window
    g_ro_w = _pwl_on_grid([ro_seg], n)
    area_ro_to_te = _held_area_grid(g_ro_w[: te + 1], dt)
    area_ro_tot = _held_area_grid(g_ro_w, dt)
    area_pre_unit = _held_area_grid(_pwl_on_grid([(pre[0], pre[1], ramp, 1.0)], n), dt)
    area_post_unit = _held_area_grid(_pwl_on_grid([(post[0], post[1], ramp, 1.0)], n), dt)
    g_pre = -area_ro_to_te / area_pre_unit
    g_post = -(area_ro_tot - area_ro_to_te) / area_post_unit
    gx = _pwl_on_grid([(pre[0], pre[1], ramp, g_pre), ro_seg, (post[0], post[1], ramp, g_post)], n)
    times = np.arange(n + 1) * dt
    total = n_pe_y * n_pe_z
    count = total if n_tr is None else min(n_tr, total)
    # Every TR has the same event times; only the phase-encode amplitudes
    # and the rf phase change. Finalize one template TR (unit phase-encode
    # amplitude, rf phase pi) and scale its lines per TR.
    unit_y = _pwl_on_grid([(pre[0], pre[1], ramp, 1.0), (post[0], post[1], ramp, -area_pre_unit / area_post_unit)], n)
    tr = psd.TR(duration=TR)
    tr.add_rf(psd.rf_rect(times[: rf_steps + 1], fa, phase=np.pi))
    tr.add_gz(times, unit_y)
    tr.add_gy(times, unit_y)
    tr.add_gx(times, gx)
    adc_k = np.arange(k_ro0 - 1, k_ro1 + 2)
    adc_v = np.ones(adc_k.size); adc_v[0] = adc_v[-1] = 0
    tr.add_adc(times[adc_k], adc_v)
    kc = rf_steps // 2
    tr.add_ext(times[[kc - 1, kc, kc + 1]], [0, 7, 0])   # rfRef at the pulse centre (Ext 7)
    tpl = tr.finalize()
    it = np.arange(count)
    jz, jy = np.divmod(it, n_pe_y)
    ay = (jy - n_pe_y // 2) / fov[1] / (gbar * area_pre_unit)
    az = (jz - n_pe_z // 2) / fov[2] / (gbar * area_pre_unit)
    par = (it % 2).astype(np.float64)          # rf phase 0 / pi alternates
    tile = lambda a: np.tile(a, count)
    seq = {
        "rfAmpLine": tile(tpl["rfAmp"]),
        "rfPhaseLine": (par[:, None] * tpl["rfPhase"][None, :]).ravel(),
        "rfFreqLine": tile(tpl["rfFreq"]),
        "rfCoilLine": np.ones(count * tpl["rfCoil"].size),
        "GzAmpLine": (az[:, None] * tpl["Gz"][None, :]).ravel(),
        "GyAmpLine": (ay[:, None] * tpl["Gy"][None, :]).ravel(),
        "GxAmpLine": tile(tpl["Gx"]),
        "ADCLine": tile(tpl["ADC"]),
        "ExtLine": tile(tpl["Ext"]),
        "utsLine": tile(tpl["uts"]),
        "tsLine": tile(tpl["ts"]),
        "flagsLine": np.asfortranarray(np.tile(tpl["flags"], (1, count))),
    }
    for k_, v_ in seq.items():
        seq[k_] = np.asfortranarray(v_, dtype=np.float64)
    # uniform +-pi intravoxel spread per TR
    dWRnd = np.zeros(shape + (spins, 1))
    for s in range(spins):
        dWRnd[..., s, 0] = (-np.pi + 2 * np.pi * (s + 0.5) / spins) / TR
    cx, cy = gaussian_coils(shape, res, n_coils, rng)
    sc = make_structs(Rho, T1, T2, dB0, res, spin_num=spins, dWRnd=dWRnd, seq=seq,
                      rx=(cx, cy), tr_num=count)
    sc.name = f"bssfp3d_{R}x{C_}x{Z}_S{spins}_Rx{n_coils}_TR{count}"
    return sc


def mt_cest(shape=(256, 256, 1), n_offsets=40, ppm_range=50.0, b0=3.0, sat_time=2.0, pulse=0.02,
            gap=0.02, sat_fa=300.0, dt=10e-6, seed=2, t2b=10e-6, spins=1, n_coils=1, per_offset=False,
            crop=None, offset_index=None):
    """Three-pool exchange with MT and CEST-style saturation.

    Pools: 0 free water, 1 myelin water (free), 2 MT bound pool
    (TypeFlag 1, T2b = 10 us, super-Lorentzian line). Pool fractions come
    from Dirichlet(8, 1, 1); exchange rates out of the larger pool are
    U[1, 30]/s and the reverse rates follow detailed balance. For each of
    ``n_offsets`` offsets (linspace over +-ppm_range ppm at ``b0`` tesla) one
    TR holds a ``sat_time`` train of Gaussian pulses (``pulse`` long,
    ``gap`` apart, ``sat_fa`` degrees each) at that offset, an ideal spoiler
    (Ext 6), a 90 deg hard readout pulse on resonance and one ADC sample.
    K is [R,C,Z,3,3]: K(a,b) is the rate from pool a into pool b.
    ``spins`` > 1 spreads each voxel's isochromats over +-25 Hz (dWRnd, same
    for every pool); ``n_coils`` > 1 adds Gaussian receive coils.
    ``per_offset``: instead of one scan whose TRs run the offsets in turn,
    a list of scans of one TR each (one offset, from equilibrium), sharing
    the phantom's VObj/VMag/VCoi/VCtl dicts, for the batched multi-pool
    entry point.
    ``crop``: (row slice, column slice) of the phantom drawn at ``shape``
    (the same tissue values as the full phantom's voxels there, receive
    coils of the cropped grid); ``offset_index``: keep only these offsets
    (indices into the full list).
    """
    rng = np.random.default_rng(seed)
    R, C_, Z = shape
    f = rng.dirichlet([8.0, 1.0, 1.0], size=shape)               # [R,C,Z,3]
    Rho = f.copy()
    T1 = np.empty(shape + (3,))
    T2 = np.empty(shape + (3,))
    T1[..., 0] = rng.uniform(0.8, 2.0, shape); T2[..., 0] = rng.uniform(0.05, 0.3, shape)
    T1[..., 1] = rng.uniform(0.3, 0.6, shape); T2[..., 1] = rng.uniform(0.01, 0.03, shape)
    T1[..., 2] = 1.0;                          T2[..., 2] = t2b
    dB0 = 2 * np.pi * rng.normal(0.0, 20.0, shape) / GYRO
    K = np.zeros(shape + (3, 3))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        k_ab = rng.uniform(1.0, 30.0, shape)                      # a -> b
        K[..., a, b] = k_ab
        K[..., b, a] = k_ab * f[..., a] / f[..., b]               # detailed balance
    gbar = GYRO / (2 * np.pi)
    offsets = np.linspace(-ppm_range, ppm_range, n_offsets) * 1e-6 * gbar * b0   # Hz
    if offset_index is not None:
        offsets = offsets[list(offset_index)]
        n_offsets = len(offsets)
    if crop is not None:
        sl = (crop[0], crop[1])
        f, Rho, T1, T2, dB0, K = f[sl], Rho[sl], T1[sl], T2[sl], dB0[sl], K[sl]
        shape = f.shape[:3]
        R, C_, Z = shape
    n_pulse = int(round(pulse / dt))
    n_gap = int(round(gap / dt))
    period = n_pulse + n_gap
    n_sat = int(round(sat_time / dt)) // period * period
    n_read = 40
    n = n_sat + n_read
    times = np.arange(n + 1) * dt
    tpl_pulse = psd.rf_gaussian(times[: n_pulse + 1], sat_fa)
    trs = []
    for off in offsets:
        tr = psd.TR(duration=n * dt)
        for k in range(n_sat // period):
            t0 = k * period
            tr.add_rf(psd.RF(times[t0: t0 + n_pulse + 1], tpl_pulse.amp, tpl_pulse.phase,
                             np.where(tpl_pulse.amp != 0, off, 0.0), tpl_pulse.coil))
        rd = psd.rf_rect(times[n_sat + 4: n_sat + 4 + 11], 90.0)
        tr.add_rf(rd)
        zero = np.zeros(n + 1)
        tr.add_gz(times[[0, n]], [0, 0])
        tr.add_gy(times[[0, n]], [0, 0])
        tr.add_gx(times, zero)
        k_adc = n_sat + 4 + 11 + 2
        tr.add_adc(times[[k_adc - 1, k_adc, k_adc + 1]], [0, 1, 0])
        tr.add_ext(times[[n_sat, n_sat + 1, n_sat + 2]], [0, 6, 0])          # ideal spoiler before readout
        tr.add_ext(times[[n - 4, n - 3, n - 2]], [0, 6, 0])                  # and after it
        trs.append(tr.finalize())
    seq = psd.concat(trs if not per_offset else trs[:1])
    res = (2e-3, 2e-3, 5e-3)
    dWRnd = np.zeros(shape + (spins, 3))
    for q in range(spins):
        dWRnd[..., q, :] = 2 * np.pi * 25.0 * (2 * (q + 0.5) / spins - 1)
    rx = gaussian_coils(shape, res, n_coils, rng) if n_coils > 1 else None
    sc = make_structs(Rho, T1, T2, dB0, res, spin_num=spins, dWRnd=dWRnd, rx=rx, seq=seq,
                      tr_num=1 if per_offset else n_offsets)
    sc.VObj["K"] = np.asfortranarray(K.astype(np.float32))
    sc.VObj["TypeFlag"] = np.array([0.0, 0.0, 1.0])
    sc.VCtl["GM_Flag"] = "on"
    sc.VCtl["CS"] = np.zeros(3)
    sc.name = f"mt_cest_{R}x{C_}x{Z}_off{n_offsets}_steps{n}"
    sc.offsets_hz = offsets
    if not per_offset:
        return sc
    scans = []
    for j, tr in enumerate(trs):
        seq_j = psd.concat([tr])
        o = Scan(sc.VObj, sc.VMag, sc.VCoi, sc.VCtl, seq_j, copy.deepcopy(sc.VVar), copy.deepcopy(sc.VSig),
                 name=f"{sc.name}_offset{j}")
        o.offset_hz = float(offsets[j])
        scans.append(o)
    return scans


def _fingerprint_template(t_rf, te, delta):
    """One finalized fingerprinting TR (hard pulse of 1 deg, ADC at TE) with
    a placeholder duration; returns it and the placeholder."""
    t_end = 0.0123456789
    tr = psd.TR(duration=t_end)
    t = np.array([0.0, delta, delta + t_rf, 2 * delta + t_rf])
    amp = np.array([0.0, 1.0, 1.0, 0.0]) * (np.pi / 180.0) / (GYRO * t_rf)
    tr.add_rf(psd.RF(t, amp, np.zeros(4), np.zeros(4), np.ones(4)))
    tr.add_gz([0.0, t_end], [0, 0])
    tr.add_gy([0.0, t_end], [0, 0])
    tr.add_gx([0.0, t_end], [0, 0])
    tr.add_adc([0.0, te, t_end], [0, 1, 0])
    tr.add_ext([0.0, t_end], [0, 0])
    return tr.finalize(), t_end


def fingerprint_set(flips_deg, tr_s, T1, T2, dB0, b1, spins=64, t_rf=0.5e-3, te=3e-3, delta=10e-6,
                    tpl=None) -> Scan:
    """One single-voxel fingerprinting scan: a train of hard pulses with
    flip ``flips_deg[k]`` and repetition time ``tr_s[k]``, one ADC sample at
    TE per TR (5 steps per TR: pulse edges, pulse, to TE, to TR). The
    ``spins`` isochromats spread uniformly over one cycle per 10 ms; B1+
    scale ``b1`` enters as TxCoilmg."""
    fl, t_end = tpl if tpl is not None else _fingerprint_template(t_rf, te, delta)
    flips = np.asarray(flips_deg, dtype=np.float64)
    trs = np.asarray(tr_s, dtype=np.float64)
    n = flips.size
    tsub = lambda a: np.where(a[None, :] == t_end, trs[:, None], a[None, :]).ravel()  # noqa: E731
    seq = {"rfAmpLine": (fl["rfAmp"][None, :] * flips[:, None]).ravel(),
           "rfPhaseLine": np.tile(fl["rfPhase"], n), "rfFreqLine": np.tile(fl["rfFreq"], n),
           "rfCoilLine": np.ones(fl["rfCoil"].size * n),
           "GzAmpLine": np.tile(fl["Gz"], n), "GyAmpLine": np.tile(fl["Gy"], n),
           "GxAmpLine": np.tile(fl["Gx"], n), "ADCLine": np.tile(fl["ADC"], n),
           "ExtLine": np.tile(fl["Ext"], n), "utsLine": tsub(fl["uts"]), "tsLine": tsub(fl["ts"]),
           "flagsLine": np.tile(fl["flags"], (1, n))}
    for k, v in seq.items():
        seq[k] = np.asfortranarray(v, dtype=np.float64)
    shape = (1, 1, 1)
    dW = np.zeros(shape + (spins, 1))
    dW[0, 0, 0, :, 0] = (-np.pi + 2 * np.pi * (np.arange(spins) + 0.5) / spins) / 0.010
    sc = make_structs(np.ones(shape + (1,)), np.full(shape + (1,), T1), np.full(shape + (1,), T2),
                      np.full(shape, dB0), (1e-3, 1e-3, 1e-3), spin_num=spins, dWRnd=dW, seq=seq, tr_num=n,
                      tx=(np.full(shape, b1), np.zeros(shape)))
    sc.name = f"fingerprint_{n}tr_S{spins}"
    return sc


def fingerprint_batch(n_sets=4096, spins=64, tr_min=500, tr_max=1500, seed=3, t_rf=0.5e-3, te=3e-3):
    """Batched uneven-length fingerprinting: ``n_sets`` single-voxel scans
    with T1~U[0.3,2] s, T2~U[0.02,0.3] s (< T1), dB0 ~ 2*pi*N(0,20 Hz)/gamma,
    B1 scale ~U[0.7,1.3]; train lengths ~U[tr_min, tr_max] TRs, flips a
    smoothed random walk in [5,70] deg, TR ~U[10,15] ms per TR."""
    rng = np.random.default_rng(seed)
    tpl = _fingerprint_template(t_rf, te, 10e-6)
    out = []
    for _ in range(n_sets):
        n = int(rng.integers(tr_min, tr_max + 1))
        walk = 5.0 + 65.0 * rng.random() + np.cumsum(rng.normal(0.0, 4.0, n))
        walk = np.abs((walk - 5.0) % 130.0 - 65.0)  # reflect into [0, 65]
        k = 25
        flips = np.convolve(np.pad(walk, (k // 2, k - 1 - k // 2), mode="edge"), np.ones(k) / k, "valid")
        flips = np.clip(70.0 - flips, 5.0, 70.0)
        trs = rng.uniform(10e-3, 15e-3, n)
        T1 = rng.uniform(0.3, 2.0)
        T2 = min(rng.uniform(0.02, 0.3), 0.99 * T1)
        dB0 = 2 * np.pi * rng.normal(0.0, 20.0) / GYRO
        b1 = rng.uniform(0.7, 1.3)
        out.append(fingerprint_set(flips, trs, T1, T2, dB0, b1, spins=spins, t_rf=t_rf, te=te, tpl=tpl))
    return out


@dataclass
class Excitation:
    """An rf-design problem for the adjoint: phantom structs, gradient
    waveform [n, 3] (Gx, Gy, Gz, T/m), target [nvox, 2], a starting rf
    [n, 2] (T) and the per-voxel parameters T1, T2 (s) and dB0 (T)."""
    VObj: dict
    VMag: dict
    VCoi: dict
    grad: np.ndarray
    target: np.ndarray
    rf: np.ndarray
    T1: np.ndarray
    T2: np.ndarray
    dB0: np.ndarray
    dt: float
    name: str = ""
    aux: tuple = ()   # VCtl, VSeq, VVar, VSig of a placeholder scan (struct packing only)

    @property
    def nvox(self) -> int:
        return int(self.target.shape[0])


def excitation(shape=(128, 128), spins=32, n_steps=2000, dt=4e-6, peak_b1=15e-6, fov=0.256, res_exc=0.02,
               n_turns=8, flip=30.0, seed=4) -> Excitation:
    """2-D selective excitation: a spiral-in gradient (k-space radius
    1/(2*res_exc) down to 0 in ``n_turns`` turns over ``n_steps`` x ``dt``),
    a smooth random complex starting rf of peak 2/3 * ``peak_b1``, and a
    seeded target: transverse magnetization sin(flip) * PD inside a random
    union of discs, 0 outside. Phantom maps as the single-slice case, with
    ``spins`` isochromats per voxel on Lorentzian quantiles (50/s)."""
    rng = np.random.default_rng(seed)
    R, C_ = shape
    shp = (R, C_, 1)
    Rho, T1, T2, dB0 = tissue_maps(shp, rng)
    res = (fov / C_, fov / R, 5e-3)
    q = np.linspace(0.01, 0.99, spins)
    dWRnd = np.zeros(shp + (spins, 1))
    for s in range(spins):
        dWRnd[..., s, 0] = 50.0 * np.tan(np.pi * (q[s] - 0.5))
    # a one-TR placeholder sequence only to assemble the structs
    tr = psd.TR(duration=1e-3)
    tr.add_rf(psd.rf_rect(np.array([0.0, 1e-4, 2e-4, 3e-4]), 1.0))
    for add in (tr.add_gz, tr.add_gy, tr.add_gx, tr.add_ext):
        add([0.0, 1e-3], [0, 0])
    tr.add_adc([0.0, 5e-4, 1e-3], [0, 1, 0])
    sc = make_structs(Rho, T1, T2, dB0, res, spin_num=spins, dWRnd=dWRnd, seq=psd.concat([tr.finalize()]))
    # spiral-in k-space trajectory and its gradient
    t = (np.arange(n_steps + 1)) * dt
    T = n_steps * dt
    kmax = 1.0 / (2 * res_exc)
    kr = kmax * (1 - t / T)
    ph = 2 * np.pi * n_turns * t / T
    kx, ky = kr * np.cos(ph), kr * np.sin(ph)
    gbar = GYRO / (2 * np.pi)
    grad = np.zeros((n_steps, 3))
    grad[:, 0] = np.diff(kx) / dt / gbar
    grad[:, 1] = np.diff(ky) / dt / gbar
    # smooth random complex rf
    n_knots = 24
    knots = rng.normal(size=(n_knots, 2))
    tk = np.linspace(0, n_steps - 1, n_knots)
    rf = np.stack([np.interp(np.arange(n_steps), tk, knots[:, 0]), np.interp(np.arange(n_steps), tk, knots[:, 1])], 1)
    rf *= (2.0 / 3.0) * peak_b1 / np.max(np.hypot(rf[:, 0], rf[:, 1]))
    # target: union of discs
    yy, xx = np.meshgrid(np.arange(R), np.arange(C_), indexing="ij")
    mask = np.zeros((R, C_), bool)
    for _ in range(4):
        cy, cx = rng.uniform(0.25, 0.75) * R, rng.uniform(0.25, 0.75) * C_
        rad = rng.uniform(0.08, 0.2) * min(R, C_)
        mask |= (yy - cy) ** 2 + (xx - cx) ** 2 < rad ** 2
    amp = np.sin(np.radians(flip)) * Rho[..., 0, 0].astype(np.float32) * mask
    target = np.zeros((R * C_, 2), np.float32)
    target[:, 1] = amp.ravel(order="F")
    f32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float32))  # noqa: E731
    return Excitation(sc.VObj, sc.VMag, sc.VCoi, f32(grad), f32(target), f32(rf),
                      f32(T1.ravel(order="F")), f32(T2.ravel(order="F")), f32(dB0.ravel(order="F")), float(dt),
                      name=f"excitation_{R}x{C_}_S{spins}_n{n_steps}", aux=(sc.VCtl, sc.VSeq, sc.VVar, sc.VSig))


def select_pools(sc: Scan, idx, cs_hz=None, t1_scale=None) -> Scan:
    """Copy of a multi-pool scan keeping (or repeating) pools ``idx``; K
    (GM layout, K[..., a, b] = rate a -> b), maps, TypeFlag, CS and the
    signal arrays follow. ``cs_hz`` replaces VCtl.CS, ``t1_scale``
    multiplies T1 per kept pool."""
    idx = list(idx)
    out = sc.copy()
    o = out.VObj
    for k in ("Mx", "My", "Mz", "Rho", "T1", "T2"):
        o[k] = np.asfortranarray(o[k][..., idx])
    if t1_scale is not None:
        o["T1"] = np.asfortranarray((o["T1"] * np.asarray(t1_scale, np.float32)).astype(np.float32))
    o["K"] = np.asfortranarray(o["K"][..., idx, :][..., idx])
    o["TypeFlag"] = np.asarray(o["TypeFlag"])[idx].astype(np.float64)
    o["TypeNum"] = np.int32(len(idx))
    out.VMag["dWRnd"] = np.asfortranarray(out.VMag["dWRnd"][..., idx])
    out.VCtl["CS"] = (np.asarray(out.VCtl["CS"])[idx] if cs_hz is None else np.asarray(cs_hz)).astype(np.float64)
    n = int(out.VSig["SignalNum"]) * len(idx) * int(out.VCoi["RxCoilNum"])
    out.VSig["Sx"], out.VSig["Sy"] = np.zeros(n), np.zeros(n)
    return out


def as_variant(sc: Scan, variant: str, self_rate: float = 0.0) -> Scan:
    """VObj.K of a GM-layout multi-pool scan rewritten in the layout of the
    MT, CEST or ME engine (BlochKernelMT.cuh:151-154,
    BlochKernelCEST.cuh:157-160/:379-384, BlochKernelME.cuh:327-333).
    MT keeps water <-> pool 1; CEST keeps water <-> every pool; ME keeps all
    rates and adds ``self_rate`` on the diagonal."""
    if variant == "GM":
        return sc
    out = sc.copy()
    K = np.asarray(sc.VObj["K"], dtype=np.float64)
    shp, Ty = K.shape[:3], K.shape[-1]
    if variant == "ME":
        Kn = np.zeros(shp + (Ty * Ty,))
        for a in range(Ty):
            for b in range(Ty):
                Kn[..., a * Ty + b] = K[..., a, b] + (self_rate if a == b else 0.0)
    elif variant == "MT":
        if Ty != 2:
            raise ValueError("the MT layout has two pools")
        Kn = np.zeros(shp + (4,))
        Kn[..., 1], Kn[..., 2] = K[..., 0, 1], K[..., 1, 0]
    elif variant == "CEST":
        if Ty < 3:
            raise ValueError("the CEST layout has at least three pools")
        Kn = np.zeros(shp + (2 * (Ty - 1),))
        Kn[..., 0], Kn[..., Ty - 1] = K[..., 0, 1], K[..., 1, 0]
        for t in range(2, Ty):
            Kn[..., t - 1], Kn[..., t + Ty - 2] = K[..., 0, t], K[..., t, 0]
    else:
        raise ValueError(variant)
    out.VObj["K"] = np.asfortranarray(Kn.astype(np.float32))
    return out


def multi_transmit(sc: Scan, scale, phase, freq=None, seed: int = 7, silent_after=None,
                   master: int = 1) -> Scan:
    """Copy of a single-Tx scan driven by len(scale) transmit coils in
    parallel (VCtl.MultiTransmit 'on'): every rf sample becomes one sample
    per coil at the same time point, coil c with amplitude scale[c] times the
    original, phase + phase[c] and frequency + freq[c] (Hz), in the order a
    PSD's rf elements of coils 1..Tx fire together (DoPulseGen.m:42-117,
    rfCoilLine = coil). Coil c's B1+ map is a smooth random magnitude in
    [0.7, 1.3] with a linear phase ramp. ``silent_after = (c, j)`` sets coil
    c's amplitude to 0 from rf sample j on (a coil switched off)."""
    out = sc.copy()
    ntx = len(scale)
    freq = np.zeros(ntx) if freq is None else np.asarray(freq, float)
    q = out.VSeq
    ts, fl = q["tsLine"], q["flagsLine"]
    cols = np.nonzero(fl[0] == 1)[0]
    n_rf = q["rfAmpLine"].size
    if cols.size != n_rf:
        raise ValueError("expected one rf flag per rf sample")
    # each rf column becomes ntx columns at the same time point
    rep = np.ones(ts.size, dtype=np.int64)
    rep[cols] = ntx
    q["tsLine"] = np.repeat(ts, rep)
    q["flagsLine"] = np.asfortranarray(np.repeat(fl, rep, axis=1))
    amp = np.empty((n_rf, ntx)); ph = np.empty((n_rf, ntx)); fr = np.empty((n_rf, ntx))
    for c in range(ntx):
        amp[:, c] = q["rfAmpLine"] * scale[c]
        ph[:, c] = q["rfPhaseLine"] + phase[c]
        fr[:, c] = q["rfFreqLine"] + freq[c]
    # a PSD's elements zero their first and last samples (rf*.m): keep the
    # phase and frequency of a zero-amplitude sample at 0 as they do
    zero = q["rfAmpLine"] == 0
    ph[zero] = 0.0
    fr[zero] = 0.0
    if silent_after is not None:
        c, j = silent_after
        amp[j:, c] = 0.0
    q["rfAmpLine"] = amp.ravel()
    q["rfPhaseLine"] = ph.ravel()
    q["rfFreqLine"] = fr.ravel()
    q["rfCoilLine"] = np.tile(np.arange(1, ntx + 1, dtype=np.float64), n_rf)
    for k in ("tsLine", "rfAmpLine", "rfPhaseLine", "rfFreqLine", "rfCoilLine"):
        q[k] = np.asfortranarray(q[k], dtype=np.float64)
    shape = out.VObj["Mz"].shape[:3] if out.VObj["Mz"].ndim >= 3 else out.VObj["Mz"].shape[:2] + (1,)
    R, C_, Z = (tuple(shape) + (1, 1, 1))[:3]
    rng = np.random.default_rng(seed)
    yy, xx, zz = np.meshgrid(np.linspace(-1, 1, R), np.linspace(-1, 1, C_), np.linspace(-1, 1, Z), indexing="ij")
    mg = np.empty((R, C_, Z, ntx))
    pe = np.empty((R, C_, Z, ntx))
    for c in range(ntx):
        a = rng.uniform(-1, 1, 4)
        mg[..., c] = 1.0 + 0.3 * np.tanh(a[0] * xx + a[1] * yy + a[2] * zz + a[3])
        b = rng.uniform(-1.5, 1.5, 3)
        pe[..., c] = b[0] * xx + b[1] * yy + b[2] * zz
    out.VCoi["TxCoilmg"] = np.asfortranarray(mg.astype(np.float32))
    out.VCoi["TxCoilpe"] = np.asfortranarray(pe.astype(np.float32))
    out.VCoi["TxCoilNum"] = np.int32(ntx)
    out.VCoi["TxCoilDefault"] = np.float64(0.0)
    out.VCtl["MultiTransmit"] = "on"
    out.VCtl["MasterTxCoil"] = int(master)
    for k in ("rfAmp", "rfPhase", "rfFreq"):
        out.VVar[k] = np.zeros(ntx)
    out.name = sc.name + f"_Tx{ntx}"
    return out


def two_species(sc: Scan, cs_hz: float = -440.0, seed: int = 11) -> Scan:
    """Copy of a single-type scan with a second species (TypeNum 2, e.g.
    fat at cs_hz from water, VCtl.CS = [0, cs_hz], as DoPreScan.m:120 sets it
    from VObj.ChemShift): its own Rho ~ U[0, 0.5], T1 ~ U[0.2, 0.4] s, T2 ~
    U[0.04, 0.1] s, the same intravoxel spread, signal per type."""
    out = sc.copy()
    rng = np.random.default_rng(seed)
    o = out.VObj
    shp = o["Rho"].shape[:3] if o["Rho"].ndim >= 3 else o["Rho"].shape[:2] + (1,)
    S = int(o["SpinNum"])
    rho2 = rng.uniform(0.0, 0.5, shp).astype(np.float32)
    t1 = rng.uniform(0.2, 0.4, shp).astype(np.float32)
    t2 = rng.uniform(0.04, 0.1, shp).astype(np.float32)
    for k, extra in (("Rho", rho2), ("T1", t1), ("T2", t2)):
        o[k] = np.asfortranarray(np.stack([o[k].reshape(shp), extra], axis=3))
    mz2 = np.repeat((rho2 / np.float32(S))[..., None], S, axis=3)
    zero = np.zeros_like(mz2)
    for k, extra in (("Mx", zero), ("My", zero), ("Mz", mz2)):
        o[k] = np.asfortranarray(np.stack([o[k].reshape(shp + (S,)), extra], axis=4).astype(np.float32))
    d = out.VMag["dWRnd"].reshape(shp + (S,))
    out.VMag["dWRnd"] = np.asfortranarray(np.stack([d, d], axis=4).astype(np.float32))
    o["TypeNum"] = np.int32(2)
    out.VCtl["CS"] = np.array([float(np.asarray(out.VCtl["CS"]).ravel()[0]), float(cs_hz)])
    n = int(out.VSig["SignalNum"]) * 2 * int(out.VCoi["RxCoilNum"])
    out.VSig["Sx"], out.VSig["Sy"] = np.zeros(n), np.zeros(n)
    out.name = sc.name + "_Ty2"
    return out


def sar_scan(base: Scan | None = None, n_tx: int = 2, n_samples: int = 8, seed: int = 5,
             body: bool = False) -> Scan:
    """A scan with the SAR engine's fields (DoCalSARAtCPU.cpp:115-166) on
    top of ``base`` (default: the 64x64 GRE): per Tx coil an E field per
    unit rf amplitude (V/m/T, Gaussian blobs of random sign per component),
    ECon = sigma/(2*density) ~ U[1e-4, 5e-4], VVar.SAR zero, and
    ``n_samples`` sample times spread over the sequence (the first at 0).
    MassDen is 1000 kg/m^3 everywhere, or with ``body`` a tissue ellipsoid
    (semi-axes 0.8, 0.65, 0.9 of the half field of view) of 1000 +- 100
    kg/m^3 in smooth variation inside and air (MassDen = ECon = 0) outside."""
    sc = (gre2d() if base is None else base).copy()
    rng = np.random.default_rng(seed)
    shape = sc.VObj["Mz"].shape[:3] if sc.VObj["Mz"].ndim >= 3 else sc.VObj["Mz"].shape[:2] + (1,)
    R, C_, Z = (tuple(shape) + (1, 1, 1))[:3]
    nvox = R * C_ * Z
    yy, xx, zz = np.meshgrid(np.linspace(-1, 1, R), np.linspace(-1, 1, C_), np.linspace(-1, 1, Z), indexing="ij")
    E = np.empty((3, R, C_, Z, n_tx))
    for c in range(n_tx):
        cy, cx = rng.uniform(-0.7, 0.7, 2)
        blob = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2 + 0.3 * zz ** 2) / (2 * 0.4 ** 2))
        for k in range(3):
            E[k, ..., c] = rng.choice([-1.0, 1.0]) * rng.uniform(2e5, 1e6) * blob
    if n_tx > 1:
        # parallel transmit: every rf event is sent on all coils at once
        # (multi-line rf: n_tx coincident rf events, one per coil, coil c
        # with amplitude weight w_c and phase offset p_c)
        q = sc.VSeq
        fl = np.asarray(q["flagsLine"]).reshape(6, -1, order="F")
        rep = np.where(fl[0] > 0, n_tx, 1)
        q["tsLine"] = np.repeat(q["tsLine"], rep)
        q["flagsLine"] = np.asfortranarray(np.repeat(fl, rep, axis=1))
        w = rng.uniform(0.5, 1.5, n_tx)
        ph = rng.uniform(-np.pi, np.pi, n_tx)
        n_rf = q["rfAmpLine"].size
        q["rfAmpLine"] = (np.repeat(q["rfAmpLine"], n_tx).reshape(n_rf, n_tx) * w).ravel()
        q["rfPhaseLine"] = (np.repeat(q["rfPhaseLine"], n_tx).reshape(n_rf, n_tx) + ph).ravel()
        q["rfFreqLine"] = np.repeat(q["rfFreqLine"], n_tx)
        q["rfCoilLine"] = np.tile(np.arange(1.0, n_tx + 1.0), n_rf)
    sc.VCoi["TxCoilNum"] = np.int32(n_tx)
    for k, name in enumerate("xyz"):
        sc.VCoi[f"TxE1{name}"] = np.asfortranarray(E[k].astype(np.float32))
    sc.VCoi["TxCoilmg"] = np.ones((R, C_, Z, n_tx), dtype=np.float32, order=F)
    sc.VCoi["TxCoilpe"] = np.zeros((R, C_, Z, n_tx), dtype=np.float32, order=F)
    econ = rng.uniform(1e-4, 5e-4, (R, C_, Z, 3))
    if body:
        inside = (yy / 0.65) ** 2 + (xx / 0.8) ** 2 + (zz / 0.9) ** 2 <= 1.0
        dens = 1000.0 + 100.0 * np.sin(3.0 * xx + 1.0) * np.cos(2.0 * yy - 0.5) * np.cos(1.5 * zz)
        sc.VObj["MassDen"] = np.asfortranarray(np.where(inside, dens, 0.0).astype(np.float32))
        econ[~inside] = 0.0
    else:
        sc.VObj["MassDen"] = np.full((R, C_, Z), 1000.0, dtype=np.float32, order=F)
    sc.VObj["ECon"] = np.asfortranarray(econ.astype(np.float32))
    for f in ("rfAmp", "rfPhase", "rfFreq"):
        sc.VVar[f] = np.zeros(n_tx)
    sc.VVar["SAR"] = np.zeros((R, C_, Z, 3), dtype=np.float32, order=F)
    sc.VVar["SARi"] = np.int32(0)
    uts = sc.VSeq["utsLine"]
    total = float(np.sum(np.diff(uts)[np.diff(uts) > 0]))
    # the last sample just inside the end: Muts sums dt step by step
    sc.VSig["tSample"] = np.linspace(0.0, total * (1 - 1e-9), n_samples)
    sc.VSig["tRealSample"] = np.zeros(n_samples)
    sc.VSig["SAR"] = np.zeros((R, C_, Z, n_samples), dtype=np.float32, order=F)
    sc.VSig["Muts"] = np.zeros(uts.size)
    sc.name = f"sar_{sc.name}_tx{n_tx}"
    return sc
