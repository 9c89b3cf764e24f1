"""Sequence lines in MRiLab's VSeq format.

MRiLab compiles a pulse-sequence XML into a MATLAB function that returns,
for one TR, the rf/Gz/Gy/Gx/ADC/Ext event lines, the sorted event times
``ts``, their unique values ``uts`` and the 6xN ``flags`` matrix
(PSD/3D/GradientEcho/PSD_GRE3D/PSD_GRE3D.m:258-359). DoPulseGen then
concatenates all TRs into the VSeq lines (Src/Main/DoPulseGen.m:26-117).
This module rebuilds that format in numpy so the engine can be driven
without MATLAB:

* :class:`TR` collects one TR's events and :meth:`TR.finalize` applies the
  PSD epilogue (drop events outside [0, TR], stable sort, unique times,
  zero the first and last sample of every line).
* :func:`concat` joins TRs into VSeq lines.
* Element helpers restate the waveform generators the PSDs call:
  ``rf_sinc`` (Macro/SeqElem/rf/rfSinc.m:2-62), ``rf_rect`` (rfRect.m),
  ``rf_gaussian`` (rfGaussian.m), ``b1_scaling`` (Src/Main/DoB1Scaling.m),
  ``std_trap`` (Macro/SeqElem/STD/StdTrap.m), ``ext_bit``
  (Macro/SeqElem/Ext/ExtBit.m with STD/StdBit.m).

Times passed in are seconds. Generators that want an exact time grid pass
integer multiples of the grid step so coincident events compare equal.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GYRO = 267538030.3797068  # rad/s/T, VObj.Gyro of Resources/VObj/*.mat


def b1_scaling(pulse: np.ndarray, dt: float, fa_deg: float, gyro: float = GYRO) -> float:
    """rf gain that gives the flip angle under the decoder's min-|.| hold.

    DoB1Scaling.m:6-10: a sample is flagged when it is the smaller end of
    a neighbouring interval; the flagged samples times dt give the area.
    """
    p = np.asarray(pulse, dtype=np.float64)
    flag = np.zeros(p.shape)
    ind = np.diff(np.abs(p))
    flag[np.nonzero(ind < 0)[0] + 1] = 1
    flag[:-1][ind >= 0] = 1
    return (fa_deg / 180.0) * np.pi / (gyro * np.sum(p * flag) * dt)


@dataclass
class RF:
    t: np.ndarray
    amp: np.ndarray
    phase: np.ndarray
    freq: np.ndarray
    coil: np.ndarray


def _rf_finish(t, shape, fa, dt, phase, freq, coil, gyro):
    shape = shape.copy()
    shape[0] = 0.0
    shape[-1] = 0.0
    amp = b1_scaling(shape, dt, fa, gyro) * shape
    ph = np.full(t.shape, float(phase))
    fr = np.full(t.shape, float(freq))
    ph[0] = ph[-1] = 0.0
    fr[0] = fr[-1] = 0.0
    return RF(t, amp, ph, fr, np.full(t.shape, float(coil)))


def rf_sinc(t: np.ndarray, fa: float, tbp: float = 4.0, apod: str = "Non", phase: float = 0.0,
            freq: float = 0.0, coil: int = 1, gyro: float = GYRO) -> RF:
    """Sinc pulse on the sample times ``t`` (rfSinc.m:25-50)."""
    t = np.asarray(t, dtype=np.float64)
    a = {"Non": 0.0, "Hamming": 0.46, "Hanning": 0.5}[apod]
    t0s, t1s = t[0], t[-1]
    tc = t - (t1s - t0s) / 2 - t0s
    t0 = (t1s - t0s) / tbp
    with np.errstate(invalid="ignore", divide="ignore"):
        shape = (t0 * np.sin(np.pi * tc / t0)) / (np.pi * tc) * ((1 - a) + a * np.cos(np.pi * tc / (tbp * t0)))
    shape[np.isnan(shape)] = np.nanmax(shape)
    dt = (t1s - t0s) / (len(t) - 1)
    return _rf_finish(t, shape, fa, dt, phase, freq, coil, gyro)


def rf_rect(t: np.ndarray, fa: float, phase: float = 0.0, freq: float = 0.0, coil: int = 1,
            gyro: float = GYRO) -> RF:
    """Hard pulse (rfRect.m:21-33)."""
    t = np.asarray(t, dtype=np.float64)
    dt = (t[-1] - t[0]) / (len(t) - 1)
    return _rf_finish(t, np.ones(t.shape), fa, dt, phase, freq, coil, gyro)


def rf_gaussian(t: np.ndarray, fa: float, phase: float = 0.0, freq: float = 0.0, coil: int = 1,
                gyro: float = GYRO) -> RF:
    """Gaussian pulse, sigma = duration / 7.434 (rfGaussian.m:18-33)."""
    t = np.asarray(t, dtype=np.float64)
    sigma = (t[-1] - t[0]) / 7.434
    shape = np.exp(-(t - (t[-1] - t[0]) / 2 - t[0]) ** 2 / (2 * sigma ** 2))
    dt = (t[-1] - t[0]) / (len(t) - 1)
    return _rf_finish(t, shape, fa, dt, phase, freq, coil, gyro)


def std_trap(t_start, t_end, t_up, t_down, amp, s_up=2, s_flat=2, s_down=2):
    """Trapezoid corner samples (StdTrap.m:2-37): up-ramp, flat, down-ramp,
    with duplicate times removed (first occurrence kept)."""
    g = np.concatenate([np.linspace(0, amp, s_up), np.full(s_flat, float(amp)), np.linspace(amp, 0, s_down)])
    t = np.concatenate([np.linspace(t_start, t_up, s_up), np.linspace(t_up, t_down, s_flat),
                        np.linspace(t_down, t_end, s_down)])
    ut, first = np.unique(t, return_index=True)
    return g[first], ut


def ext_bit(t_start, min_upd_rate, code):
    """Ext bit (ExtBit.m:272-277, StdBit.m): three samples [0 code 0]."""
    t = np.array([t_start, t_start + min_upd_rate, t_start + 2 * min_upd_rate])
    return np.array([0.0, float(code), 0.0]), t


@dataclass
class TR:
    """Events of one TR; times relative to the TR start, in seconds."""
    duration: float
    rf_t: list = field(default_factory=list)
    rf_amp: list = field(default_factory=list)
    rf_phase: list = field(default_factory=list)
    rf_freq: list = field(default_factory=list)
    rf_coil: list = field(default_factory=list)
    gz: list = field(default_factory=lambda: [[], []])
    gy: list = field(default_factory=lambda: [[], []])
    gx: list = field(default_factory=lambda: [[], []])
    adc: list = field(default_factory=lambda: [[], []])
    ext: list = field(default_factory=lambda: [[], []])

    def add_rf(self, rf: RF):
        self.rf_t.append(rf.t); self.rf_amp.append(rf.amp); self.rf_phase.append(rf.phase)
        self.rf_freq.append(rf.freq); self.rf_coil.append(rf.coil)

    @staticmethod
    def _add(line, t, v):
        line[0].append(np.asarray(t, dtype=np.float64).ravel())
        line[1].append(np.asarray(v, dtype=np.float64).ravel())

    def add_gz(self, t, v): self._add(self.gz, t, v)
    def add_gy(self, t, v): self._add(self.gy, t, v)
    def add_gx(self, t, v): self._add(self.gx, t, v)
    def add_adc(self, t, v): self._add(self.adc, t, v)
    def add_ext(self, t, v): self._add(self.ext, t, v)

    def finalize(self) -> dict:
        """The PSD epilogue of PSD_GRE3D.m:258-359 for one Pulses section
        spanning [0, duration]."""
        T = float(self.duration)

        def cat(lst):
            return np.concatenate(lst) if lst else np.zeros(0)

        rf_t, rf_a, rf_p = cat(self.rf_t), cat(self.rf_amp), cat(self.rf_phase)
        rf_f, rf_c = cat(self.rf_freq), cat(self.rf_coil)
        lines = {}
        keep = (rf_t >= 0) & (rf_t <= T)
        rf_t, rf_a, rf_p, rf_f, rf_c = rf_t[keep], rf_a[keep], rf_p[keep], rf_f[keep], rf_c[keep]
        rf_a = np.where(np.abs(rf_a) < np.finfo(float).eps, 0.0, rf_a)
        for name, ln in (("Gz", self.gz), ("Gy", self.gy), ("Gx", self.gx), ("ADC", self.adc), ("Ext", self.ext)):
            t, v = cat(ln[0]), cat(ln[1])
            k = (t >= 0) & (t <= T)
            lines[name] = (t[k], v[k])
        for name in ("Gz", "Gy", "Gx", "ADC", "Ext"):
            if lines[name][0].size == 0:
                raise ValueError(f"{name} sequence line can not be empty")
        if rf_t.size == 0:
            raise ValueError("rf sequence line can not be empty")
        order = ["rf", "Gz", "Gy", "Gx", "ADC", "Ext"]
        times = [np.array([0.0, T]), rf_t] + [lines[n][0] for n in order[1:]]
        flags = [np.zeros((6, 2))]
        for j, tt in enumerate(times[1:]):
            f = np.zeros((6, tt.size))
            f[j, :] = 1
            flags.append(f)
        ts = np.concatenate(times)
        fl = np.concatenate(flags, axis=1)
        idx = np.argsort(ts, kind="stable")
        ts = ts[idx]
        fl = fl[:, idx]
        uts = np.unique(ts)
        o = np.argsort(rf_t, kind="stable")
        out = {"rfAmp": rf_a[o], "rfPhase": rf_p[o], "rfFreq": rf_f[o], "rfCoil": rf_c[o]}
        for n in order[1:]:
            t, v = lines[n]
            out[n] = v[np.argsort(t, kind="stable")].copy()
        for n in ("rfAmp", "rfPhase", "rfFreq", "Gz", "Gy", "Gx", "ADC", "Ext"):
            out[n][0] = 0.0
            out[n][-1] = 0.0
        out["uts"], out["ts"], out["flags"] = uts, ts, fl
        return out


def concat(trs: list[dict], multitransmit: bool = False) -> dict:
    """Join finalized TRs into the VSeq lines (DoPulseGen.m:42-117)."""
    seq = {
        "rfAmpLine": np.concatenate([d["rfAmp"] for d in trs]),
        "rfPhaseLine": np.concatenate([d["rfPhase"] for d in trs]),
        "rfFreqLine": np.concatenate([d["rfFreq"] for d in trs]),
        "rfCoilLine": np.concatenate([d["rfCoil"] for d in trs]),
        "GzAmpLine": np.concatenate([d["Gz"] for d in trs]),
        "GyAmpLine": np.concatenate([d["Gy"] for d in trs]),
        "GxAmpLine": np.concatenate([d["Gx"] for d in trs]),
        "ADCLine": np.concatenate([d["ADC"] for d in trs]),
        "ExtLine": np.concatenate([d["Ext"] for d in trs]),
        "utsLine": np.concatenate([d["uts"] for d in trs]),
        "tsLine": np.concatenate([d["ts"] for d in trs]),
        "flagsLine": np.asfortranarray(np.concatenate([d["flags"] for d in trs], axis=1)),
    }
    if not multitransmit:
        seq["rfCoilLine"] = np.ones_like(seq["rfCoilLine"])  # DoPulseGen.m:111-113
    for k, v in seq.items():
        seq[k] = np.asfortranarray(v, dtype=np.float64)
    return seq


def held_area(t: np.ndarray, g: np.ndarray) -> float:
    """Area of a sampled line as the decoder holds it: each interval takes the
    endpoint of smaller magnitude (DoScanAtCPU.cpp:322-329)."""
    a = np.where(np.abs(g[:-1]) <= np.abs(g[1:]), g[:-1], g[1:])
    return float(np.sum(a * np.diff(t)))


def sample_pwl(breaks_t, breaks_g, t):
    """A piecewise-linear waveform given by breakpoints, sampled at ``t``."""
    return np.interp(t, breaks_t, breaks_g, left=0.0, right=0.0)
