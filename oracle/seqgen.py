"""Sequence compiler oracle -- TEST INFRASTRUCTURE ONLY.

A Python restatement of MRiLab's PSD pipeline, the checker of
mrs_seq_compile: DoWriteXML2m.m:148-337 (the generated PSD function for one
TR), DoPulseGen.m:18-117 (the slice x phase loop, dummy pulses,
concatenation) and the Macro/SeqElem elements rfSinc.m, rfRect.m,
rfGaussian.m, rfFermi.m, GzCartesian.m, GyCartesian.m, GzSelective.m,
GxCartesian.m, ADCCartesian.m, the EPI (GxEPI.m, GyEPI.m, ADCEPI.m), FSE
(GxFSE.m, GyFSE.m, ADCFSE.m), radial (GxRadial.m, GyRadial.m, ADCRadial.m)
and spiral (GxSpiral.m, GySpiral.m, ADCSpiral.m) elements, ExtBit.m, with
StdTrap.m, StdBit.m, ADCBlock.m, the area trapezoids and
Src/Main/DoB1Scaling.m. It parses the XML with xml.etree, evaluates
attribute expressions with Python's own arithmetic (MATLAB's precedence and
left-to-right order coincide with Python's for + - * / and unary minus; ^
becomes **), and builds arrays with numpy. MATLAB semantics it keeps:
linspace as d1 + (k*(d2-d1))/(n-1) with exact ends, unique() keeping the
first occurrence, stable sorts, round half away from zero.
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET

import numpy as np

EPS = 2.220446049250313e-16


def linspace(d1, d2, n):
    """MATLAB linspace."""
    n = int(math.floor(n))
    if n < 1:
        return np.zeros(0)
    if n == 1:
        return np.array([float(d2)])
    k = np.arange(n, dtype=np.float64)
    y = d1 + (k * (d2 - d1)) / (n - 1)
    y[0] = d1
    y[-1] = d2
    return y


def unique_first(t, *vals):
    """[t, m] = unique(t); vals(m)."""
    t = np.asarray(t, dtype=np.float64)
    order = np.argsort(t, kind="stable")
    ts = t[order]
    keep = np.ones(ts.size, dtype=bool)
    keep[1:] = ts[1:] != ts[:-1]
    idx = order[keep]
    return (t[idx],) + tuple(np.asarray(v)[idx] for v in vals)


def std_trap(t0, t1, tu, td, amp, su, sf, sd):
    t = np.concatenate([linspace(t0, tu, su), linspace(tu, td, sf), linspace(td, t1, sd)])
    g = np.concatenate([linspace(0.0, amp, su), np.full(int(math.floor(sf)), float(amp)), linspace(amp, 0.0, sd)])
    t, g = unique_first(t, g)
    return g, t


def b1_gain(p, dt, fa, gyro):
    """DoB1Scaling.m."""
    p = np.asarray(p, dtype=np.float64)
    flag = np.zeros(p.size)
    d = np.diff(np.abs(p))
    flag[np.nonzero(d < 0)[0] + 1] = 1
    flag[:-1][d >= 0] = 1
    s = 0.0
    for x in (p * flag):
        s = s + x
    return ((fa / 180) * math.pi) / (gyro * s * dt)


def duplicate(vals, t, P):
    dups = max(1.0, P["Duplicates"])
    sp = max(0.0, P["DupSpacing"])
    if dups == 1 or sp == 0:
        return vals, t
    n = int(dups)
    vals = [np.tile(v, n) for v in vals]
    t = np.concatenate([t + d * sp for d in range(n)])
    return vals, t


class Ctx:
    def __init__(self, vctl, gyro):
        self.vctl = dict(vctl)
        self.vctl.setdefault("TEAnchorTime", 0.0)
        self.vctl.setdefault("MasterTxCoil", 1.0)
        self.gyro = gyro
        self.cv = {}
        self.tr = self.slice = self.phase = 0.0

    def num(self, k):
        v = self.vctl[k]
        if isinstance(v, str):
            raise ValueError(f"VCtl.{k} is text")
        return float(np.atleast_1d(v)[0])

    def text(self, k, default):
        v = self.vctl.get(k, default)
        return v if isinstance(v, str) else default


def _mround(x):
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def evaluate(expr, c: Ctx):
    """A MATLAB scalar expression of the PSD attributes."""
    src = expr.replace("^", "**")

    def vctl_ref(m):
        name, idx = m.group(1), m.group(2)
        return f"_V({name!r}, {idx})" if idx else f"_V({name!r}, 1)"

    src = re.sub(r"VCtl\.(\w+)(?:\(([^()]+)\))?", vctl_ref, src)
    src = re.sub(r"VVar\.(\w+)", lambda m: f"_VV({m.group(1)!r})", src)
    src = src.replace("VObj.Gyro", "_gyro")

    def _V(name, idx):
        v = np.atleast_1d(c.vctl[name])
        return float(v[int(idx) - 1])

    def _VV(name):
        return {"TRCount": c.tr, "SliceCount": c.slice, "PhaseCount": c.phase}[name]

    def _rem(a, b):
        return math.fmod(a, b) if b != 0 else math.nan

    def _mod(a, b):
        if b == 0:
            return a
        r = math.fmod(a, b)
        return r + b if r != 0 and ((r < 0) != (b < 0)) else r

    ns = {"_V": _V, "_VV": _VV, "_gyro": c.gyro, "pi": math.pi, "Inf": math.inf, "inf": math.inf,
          "NaN": math.nan, "eps": EPS, "abs": abs, "floor": math.floor, "ceil": math.ceil, "round": _mround,
          "sqrt": math.sqrt, "exp": math.exp, "rem": _rem, "mod": _mod, "max": max, "min": min}
    ns.update(c.cv)
    return float(eval(src, {"__builtins__": {}}, ns))  # noqa: S307 - fixture expressions only


def attr_value(name, v, c):
    if v == "":
        return None
    if v[0] == "$":
        opts = re.findall(r"'([^']*)'", v[2:])
        return opts[int(v[1]) - 1]
    if name == "Notes":
        return v
    return evaluate(v, c)


def params(el, c):
    return {k: attr_value(k, v, c) for k, v in el.attrib.items()}


# ------------------------------------------------------------- elements
def rf_element(name, P, c):
    if name in ADIABATIC:
        return rf_adiabatic(name, P)
    ts, te, dt = P["tStart"], P["tEnd"], P["dt"]
    t = linspace(ts, te, math.ceil((te - ts) / dt) + 1)
    if name == "rfSinc":
        a = {"Non": 0.0, "Hamming": 0.46, "Hanning": 0.5}[P["Apod"]]
        t = t - (te - ts) / 2 - ts
        t0 = (te - ts) / P["TBP"]
        with np.errstate(invalid="ignore", divide="ignore"):
            amp = ((t0 * np.sin(math.pi * t / t0)) / (math.pi * t)) * ((1 - a) + a * np.cos((math.pi * t) / (P["TBP"] * t0)))
        amp[np.isnan(amp)] = np.nanmax(amp)
        t = t + (te - ts) / 2 + ts
    elif name == "rfRect":
        amp = np.ones(t.size)
    elif name == "rfGaussian":
        sigma = (te - ts) / 7.434
        d = t - (te - ts) / 2 - ts
        amp = np.exp(-(d * d) / (2 * (sigma * sigma)))
    elif name == "rfFermi":
        tw = ((te - ts) - 2 * P["PW"]) / 13.81
        amp = 1 / (1 + np.exp((np.abs(t - (te - ts) / 2 - ts) - P["PW"]) / tw))
    else:
        raise ValueError(f"rf element {name} is not supported")
    amp[0] = 0
    amp[-1] = 0
    amp = b1_gain(amp, dt, P["FA"], c.gyro) * amp
    n = t.size
    phase = np.full(n, float(P["rfPhase"]))
    freq = np.full(n, float(P["rfFreq"]))
    coil = np.full(n, float(P["CoilID"]))
    phase[[0, -1]] = 0
    freq[[0, -1]] = 0
    (amp, freq, phase, coil), t = duplicate([amp, freq, phase, coil], t, P)
    return amp, phase, freq, coil, t


ADIABATIC = ("rfHyperbolicSecant", "rfTanhTan", "rfBIR", "rfBIREF")


def rf_adiabatic(name, P):
    """rfHyperbolicSecant.m, rfTanhTan.m, rfBIR.m, rfBIREF.m: MaxB1 amplitude,
    no flip-angle scaling."""
    ts, te, dt, b1 = P["tStart"], P["tEnd"], P["dt"], P["MaxB1"]
    T = te - ts
    t = linspace(ts, te, math.ceil((te - ts) / dt) + 1) if name == "rfHyperbolicSecant" else \
        linspace(0.0, T, math.ceil(T / dt) + 1)
    n = t.size
    amp, freq, phase = np.zeros(n), np.zeros(n), np.zeros(n)
    for k, x in enumerate(t):
        if name == "rfHyperbolicSecant":
            beta = (P["TBP"] / (te - ts)) * math.pi / P["Adiab"]
            a = beta * (x - (te - ts) / 2 - ts)
            amp[k] = b1 * (1 / math.cosh(a))
            freq[k] = -beta * P["Adiab"] * math.tanh(a) / (2 * math.pi)
            phase[k] = P["rfPhase"]
        elif name == "rfTanhTan":
            A = P["TBP"] * math.pi / T
            first = x < T / 2
            u = x if first else T - x
            amp[k] = b1 * math.tanh((2 * 10 * u) / T)
            f = math.tan(math.atan(20) * (1 - 2 * u / T)) / math.tan(math.atan(20))
            freq[k] = (A if first else -A) * f / (2 * math.pi)
            phase[k] = P["rfPhase"]
        elif name == "rfBIR":
            mf, flag = P["MaxFreq"], P["BIRFlag"]
            if flag == "BIR-1":
                z = math.pi / T
                freq[k] = (mf if x < T / 2 else -mf) * math.sin(z * x)
                phase[k] = 0.0 if x < T / 2 else math.pi / 2
                amp[k] = b1 * math.cos(z * x)
            elif flag == "BIR-2":
                z = 2 * math.pi / T
                phase[k] = 0.0 if x < T / 4 else (math.pi / 2 if x < T / 2 else math.pi / 2 + math.pi)
                amp[k] = abs(b1 * math.cos(z * x))
                freq[k] = abs(mf * math.sin(z * x))
            elif flag == "BIR-4":
                lam, beta = P["Lambda"], P["Beta"]
                den = math.tan(beta) * 2 * math.pi
                q = 4 * x / T
                if x < T / 4:
                    amp[k], freq[k] = b1 * math.tanh(lam * (1 - q)), math.tan(4 * beta * x / T) / den
                elif x < T / 2:
                    amp[k], freq[k] = b1 * math.tanh(lam * (q - 1)), math.tan(beta * (q - 2)) / den
                elif x < 3 * T / 4:
                    amp[k], freq[k] = b1 * math.tanh(lam * (3 - q)), math.tan(beta * (q - 2)) / den
                else:
                    amp[k], freq[k] = b1 * math.tanh(lam * (q - 3)), math.tan(beta * (q - 4)) / den
            else:
                raise ValueError(f"rfBIR: unknown BIRFlag {flag}")
        else:
            mf, flag = P["MaxFreq"], P["BIREFFlag"]
            if flag == "BIREF-1":
                z = math.pi / T
                amp[k] = (b1 if x < T / 2 else -b1) * math.sin(z * x)
                freq[k] = mf * abs(math.cos(z * x))
            elif flag == "BIREF-2a":
                z = math.pi / T
                freq[k] = (mf if x < T / 2 else -mf) * math.sin(z * x)
                amp[k] = b1 * abs(math.cos(z * x))
            elif flag == "BIREF-2b":
                z = 2 * math.pi / T
                if x < T / 2:
                    amp[k] = b1 * abs(math.cos(z * x))
                    freq[k] = (mf if x < T / 4 else -mf) * math.sin(z * x)
                else:
                    amp[k] = -b1 * math.cos(z * x)
                    freq[k] = (-mf if x < 3 * T / 4 else mf) * math.sin(z * x)
            else:
                raise ValueError(f"rfBIREF: unknown BIREFFlag {flag}")
    if name != "rfHyperbolicSecant":
        t = t + ts
    coil = np.full(n, float(P["CoilID"]))
    for v in (amp, freq, phase):
        v[[0, -1]] = 0
    (amp, freq, phase, coil), t = duplicate([amp, freq, phase, coil], t, P)
    return amp, phase, freq, coil, t


def g_trapezoid(name, ax, P, c):
    """{Gx,Gy,Gz}Trapezoid.m, *AreaTrapezoid.m, *AreaTrapezoid2.m as elements."""
    if name == ax + "Trapezoid":
        sr = max(2.0, P["sRamp"])
        g, t = std_trap(P["tStart"] - P["tRamp"], P["tEnd"] + P["tRamp"], P["tStart"], P["tEnd"], P[ax + "Amp"],
                        sr, 2, sr)
    elif name == ax + "AreaTrapezoid":
        g, t = area_trap(P["tStart"], P["tEnd"], P["Area"], c)
    elif name == ax + "AreaTrapezoid2":
        g, t = area_trap2(P["tStart"], P["Area"], c)
    else:
        raise ValueError(f"{name} is not supported")
    (g,), t = duplicate([g], t, P)
    return g, t


def gz_selective2(P):
    """GzSelective2.m."""
    t2s, t2e, tr = P["t2Start"], P["t2End"], P["tRamp"]
    a1, a2, a3 = P["Gz1Amp"], P["Gz2Amp"], P["Gz3Amp"]
    g1, tt1 = std_trap(0, P["tGz1"] + 2 * tr, tr, P["tGz1"] + tr, a1, 2, 2, 2)
    g3, tt3 = std_trap(0, P["tGz3"] + 2 * tr, tr, P["tGz3"] + tr, a3, 2, 2, 2)
    t2 = linspace(t2s, t2e, 2)
    g, t = np.array([a2, a2]), t2
    if a1 != 0 and P["tGz1"] != 0:
        g1 = g1.copy()
        g1[-1] = a2
        g, t = np.r_[g1, g], np.r_[t2s - tt1[-1] + tt1, t]
    else:
        g, t = np.r_[0.0, g], np.r_[t[0] - tr, t]
    if a3 != 0 and P["tGz3"] != 0:
        g3 = g3.copy()
        g3[0] = a2
        g, t = np.r_[g, g3], np.r_[t, t2e + tt3]
    else:
        g, t = np.r_[g, 0.0], np.r_[t, t[-1] + tr]
    t, g = unique_first(t, g)
    (g,), t = duplicate([g], t, P)
    return g, t


def g_pe(P, c, z):
    gbar = c.gyro / (2 * math.pi)
    if z:
        step = c.slice - math.floor(c.num("SliceNum") / 2) - 1
        fov, s1, s2 = c.num("FOVSlice"), P["Gz1Sign"], P["Gz2Sign"]
    else:
        step = c.phase - c.num("ResPhase") / 2 - 1
        fov, s1, s2 = c.num("FOVPhase"), P["Gy1Sign"], P["Gy2Sign"]
    t1s, t1e, t2s, t2e, tr = P["t1Start"], P["t1End"], P["t2Start"], P["t2End"], P["tRamp"]
    g1, tt1 = std_trap(t1s - tr, t1e + tr, t1s, t1e, step * (1 / (gbar * (t1e - t1s) * fov)) * s1, 2, 2, 2)
    g2, tt2 = std_trap(t2s - tr, t2e + tr, t2s, t2e, step * (1 / (gbar * (t2e - t2s) * fov)) * s2, 2, 2, 2)
    if s2 != 0:
        g1, tt1 = np.concatenate([g1, g2]), np.concatenate([tt1, tt2])
    t, g = unique_first(tt1, g1)
    (g,), t = duplicate([g], t, P)
    return g, t


def readout_geometry(c):
    gbar = c.gyro / (2 * math.pi)
    gx = (1 / c.num("FOVFreq")) / (gbar * (1 / c.num("BandWidth")))
    return gx, 1 / (2 * gbar * gx * c.num("RFreq"))


def gx_cartesian(P, c):
    gx, th = readout_geometry(c)
    ta, tr = c.num("TEAnchorTime"), P["tRamp"]
    t1s, t2m, t3s = P["t1Start"], P["t2Middle"], P["t3Start"]
    parts = []
    if P["Gx1Sign"] != 0:
        parts.append(std_trap(t1s - tr, t1s + th + tr, t1s, t1s + th, gx * P["Gx1Sign"], 2, 2, 2))
    parts.append(std_trap(t2m + ta - th - tr, t2m + ta + th + tr, t2m + ta - th, t2m + ta + th, gx * P["Gx2Sign"],
                          2, 2, 2))
    if P["Gx3Sign"] != 0:
        parts.append(std_trap(t3s - tr, t3s + th + tr, t3s, t3s + th, gx * P["Gx3Sign"], 2, 2, 2))
    t, g = unique_first(np.concatenate([p[1] for p in parts]), np.concatenate([p[0] for p in parts]))
    (g,), t = duplicate([g], t, P)
    return g, t


def gz_selective(P):
    t2s, t2e, tr, ga = P["t2Start"], P["t2End"], P["tRamp"], P["GzAmp"]
    half = (t2e - t2s) / 2
    g1, tt1 = std_trap(0, half + 2 * tr, tr, half + tr, P["Gz1Sign"] * ga, 2, 2, 2)
    g2, tt2 = std_trap(t2s - tr, t2e + tr, t2s, t2e, P["Gz2Sign"] * ga, 2, 2, 2)
    g3, tt3 = std_trap(0, half + 2 * tr, tr, half + tr, P["Gz3Sign"] * ga, 2, 2, 2)
    g, t = g2, tt2
    if P["Gz1Sign"] != 0:
        g, t = np.concatenate([g1, g]), np.concatenate([t2s - tr - tt1[-1] + tt1, t])
    if P["Gz3Sign"] != 0:
        g, t = np.concatenate([g, g3]), np.concatenate([t, t2e + tr + tt3])
    t, g = unique_first(t, g)
    (g,), t = duplicate([g], t, P)
    return g, t


def adc_cartesian(P, c):
    _, th = readout_geometry(c)
    tm, ta, mu = P["tMiddle"], c.num("TEAnchorTime"), c.num("MinUpdRate")
    g, t = std_trap(tm + ta - th - mu, tm + ta + th + mu, tm + ta - th, tm + ta + th, 1, 2, c.num("ResFreq"), 2)
    (g,), t = duplicate([g], t, P)
    return g, t


def ext_bit(P, c):
    t = linspace(P["tStart"], P["tStart"] + 2 * c.num("MinUpdRate"), 3)
    g = np.full(3, float(P["Ext"]))
    g[[0, -1]] = 0
    (g,), t = duplicate([g], t, P)
    return g, t


# ------------------------------------------------- EPI, FSE, radial, spiral
def area_trap2(ts, area, c):
    """GxAreaTrapezoid2.m / GyAreaTrapezoid2.m."""
    a = abs(area) / (c.gyro / (2 * math.pi))
    mu = c.num("MinUpdRate")
    amp = c.num("MaxGrad")
    if a / amp < mu:
        te = mu + ts
        amp = a / mu
    else:
        te = a / amp + ts
    tr = max(mu, amp / c.num("MaxSlewRate"))
    return std_trap(ts - tr, te + tr, ts, te, amp * np.sign(area), 2, 2, 2)


def area_trap(ts, te, area, c):
    """GyAreaTrapezoid.m."""
    amp = (abs(area) / (c.gyro / (2 * math.pi))) / (te - ts)
    tr = max(c.num("MinUpdRate"), amp / c.num("MaxSlewRate"))
    return std_trap(ts - tr, te + tr, ts, te, amp * np.sign(area), 2, 2, 2)


def adc_block(ts, te, n, c):
    """ADCBlock.m."""
    if n == 1:
        return np.array([0.0, 1.0, 0.0]), linspace(ts, te, 3)
    mu = c.num("MinUpdRate")
    return std_trap(ts - mu, te + mu, ts, te, 1, 2, n, 2)


def _join(parts):
    g = np.concatenate([p[0] for p in parts])
    t = np.concatenate([p[1] for p in parts])
    t, g = unique_first(t, g)
    return g, t


def _epi_offset(tm, c, half):
    esp, etl = c.num("EPI_ESP"), c.num("EPI_ETL")
    off = (tm + c.num("TEAnchorTime")) - (math.floor(etl / 2) + half) * esp
    if c.text("EPI_EchoShifting", "off") == "on":
        off = off + (c.phase - 1) * (esp / c.num("EPI_ShotNum"))
    return off


def gx_epi(P, c):
    gx, th = readout_geometry(c)
    esp, etl = c.num("EPI_ESP"), c.num("EPI_ETL")
    pre = area_trap2(P["t1Start"], (1 / c.num("RFreq")) / 2, c)
    off = _epi_offset(P["t2Middle"], c, 0.5)
    tr = esp / 2 - th
    parts = [(pre[0] * P["Gx1Sign"], pre[1])]
    for i in range(1, int(etl) + 1):
        ts = off + (i - 1) * esp + tr
        g, t = std_trap(ts - tr, ts + 2 * th + tr, ts, ts + 2 * th, gx, 2, 2, 2)
        parts.append((g * ((-1.0) ** (i - 1) * P["Gx2Sign"]), t))
    return _join(parts)


def gy_epi(P, c):
    kmax = (1 / c.num("RPhase")) / 2
    etl, esp = c.num("EPI_ETL"), c.num("EPI_ESP")
    dk = (2 * kmax) / etl
    pre = area_trap2(P["t1Start"], kmax - (c.phase - 1) * (dk / c.num("EPI_ShotNum")), c)
    parts = [(pre[0] * P["Gy1Sign"], pre[1])]
    off = _epi_offset(P["t2Middle"], c, -0.5)
    for i in range(1, int(etl)):
        g, t = area_trap2(off + (i - 1) * esp, dk, c)
        parts.append((g * P["Gy2Sign"], t))
    return _join(parts)


def adc_epi(P, c):
    _, th = readout_geometry(c)
    esp, etl = c.num("EPI_ESP"), c.num("EPI_ETL")
    off = _epi_offset(P["tMiddle"], c, 0.5)
    tr = esp / 2 - th
    parts = []
    for i in range(1, int(etl) + 1):
        ts = off + (i - 1) * esp + tr
        parts.append(adc_block(ts, ts + 2 * th, c.num("ResFreq"), c))
    return _join(parts)


def _fse_offset(tm, c):
    return (tm + c.num("TEAnchorTime")) - math.floor(c.num("FSE_ETL") / 2) * c.num("FSE_ESP")


def gx_fse(P, c):
    gx, th = readout_geometry(c)
    tr = gx / c.num("MaxSlewRate")
    t1s = P["t1Start"]
    parts = [std_trap(t1s - tr, t1s + th + tr, t1s, t1s + th, gx * P["Gx1Sign"], 2, 2, 2)]
    off = _fse_offset(P["t2Middle"], c)
    for i in range(1, int(c.num("FSE_ETL")) + 1):
        m = off + (i - 1) * c.num("FSE_ESP")
        parts.append(std_trap(m - th - tr, m + th + tr, m - th, m + th, gx * P["Gx2Sign"], 2, 2, 2))
    return _join(parts)


def gy_fse(P, c):
    kmax = (1 / c.num("RPhase")) / 2
    etl, esp = c.num("FSE_ETL"), c.num("FSE_ESP")
    dk = (2 * kmax) / etl
    off = _fse_offset(P["tMiddle"], c)
    parts = []
    for i in range(1, int(etl) + 1):
        k = -(kmax - (c.phase - 1) * (dk / c.num("FSE_ShotNum"))) + (i - 1) * dk
        ts = off + (i - 1) * esp - P["tOffset"]
        parts.append(area_trap(ts, ts + P["tGy1"], k * P["Gy1Sign"], c))
        ts = off + (i - 1) * esp + P["tOffset"] - P["tGy2"]
        parts.append(area_trap(ts, ts + P["tGy2"], k * P["Gy2Sign"], c))
    return _join(parts)


def adc_fse(P, c):
    _, th = readout_geometry(c)
    tr = c.num("MinUpdRate")
    off = _fse_offset(P["tMiddle"], c)
    parts = []
    for i in range(1, int(c.num("FSE_ETL")) + 1):
        m = off + (i - 1) * c.num("FSE_ESP")
        parts.append(std_trap(m - th - tr, m + th + tr, m - th, m + th, 1, 2, c.num("ResFreq"), 2))
    return _join(parts)


def _radial_geometry(c):
    fov, res = c.num("FOVFreq"), c.num("ResFreq")
    gbar = c.gyro / (2 * math.pi)
    gx = (1 / fov) / (gbar * (1 / c.num("BandWidth")))
    return gx, 1 / (2 * gbar * gx * (fov / res))


def g_radial(P, c, y):
    pat = c.text("R_AngPattern", "Linear")
    if pat == "Linear":
        inc = evaluate(c.text("R_AngRange", ""), c) / c.num("R_SpokeNum")
    elif pat == "Golden":
        inc = 111.246 * (math.pi / 180)
    else:
        raise ValueError("VCtl.R_AngPattern must be Linear or Golden")
    gx, th = _radial_geometry(c)
    ang = inc * (c.phase - 1)
    amp = gx * (math.sin(ang) if y else math.cos(ang))
    ax = "Gy" if y else "Gx"
    t1s, t2m, t3s, tr = P["t1Start"], P["t2Middle"], P["t3Start"], P["tRamp"]
    ta = c.num("TEAnchorTime")
    parts = []
    if P[ax + "1Sign"] != 0:
        parts.append(std_trap(t1s - tr, t1s + th + tr, t1s, t1s + th, amp * P[ax + "1Sign"], 2, 2, 2))
    parts.append(std_trap(t2m + ta - th - tr, t2m + ta + th + tr, t2m + ta - th, t2m + ta + th,
                          amp * P[ax + "2Sign"], 2, 2, 2))
    if P[ax + "3Sign"] != 0:
        parts.append(std_trap(t3s - tr, t3s + th + tr, t3s, t3s + th, amp * P[ax + "3Sign"], 2, 2, 2))
    return _join(parts)


def adc_radial(P, c):
    _, th = _radial_geometry(c)
    tm, ta, mu = P["tMiddle"], c.num("TEAnchorTime"), c.num("MinUpdRate")
    return std_trap(tm + ta - th - mu, tm + ta + th + mu, tm + ta - th, tm + ta + th, 1, 2,
                    c.num("R_SampPerSpoke"), 2)


def colon(a, d, b):
    """a:d:b with a relative tolerance on the end point."""
    q = (b - a) / d
    if not q >= 0 or not math.isfinite(q):
        return np.array([a]) if q == 0 else np.zeros(0)
    n = int(math.floor(q + 1e-10)) + 1
    return a + np.arange(n, dtype=np.float64) * d


def spiral_shape(c, dt, need_samples):
    fov, res, g = c.num("FOVFreq"), c.num("ResFreq"), c.gyro
    lam, sr, grad = c.num("S_Lamda"), c.num("S_SlewRate"), c.num("S_Gradient")
    thmax = (res / (2 * fov)) / lam
    beta = (g / (2 * math.pi)) * (sr / lam)
    a2 = (9 * beta / 4) ** (1.0 / 3)
    cl = sr / c.num("S_SlewRate0")
    ts = ((3 * g * grad) / (4 * math.pi * lam * (a2 * a2))) ** 3

    def th1(t):  # libm's pow, element by element, as the compiled library calls it
        if np.ndim(t):
            return np.array([th1(x) for x in t])
        t = float(t)
        return (0.5 * beta * (t * t)) / (cl + (beta / (2 * a2)) * math.pow(t, 4.0 / 3))

    if th1(ts) < thmax:
        t1 = colon(0.0, dt, ts) if need_samples else np.array([0.0, ts])
        t1e = t1[-1] if need_samples else ts
        th1e = th1(t1e)
        tacq = t1e + ((math.pi * lam) / (g * grad)) * ((thmax * thmax) - (th1e * th1e))
        if not need_samples:
            return None, None, tacq
        t2 = colon(t1e, dt, tacq)
        th2 = np.array([math.sqrt((th1e * th1e) + (g / (math.pi * lam)) * grad * (x - t1e)) for x in t2])
        return np.r_[t1[:-1], t2], np.r_[th1(t1[:-1]), th2], tacq
    tacq = ((2 * math.pi * fov) / (3 * c.num("S_ShotNum"))) * math.sqrt(math.pi / (g * sr * c.num("RFreq") ** 3))
    if not need_samples:
        return None, None, tacq
    t = colon(0.0, dt, tacq)
    return t, th1(t), tacq


def g_spiral(P, c, y):
    t, th, _ = spiral_shape(c, P["dt"], True)
    rot = (c.phase - 1) * (2 * math.pi / c.num("S_ShotNum"))
    dth = np.r_[0.0, np.diff(th) / np.diff(t)]
    a = th + rot
    sa = np.array([math.sin(x) for x in a])
    ca = np.array([math.cos(x) for x in a])
    f = (sa + th * ca) if y else (ca - th * sa)
    g = ((2 * math.pi) / c.gyro) * c.num("S_Lamda") * dth * f
    tt = P["tStart"] + c.num("TEAnchorTime") + t
    tt = np.r_[tt, tt[-1] + max(c.num("MinUpdRate"), abs(g[-1] / c.num("S_SlewRate")))]
    g = np.r_[g, 0.0]
    tt, g = unique_first(tt, g)
    return g, tt


def adc_spiral(P, c):
    _, _, t = spiral_shape(c, 0.0, False)
    ts, ta, mu = P["tStart"], c.num("TEAnchorTime"), c.num("MinUpdRate")
    return std_trap(ts + ta - mu, ts + ta + t + mu, ts + ta, ts + ta + t, 1, 2,
                    math.floor(t * c.num("BandWidth")), 2)


GY = {"GyCartesian": lambda P, c: g_pe(P, c, False), "GyEPI": gy_epi, "GyFSE": gy_fse,
      "GyRadial": lambda P, c: g_radial(P, c, True), "GySpiral": lambda P, c: g_spiral(P, c, True)}
GX = {"GxCartesian": lambda P, c: gx_cartesian(P, c), "GxEPI": gx_epi, "GxFSE": gx_fse,
      "GxRadial": lambda P, c: g_radial(P, c, False), "GxSpiral": lambda P, c: g_spiral(P, c, False)}
ADCS = {"ADCCartesian": lambda P, c: adc_cartesian(P, c), "ADCEPI": adc_epi, "ADCFSE": adc_fse,
        "ADCRadial": adc_radial, "ADCSpiral": adc_spiral}


# ------------------------------------------------------------- one TR
def psd_tr(root, c):
    for k in root:  # the CVs are assigned on every call of the PSD function
        if k.tag == "CVs":
            for n, v in k.attrib.items():
                c.cv[n] = evaluate(v, c) if v else math.nan
    L = {k: [] for k in ("rfA", "rfP", "rfF", "rfC", "rfT", "Gz", "GzT", "Gy", "GyT", "Gx", "GxT", "ADC", "ADCT",
                         "Ext", "ExtT")}
    se = []
    for sec, pulses in enumerate([k for k in root if k.tag == "Pulses"], start=1):
        S = params(pulses, c)
        tS, tE = S["tS"], S["tE"]
        if not tS < tE:
            raise ValueError(f"SE setting is incorrect for Pulses {sec}!")
        if (c.tr < S["TRStart"] or c.tr > S["TREnd"] or math.fmod(c.tr - S["TRStart"], S["Freq"]) != 0
                or S["Switch"] == "off"):
            continue
        sl = {k: [] for k in L}
        for grp in pulses:
            for el in grp:
                P = params(el, c)
                if P["Switch"] != "on":
                    continue
                if grp.tag == "rf":
                    if P["AnchorTE"] == "on":
                        a = c.text("TEAnchor", "Middle")
                        c.vctl["TEAnchorTime"] = {"Start": P["tStart"], "Middle": (P["tStart"] + P["tEnd"]) / 2,
                                                  "End": P["tEnd"]}.get(a, c.vctl["TEAnchorTime"])
                    amp, ph, fr, co, t = rf_element(el.tag, P, c)
                    if c.text("MultiTransmit", "off") != "off" or c.num("MasterTxCoil") == co[0]:
                        for k, v in zip(("rfA", "rfP", "rfF", "rfC", "rfT"), (amp, ph, fr, co, t)):
                            sl[k].append(v)
                    continue
                if grp.tag == "GzSS":
                    if el.tag == "GzCartesian":
                        g, t = g_pe(P, c, True)
                    elif el.tag == "GzSelective":
                        g, t = gz_selective(P)
                    elif el.tag == "GzSelective2":
                        g, t = gz_selective2(P)
                    else:
                        g, t = g_trapezoid(el.tag, "Gz", P, c)
                    key = "Gz"
                elif grp.tag == "GyPE":
                    g, t = GY[el.tag](P, c) if el.tag in GY else g_trapezoid(el.tag, "Gy", P, c)
                    key = "Gy"
                elif grp.tag == "GxR":
                    g, t = GX[el.tag](P, c) if el.tag in GX else g_trapezoid(el.tag, "Gx", P, c)
                    key = "Gx"
                elif grp.tag == "ADC":
                    g, t = ADCS[el.tag](P, c)
                    key = "ADC"
                else:
                    g, t = ext_bit(P, c)
                    key = "Ext"
                sl[key].append(g)
                sl[key + "T"].append(t)
        cat = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in sl.items()}
        span = tE - tS
        for tkey, vkeys in (("rfT", ("rfA", "rfP", "rfF", "rfC")), ("GzT", ("Gz",)), ("GyT", ("Gy",)),
                            ("GxT", ("Gx",)), ("ADCT", ("ADC",)), ("ExtT", ("Ext",))):
            keep = ~((cat[tkey] < 0) | (cat[tkey] > span))
            for vk in vkeys:
                cat[vk] = cat[vk][keep]
            cat[tkey] = cat[tkey][keep] + tS
        cat["rfA"][np.abs(cat["rfA"]) < EPS] = 0
        for k in L:
            L[k].append(cat[k])
        se += [tS, tE]
    A = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in L.items()}
    t0, t1 = min(se), max(se)
    ts = np.concatenate([[t0, t1], A["rfT"], A["GzT"], A["GyT"], A["GxT"], A["ADCT"], A["ExtT"]]) - t0
    line = np.concatenate([[-1, -1]] + [np.full(A[k].size, i) for i, k in
                                         enumerate(("rfT", "GzT", "GyT", "GxT", "ADCT", "ExtT"))])
    order = np.argsort(ts, kind="stable")
    ts = ts[order]
    line = line[order]
    flags = np.zeros((6, ts.size))
    for i in range(6):
        flags[i, line == i] = 1
    uts = np.unique(ts)
    out = {"uts": uts, "ts": ts, "flags": flags}
    for tkey, vkeys in (("rfT", ("rfA", "rfP", "rfF", "rfC")), ("GzT", ("Gz",)), ("GyT", ("Gy",)),
                        ("GxT", ("Gx",)), ("ADCT", ("ADC",)), ("ExtT", ("Ext",))):
        o = np.argsort(A[tkey] - t0, kind="stable")
        for vk in vkeys:
            out[vk] = A[vk][o]
    for vk in ("rfA", "rfP", "rfF", "Gz", "Gy", "Gx", "ADC", "Ext"):
        out[vk][[0, -1]] = 0
    return out


def compile_psd(xml_text: str, vctl: dict, gyro: float = 267538030.3797068) -> dict:
    """VSeq lines (same keys as mrilab_amd.engine.compile_psd)."""
    root = ET.fromstring(xml_text)
    c = Ctx(vctl, gyro)
    tr0, fa0 = c.num("TR"), c.num("FlipAng")
    has_dp = "DP_Flag" in c.vctl
    dp_on = c.text("DP_Flag", "off") == "on"
    keys = ("rfA", "rfP", "rfF", "rfC", "Gz", "Gy", "Gx", "ADC", "Ext", "uts", "ts")
    acc = {k: [] for k in keys}
    flags = []
    dp, s, j = 0, 1, 1
    n2, n1 = c.num("SecondPhNum"), c.num("FirstPhNum")
    n_tr = 0
    while s <= n2:
        c.slice = s
        restarted = False
        while j <= n1:
            c.phase = j
            c.tr += 1
            if has_dp:
                c.vctl["TR"], c.vctl["FlipAng"] = ((c.num("DP_TR"), c.num("DP_FlipAng")) if dp_on and dp == 0
                                                   else (tr0, fa0))
            o = psd_tr(root, c)
            for k in keys:
                acc[k].append(o[k])
            flags.append(o["flags"])
            n_tr += 1
            if has_dp and dp_on and dp == 0:
                for k in ("Gz", "Gy", "Gx", "ADC"):
                    acc[k] = [a * 0 for a in acc[k]]
                if c.tr == c.num("DP_Num"):
                    s, j, dp, restarted = 1, 1, 1, True
                    break
            j += 1
        if restarted:
            continue
        j = 1
        s += 1
    cat = {k: np.concatenate(v) for k, v in acc.items()}
    if c.text("MultiTransmit", "off") == "off":
        cat["rfC"] = np.ones_like(cat["rfC"])
    return {"utsLine": cat["uts"], "tsLine": cat["ts"], "rfAmpLine": cat["rfA"], "rfPhaseLine": cat["rfP"],
            "rfFreqLine": cat["rfF"], "rfCoilLine": cat["rfC"], "GzAmpLine": cat["Gz"], "GyAmpLine": cat["Gy"],
            "GxAmpLine": cat["Gx"], "ADCLine": cat["ADC"], "ExtLine": cat["Ext"],
            "flagsLine": np.asfortranarray(np.concatenate(flags, axis=1)), "TRNum": n_tr}


def check(q: dict, max_grad: float, max_slew: float, min_upd: float, tx_coil_num: int = 0, vctl=None):
    """DoMultiTransmitChk.m (with ``vctl`` and tx_coil_num > 0), DoGradChk.m
    then DoUpdRateChk.m on compiled lines: None, or the first failing
    check's message."""
    ts, fl = q["tsLine"], q["flagsLine"]
    if vctl is not None and tx_coil_num > 0 and vctl.get("MultiTransmit") == "on" and q["rfCoilLine"].size:
        if tx_coil_num < np.max(q["rfCoilLine"]):
            return "Multi-transmit check failed: the sequence uses more coil elements than the Tx coil provides"
        if not np.any(q["rfCoilLine"] == float(vctl.get("MasterTxCoil", 1.0))):
            return "Multi-transmit check failed: the master Tx coil element is not used"
    names = (("GzSS", "GzAmpLine", 1), ("GyP", "GyAmpLine", 2), ("GxR", "GxAmpLine", 3))
    gtol, stol = abs(max_grad * 0.005), abs(max_slew * 0.005)
    with np.errstate(divide="ignore", invalid="ignore"):
        for name, line, row in names:
            t, g = ts[fl[row] != 0], q[line]
            sr = np.diff(g) / np.diff(t)
            if np.any((np.abs(sr) - max_slew) > stol):
                return f"Gradient check failed: {name} exceeds maximum allowed gradient slew rate"
            if np.any((np.abs(g) - max_grad) > gtol):
                return f"Gradient check failed: {name} exceeds maximum allowed gradient amplitude"
    utol = -abs(min_upd * 0.01)

    def too_fast(t):
        dt = np.diff(t)
        dt = dt[~(dt < 0.5 * EPS)]
        return bool(np.any((np.abs(dt) - min_upd) < utol))

    t = ts[fl[0] != 0]
    for cid in np.unique(q["rfCoilLine"]):
        if too_fast(t[q["rfCoilLine"] == cid]):
            return f"Update rate check failed: rf{cid:g} sequence line exceeds minimum update rate"
    for name, line, row in names:
        if too_fast(ts[fl[row] != 0]):
            return f"Update rate check failed: {name} sequence line exceeds minimum update rate"
    return None
