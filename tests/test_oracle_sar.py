"""Pins the SAR oracle (oracle/sar_oracle.c, DoCalSARAtCPU.cpp restated)
by closed forms on hand-built sequences and by exact scaling laws.

* A rect pulse of amplitude A on one coil over N steps of dt gives
  VVar.SAR_k = N dt (A E_k)^2 and VSig.SAR = sum_k SAR_k ECon_k at the
  sample times (start, middle, end).
* Two coils transmitting at the same time points (multi-line rf, flag
  count 2) superpose amplitudes before squaring: N dt (A1 E1 + A2 E2)^2.
* A TR boundary (time going back) counts a TR and keeps Muts continuous.
* Doubling E or every rf amplitude multiplies SAR by exactly 4; doubling
  ECon doubles the samples exactly.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import workloads


def hand_scan(events, n_tx=1, n_samples=3, tsample=None, shape=(3, 2, 1), seed=0):
    """events: list of (t, [(coil, amp), ...]) rf events per time point, in
    time order; returns a Scan with SAR fields and these lines only."""
    sc = workloads.sar_scan(workloads.gre2d(n_pe=4, n_ro=4, tr_steps=40, te_steps=20, rf_steps=5), n_tx=n_tx,
                            n_samples=n_samples, seed=seed)
    R, C_, Z = shape
    nvox = R * C_ * Z
    rng = np.random.default_rng(seed + 1)
    ts, amp, coil, flags = [], [], [], []
    for t, rfs in events:
        for c, a in rfs:
            ts.append(t)
            amp.append(a)
            coil.append(c)
            flags.append([1, 0, 0, 0, 0, 0])
    ts = np.array(ts, dtype=np.float64)
    uts = np.unique(ts)
    z = np.zeros(0)
    sc.VSeq = {"tsLine": ts, "utsLine": uts, "rfAmpLine": np.array(amp, dtype=np.float64),
               "rfPhaseLine": np.zeros(len(amp)), "rfFreqLine": np.zeros(len(amp)),
               "rfCoilLine": np.array(coil, dtype=np.float64), "GzAmpLine": z.copy(), "GyAmpLine": z.copy(),
               "GxAmpLine": z.copy(), "ADCLine": z.copy(), "ExtLine": z.copy(),
               "flagsLine": np.asfortranarray(np.array(flags, dtype=np.float64).T)}
    E = rng.uniform(-1e6, 1e6, (3, R, C_, Z, n_tx)).astype(np.float32)
    for k, name in enumerate("xyz"):
        sc.VCoi[f"TxE1{name}"] = np.asfortranarray(E[k])
    sc.VObj["MassDen"] = np.full(shape, 1000.0, dtype=np.float32, order="F")
    sc.VObj["ECon"] = np.asfortranarray(rng.uniform(1e-4, 5e-4, shape + (3,)).astype(np.float32))
    sc.VVar["SAR"] = np.zeros(shape + (3,), dtype=np.float32, order="F")
    sc.VSig["tSample"] = np.asarray(tsample if tsample is not None else np.linspace(0, uts[-1], n_samples),
                                    dtype=np.float64)
    sc.VSig["tRealSample"] = np.zeros(sc.VSig["tSample"].size)
    sc.VSig["SAR"] = np.zeros(shape + (sc.VSig["tSample"].size,), dtype=np.float32, order="F")
    sc.VSig["Muts"] = np.zeros(uts.size)
    return sc, E


def test_rect_pulse_closed_form():
    N, dt, A = 40, 1e-5, 7e-6
    ev = [(k * dt, [(1, A)]) for k in range(N + 1)]
    sc, E = hand_scan(ev, tsample=[0.0, 0.5 * N * dt, N * dt * (1 - 1e-9)])
    out = oracle.sar(sc)
    base = out.VVar["SAR"]
    want = N * dt * (A * E[..., 0].astype(np.float64)) ** 2            # [3, R, C, Z]
    np.testing.assert_allclose(np.moveaxis(base, -1, 0), want, rtol=1e-5)
    econ = np.moveaxis(out.VObj["ECon"].astype(np.float64), -1, 0)
    S = out.VSig["SAR"]
    assert np.all(S[..., 0] == 0)                                       # sampled at t = 0
    np.testing.assert_allclose(S[..., 2], (want * econ).sum(axis=0), rtol=1e-5)
    np.testing.assert_allclose(S[..., 1], (want * econ).sum(axis=0) * 20 / 40, rtol=1e-5)
    np.testing.assert_allclose(out.VSig["tRealSample"], [0, 20 * dt, 40 * dt], rtol=1e-12)
    assert int(out.VVar["SARi"]) == 3


def test_two_coils_superpose_before_squaring():
    N, dt, A1, A2 = 25, 2e-5, 5e-6, -3e-6
    ev = [(k * dt, [(1, A1), (2, A2)]) for k in range(N + 1)]
    sc, E = hand_scan(ev, n_tx=2, tsample=[N * dt * (1 - 1e-9)])
    out = oracle.sar(sc)
    Ed = E.astype(np.float64)
    want = N * dt * (A1 * Ed[..., 0] + A2 * Ed[..., 1]) ** 2
    np.testing.assert_allclose(np.moveaxis(out.VVar["SAR"], -1, 0), want, rtol=2e-5)


def test_tr_boundary_keeps_time_continuous():
    dt, A = 1e-5, 4e-6
    # two TRs whose local time restarts at 0
    ev = [(k * dt, [(1, A)]) for k in range(11)] + [(k * dt, [(1, A)]) for k in range(11)]
    sc, E = hand_scan(ev, tsample=[0.0])
    # restore the per-TR time order the sequence compiler produces (utsLine restarts)
    ts = np.concatenate([np.arange(11) * dt, np.arange(11) * dt])
    sc.VSeq["tsLine"] = ts
    sc.VSeq["utsLine"] = ts.copy()
    sc.VSig["Muts"] = np.zeros(ts.size)
    sc.VSig["tSample"] = np.array([0.0, 19.5 * dt])
    sc.VSig["tRealSample"] = np.zeros(2)
    sc.VSig["SAR"] = np.zeros(sc.VSig["SAR"].shape[:3] + (2,), dtype=np.float32, order="F")
    tr0 = int(sc.VVar["TRCount"])
    out = oracle.sar(sc)
    assert int(out.VVar["TRCount"]) == tr0 + 1
    # the boundary point repeats the time before it: the sample at 19.5 dt lands after 20 steps
    np.testing.assert_allclose(out.VSig["tRealSample"][1], 20 * dt, rtol=1e-12)


def test_scaling_laws_are_exact():
    sc = workloads.sar_scan(workloads.gre2d(n_pe=8, n_ro=8, tr_steps=200, te_steps=100, rf_steps=40), n_tx=2)
    a = oracle.sar(sc)
    s2 = sc.copy()
    for k in "xyz":
        s2.VCoi[f"TxE1{k}"] = s2.VCoi[f"TxE1{k}"] * np.float32(2)
    b = oracle.sar(s2)
    assert np.array_equal(b.VVar["SAR"], 4 * a.VVar["SAR"])
    s3 = sc.copy()
    s3.VSeq["rfAmpLine"] = s3.VSeq["rfAmpLine"] * 2
    c = oracle.sar(s3)
    assert np.array_equal(c.VVar["SAR"], 4 * a.VVar["SAR"])
    s4 = sc.copy()
    s4.VObj["ECon"] = s4.VObj["ECon"] * np.float32(2)
    d = oracle.sar(s4)
    assert np.array_equal(d.VSig["SAR"], 2 * a.VSig["SAR"])
    assert float(np.abs(a.VSig["SAR"]).max()) > 0
    assert int(a.VVar["SARi"]) == sc.VSig["tSample"].size
