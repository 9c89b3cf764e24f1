"""The CPU oracle (oracle/bloch_oracle.c) pinned by closed forms.

The reference (MRiLab's DoScanAtCPU with Intel IPP and MATLAB's MEX API)
cannot be built or run here and ships no tests or golden outputs, so every
oracle function is pinned by what the mathematics fixes: closed-form
precession, relaxation and rotation, invariants of the discrete scheme,
and the decoder's bookkeeping on hand-built sequences. Each case is chosen
so that a plausible mistake (a sign, a dropped term, a swapped operand or
index) fails it.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import psd, workloads
from mrilab_amd.psd import GYRO
from tests.conftest import adc_capture_slots

PI = 3.14159265359


def one_voxel_scan(events, T, T1=1e9, T2=1e9, rho=1.0, dB0=0.0, dwrnd=0.0, M0=(0.0, 0.0, 1.0),
                   x=0.0, y=0.0, z=0.0, txpe=0.0, txmg=1.0, cs=0.0):
    """A 1x1x1 phantom and one TR whose lines are given as dicts of
    (times, values) per line name."""
    tr = psd.TR(duration=T)
    rf = events.get("rf")
    if rf is None:
        rf = psd.RF(np.array([0.0, T]), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2))
    tr.add_rf(rf)
    for name, add in (("Gz", tr.add_gz), ("Gy", tr.add_gy), ("Gx", tr.add_gx), ("ADC", tr.add_adc),
                      ("Ext", tr.add_ext)):
        t, v = events.get(name, (np.array([0.0, T]), np.zeros(2)))
        add(t, v)
    seq = psd.concat([tr.finalize()])
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.full(shape + (1,), rho), np.full(shape + (1,), T1), np.full(shape + (1,), T2),
                                np.full(shape, dB0), (1e-3, 1e-3, 1e-3),
                                dWRnd=np.full(shape + (1, 1), dwrnd), seq=seq, cs=[cs],
                                tx=(np.full(shape, txmg), np.full(shape, txpe)))
    sc.VObj["Mx"][:] = M0[0]
    sc.VObj["My"][:] = M0[1]
    sc.VObj["Mz"][:] = M0[2]
    sc.VMag["Gxgrid"][:] = x
    sc.VMag["Gygrid"][:] = y
    sc.VMag["Gzgrid"][:] = z
    return sc


def M(sc):
    return np.array([sc.VObj["Mx"].ravel()[0], sc.VObj["My"].ravel()[0], sc.VObj["Mz"].ravel()[0]], dtype=np.float64)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_free_precession_phase_and_t2_decay(prec):
    """Off-resonance w = gamma*dB0 + dWRnd + 2*pi*CS rotates M about -z by w*t
    (Mx' = Mx cos + My sin, BlochKernel.h:314-343); T2 decays |Mxy|."""
    dB0, dwrnd, cs, T2, n, dt = 1e-6, 40.0, 3.0, 0.05, 200, 1e-5
    t = np.arange(n + 1) * dt
    # gradient samples at every grid point keep the lines "fired"
    sc = one_voxel_scan({"Gx": (t, np.zeros(n + 1))}, n * dt, T2=T2, dB0=dB0, dwrnd=dwrnd, cs=cs,
                        M0=(1.0, 0.0, 0.0))
    out = oracle.scan(sc, prec)
    w = GYRO * float(np.float32(dB0)) + float(np.float32(dwrnd)) + 2 * PI * cs   # maps are fp32
    T = n * dt
    expect = np.exp(-T / float(np.float32(T2))) * np.array([np.cos(w * T), -np.sin(w * T)])
    tol = 1e-4 if prec == "f32" else 1e-9
    np.testing.assert_allclose(M(out)[:2], expect, atol=tol)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_t1_recovery_closed_form(prec):
    """Mz(t) = M0 + (Mz0 - M0) exp(-t/T1) with M0 = Rho/SpinNum (BlochKernel.h:354-368)."""
    T1, rho, n, dt = 0.01, 0.8, 500, 1e-5
    t = np.arange(n + 1) * dt
    sc = one_voxel_scan({"Gx": (t, np.zeros(n + 1))}, n * dt, T1=T1, rho=rho, M0=(0.0, 0.0, -0.5))
    out = oracle.scan(sc, prec)
    T1f, rhof = float(np.float32(T1)), float(np.float32(rho))   # maps are fp32
    expect = rhof + (-0.5 - rhof) * np.exp(-n * dt / T1f)
    tol = 1e-4 if prec == "f32" else 1e-9
    assert abs(M(out)[2] - expect) < tol


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("phase,expect", [(0.0, (0.0, 1.0, 0.0)), (PI / 2, (-1.0, 0.0, 0.0))])
def test_hard_pulse_90deg(prec, phase, expect):
    """On resonance a 90 deg pulse of phase 0 takes +z to +y; phase pi/2 to -x
    (rotation about -B1, BlochKernel.h:146-306). The rf gain comes from
    DoB1Scaling (DoB1Scaling.m:6-10), so the flip angle is exact."""
    n, dt = 10, 1e-5
    t = np.arange(n + 1) * dt
    rf = psd.rf_rect(t, 90.0, phase=phase)
    sc = one_voxel_scan({"rf": rf, "Gx": (t, np.zeros(n + 1))}, n * dt)
    out = oracle.scan(sc, prec)
    tol = 1e-5 if prec == "f32" else 1e-12
    np.testing.assert_allclose(M(out), expect, atol=tol)


def rodrigues(m, bx, by, w, dt):
    """Rotation by |Omega|*dt about -Omega, Omega = (bx, by, w)."""
    om = np.sqrt(bx * bx + by * by + w * w)
    n = np.array([bx, by, w]) / om
    c, s = np.cos(om * dt), np.sin(om * dt)
    return m * c - np.cross(n, m) * s + n * np.dot(n, m) * (1 - c)


@pytest.mark.parametrize("seed", range(6))
def test_rf_step_matches_rodrigues(seed):
    """One rf step of the literal reference matrix (fp64 oracle) equals the
    Rodrigues rotation about -(gamma*B1*cos(pe+phase), gamma*B1*sin(pe+phase), w)
    for random off-resonance, B1+ phase and rf phase (also negative
    amplitudes, folded to |A| and phase + pi)."""
    rng = np.random.default_rng(seed)
    dt = 1e-5
    A = rng.uniform(-3e-6, 3e-6)
    ph = rng.uniform(-3, 3)
    pe = rng.uniform(-3, 3)
    dB0 = rng.uniform(-2e-6, 2e-6)
    m0 = rng.normal(size=3)
    t = np.array([0.0, dt, 2 * dt])
    rf = psd.RF(t, np.array([A, A, 0.0]), np.array([ph, ph, 0.0]), np.zeros(3), np.ones(3))
    tr = psd.TR(duration=2 * dt)
    tr.add_rf(rf)
    # keep rf as the first/last sample zeroing would remove it: build lines directly
    seq = {"rfAmpLine": np.array([A, A, 0.0]), "rfPhaseLine": np.array([ph, ph, 0.0]),
           "rfFreqLine": np.zeros(3), "rfCoilLine": np.ones(3),
           "GzAmpLine": np.zeros(2), "GyAmpLine": np.zeros(2), "GxAmpLine": np.zeros(2),
           "ADCLine": np.zeros(2), "ExtLine": np.zeros(2),
           "utsLine": np.array([0.0, dt]), "tsLine": np.array([0.0, 0.0, dt, dt]),
           "flagsLine": np.asfortranarray(np.array([[1, 0, 0, 0, 0, 0], [0, 1, 1, 1, 0, 0],
                                                    [1, 0, 0, 0, 0, 0], [0, 1, 1, 1, 0, 0]], float).T)}
    seq = {k: np.asfortranarray(v, dtype=np.float64) for k, v in seq.items()}
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.ones(shape + (1,)), np.full(shape + (1,), 1e9), np.full(shape + (1,), 1e9),
                                np.full(shape, dB0), (1e-3, 1e-3, 1e-3), seq=seq,
                                tx=(np.ones(shape), np.full(shape, pe)))
    for i, k in enumerate(("Mx", "My", "Mz")):
        sc.VObj[k][:] = m0[i]
    out = oracle.scan(sc, "f64")
    Af, phf = abs(A), ph + (PI if A < 0 else 0.0)
    b = GYRO * Af
    psi = float(np.float32(pe)) + phf
    expect = rodrigues(m0.astype(np.float32).astype(np.float64), b * np.cos(psi), b * np.sin(psi),
                       GYRO * float(np.float32(dB0)), dt)
    np.testing.assert_allclose(M(out), expect, atol=1e-9)
    assert out.VVar["utsi"] == 1 and out.VVar["rfi"] == 2


def test_rotation_preserves_norm_without_relaxation():
    """|M| is invariant under precession and rf when T1, T2 are huge."""
    sc = workloads.gre2d(n_pe=2, spins=1, spoil=False)
    sc.VObj["T1"][:] = 1e12
    sc.VObj["T2"][:] = 1e12
    out = oracle.scan(sc, "f64")
    n0 = np.sqrt(sc.VObj["Mx"] ** 2 + sc.VObj["My"] ** 2 + sc.VObj["Mz"] ** 2)
    n1 = np.sqrt(out.VObj["Mx"].astype(np.float64) ** 2 + out.VObj["My"] ** 2 + out.VObj["Mz"] ** 2)
    np.testing.assert_allclose(n1, n0, atol=1e-6)


def test_b1_scaling_flip_angle_is_exact():
    """DoB1Scaling weights exactly the samples the decoder holds (the smaller
    end of each interval), so for shapes without interior minima (rect,
    gaussian) the held area gives the nominal flip angle."""
    t = np.arange(101) * 1e-5
    for rf in (psd.rf_rect(t, 37.0), psd.rf_gaussian(t, 37.0)):
        a = rf.amp
        held = np.where(np.abs(a[:-1]) <= np.abs(a[1:]), a[:-1], a[1:])
        assert abs(np.degrees(GYRO * np.sum(held) * 1e-5) - 37.0) < 1e-9
    # a sinc's valley samples are the smaller end of two intervals but are
    # weighted once: the held flip angle is close to, not exactly, nominal
    a = psd.rf_sinc(t, 37.0).amp
    held = np.where(np.abs(a[:-1]) <= np.abs(a[1:]), a[:-1], a[1:])
    assert abs(np.degrees(GYRO * np.sum(held) * 1e-5) - 37.0) < 1.0


def test_ernst_steady_state_spoiled_gre():
    """A spoiled GRE (ideal spoiler every TR) on one isochromat reaches the
    Ernst steady state Mz = M0 (1 - E1) / (1 - E1 cos a) before each pulse."""
    T1, TR, fa = 0.05, 5e-3, 30.0
    n = 500
    dt = TR / n
    trs = []
    t = np.arange(n + 1) * dt
    for _ in range(60):
        tr = psd.TR(duration=TR)
        tr.add_rf(psd.rf_rect(t[:4], fa))    # samples 0 A A 0: one held step of dt
        tr.add_gx(t, np.zeros(n + 1))
        tr.add_gy(t[[0, n]], [0, 0])
        tr.add_gz(t[[0, n]], [0, 0])
        tr.add_adc(t[[0, n]], [0, 0])
        tr.add_ext(t[[n - 3, n - 2, n - 1]], [0, 6, 0])
        trs.append(tr.finalize())
    seq = psd.concat(trs)
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.ones(shape + (1,)), np.full(shape + (1,), T1), np.full(shape + (1,), 1e-3),
                                np.zeros(shape), (1e-3,) * 3, seq=seq, tr_num=60)
    out = oracle.scan(sc, "f64")
    E1 = np.exp(-TR / T1)
    ernst = (1 - E1) / (1 - E1 * np.cos(np.radians(fa)))
    # state at the end of the last TR = just before the next pulse
    assert abs(M(out)[2] - ernst) < 2e-3


def test_kspace_centre_and_decoder_counters():
    """The GRE generator puts k = 0 at TE; the oracle decoder ends with the
    counters the sequence implies (utsi = n_uts-1, rfi = n_rf, ...)."""
    sc = workloads.gre2d(n_pe=4)
    out = oracle.scan(sc, "f64")
    kx = out.VSig["Kx"].reshape(4, 64)
    ky = out.VSig["Ky"].reshape(4, 64)
    assert np.allclose(kx[:, 32], 0.0, atol=1e-9)
    np.testing.assert_allclose(ky[:, 0], (np.arange(4) - 2) / 0.256, atol=1e-9)
    np.testing.assert_allclose(np.diff(kx, axis=1), 1 / 0.256, atol=1e-9)
    assert out.VVar["utsi"] == sc.VSeq["utsLine"].size - 1
    assert out.VVar["rfi"] == sc.VSeq["rfAmpLine"].size
    assert out.VVar["ADCi"] == sc.VSeq["ADCLine"].size
    assert out.VVar["TRCount"] == 4


def test_f32_oracle_tracks_f64():
    sc = workloads.gre2d(n_pe=4)
    a, b = oracle.scan(sc, "f32"), oracle.scan(sc, "f64")
    sa, sb = a.VSig["Sx"] + 1j * a.VSig["Sy"], b.VSig["Sx"] + 1j * b.VSig["Sy"]
    assert np.linalg.norm(sa - sb) / np.linalg.norm(sb) < 1e-3


def test_signal_is_brute_force_coil_sum():
    """The ADC sample equals sum over voxels of (Mx + iMy) conj(C) exp(-i rfRef),
    checked against a brute-force fp64 sum from the magnetization captured
    in RunMode 1 (DoScanAtCPU.cpp:445-491, 544-552)."""
    sc = workloads.bssfp3d(shape=(4, 4, 2), n_ro=8, spins=2, n_coils=3, n_tr=2, tr_steps=80,
                           fov=(0.016, 0.016, 0.008))
    seq = sc.VSeq
    sc2 = workloads.with_run_mode1(sc)
    out = oracle.scan(sc2, "f64")
    adc_slots = adc_capture_slots(seq)
    N = int(out.VSig["SignalNum"])
    mx, my = out.VSig["Mx"], out.VSig["My"]
    C = sc2.VCoi["RxCoilx"].astype(np.float64) + 1j * sc2.VCoi["RxCoily"].astype(np.float64)
    s = out.VSig["Sx"] + 1j * out.VSig["Sy"]
    s = s.reshape(3, N)
    assert len(adc_slots) == N
    rfref = np.pi * (np.arange(2) % 2)
    for n, u in enumerate(adc_slots):
        m = (mx[..., u] + 1j * my[..., u]).astype(np.complex128)  # [R,C,Z,S,1]
        msum = m.sum(axis=(3, 4))
        ref = rfref[n // 8]
        for c in range(3):
            brute = np.sum(msum * np.conj(C[..., c])) * np.exp(-1j * ref)
            # the captured state is fp32 (VSig.M* are single), the oracle sums fp64
            assert abs(brute - s[c, n]) < 1e-6 * max(1.0, abs(brute)), (n, c)
