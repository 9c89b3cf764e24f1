"""The multi-pool oracle (GM model) pinned by closed forms.

The reference's multi-pool engines exist only as CUDA kernels (no NVIDIA
toolchain or GPU here) and ship no outputs, so the oracle's GM model
(oracle/bloch_oracle.c, restating BlochKernelGM.cuh) is pinned by:
  * the super-Lorentzian lineshape sum vs adaptive quadrature of the
    integral it discretizes;
  * two free pools exchanging without rf vs the analytic solution of the
    2x2 rate system (pins the K(a,b) = rate a -> b index convention);
  * a bound pool under CW saturation vs its exponential approach to
    R1*M0/(R1+W) (pins W = pi*(gamma*B1)^2*G);
  * a single free pool vs closed-form precession and relaxation;
  * first-order convergence of the operator split to the exact scheme.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle
from mrilab_amd import psd, workloads
from mrilab_amd.psd import GYRO

PI = 3.14159265359


def sl_integral(dw, t2b):
    f = lambda u: np.sqrt(2 / np.pi) * t2b / abs(3 * u * u - 1) * np.exp(-2 * (dw * t2b / (3 * u * u - 1)) ** 2)
    a, _ = integrate.quad(f, 0, 1 / np.sqrt(3), limit=400)
    b, _ = integrate.quad(f, 1 / np.sqrt(3), 1, limit=400)
    return a + b


@pytest.mark.parametrize("dw_hz", [200.0, 1000.0, 3000.0, 6000.0])
def test_superlorentzian_matches_quadrature(dw_hz):
    t2b = 10e-6
    dw = 2 * np.pi * dw_hz
    g = oracle.superlorentzian(dw, t2b, "f64")
    ref = sl_integral(dw, t2b)
    assert abs(g - ref) / ref < 2e-3


def pool_scan(n_types, flags, T1, T2, rho, K, M0, events, T, dt, b1_mg=1.0, dB0=0.0):
    """1x1x1 phantom with n_types pools, one TR built from ``events``."""
    tr = psd.TR(duration=T)
    rf = events.get("rf") or psd.RF(np.array([0.0, T]), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2))
    tr.add_rf(rf)
    t = np.arange(int(round(T / dt)) + 1) * dt
    tr.add_gz([0.0, T], [0, 0])
    tr.add_gy([0.0, T], [0, 0])
    tr.add_gx(t, np.zeros(t.size))
    tr.add_adc([0.0, T], [0, 0])
    tr.add_ext([0.0, T], [0, 0])
    seq = psd.concat([tr.finalize()])
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.array(rho, float).reshape(shape + (n_types,)),
                                np.array(T1, float).reshape(shape + (n_types,)),
                                np.array(T2, float).reshape(shape + (n_types,)),
                                np.full(shape, dB0), (1e-3,) * 3, seq=seq,
                                tx=(np.full(shape, b1_mg), np.zeros(shape)))
    for t_ in range(n_types):
        sc.VObj["Mx"][..., 0, t_] = M0[t_][0]
        sc.VObj["My"][..., 0, t_] = M0[t_][1]
        sc.VObj["Mz"][..., 0, t_] = M0[t_][2]
    sc.VObj["K"] = np.asfortranarray(np.array(K, np.float32).reshape(shape + (n_types, n_types)))
    sc.VObj["TypeFlag"] = np.array(flags, float)
    sc.VCtl["CS"] = np.zeros(n_types)
    return sc


def mz(sc, t):
    return float(sc.VObj["Mz"][0, 0, 0, 0, t])


def test_two_pool_exchange_no_rf_matches_analytic():
    """dMz/dt = -(R1a + kab) Mz_a + kba Mz_b + R1a M0a (and symmetric) with
    K(a,b) = rate a -> b; compare with expm of the 3x3 affine system."""
    T1 = [0.8, 0.3]; T2 = [0.1, 0.02]; rho = [0.7, 0.3]
    kab, kba = 20.0, 20.0 * 0.7 / 0.3
    K = [[0, kab], [kba, 0]]
    M0 = [(0, 0, 0.1), (0, 0, -0.2)]
    T, dt = 0.05, 1e-4
    sc = pool_scan(2, [0, 0], T1, T2, rho, K, M0, {}, T, dt)
    out = oracle.gm_scan(sc, "f64", 1)
    A = np.array([[-(1 / np.float32(T1[0]) + kab), kba, 1 / np.float32(T1[0]) * np.float32(rho[0])],
                  [kab, -(1 / np.float32(T1[1]) + kba), 1 / np.float32(T1[1]) * np.float32(rho[1])],
                  [0, 0, 0]], dtype=np.float64)
    from scipy.linalg import expm
    x = expm(A * T) @ np.array([0.1, -0.2, 1.0])
    assert abs(mz(out, 0) - x[0]) < 1e-6 and abs(mz(out, 1) - x[1]) < 1e-6


def test_bound_pool_cw_saturation_closed_form():
    """A bound pool (no exchange) under constant rf at offset F relaxes as
    Mz(t) = Minf + (Mz0 - Minf) exp(-(R1 + W) t), Minf = R1 M0 / (R1 + W),
    W = pi (gamma B1)^2 G(2 pi F + gamma dB0)."""
    T, dt = 0.02, 1e-4
    n = int(round(T / dt))
    t = np.arange(n + 1) * dt
    B1, F = 2e-6, 2000.0
    rf = psd.RF(t, np.full(n + 1, B1), np.zeros(n + 1), np.full(n + 1, F), np.ones(n + 1))
    rf.amp[-1] = B1
    sc = pool_scan(2, [0, 1], [1.0, 1.0], [0.05, 10e-6], [0.0, 0.5], [[0, 0], [0, 0]],
                   [(0, 0, 0), (0, 0, 0.5)], {"rf": rf}, T, dt)
    # pool 0 has Rho = 0: it is skipped, leaving the bound pool alone
    out = oracle.gm_scan(sc, "f64", 1)
    G = sl_integral(2 * np.pi * F, 10e-6)
    W = np.pi * (GYRO * B1) ** 2 * G
    R1 = 1.0
    minf = R1 * 0.5 / (R1 + W)
    expect = minf + (0.5 - minf) * np.exp(-(R1 + W) * T)
    assert abs(mz(out, 1) - expect) / 0.5 < 2e-3   # lineshape sum vs integral


def test_single_free_pool_exact_scheme_is_bloch():
    """With one free pool and a bound pool of zero density, the exact scheme
    is the closed-form Bloch solution."""
    T, dt = 0.01, 1e-5
    dB0 = 1e-6
    sc = pool_scan(2, [0, 1], [0.5, 1.0], [0.04, 1e-5], [1.0, 0.0], [[0, 0], [0, 0]],
                   [(1.0, 0, 0.2), (0, 0, 0)], {}, T, dt, dB0=dB0)
    out = oracle.gm_scan(sc, "f64", 1)
    w = GYRO * float(np.float32(dB0))
    e2 = np.exp(-T / float(np.float32(0.04)))
    e1 = np.exp(-T / float(np.float32(0.5)))
    np.testing.assert_allclose([out.VObj["Mx"][0, 0, 0, 0, 0], out.VObj["My"][0, 0, 0, 0, 0]],
                               [e2 * np.cos(w * T), -e2 * np.sin(w * T)], atol=1e-7)
    assert abs(mz(out, 0) - (1 + (0.2 - 1) * e1)) < 1e-7


def test_split_converges_to_exact_first_order():
    """The GM split scheme approaches the exact scheme with O(dt) error."""
    errs = []
    for dt in (4e-5, 2e-5, 1e-5):
        sc = workloads.mt_cest(shape=(2, 2, 1), n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.0,
                               sat_fa=200.0, dt=dt, seed=5)
        a = oracle.gm_scan(sc, "f64", 0)
        b = oracle.gm_scan(sc, "f64", 1)
        ma = np.concatenate([a.VObj[k].ravel() for k in ("Mx", "My", "Mz")])
        mb = np.concatenate([b.VObj[k].ravel() for k in ("Mx", "My", "Mz")])
        errs.append(np.linalg.norm(ma - mb))
    assert errs[1] < 0.65 * errs[0] and errs[2] < 0.65 * errs[1]


def test_split_scheme_fp32_tracks_fp64():
    sc = workloads.mt_cest(shape=(3, 3, 1), n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.02)
    a, b = oracle.gm_scan(sc, "f32", 0), oracle.gm_scan(sc, "f64", 0)
    sa, sb = a.VSig["Sx"] + 1j * a.VSig["Sy"], b.VSig["Sx"] + 1j * b.VSig["Sy"]
    assert np.linalg.norm(sa - sb) / np.linalg.norm(sb) < 1e-3
