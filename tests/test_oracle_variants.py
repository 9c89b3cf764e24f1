"""The MT, CEST and ME layouts of the multi-pool oracle, pinned by closed
forms and by the GM model they specialize.

  * ME is the GM model with K transposed: with a zero diagonal the oracle's
    split and exact schemes give the GM results bit for bit;
  * MT water keeps its transverse magnetization out of the exchange (free
    decay at R2 only) and exchanges longitudinally with the bound pool as the
    analytic 2x2 system; CEST pools exchange every component with water;
  * the exact scheme of CEST equals scipy's expm of the hand-built
    generator; the split of every variant converges to it at first order.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from mrilab_amd import workloads
from mrilab_amd.psd import GYRO
from tests.test_oracle_multipool import pool_scan


def mag(sc):
    return np.concatenate([sc.VObj[k].ravel() for k in ("Mx", "My", "Mz")])


def base3(**kw):
    return workloads.mt_cest(shape=(3, 3, 1), n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.0, **kw)


@pytest.mark.parametrize("scheme", [0, 1])
def test_me_is_gm_with_transposed_k(scheme):
    sc = workloads.select_pools(base3(), [0, 1])            # two free pools
    gm = oracle.gm_scan(sc, "f32" if scheme == 0 else "f64", scheme)
    me = oracle.gm_scan(workloads.as_variant(sc, "ME"), "f32" if scheme == 0 else "f64", scheme, variant="ME")
    np.testing.assert_array_equal(mag(me), mag(gm))
    np.testing.assert_array_equal(me.VSig["Sx"], gm.VSig["Sx"])


def test_me_self_exchange_is_first_order_small():
    """A diagonal rate exchanges a pool with its own start-of-step state: an
    O(dt) change in the split, none in the exact scheme."""
    sc = workloads.select_pools(base3(), [0, 1])
    a = oracle.gm_scan(workloads.as_variant(sc, "ME"), "f64", 1, variant="ME")
    b = oracle.gm_scan(workloads.as_variant(sc, "ME", self_rate=50.0), "f64", 1, variant="ME")
    np.testing.assert_array_equal(mag(a), mag(b))
    c = oracle.gm_scan(workloads.as_variant(sc, "ME", self_rate=50.0), "f64", 0, variant="ME")
    d = oracle.gm_scan(workloads.as_variant(sc, "ME"), "f64", 0, variant="ME")
    sc_, sd = c.VSig["Sx"] + 1j * c.VSig["Sy"], d.VSig["Sx"] + 1j * d.VSig["Sy"]
    assert 0 < np.linalg.norm(sc_ - sd) < 1e-3 * np.linalg.norm(sd)


def mt_two_pool(M0, T=0.05, dt=1e-4, kwb=20.0, kbw=30.0):
    sc = pool_scan(2, [0, 1], [0.8, 1.0], [0.1, 10e-6], [0.7, 0.3], [[0, kwb], [kbw, 0]], M0, {}, T, dt)
    return workloads.as_variant(sc, "MT")


@pytest.mark.parametrize("scheme", [0, 1])
def test_mt_longitudinal_exchange_matches_analytic(scheme):
    T, kwb, kbw = 0.05, 20.0, 30.0
    sc = mt_two_pool([(0, 0, 0.1), (0, 0, -0.2)], T, 1e-5 if scheme == 0 else 1e-4, kwb, kbw)
    out = oracle.gm_scan(sc, "f64", scheme, variant="MT")
    R1w, R1b = 1 / np.float32(0.8), 1 / np.float32(1.0)
    A = np.array([[-(R1w + kwb), kbw, R1w * np.float32(0.7)],
                  [kwb, -(R1b + kbw), R1b * np.float32(0.3)], [0, 0, 0]])
    x = expm(A * T) @ np.array([0.1, -0.2, 1.0])
    got = [out.VObj["Mz"][0, 0, 0, 0, 0], out.VObj["Mz"][0, 0, 0, 0, 1]]
    tol = 1e-7 if scheme == 1 else 2e-4   # split: O(dt) away from the exact solution
    np.testing.assert_allclose(got, x[:2], atol=tol)


@pytest.mark.parametrize("scheme", [0, 1])
def test_mt_water_transverse_not_exchanged(scheme):
    T = 0.02
    sc = mt_two_pool([(1.0, 0, 0), (0, 0, 0.3)], T, 1e-4)
    out = oracle.gm_scan(sc, "f64", scheme, variant="MT")
    mxy = np.hypot(out.VObj["Mx"][0, 0, 0, 0, 0], out.VObj["My"][0, 0, 0, 0, 0])
    assert mxy == pytest.approx(np.exp(-T / np.float32(0.1)), rel=1e-9)


def cest_scan():
    sc = workloads.select_pools(base3(), [0, 2, 1], cs_hz=[0.0, 0.0, 447.0])  # water, MT, CEST
    return workloads.as_variant(sc, "CEST")


def test_cest_exact_matches_expm_of_generator():
    """No rf, no gradients: one step of the exact scheme is expm of the
    CEST generator built from the plane layout (per voxel and spin)."""
    sc = cest_scan()
    T, dt = 2e-3, 1e-4
    sc = pool_scan(3, [0, 1, 0], [1.2, 1.0, 0.9], [0.08, 10e-6, 0.05], [0.8, 0.15, 0.05],
                   [[0, 15, 40], [25, 0, 0], [300, 0, 0]], [(0.3, 0.1, 0.6), (0, 0, 0.1), (0.02, 0, 0.04)],
                   {}, T, dt)
    sc.VCtl["CS"] = np.array([0.0, 0.0, 447.0])
    sc = workloads.as_variant(sc, "CEST")
    out = oracle.gm_scan(sc, "f64", 1, variant="CEST")
    f = np.float32
    R1 = [1 / f(1.2), 1 / f(1.0), 1 / f(0.9)]
    R2 = [1 / f(0.08), None, 1 / f(0.05)]
    rho = [f(0.8), f(0.15), f(0.05)]
    kwm, kmw, kwc, kcw = 15.0, 25.0, 40.0, 300.0
    w2 = 2 * np.pi * 447.0
    # state: wx wy wz mz cx cy cz 1
    A = np.zeros((8, 8))
    A[0, 0] = A[1, 1] = -(R2[0] + kwc)
    A[0, 4] = A[1, 5] = kcw
    A[2, 2] = -(R1[0] + kwm + kwc); A[2, 3] = kmw; A[2, 6] = kcw; A[2, 7] = R1[0] * rho[0]
    A[3, 3] = -(R1[1] + kmw); A[3, 2] = kwm; A[3, 7] = R1[1] * rho[1]
    A[4, 4] = A[5, 5] = -(R2[2] + kcw)
    A[4, 0] = A[5, 1] = kwc
    A[4, 5] += w2; A[5, 4] -= w2           # dM/dt = M x (0, 0, w)
    A[6, 6] = -(R1[2] + kcw); A[6, 2] = kwc; A[6, 7] = R1[2] * rho[2]
    x = expm(A * T) @ np.array([0.3, 0.1, 0.6, 0.1, 0.02, 0.0, 0.04, 1.0])
    o = out.VObj
    got = [o["Mx"][0, 0, 0, 0, 0], o["My"][0, 0, 0, 0, 0], o["Mz"][0, 0, 0, 0, 0], o["Mz"][0, 0, 0, 0, 1],
           o["Mx"][0, 0, 0, 0, 2], o["My"][0, 0, 0, 0, 2], o["Mz"][0, 0, 0, 0, 2]]
    np.testing.assert_allclose(got, x[:7], atol=1e-7)


@pytest.mark.parametrize("variant", ["MT", "CEST", "ME"])
def test_split_converges_to_exact(variant):
    errs = []
    for dt in (4e-5, 2e-5, 1e-5):
        sc = workloads.mt_cest(shape=(2, 2, 1), n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.0, sat_fa=200.0,
                               dt=dt, seed=5)
        if variant == "MT":
            sc = workloads.select_pools(sc, [0, 2])
        elif variant == "CEST":
            sc = workloads.select_pools(sc, [0, 2, 1], cs_hz=[0.0, 0.0, 447.0])
        sc = workloads.as_variant(sc, variant)
        a = oracle.gm_scan(sc, "f64", 0, variant=variant)
        b = oracle.gm_scan(sc, "f64", 1, variant=variant)
        errs.append(np.linalg.norm(mag(a) - mag(b)))
    assert errs[1] < 0.65 * errs[0] and errs[2] < 0.65 * errs[1]


@pytest.mark.parametrize("variant", ["MT", "CEST"])
def test_water_only_signal(variant):
    sc = workloads.mt_cest(shape=(2, 2, 1), n_offsets=3, sat_time=0.04)
    sc = workloads.select_pools(sc, [0, 2] if variant == "MT" else [0, 2, 1])
    out = oracle.gm_scan(workloads.as_variant(sc, variant), "f32", 0, variant=variant)
    n = int(out.VSig["SignalNum"])
    s = (out.VSig["Sx"] + 1j * out.VSig["Sy"]).reshape(-1, n)
    assert np.all(np.abs(s[0]) > 0) and np.all(s[1:] == 0)
