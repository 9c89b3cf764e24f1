"""Pins of the oracle branches that the single-coil, single-pool closed forms
leave open: multi-transmit rf superposition (BlochKernel.h:73-121 and
BlochKernelGM.cuh:348-380), the one-step form of the multi-pool operator
split (BlochKernelGM.cuh:159-167,408-433; BlochKernelMT.cuh:325-333,
:400-407) and k-space lock/release (Plugin_LockK.m, Plugin_ReleaseK.m).

Each pin compares the oracle with an answer fixed independently of the
branch under test: the multi-transmit rf step against the single-coil step
of the summed complex field, the averaged saturation rate against the
closed-form CW saturation of a bound pool, the split step with no exchange
against the single-pool scan, one split step against a hand evaluation with
the partner pools' start-of-step state, and the locked k-space against the
gradient area. ``test_planted_mistakes_fail_a_pin`` compiles the oracle
with each mistake of its ORC_MUTATION list and checks that a pin fails.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import psd, workloads
from mrilab_amd.psd import GYRO

PI = 3.14159265359


# --------------------------------------------------------------------------
# builders
# --------------------------------------------------------------------------

def _pulse(n_rf, dt, b1, phase, freq, coil, n):
    """A rect rf of amplitude b1 (T) on grid points 0..n_rf of an n-step TR."""
    t = np.arange(n_rf + 1) * dt
    amp = np.full(n_rf + 1, float(b1))
    amp[0] = amp[-1] = 0.0
    ph = np.full(n_rf + 1, float(phase))
    fr = np.full(n_rf + 1, float(freq))
    ph[0] = ph[-1] = fr[0] = fr[-1] = 0.0
    return psd.RF(t, amp, ph, fr, np.full(n_rf + 1, float(coil)))


def tx_scan(coils, mg, pe, n=60, n_rf=30, dt=1e-5, T1=0.5, T2=0.05, dB0=2e-7, dwrnd=30.0, cs=5.0,
            M0=(0.1, -0.2, 0.9), multitransmit=True, master=1, ext7=False):
    """One voxel, one TR of n steps; ``coils`` lists (b1 (T), phase, freq) of
    rect pulses fired together on grid points 0..n_rf, coil c+1 with B1+ map
    mg[c] * exp(i pe[c]); a gradient sample on every grid point keeps the
    lines fired and moves the voxel (at x, y, z != 0) off resonance."""
    tr = psd.TR(duration=n * dt)
    for c, (b1, ph, fr) in enumerate(coils):
        tr.add_rf(_pulse(n_rf, dt, b1, ph, fr, c + 1, n))
    t = np.arange(n + 1) * dt
    tr.add_gz(t, np.full(n + 1, 1e-3))
    tr.add_gy(t, np.full(n + 1, -2e-3))
    tr.add_gx(t, np.full(n + 1, 0.5e-3))
    tr.add_adc([0.0, n * dt], [0, 0])
    if ext7:
        tr.add_ext(t[[n_rf // 2 - 1, n_rf // 2, n_rf // 2 + 1]], [0, 7, 0])
    else:
        tr.add_ext([0.0, n * dt], [0, 0])
    seq = psd.concat([tr.finalize()], multitransmit=multitransmit)
    shape = (1, 1, 1)
    ntx = len(coils)
    sc = workloads.make_structs(np.full(shape + (1,), 1.0), np.full(shape + (1,), T1), np.full(shape + (1,), T2),
                                np.full(shape, dB0), (1e-3, 1e-3, 1e-3), dWRnd=np.full(shape + (1, 1), dwrnd),
                                seq=seq, cs=[cs],
                                tx=(np.asarray(mg, float).reshape(shape + (ntx,)),
                                    np.asarray(pe, float).reshape(shape + (ntx,))),
                                multitransmit=multitransmit, master_tx=master)
    sc.VObj["Mx"][:] = M0[0]
    sc.VObj["My"][:] = M0[1]
    sc.VObj["Mz"][:] = M0[2]
    sc.VMag["Gxgrid"][:] = 0.004
    sc.VMag["Gygrid"][:] = -0.003
    sc.VMag["Gzgrid"][:] = 0.002
    return sc


def M(sc):
    return np.array([sc.VObj[k].ravel()[0] for k in ("Mx", "My", "Mz")], dtype=np.float64)


def f32(x):
    return float(np.float32(x))


# The engine's state arrays (VObj.Mx/My/Mz) are fp32, so the fp64 oracle's
# result is rounded once on output: results that agree in exact arithmetic
# agree to one fp32 ulp of values below 1.
ULP = 2.0 ** -23


def assert_f32_close(got, want):
    np.testing.assert_allclose(got, want, rtol=0, atol=ULP)


# --------------------------------------------------------------------------
# multi-transmit superposition
# --------------------------------------------------------------------------

MTX_CASES = {
    # two coils of unequal amplitude and phase on unequal maps: the step is
    # the single-coil step of the summed complex field (PolarToCart, sum,
    # CartToPolar with atan2(im, re))
    "superposition": dict(coils=[(3e-6, 0.4, 0.0), (5e-6, -1.9, 0.0)], mg=[0.9, 1.2], pe=[0.3, -0.7]),
    # per-coil rf frequency: each active coil adds 2*pi*f to the off-resonance
    "frequencies": dict(coils=[(4e-6, 0.2, 150.0), (2e-6, 1.1, -40.0)], mg=[1.0, 0.8], pe=[0.0, 0.5]),
    # a coil with amplitude 0 contributes neither field nor frequency
    "silent_coil": dict(coils=[(4e-6, 0.6, 120.0), (0.0, 0.0, 300.0)], mg=[1.1, 0.9], pe=[-0.2, 0.4]),
    # a negative amplitude is |A| at phase + pi
    "negative_amplitude": dict(coils=[(-3e-6, 0.7, 0.0), (2e-6, 0.1, 0.0), (1e-6, 2.5, 60.0)],
                               mg=[1.0, 0.7, 1.3], pe=[0.1, -0.3, 0.9]),
}


def equivalent_single(case):
    """The single-coil pulse with the summed complex field and summed
    frequency of the active coils; fp64 numpy, independent of the oracle.
    The B1+ maps are fp32 arrays, the sequence lines fp64."""
    b = 0j
    f = 0.0
    for (a, ph, fr), m, p in zip(case["coils"], case["mg"], case["pe"]):
        m, p = f32(m), f32(p)
        if a == 0:
            continue
        if a < 0:
            a, ph = -a, ph + PI
        b += a * m * np.exp(1j * (p + ph))
        f += fr
    return abs(b), float(np.angle(b)), f


def run_mtx(case, lib=None):
    sc = tx_scan(case["coils"], case["mg"], case["pe"])
    b1, ph, fr = equivalent_single(case)
    one = tx_scan([(b1, ph, fr)], [1.0], [0.0], multitransmit=False)
    return M(oracle.scan(sc, "f64", lib=lib)), M(oracle.scan(one, "f64"))


@pytest.mark.parametrize("name", list(MTX_CASES))
def test_multi_tx_step_is_single_coil_step_of_summed_field(name):
    got, want = run_mtx(MTX_CASES[name])
    # the rf rotates M far from its start (a pin, not a no-op)
    assert np.linalg.norm(want - np.array([0.1, -0.2, 0.9])) > 0.1
    assert_f32_close(got, want)


def test_multi_tx_fp32_build_agrees():
    for case in MTX_CASES.values():
        sc = tx_scan(case["coils"], case["mg"], case["pe"])
        a, b = M(oracle.scan(sc, "f32")), M(oracle.scan(sc, "f64"))
        np.testing.assert_allclose(a, b, atol=2e-5)


def test_multi_tx_rfref_uses_master_coil_phase():
    """Ext 7 with MultiTransmit on takes rfRef = rfPhase(MasterTxCoil)
    (Plugin_rfRef.m:7-12); the held phase is the pulse's phase."""
    coils = [(3e-6, 0.4, 0.0), (2e-6, -1.3, 0.0)]
    for master, want in ((1, 0.4), (2, -1.3)):
        sc = tx_scan(coils, [1.0, 1.0], [0.0, 0.0], master=master, ext7=True)
        out = oracle.scan(sc, "f64")
        assert float(out.VVar["rfRef"]) == want


def gm_tx_scan(coils, mg, T=0.02, dt=1e-4, F=2000.0, rho=(0.0, 0.5), T2=(0.05, 10e-6), M0=((0, 0, 0), (0, 0, 0.5)),
               K=((0, 0), (0, 0)), dB0=0.0, free_pool_only=False):
    """One voxel, pool 0 free and pool 1 bound (TypeFlag 1), rect rf of
    every coil at frequency F over the whole TR."""
    n = int(round(T / dt))
    t = np.arange(n + 1) * dt
    tr = psd.TR(duration=T)
    for c, (b1, ph) in enumerate(coils):
        amp = np.full(n + 1, float(b1))
        tr.add_rf(psd.RF(t, amp, np.full(n + 1, float(ph)), np.full(n + 1, F), np.full(n + 1, float(c + 1))))
    tr.add_gz([0.0, T], [0, 0])
    tr.add_gy([0.0, T], [0, 0])
    tr.add_gx(t, np.zeros(n + 1))
    tr.add_adc([0.0, T], [0, 0])
    tr.add_ext([0.0, T], [0, 0])
    seq = psd.concat([tr.finalize()], multitransmit=True)
    # finalize zeroes the first and last samples; hold the amplitude there
    shape = (1, 1, 1)
    ntx = len(coils)
    sc = workloads.make_structs(np.array(rho, float).reshape(shape + (2,)), np.array([0.8, 1.0]).reshape(shape + (2,)),
                                np.array(T2, float).reshape(shape + (2,)), np.full(shape, dB0), (1e-3,) * 3, seq=seq,
                                tx=(np.asarray(mg, float).reshape(shape + (ntx,)), np.zeros(shape + (ntx,))),
                                multitransmit=True)
    for p in range(2):
        for k, comp in enumerate(("Mx", "My", "Mz")):
            sc.VObj[comp][..., 0, p] = M0[p][k]
    sc.VObj["K"] = np.asfortranarray(np.array(K, np.float32).reshape(shape + (2, 2)))
    sc.VObj["TypeFlag"] = np.array([0.0, 1.0])
    sc.VCtl["CS"] = np.zeros(2)
    return sc


def sl_integral(dw, t2b):
    from scipy import integrate
    f = lambda u: np.sqrt(2 / np.pi) * t2b / abs(3 * u * u - 1) * np.exp(-2 * (dw * t2b / (3 * u * u - 1)) ** 2)
    a, _ = integrate.quad(f, 0, 1 / np.sqrt(3), limit=400)
    b, _ = integrate.quad(f, 1 / np.sqrt(3), 1, limit=400)
    return a + b


def gm_two_coil_saturation(scheme, lib=None):
    """Bound pool alone (free pool of zero density) under two coils of B1
    2 uT and 1 uT (maps 1.0 and 1.5): W is the mean of the coils' rates,
    pi*gamma^2*G*(B1a^2 + B1b^2)/2 (BlochKernelGM.cuh:348-380), and Mz
    relaxes to R1 M0 / (R1 + W) at rate R1 + W."""
    T, F = 0.02, 2000.0
    sc = gm_tx_scan([(2e-6, 0.0), (1e-6 / 1.5, 1.0)], [1.0, 1.5], T=T, F=F)
    out = oracle.gm_scan(sc, "f64", scheme, lib=lib)
    G = sl_integral(2 * np.pi * F, 10e-6)
    n = int(round(T / 1e-4))
    # the decoder holds min(|A_k|, |A_k+1|): finalize zeroed the last sample,
    # so rf is on for n - 1 of the n steps
    W = np.pi * GYRO ** 2 * G * ((2e-6) ** 2 + (1e-6) ** 2) / 2
    R1 = 1.0
    ton = (n - 1) * 1e-4
    minf = R1 * 0.5 / (R1 + W)
    mz = minf + (0.5 - minf) * np.exp(-(R1 + W) * ton)
    mz = 0.5 + (mz - 0.5) * np.exp(-R1 * 1e-4)   # last step: relaxation only
    return float(out.VObj["Mz"][0, 0, 0, 0, 1]), mz


@pytest.mark.parametrize("scheme,tol", [(1, 2e-3), (0, 6e-3)])
def test_gm_two_coil_saturation_rate_is_mean(scheme, tol):
    got, want = gm_two_coil_saturation(scheme)
    assert abs(got - want) / 0.5 < tol


def test_gm_multi_tx_free_pool_is_single_coil_step():
    """The free pool of the multi-pool split under two coils equals the
    single-pool multi-transmit step (BlochKernelGM.cuh:240-319 with
    BlochKernel.h:73-121's field), bound pool of zero density."""
    case = MTX_CASES["superposition"]
    b1, ph, fr = equivalent_single(case)
    a = tx_scan(case["coils"], case["mg"], case["pe"])
    b = tx_scan([(b1, ph, fr)], [1.0], [0.0], multitransmit=False)
    for sc in (a, b):   # the same voxel as pool 0 of a two-pool phantom
        _as_two_pool(sc)
    ga = M(oracle.gm_scan(a, "f64", 0))
    gb = M(oracle.gm_scan(b, "f64", 0))
    assert_f32_close(ga, gb)


def _as_two_pool(sc):
    """Adds a bound pool of zero density to a one-voxel single-pool scan."""
    for k in ("Mx", "My", "Mz", "dWRnd"):
        src = sc.VObj if k != "dWRnd" else sc.VMag
        a = src[k]
        src[k] = np.asfortranarray(np.concatenate([a, np.zeros_like(a)], axis=4))
    for k in ("Rho", "T1", "T2"):
        a = sc.VObj[k]
        extra = np.zeros_like(a) if k == "Rho" else np.full_like(a, 1.0 if k == "T1" else 10e-6)
        sc.VObj[k] = np.asfortranarray(np.concatenate([a, extra], axis=3))
    sc.VObj["TypeNum"] = np.int32(2)
    sc.VObj["K"] = np.zeros((1, 1, 1, 2, 2), dtype=np.float32, order="F")
    sc.VObj["TypeFlag"] = np.array([0.0, 1.0])
    sc.VCtl["CS"] = np.array([sc.VCtl["CS"][0], 0.0])
    n = sc.VSig["SignalNum"]
    sc.VSig["Sx"] = np.zeros(n * 2)
    sc.VSig["Sy"] = np.zeros(n * 2)


# --------------------------------------------------------------------------
# the split step
# --------------------------------------------------------------------------

def two_free_pools(K, T1=(0.7, 0.25), T2=(0.08, 0.02), rho=(0.6, 0.4), cs=(0.0, -35.0),
                   M0=((0.3, -0.1, 0.5), (-0.2, 0.25, 0.1)), n=80, dt=1e-5, rf=True, dB0=3e-7):
    """One voxel with two free pools (GM), a sinc pulse and gradients."""
    T = n * dt
    t = np.arange(n + 1) * dt
    tr = psd.TR(duration=T)
    if rf:
        tr.add_rf(psd.rf_sinc(t[:41], 60.0))
    else:
        tr.add_rf(psd.RF(np.array([0.0, T]), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2)))
    tr.add_gz(t, np.full(n + 1, 1e-3))
    tr.add_gy(t, np.linspace(-2e-3, 2e-3, n + 1))
    tr.add_gx(t, np.full(n + 1, 0.5e-3))
    tr.add_adc([0.0, T], [0, 0])
    tr.add_ext([0.0, T], [0, 0])
    seq = psd.concat([tr.finalize()])
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.array(rho, float).reshape(shape + (2,)), np.array(T1, float).reshape(shape + (2,)),
                                np.array(T2, float).reshape(shape + (2,)), np.full(shape, dB0), (1e-3,) * 3,
                                dWRnd=np.array([[[[[20.0, -15.0]]]]]), seq=seq, cs=list(cs))
    for p in range(2):
        for k, comp in enumerate(("Mx", "My", "Mz")):
            sc.VObj[comp][..., 0, p] = M0[p][k]
    sc.VMag["Gxgrid"][:] = 0.004
    sc.VMag["Gygrid"][:] = -0.003
    sc.VMag["Gzgrid"][:] = 0.002
    sc.VObj["K"] = np.asfortranarray(np.array(K, np.float32).reshape(shape + (2, 2)))
    sc.VObj["TypeFlag"] = np.array([0.0, 0.0])
    return sc


def single_pool_of(sc, p):
    """Pool p of a two-pool one-voxel scan as a single-pool scan."""
    out = sc.copy()
    for src, keys in ((out.VObj, ("Mx", "My", "Mz")), (out.VMag, ("dWRnd",))):
        for k in keys:
            src[k] = np.asfortranarray(src[k][..., p:p + 1])
    for k in ("Rho", "T1", "T2"):
        out.VObj[k] = np.asfortranarray(out.VObj[k][..., p:p + 1])
    out.VObj["TypeNum"] = np.int32(1)
    out.VCtl["CS"] = np.array([out.VCtl["CS"][p]])
    del out.VObj["K"], out.VObj["TypeFlag"]
    n = out.VSig["SignalNum"]
    out.VSig["Sx"] = np.zeros(n)
    out.VSig["Sy"] = np.zeros(n)
    return out


def test_split_without_exchange_is_single_pool_scan():
    """K = 0: each free pool of the split steps exactly as the single-pool
    engine steps that pool alone (rf rotation, off-resonance with its own
    CS and dWRnd, relaxation), to fp64 rounding; the two kernels round the
    relaxation factors in different orders."""
    sc = two_free_pools(((0, 0), (0, 0)))
    gm = oracle.gm_scan(sc, "f64", 0)
    for p in range(2):
        one = oracle.scan(single_pool_of(sc, p), "f64")
        got = np.array([gm.VObj[k][0, 0, 0, 0, p] for k in ("Mx", "My", "Mz")], dtype=np.float64)
        assert_f32_close(got, M(one))


def split_one_step(lib=None):
    """One step (dt 10 ms) of two free pools without rf, T1 = T2 = inf, k*dt
    = 0.5 and 0.2, pool 0 off resonance: the split rotates each pool, then
    M_t' = R_t M_t e^(-dt k_t->p) + (1 - e^(-dt k_p->t)) M_p(start)
    (BlochKernelGM.cuh:159-167 buffer the start-of-step state, :408-433)."""
    dt = 0.01
    kab, kba = 50.0, 20.0        # K(a, b) = rate a -> b
    w = 40.0                     # rad/s, pool 0 only (dWRnd)
    Ma = np.array([0.6, -0.3, 0.2], dtype=np.float32).astype(np.float64)   # stored as fp32
    Mb = np.array([-0.1, 0.4, 0.8], dtype=np.float32).astype(np.float64)
    tr = psd.TR(duration=dt)
    tr.add_rf(psd.RF(np.array([0.0, dt]), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2)))
    for add in (tr.add_gz, tr.add_gy, tr.add_gx, tr.add_adc, tr.add_ext):
        add([0.0, dt], [0, 0])
    seq = psd.concat([tr.finalize()])
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.array([1.0, 1.0]).reshape(shape + (2,)), np.full(shape + (2,), 1e30),
                                np.full(shape + (2,), 1e30), np.zeros(shape), (1e-3,) * 3,
                                dWRnd=np.array([[[[[w, 0.0]]]]]), seq=seq, cs=[0.0, 0.0])
    for p, m0 in enumerate((Ma, Mb)):
        for k, comp in enumerate(("Mx", "My", "Mz")):
            sc.VObj[comp][..., 0, p] = m0[k]
    sc.VObj["K"] = np.asfortranarray(np.array([[0, kab], [kba, 0]], np.float32).reshape(shape + (2, 2)))
    sc.VObj["TypeFlag"] = np.array([0.0, 0.0])
    out = oracle.gm_scan(sc, "f64", 0, lib=lib)
    wf = f32(w)
    c, s = np.cos(wf * dt), np.sin(wf * dt)
    Ra = np.array([Ma[0] * c + Ma[1] * s, Ma[1] * c - Ma[0] * s, Ma[2]])
    ea, eb = np.exp(-dt * kab), np.exp(-dt * kba)
    want_a = Ra * ea + (1 - eb) * Mb
    want_b = Mb * eb + (1 - ea) * Ma
    got = [np.array([out.VObj[k][0, 0, 0, 0, p] for k in ("Mx", "My", "Mz")], dtype=np.float64) for p in (0, 1)]
    return got, (want_a, want_b)


def test_split_one_step_uses_start_of_step_partners():
    (ga, gb), (wa, wb) = split_one_step()
    assert_f32_close(ga, wa)
    assert_f32_close(gb, wb)


def mt_one_step(lib=None):
    """One MT split step (BlochKernelMT.cuh:318-333, :395-407): water
    Mz' = Mz e^(-dt/T1 - dt k12) + Rho(1 - e^(-dt/T1))/S + Mzb (1 - e^(-dt k21)),
    bound Mzb' = Mzb e^(-dt/T1b - dt k21) + Rhob(1 - e^(-dt/T1b))/S + Mz (1 - e^(-dt k12)),
    both with the other pool's start-of-step Mz; no rf, water on resonance."""
    dt, k12, k21 = 0.01, 40.0, 70.0
    T1w, T1b, rw, rb = 0.9, 1.0, 0.8, 0.2
    mzw, mzb = -0.3, 0.15
    tr = psd.TR(duration=dt)
    tr.add_rf(psd.RF(np.array([0.0, dt]), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2)))
    for add in (tr.add_gz, tr.add_gy, tr.add_gx, tr.add_adc, tr.add_ext):
        add([0.0, dt], [0, 0])
    seq = psd.concat([tr.finalize()])
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.array([rw, rb]).reshape(shape + (2,)), np.array([T1w, T1b]).reshape(shape + (2,)),
                                np.array([0.05, 10e-6]).reshape(shape + (2,)), np.zeros(shape), (1e-3,) * 3,
                                seq=seq, cs=[0.0, 0.0])
    sc.VObj["Mz"][..., 0, 0] = mzw
    sc.VObj["Mz"][..., 0, 1] = mzb
    K = np.zeros(shape + (4,), np.float32)
    K[..., 1] = k12
    K[..., 2] = k21
    sc.VObj["K"] = np.asfortranarray(K)
    out = oracle.gm_scan(sc, "f64", 0, variant="MT", lib=lib)
    T1w, T1b, rw, rb, mzw, mzb = (f32(x) for x in (T1w, T1b, rw, rb, mzw, mzb))
    want_w = mzw * np.exp(-dt / T1w - dt * k12) + rw * (1 - np.exp(-dt / T1w)) + mzb * (1 - np.exp(-dt * k21))
    want_b = mzb * np.exp(-dt / T1b - dt * k21) + rb * (1 - np.exp(-dt / T1b)) + mzw * (1 - np.exp(-dt * k12))
    return (float(out.VObj["Mz"][0, 0, 0, 0, 0]), float(out.VObj["Mz"][0, 0, 0, 0, 1])), (want_w, want_b)


def test_mt_one_step_hand_evaluated():
    (gw, gb), (ww, wb) = mt_one_step()
    assert_f32_close([gw, gb], [ww, wb])


# --------------------------------------------------------------------------
# k-space lock and release
# --------------------------------------------------------------------------

def lock_release(lib=None):
    """Gx = G on every grid point; Ext 3 (LockK) at point 20, Ext 4
    (ReleaseK) at point 60, ADC at points 10, 30, 60 and 70. k is read
    before the step's integration, so the samples sit at kx = 10, 30, 20
    and 30 grid steps of gbar*G*dt."""
    n, dt, G = 100, 1e-5, 4e-3
    t = np.arange(n + 1) * dt
    tr = psd.TR(duration=n * dt)
    tr.add_rf(psd.RF(np.array([0.0, n * dt]), np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2)))
    tr.add_gz([0.0, n * dt], [0, 0])
    tr.add_gy([0.0, n * dt], [0, 0])
    g = np.full(n + 1, G)
    tr.add_gx(t, g)
    adc_k = [10, 30, 60, 70]
    at = np.concatenate([t[[k - 1, k, k + 1]] for k in adc_k])
    av = np.tile([0.0, 1.0, 0.0], len(adc_k))
    tr.add_adc(at, av)
    tr.add_ext(np.concatenate([t[[19, 20, 21]], t[[59, 60, 61]]]), [0, 3, 0, 0, 4, 0])
    seq = psd.concat([tr.finalize()])
    shape = (1, 1, 1)
    sc = workloads.make_structs(np.ones(shape + (1,)), np.ones(shape + (1,)), np.full(shape + (1,), 0.1),
                                np.zeros(shape), (1e-3,) * 3, seq=seq)
    out = oracle.scan(sc, "f64", lib=lib)
    # the first point's sample is held at min(|g0|, |g1|) = 0 (finalize
    # zeroes the line's ends), so k grows by one step per point from point 1
    step = GYRO / (2 * PI) * G * dt
    want = np.array([9, 29, 19, 29]) * step
    return np.asarray(out.VSig["Kx"]), want


def test_lock_release_restores_the_locked_k():
    got, want = lock_release()
    np.testing.assert_allclose(got, want, rtol=1e-12)


# --------------------------------------------------------------------------
# planted mistakes
# --------------------------------------------------------------------------

def _fails(fn):
    try:
        fn()
    except AssertionError:
        return True
    return False


def _mtx_pin(lib, name):
    got, want = run_mtx(MTX_CASES[name], lib=lib)
    assert_f32_close(got, want)


def _mtx_gm_pin(lib):
    case = MTX_CASES["superposition"]
    b1, ph, fr = equivalent_single(case)
    a = tx_scan(case["coils"], case["mg"], case["pe"])
    b = tx_scan([(b1, ph, fr)], [1.0], [0.0], multitransmit=False)
    _as_two_pool(a)
    _as_two_pool(b)
    assert_f32_close(M(oracle.gm_scan(a, "f64", 0, lib=lib)), M(oracle.gm_scan(b, "f64", 0)))


def _sat_pin(lib):
    got, want = gm_two_coil_saturation(1, lib=lib)
    assert abs(got - want) / 0.5 < 2e-3


def _split_pin(lib):
    (ga, gb), (wa, wb) = split_one_step(lib=lib)
    assert_f32_close(ga, wa)
    assert_f32_close(gb, wb)


def _mt_pin(lib):
    (gw, gb), (ww, wb) = mt_one_step(lib=lib)
    assert_f32_close([gw, gb], [ww, wb])


def _lock_pin(lib):
    got, want = lock_release(lib=lib)
    np.testing.assert_allclose(got, want, rtol=1e-12)


MUTANT_PINS = {
    1: [lambda lib: _mtx_pin(lib, "superposition"), _mtx_gm_pin],
    2: [lambda lib: _mtx_pin(lib, "frequencies")],
    3: [lambda lib: _mtx_pin(lib, "silent_coil")],
    4: [_sat_pin],
    5: [_split_pin],
    6: [_mt_pin],
    7: [_lock_pin],
}


@pytest.mark.parametrize("mutation", sorted(MUTANT_PINS))
def test_planted_mistakes_fail_a_pin(mutation, tmp_path):
    lib = oracle.build_mutant(mutation, str(tmp_path))
    for pin in MUTANT_PINS[mutation]:
        assert _fails(lambda: pin(lib)), f"mutation {mutation} passes a pin"
    # and the unmutated oracle passes the same pins
    for pin in MUTANT_PINS[mutation]:
        pin(oracle._lib("f64"))
