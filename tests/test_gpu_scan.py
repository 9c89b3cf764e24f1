"""GPU parity of the single-pool scan against the CPU oracle.

Every test calls the HIP engine through the C ABI (mrilab_amd.engine) and
compares with oracle/bloch_oracle.c on the same seeded phantom and
sequence:
  * "f32" oracle = DoScanAtCPU's fp32 operation order; bar 1e-4 rel-L2
    for signal and magnetization (precision_mode 0);
  * "f64" oracle = same discrete scheme in fp64; precision_mode 1 must be
    within 1e-4 of it.
The decoder outputs (VVar counters, k-space) are compared exactly.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import engine, workloads
from tests.conftest import adc_capture_slots, rel_l2

pytestmark = pytest.mark.gpu


def run_gpu(sc, precision=0, entry="DoScanAtGPU", **opt):
    out = sc.copy()
    options = engine.make_options(precision_mode=precision, **opt)
    getattr(engine, entry)(*out.structs, options=options)
    return out


def sig(sc):
    return sc.VSig["Sx"] + 1j * sc.VSig["Sy"]


def mag(sc):
    return np.concatenate([sc.VObj["Mx"].ravel(), sc.VObj["My"].ravel(), sc.VObj["Mz"].ravel()])


def assert_decoder_equal(a, b):
    for k in ("utsi", "rfi", "Gzi", "Gyi", "Gxi", "ADCi", "Exti", "TRCount"):
        assert int(a.VVar[k]) == int(b.VVar[k]), k
    for k in ("t", "dt", "Kx", "Ky", "Kz", "rfRef", "GzAmp", "GyAmp", "GxAmp"):
        assert float(a.VVar[k]) == float(b.VVar[k]), k
    for k in ("rfAmp", "rfPhase", "rfFreq"):
        np.testing.assert_array_equal(a.VVar[k], b.VVar[k])
    for k in ("Kx", "Ky", "Kz"):
        np.testing.assert_array_equal(a.VSig[k], b.VSig[k])


@pytest.fixture(scope="module")
def gre():
    sc = workloads.gre2d()
    print(f"case {sc.name}: iso={sc.n_iso} steps={sc.n_steps()}", flush=True)
    return sc, oracle.scan(sc, "f32"), oracle.scan(sc, "f64")


def test_gre2d_reference_mode(gpu, gre):
    sc, o32, o64 = gre
    g = run_gpu(sc, precision=0)
    e_sig, e_mag = rel_l2(sig(g), sig(o32)), rel_l2(mag(g), mag(o32))
    print(f"gre2d ref-mode vs f32 oracle: signal {e_sig:.3e} magnetization {e_mag:.3e}; "
          f"vs f64: {rel_l2(sig(g), sig(o64)):.3e}", flush=True)
    assert e_sig < 1e-4 and e_mag < 1e-4
    assert_decoder_equal(g, o32)


def test_gre2d_accurate_mode(gpu, gre):
    sc, o32, o64 = gre
    g = run_gpu(sc, precision=1)
    e_sig, e_mag = rel_l2(sig(g), sig(o64)), rel_l2(mag(g), mag(o64))
    print(f"gre2d accurate-mode vs f64 oracle: signal {e_sig:.3e} magnetization {e_mag:.3e}", flush=True)
    assert e_sig < 1e-4 and e_mag < 1e-4


def test_cpu_entry_point_same_result(gpu, gre):
    sc, o32, _ = gre
    a = run_gpu(sc, entry="DoScanAtCPU")
    b = run_gpu(sc, entry="DoScanAtGPU")
    np.testing.assert_array_equal(sig(a), sig(b))


def test_deterministic_and_plan_rerun(gpu):
    sc = workloads.gre2d(n_pe=8)
    a = run_gpu(sc)
    b = run_gpu(sc)
    np.testing.assert_array_equal(sig(a), sig(b))
    np.testing.assert_array_equal(mag(a), mag(b))
    out = sc.copy()
    with engine.Plan(*out.structs) as p:
        for _ in range(3):
            p.run()
        r = sc.copy()
        p.fetch(VObj=r.VObj, VVar=r.VVar, VSig=r.VSig)
    np.testing.assert_array_equal(sig(r), sig(a))
    np.testing.assert_array_equal(mag(r), mag(a))


@pytest.mark.parametrize("free_steps", [2, 1])
@pytest.mark.parametrize("shape,spins,coils,ntr", [((16, 16, 8), 4, 8, 6), ((12, 10, 4), 2, 3, 5),
                                                   ((8, 8, 2), 8, 1, 4), ((8, 8, 2), 1, 16, 3)])
def test_bssfp3d_multicoil(gpu, shape, spins, coils, ntr, free_steps):
    sc = workloads.bssfp3d(shape=shape, n_ro=16, spins=spins, n_coils=coils, n_tr=ntr, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    print(f"case {sc.name} free_steps {free_steps}", flush=True)
    o32 = oracle.scan(sc, "f32")
    g = run_gpu(sc, free_steps=free_steps)
    e_sig, e_mag = rel_l2(sig(g), sig(o32)), rel_l2(mag(g), mag(o32))
    print(f"  signal {e_sig:.3e} magnetization {e_mag:.3e}", flush=True)
    assert e_sig < 1e-4 and e_mag < 1e-4
    assert_decoder_equal(g, o32)


@pytest.mark.parametrize("pe,ntr", [((16, 8), 128), ((64, 32), 512)])
def test_long_bssfp_free_step_modes(gpu, pe, ntr):
    """128 and 512 TRs of the 3-D bSSFP sequence at its full field of view on
    a 16 x 16 x 8 phantom (4 isochromats, 8 coils), both free-step modes
    against the fp32 oracle. bSSFP turns every isochromat's phase per TR into
    signal, so fp32 rounding of the step angles builds up over the TRs: both
    modes compute each step's angle in the reference's order and rounding,
    and the deferred mode (free_steps 2) applies the fp64 sum of a free
    stretch's angles as one rotation; it has to stay as close to the oracle
    as stepping every rotation does. The default (0) defers here."""
    sc = workloads.bssfp3d(shape=(16, 16, 8), n_pe_y=pe[0], n_pe_z=pe[1], n_tr=ntr)
    print(f"case {sc.name}, {ntr} TRs", flush=True)
    o32 = oracle.scan(sc, "f32")
    d = run_gpu(sc, free_steps=2)
    s = run_gpu(sc, free_steps=1)
    a = sc.copy()
    with engine.Plan(*a.structs) as p:
        assert p.info()["deferred_free_steps"] == 1
        p.run()
        p.fetch(VObj=a.VObj, VVar=a.VVar, VSig=a.VSig)
    np.testing.assert_array_equal(sig(a), sig(d))
    e = [rel_l2(sig(d), sig(o32)), rel_l2(mag(d), mag(o32)), rel_l2(sig(s), sig(o32)), rel_l2(mag(s), mag(o32)),
         rel_l2(sig(d), sig(s))]
    print("  deferred: signal %.3e magnetization %.3e; stepwise: %.3e %.3e; between %.3e" % tuple(e), flush=True)
    assert max(e) < 1e-4
    assert e[0] < 1.5 * e[2] + 1e-6
    assert_decoder_equal(d, o32)


def test_free_steps_option_checked(gpu):
    sc = workloads.gre2d(n_pe=2)
    with pytest.raises(engine.MRSError, match="free_steps"):
        run_gpu(sc, free_steps=3)


@pytest.mark.parametrize("coils,spins", [(8, 4), (3, 2), (5, 1)])
def test_mfma_coil_sum_matches_butterfly(gpu, coils, spins):
    """The MFMA projection of staged samples onto the coil maps (4 or 8
    padded coils, the default) and the per-lane products with a wavefront
    butterfly (coil_sum = 1) give the same signal and magnetization to fp32
    rounding (the two kernel instances are compiled apart, so the spin
    updates may round differently), and each matches the oracle; the scan
    ends mid-batch of 8 samples."""
    sc = workloads.bssfp3d(shape=(10, 9, 3), n_ro=13, spins=spins, n_coils=coils, n_tr=3, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    print(f"case {sc.name}", flush=True)
    a = run_gpu(sc)
    b = run_gpu(sc, coil_sum=1)
    o32 = oracle.scan(sc, "f32")
    assert rel_l2(sig(a), sig(b)) < 1e-6
    assert rel_l2(mag(a), mag(b)) < 1e-6
    assert rel_l2(sig(a), sig(o32)) < 1e-4 and rel_l2(sig(b), sig(o32)) < 1e-4


@pytest.mark.parametrize("max_adc", [16, 7, 5])
def test_small_launches_match_one_launch(gpu, max_adc):
    """Cutting the scan into many launches (state through HBM, fresh rotation
    caches, partial buffers per launch) gives the same signal. Budgets below
    the 16-sample readout also cap the folded ADC runs mid-readout."""
    sc = workloads.bssfp3d(shape=(8, 8, 4), n_ro=16, spins=4, n_coils=4, n_tr=6, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    a = run_gpu(sc)
    b = run_gpu(sc, max_adc_per_launch=max_adc)
    assert rel_l2(sig(b), sig(a)) < 1e-6
    assert rel_l2(mag(b), mag(a)) < 1e-6


def test_full_phantom_first_tr(gpu):
    """The 3-D bSSFP workload at its full size (128 x 128 x 64 voxels, 4
    isochromats, 8 coils, 128 readout samples) for its first TR, against the
    fp32 oracle: parity at the benchmarked shape, 2.1e9 isochromat-steps."""
    sc = workloads.bssfp3d(n_tr=1)
    print(f"case {sc.name} full phantom, 1 TR", flush=True)
    g = run_gpu(sc)
    o32 = oracle.scan(sc, "f32")
    e_sig, e_mag = rel_l2(sig(g), sig(o32)), rel_l2(mag(g), mag(o32))
    print(f"  signal {e_sig:.3e} magnetization {e_mag:.3e}", flush=True)
    assert e_sig < 1e-4 and e_mag < 1e-4
    assert_decoder_equal(g, o32)
    # error against the fp64 oracle, reference and accurate modes
    o64 = oracle.scan(sc, "f64")
    a = run_gpu(sc, precision=1)
    e = [rel_l2(sig(g), sig(o64)), rel_l2(mag(g), mag(o64)), rel_l2(sig(a), sig(o64)), rel_l2(mag(a), mag(o64))]
    print("  vs fp64 oracle: reference mode signal %.3e magnetization %.3e; accurate mode %.3e %.3e" % tuple(e),
          flush=True)
    assert max(e) < 1e-4


def test_coil_sum_error_bound(gpu):
    """Error of the in-block fp32 coil sum (the MFMA projection, 8 coils, 4
    isochromats per thread, 256-thread blocks), measured against an fp64
    sum of the very fp32 states the kernel summed (RunMode-1 capture).

    Within a block the sum is plain fp32: 3 adds of a lane's isochromats, one
    product rounding, at most 64 accumulations in the MFMA chain over the
    wave's 64 voxels (one chain each for the two real products of Re and of
    Im), the add combining those two, and 3 adds of the block's 4 waves in
    LDS; across blocks the partials are added with Neumaier compensation in
    fp64. So |S - S_exact| <= 72 u sum_terms |m| |C| (u = 2^-24, first
    order), about 4.3e-6 of the sum of the terms' magnitudes."""
    sc = workloads.bssfp3d(shape=(16, 16, 8), n_ro=16, spins=4, n_coils=8, n_tr=4, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    rm = workloads.with_run_mode1(sc)
    g = rm.copy()
    engine.DoScanAtGPU(*g.structs)
    slots = adc_capture_slots(sc.VSeq)
    N = int(g.VSig["SignalNum"])
    assert slots.size == N
    C = (sc.VCoi["RxCoilx"].astype(np.float64) + 1j * sc.VCoi["RxCoily"].astype(np.float64))  # [R,C,Z,8]
    s = (g.VSig["Sx"] + 1j * g.VSig["Sy"]).reshape(8, N)
    mx, my = g.VSig["Mx"], g.VSig["My"]
    worst = 0.0
    for n, u in enumerate(slots):
        m = mx[..., u].astype(np.float64) + 1j * my[..., u].astype(np.float64)  # [R,C,Z,S,1]
        msum = m.sum(axis=(3, 4))
        mabs = np.abs(m).sum(axis=(3, 4))
        ref = np.pi * ((n // 16) % 2)
        for c in range(8):
            exact = np.sum(msum * np.conj(C[..., c])) * np.exp(-1j * ref)
            scale = np.sum(mabs * np.abs(C[..., c]))
            worst = max(worst, abs(s[c, n] - exact) / scale)
    bound = 72 * 2.0 ** -24
    print(f"coil sum: worst |S - S_exact| / sum|terms| = {worst:.2e} over {N} samples x 8 coils "
          f"(bound {bound:.2e})", flush=True)
    assert worst < bound
