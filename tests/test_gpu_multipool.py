"""GPU parity of the multi-pool (GM) scan against the CPU oracle.

mrs_DoGMScanAtGPU (the drop-in for DoGMScanAtGPU.cu:76) is called through
the C ABI and compared with oracle/bloch_oracle.c's GM model on the same
seeded MT/CEST phantom and saturation sequence:
  * scheme 0 (operator split, fp32, BlochKernelGM.cuh's order) against the
    fp32 oracle split: 1e-4 rel-L2 on signal and magnetization;
  * scheme 1 (exact matrix exponential, fp64 inside a step) against the
    fp64 oracle's exact scheme: signal 1e-5 rel-L2, magnetization 1e-4.
    The engine steps with the fp32 dt and field values of its records (as
    the reference engines do) while the oracle uses the fp64 ones; the
    final state of these scans is near zero (spoiled, just after a 90 deg
    pulse), so that input rounding shows in its relative error (~1e-5).
Pool counts 2, 3 and 4, spins per voxel 1/2/4 and 1-4 receive coils.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import engine, workloads
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def sig(sc):
    return sc.VSig["Sx"] + 1j * sc.VSig["Sy"]


def mag(sc):
    return np.concatenate([sc.VObj["Mx"].ravel(), sc.VObj["My"].ravel(), sc.VObj["Mz"].ravel()])


def with_pools(sc, idx, t1_scale=None):
    """Keep (or repeat) pools ``idx`` of a multi-pool scan."""
    idx = list(idx)
    out = sc.copy()
    o = out.VObj
    for k in ("Mx", "My", "Mz"):
        o[k] = np.asfortranarray(o[k][..., idx])
    for k in ("Rho", "T1", "T2"):
        o[k] = np.asfortranarray(o[k][..., idx])
    if t1_scale is not None:
        o["T1"] = np.asfortranarray((o["T1"] * np.asarray(t1_scale, np.float32)).astype(np.float32))
    o["K"] = np.asfortranarray(o["K"][..., idx, :][..., idx])
    o["TypeFlag"] = np.asarray(o["TypeFlag"])[idx].astype(np.float64)
    o["TypeNum"] = np.int32(len(idx))
    out.VMag["dWRnd"] = np.asfortranarray(out.VMag["dWRnd"][..., idx])
    out.VCtl["CS"] = np.asarray(out.VCtl["CS"])[idx].astype(np.float64)
    n = int(out.VSig["SignalNum"]) * len(idx) * int(out.VCoi["RxCoilNum"])
    out.VSig["Sx"], out.VSig["Sy"] = np.zeros(n), np.zeros(n)
    return out


def run_gm(sc, scheme, **opt):
    out = sc.copy()
    engine.DoGMScanAtGPU(*out.structs, options=engine.make_options(multipool_scheme=scheme, **opt))
    return out


CASES = [
    ("3pool", dict(shape=(6, 5, 1), n_offsets=3, sat_time=0.08, spins=1, n_coils=1), None, None),
    ("3pool_s2_c3", dict(shape=(4, 4, 2), n_offsets=3, sat_time=0.08, spins=2, n_coils=3), None, None),
    ("2pool_s4_c4", dict(shape=(4, 3, 1), n_offsets=2, sat_time=0.08, spins=4, n_coils=4), [0, 2], None),
    ("4pool_c2", dict(shape=(3, 3, 1), n_offsets=2, sat_time=0.08, spins=1, n_coils=2), [0, 1, 1, 2],
     [1.0, 1.0, 2.0, 1.0]),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    name, kw, idx, t1s = request.param
    sc = workloads.mt_cest(**kw)
    if idx is not None:
        sc = with_pools(sc, idx, t1s)
    print(f"case {name}: {sc.name} pools={int(sc.VObj['TypeNum'])} iso={sc.n_iso}", flush=True)
    return sc


def test_split_scheme_matches_f32_oracle(gpu, case):
    ref = oracle.gm_scan(case, "f32", 0)
    g = run_gm(case, 0)
    e_sig, e_mag = rel_l2(sig(g), sig(ref)), rel_l2(mag(g), mag(ref))
    print(f"  split vs f32 oracle: signal {e_sig:.3e} magnetization {e_mag:.3e}", flush=True)
    assert e_sig < 1e-4 and e_mag < 1e-4
    for k in ("ADCi", "TRCount", "utsi"):
        assert int(g.VVar[k]) == int(ref.VVar[k]), k


def test_exact_scheme_matches_f64_oracle(gpu, case):
    ref = oracle.gm_scan(case, "f64", 1)
    g = run_gm(case, 1)
    e_sig, e_mag = rel_l2(sig(g), sig(ref)), rel_l2(mag(g), mag(ref))
    print(f"  exact vs f64 oracle: signal {e_sig:.3e} magnetization {e_mag:.3e}", flush=True)
    assert e_sig < 1e-5 and e_mag < 1e-4


def test_gm_deterministic_and_launch_split(gpu):
    sc = workloads.mt_cest(shape=(4, 4, 1), n_offsets=4, sat_time=0.04, spins=2, n_coils=2)
    a = run_gm(sc, 0)
    b = run_gm(sc, 0)
    np.testing.assert_array_equal(sig(a), sig(b))
    c = run_gm(sc, 0, max_adc_per_launch=1)
    assert rel_l2(sig(c), sig(a)) < 1e-6 and rel_l2(mag(c), mag(a)) < 1e-6


def test_gm_z_spectrum_shape(gpu):
    """Signal is stored per pool ([type][sample]); the bound pool (TypeFlag 1)
    contributes none, and the free-water dip sits at the water resonance."""
    sc = workloads.mt_cest(shape=(2, 2, 1), n_offsets=5, sat_time=0.2, ppm_range=10.0)
    g = run_gm(sc, 1)
    s = np.abs(sig(g)).reshape(3, 5)
    assert np.all(s[2] == 0)
    assert np.argmin(s[0]) == 2 and s[0, 2] < 0.5 * min(s[0, 1], s[0, 3])


def variant_case(name):
    base = dict(n_offsets=3, sat_time=0.08)
    if name == "MT":
        sc = workloads.select_pools(workloads.mt_cest(shape=(6, 5, 1), **base), [0, 2])
    elif name == "MT_s2_c3":
        sc = workloads.select_pools(workloads.mt_cest(shape=(4, 4, 1), spins=2, n_coils=3, **base), [0, 2])
    elif name == "CEST":
        sc = workloads.select_pools(workloads.mt_cest(shape=(5, 4, 1), **base), [0, 2, 1], cs_hz=[0, 0, 447.0])
    elif name == "CEST4_c2":
        sc = workloads.select_pools(workloads.mt_cest(shape=(3, 3, 1), n_coils=2, **base), [0, 2, 1, 1],
                                    cs_hz=[0, 0, 447.0, -380.0], t1_scale=[1, 1, 1, 1.5])
    else:  # ME: every pool free, with a self-exchange rate on the diagonal
        sc = workloads.select_pools(workloads.mt_cest(shape=(4, 4, 1), spins=2, **base), [0, 1, 2])
    variant = name.split("_")[0].rstrip("4")
    return workloads.as_variant(sc, variant, self_rate=20.0 if variant == "ME" else 0.0), variant


ENTRY = {"MT": engine.DoMTScanAtGPU, "CEST": engine.DoCESTScanAtGPU, "ME": engine.DoMEScanAtGPU}


@pytest.mark.parametrize("name", ["MT", "MT_s2_c3", "CEST", "CEST4_c2", "ME"])
@pytest.mark.parametrize("scheme", [0, 1])
def test_variant_engines_match_oracle(gpu, name, scheme):
    sc, variant = variant_case(name)
    print(f"case {name} ({variant}) scheme {scheme}: {sc.name} pools={int(sc.VObj['TypeNum'])}", flush=True)
    ref = oracle.gm_scan(sc, "f32" if scheme == 0 else "f64", scheme, variant=variant)
    g = sc.copy()
    ENTRY[variant](*g.structs, options=engine.make_options(multipool_scheme=scheme))
    e_sig, e_mag = rel_l2(sig(g), sig(ref)), rel_l2(mag(g), mag(ref))
    print(f"  vs oracle: signal {e_sig:.3e} magnetization {e_mag:.3e}", flush=True)
    if scheme == 0:
        assert e_sig < 1e-4 and e_mag < 1e-4
    else:
        assert e_sig < 1e-5 and e_mag < 1e-4
