"""Receive coils beyond one kernel pass (32 for the single-pool kernel, 8
for the multi-pool kernel, 4 with 5 or 6 pools): the engine runs the scan's
program once per group of coils, each pass from the same starting state, so
every coil's signal is the one a scan of that coil alone gives; checked
against the oracle, which sums every coil in one pass
(DoScanAtCPU.cpp:445-491 and BlochKernelGM.cuh:192 loop over RxCoilNum)."""
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


@pytest.mark.parametrize("coils", [33, 40, 64])
def test_single_pool_more_than_32_coils(gpu, coils):
    sc = workloads.bssfp3d(shape=(8, 8, 2), n_ro=8, spins=2, n_coils=coils, n_tr=3, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    print(f"case {sc.name}", flush=True)
    g = sc.copy()
    engine.DoScanAtGPU(*g.structs)
    ref = oracle.scan(sc, "f32")
    e_s, e_m = rel_l2(sig(g), sig(ref)), rel_l2(mag(g), mag(ref))
    N = int(sc.VSig["SignalNum"])
    per = [rel_l2(sig(g)[c * N:(c + 1) * N], sig(ref)[c * N:(c + 1) * N]) for c in range(coils)]
    print(f"  signal {e_s:.3e} (worst coil {max(per):.3e}) magnetization {e_m:.3e}", flush=True)
    assert e_s < 1e-4 and e_m < 1e-4 and max(per) < 1e-4
    # the first 32 coils equal a scan with only those coils
    head = sc.copy()
    for k in ("RxCoilx", "RxCoily"):
        head.VCoi[k] = np.asfortranarray(sc.VCoi[k][..., :32])
    head.VCoi["RxCoilNum"] = np.int32(32)
    head.VSig["Sx"] = np.zeros(N * 32)
    head.VSig["Sy"] = np.zeros(N * 32)
    engine.DoScanAtGPU(*head.structs)
    np.testing.assert_array_equal(sig(head), sig(g)[:32 * N])


@pytest.mark.parametrize("idx,coils,scheme", [([0, 1, 2], 12, 0), ([0, 1, 2], 12, 1), ([0, 1, 1, 2, 2], 8, 1)],
                         ids=["GM3_c12_split", "GM3_c12_exact", "GM5_c8_exact"])
def test_multipool_coil_groups(gpu, idx, coils, scheme):
    base = workloads.mt_cest(shape=(4, 3, 1), n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.0, sat_fa=200.0,
                             n_coils=coils)
    sc = workloads.select_pools(base, idx) if idx != [0, 1, 2] else base
    print(f"case TypeNum {int(sc.VObj['TypeNum'])}, Rx {coils}, scheme {scheme}", flush=True)
    g = sc.copy()
    engine.DoGMScanAtGPU(*g.structs, options=engine.make_options(multipool_scheme=scheme))
    ref = oracle.gm_scan(sc, "f32" if scheme == 0 else "f64", scheme)
    e_s = rel_l2(sig(g), sig(ref))
    d = np.linalg.norm(mag(g) - mag(ref)) / max(np.linalg.norm(mag(ref)), np.linalg.norm(mag(sc)))
    print(f"  signal {e_s:.3e} magnetization {d:.3e}", flush=True)
    assert e_s < 1e-4 and d < 1e-4


def test_coil_groups_with_a_motion_callback(gpu):
    """A scan of 36 coils (two passes) with an Ext callback that moves the
    object mid-scan: each pass restarts a chunk from the same fetched and
    re-uploaded state, so every coil matches the oracle's one-pass scan
    with the same callback."""
    from tests.test_gpu_paths import Recorder
    sc = workloads.bssfp3d(shape=(8, 8, 2), n_ro=8, spins=2, n_coils=36, n_tr=4, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    ext = sc.VSeq["ExtLine"]
    zero = [i for i in range(ext.size) if ext[i] == 0]
    ext[zero[len(zero) // 2]] = 8  # Plugin_ExecuteMotion in the middle of the scan
    ref = sc.copy()
    cb, keep = engine.make_callbacks(Recorder(ref))
    oracle.scan(ref, "f32", callbacks=cb, inplace=True)
    g = sc.copy()
    rec = Recorder(g)
    engine.DoScanAtGPU(*g.structs, ext=rec)
    assert [c[0] for c in rec.calls] == [8.0] and rec.calls[0][2]
    e_s, e_m = rel_l2(sig(g), sig(ref)), rel_l2(mag(g), mag(ref))
    print(f"  36 coils with motion: signal {e_s:.3e} magnetization {e_m:.3e}", flush=True)
    assert e_s < 1e-4 and e_m < 1e-4
    plain = sc.copy()
    engine.DoScanAtGPU(*plain.structs)
    assert rel_l2(sig(plain), sig(g)) > 1e-3  # the motion took effect
