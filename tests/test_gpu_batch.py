"""GPU parity of the batched engine (mrs_batch_* / mrs_DoScanBatchAtGPU).

Each scan of a batch must give what the single-scan entry point gives for
it alone (same step code; only the fp32 order of the coil sums differs,
bar 1e-5), and what the fp32 oracle gives (bar 1e-4), with identical
decoder outputs. Batches mix uneven sequence lengths, spin counts, voxel
counts (several tiles per scan) and receive-coil counts; the graph replay
is checked against eager runs with host work in between.
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


def mixed_batch():
    scans = workloads.fingerprint_batch(n_sets=6, spins=64, tr_min=40, tr_max=90, seed=11)
    scans += workloads.fingerprint_batch(n_sets=3, spins=16, tr_min=20, tr_max=60, seed=12)
    scans.append(workloads.bssfp3d(shape=(8, 8, 2), n_ro=16, spins=4, n_coils=3, n_tr=4, tr_steps=120,
                                   fov=(0.064, 0.064, 0.032)))
    scans.append(workloads.gre2d(n_pe=4, n_ro=16, tr_steps=200, te_steps=100, rf_steps=20, spins=4))
    return scans


def run_batch(scans, **opt):
    outs = [s.copy() for s in scans]
    engine.DoScanBatchAtGPU([o.structs for o in outs], options=engine.make_options(**opt))
    return outs


@pytest.mark.parametrize("free_steps", [2, 1])
def test_batch_matches_single_scans_and_oracle(gpu, free_steps):
    """Both free-step modes (2: deferred rotation, 1: every step rotates);
    each batched scan against the single-scan entry point in the same mode."""
    scans = mixed_batch()
    for s in scans:
        print(f"case {s.name}: iso={s.n_iso} steps={s.n_steps()} free_steps {free_steps}", flush=True)
    outs = run_batch(scans, free_steps=free_steps)
    with engine.Batch([s.copy().structs for s in scans], options=engine.make_options(free_steps=free_steps)) as bb:
        assert bb.info()["deferred_free_steps"] == (1 if free_steps == 2 else 0)
    for s, b in zip(scans, outs):
        single = s.copy()
        engine.DoScanAtGPU(*single.structs, options=engine.make_options(free_steps=free_steps))
        ref = oracle.scan(s, "f32")
        e1, e2 = rel_l2(sig(b), sig(single)), rel_l2(mag(b), mag(single))
        e3, e4 = rel_l2(sig(b), sig(ref)), rel_l2(mag(b), mag(ref))
        print(f"  {s.name}: vs single {e1:.2e}/{e2:.2e}  vs oracle {e3:.2e}/{e4:.2e}", flush=True)
        assert e1 < 1e-5 and e2 < 1e-5
        assert e3 < 1e-4 and e4 < 1e-4
        for k in ("utsi", "rfi", "ADCi", "TRCount"):
            assert int(b.VVar[k]) == int(ref.VVar[k]), k
        for k in ("Kx", "Ky", "Kz"):
            np.testing.assert_array_equal(b.VSig[k], ref.VSig[k])


@pytest.mark.parametrize("spt,block", [(1, 64), (2, 128), (4, 256)])
def test_batch_fixed_shapes_agree(gpu, spt, block):
    scans = workloads.fingerprint_batch(n_sets=12, spins=64, tr_min=30, tr_max=70, seed=5)
    a = run_batch(scans, spins_per_thread=1, block_size=64)
    b = run_batch(scans, spins_per_thread=spt, block_size=block)
    for x, y in zip(a, b):
        assert rel_l2(sig(y), sig(x)) < 1e-6
        assert rel_l2(mag(y), mag(x)) < 1e-6


def test_batch_graph_replays_match_eager(gpu):
    """Replays of the captured graph, with fetches and host work between
    them, are identical to each other (state is re-read from M0 every run)
    and agree with an eager single-scan run of every scan."""
    scans = workloads.fingerprint_batch(n_sets=20, spins=64, tr_min=30, tr_max=120, seed=9)
    eager = []
    for s in scans:
        o = s.copy()
        engine.DoScanAtGPU(*o.structs)
        eager.append(o)
    first = None
    with engine.Batch([s.copy().structs for s in scans], options=engine.make_options(timing=1)) as b:
        info = b.info()
        print(f"batch info: {info}", flush=True)
        assert info["n_scans"] == 20 and info["tuned"] == 1 and info["n_candidates"] >= 2
        # fingerprint TRs hold a few free steps per ADC sample and rf pulse:
        # by default they rotate every step
        assert info["deferred_free_steps"] == 0
        for rep in range(3):
            b.run()
            b.sync()
            outs = [s.copy() for s in scans]
            b.fetch([o.structs for o in outs])
            _ = np.linalg.norm(np.random.default_rng(rep).normal(size=100000))  # host work between replays
            if first is None:
                first = outs
            for o, f, e in zip(outs, first, eager):
                np.testing.assert_array_equal(sig(o), sig(f))
                np.testing.assert_array_equal(mag(o), mag(f))
                assert rel_l2(sig(o), sig(e)) < 1e-5 and rel_l2(mag(o), mag(e)) < 1e-5
        assert b.info()["last_run_ms"] > 0


@pytest.mark.parametrize("prec,bar", [(0, 2e-3), (1, 1e-4)])
def test_fingerprints_against_fp64_oracle(gpu, prec, bar):
    """The fingerprinting workload's error against the fp64 oracle (the same
    discrete scheme in fp64), full-length trains of 500-1500 TRs: reference
    mode carries the reference's fp32 relaxation rounding, accurate mode
    does not (a single-voxel scan reads the host-computed factors of each
    record; the result equals the one computed on the device)."""
    scans = workloads.fingerprint_batch(n_sets=6, spins=64, tr_min=500, tr_max=1500, seed=21)
    outs = run_batch(scans, precision_mode=prec)
    worst = 0.0
    for s, b in zip(scans, outs):
        ref = oracle.scan(s, "f64")
        e_s, e_m = rel_l2(sig(b), sig(ref)), rel_l2(mag(b), mag(ref))
        single = s.copy()
        engine.DoScanAtGPU(*single.structs, options=engine.make_options(precision_mode=prec))
        e_1 = rel_l2(sig(b), sig(single))
        print(f"  {s.name} precision_mode {prec}: vs fp64 oracle signal {e_s:.2e} magnetization {e_m:.2e}; "
              f"vs single scan {e_1:.1e}", flush=True)
        assert e_1 < 1e-5
        worst = max(worst, e_s, e_m)
    assert worst < bar
