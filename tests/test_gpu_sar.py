"""GPU local SAR (mrs_DoCalSARAtGPU) against the oracle (oracle/sar_oracle.c,
pinned by tests/test_oracle_sar.py).

Both evaluate the reference's fp32 statements in the same order with FP
contraction off, so VVar.SAR and VSig.SAR must agree bit for bit; the
decoder state (Muts, tRealSample, SARi, utsi, rfi, TRCount) exactly.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import engine, workloads

pytestmark = pytest.mark.gpu


def _run_both(sc):
    a = sc.copy()
    engine.DoCalSARAtGPU(*a.structs)
    b = oracle.sar(sc)
    return a, b


def _compare(a, b, tag):
    for k in ("SAR",):
        x, y = a.VVar[k], b.VVar[k]
        assert np.array_equal(x, y), f"{tag} VVar.SAR max diff {np.abs(x - y).max():.3e}"
    x, y = a.VSig["SAR"], b.VSig["SAR"]
    assert np.array_equal(x, y), f"{tag} VSig.SAR max diff {np.abs(x - y).max():.3e}"
    for k in ("Muts", "tRealSample"):
        assert np.array_equal(a.VSig[k], b.VSig[k]), f"{tag} {k}"
    assert a.VSig["SAR_rf_steps"] == b.VSig["SAR_rf_steps"] > 0, tag
    for k in ("SARi", "utsi", "rfi", "TRCount"):
        assert int(a.VVar[k]) == int(b.VVar[k]), f"{tag} {k}"
    for k in ("rfAmp", "rfPhase", "rfFreq"):
        assert np.array_equal(a.VVar[k], b.VVar[k]), f"{tag} {k}"
    assert float(np.abs(b.VSig["SAR"]).max()) > 0, tag


@pytest.mark.parametrize("n_tx", [1, 2, 5])
def test_sar_matches_oracle_gre(gpu, n_tx):
    sc = workloads.sar_scan(workloads.gre2d(n_pe=16, n_ro=16), n_tx=n_tx, n_samples=6, seed=n_tx)
    print(f"case {sc.name}", flush=True)
    a, b = _run_both(sc)
    _compare(a, b, sc.name)


def test_sar_matches_oracle_3d(gpu):
    base = workloads.bssfp3d(shape=(12, 10, 4), n_ro=10, n_coils=1, spins=1, n_tr=24)
    sc = workloads.sar_scan(base, n_tx=3, n_samples=5, seed=9)
    print(f"case {sc.name}", flush=True)
    a, b = _run_both(sc)
    _compare(a, b, sc.name)


def test_sar_stops_at_the_last_sample(gpu):
    sc = workloads.sar_scan(workloads.gre2d(n_pe=8, n_ro=8), n_tx=2, n_samples=3, seed=2)
    total = float(sc.VSig["tSample"][-1])
    sc.VSig["tSample"] = np.array([0.0, 0.3 * total, 0.6 * total])  # the loop returns after the third
    a, b = _run_both(sc)
    _compare(a, b, "early stop")
    assert int(a.VVar["SARi"]) == 3
    assert int(a.VVar["utsi"]) < sc.VSeq["utsLine"].size - 1


def test_sar_accumulates_onto_existing_values(gpu):
    # a second pass from a reset decoder state adds to VVar.SAR and VSig.SAR (+=)
    sc = workloads.sar_scan(workloads.gre2d(n_pe=8, n_ro=8), n_tx=2, n_samples=4, seed=3)
    a, b = _run_both(sc)
    for s in (a, b):
        for k in ("utsi", "rfi", "SARi"):
            s.VVar[k] = np.int32(0)
        s.VSig["Muts"][:] = 0
    engine.DoCalSARAtGPU(*a.structs)
    b = oracle.sar(b)
    _compare(a, b, "second pass")
