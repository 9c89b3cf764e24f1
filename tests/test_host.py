"""Host-side checks that need no GPU: sequence-line format, the dtype
contract of the struct packer, and the C-ABI library's exported symbols."""
import ctypes
import os
import re

import numpy as np
import pytest

from mrilab_amd import abi, engine, psd, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported_by_library():
    """Every function declared in include/mrilab_amd.h is exported by
    libmrilab_amd.so (loading needs no GPU; nothing is called)."""
    hdr = open(os.path.join(ROOT, "include", "mrilab_amd.h")).read()
    declared = set(re.findall(r"\b(mrs_[A-Za-z_]+)\s*\(", hdr)) - {"mrs_status"}
    assert declared == set(engine.EXPORTS)
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(engine.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_tr_finalize_matches_psd_epilogue():
    """Per TR: ts holds [0, T] plus every event sorted stably, uts = unique(ts),
    flags columns follow ts, each line's first and last samples are zeroed
    (PSD_GRE3D.m:317-359)."""
    T = 1e-3
    tr = psd.TR(duration=T)
    tr.add_rf(psd.rf_rect(np.array([0, 1, 2, 3]) * 1e-4, 10.0))
    tr.add_gz([0, 5e-4, T], [1.0, 2.0, 3.0])
    tr.add_gy([0, T], [0, 0])
    tr.add_gx([2e-4, 5e-4], [4.0, 5.0])
    tr.add_adc([5e-4, 6e-4], [1, 1])
    tr.add_ext([1e-4, 2e-4, 3e-4], [0, 6, 0])
    d = tr.finalize()
    assert np.all(np.diff(d["ts"]) >= 0)
    np.testing.assert_array_equal(d["uts"], np.unique(d["ts"]))
    assert d["flags"].shape == (6, d["ts"].size)
    # columns of the two SE points are all-zero and come first at t = 0
    assert np.all(d["flags"][:, 0] == 0)
    assert d["Gz"][0] == 0 and d["Gz"][-1] == 0 and d["Gz"][1] == 2.0
    assert d["ADC"][0] == 0 and d["ADC"][-1] == 0
    # stable order at t = 0: SE, rf, Gz, Gy (rf before gradients)
    at0 = d["flags"][:, d["ts"] == 0]
    assert [int(np.argmax(c)) if c.any() else -1 for c in at0.T] == [-1, 0, 1, 2]


def test_concat_forces_rf_coil_one_for_single_transmit():
    sc = workloads.gre2d(n_pe=2)
    assert np.all(sc.VSeq["rfCoilLine"] == 1)
    assert sc.VSeq["flagsLine"].shape == (6, sc.VSeq["tsLine"].size)
    assert int(sc.VSig["SignalNum"]) == int(np.count_nonzero(sc.VSeq["ADCLine"] == 1))


def test_pack_scan_rejects_wrong_dtypes():
    """Where the MEX would reinterpret a wrongly typed array, the packer raises."""
    sc = workloads.gre2d(n_pe=2)
    abi.pack_scan(*sc.structs)  # the correct contract packs
    bad = sc.copy()
    bad.VObj["T1"] = bad.VObj["T1"].astype(np.float64)
    with pytest.raises(TypeError, match="VObj.T1"):
        abi.pack_scan(*bad.structs)
    bad = sc.copy()
    bad.VVar["utsi"] = np.int64(0)
    with pytest.raises(TypeError, match="VVar.utsi"):
        abi.pack_scan(*bad.structs)
    bad = sc.copy()
    bad.VSeq["tsLine"] = bad.VSeq["tsLine"].astype(np.float32)
    with pytest.raises(TypeError, match="VSeq.tsLine"):
        abi.pack_scan(*bad.structs)
    bad = sc.copy()
    bad.VSig["Sx"] = np.zeros(3)
    with pytest.raises(ValueError, match="VSig.Sx"):
        abi.pack_scan(*bad.structs)


def test_struct_layout_matches_header_sizes():
    """ctypes mirrors have the field count the header declares."""
    hdr = open(os.path.join(ROOT, "include", "mrilab_amd.h")).read()
    body = hdr[hdr.index("typedef struct mrs_vvar"):hdr.index("} mrs_vvar;")]
    names = re.findall(r"\b([A-Za-z]+)\b\s*[,;]", body.replace("*", " "))
    py = [f for f, _ in abi.VVarC._fields_]
    for n in py:
        assert n in body, n


def test_bssfp_workload_shape_and_counts():
    sc = workloads.bssfp3d(shape=(8, 8, 4), n_ro=8, spins=2, n_coils=3, n_tr=5, tr_steps=100,
                           fov=(0.032, 0.032, 0.016))
    assert sc.VObj["Mz"].shape == (8, 8, 4, 2, 1)
    assert sc.VCoi["RxCoilx"].shape == (8, 8, 4, 3)
    assert int(sc.VSig["SignalNum"]) == 5 * 8
    assert sc.n_steps() == 5 * 100
    # uniform +-pi spread of intravoxel frequencies over one TR
    w = sc.VMag["dWRnd"][0, 0, 0, :, 0] * 100 * 10e-6
    np.testing.assert_allclose(w, [-np.pi / 2, np.pi / 2], atol=1e-5)


def test_fingerprint_batch_shape():
    """Uneven trains: lengths in range, 5 steps and one ADC sample per TR,
    flips in [5, 70] deg (area of the held pulse), TR in [10, 15] ms."""
    scans = workloads.fingerprint_batch(n_sets=5, spins=64, tr_min=20, tr_max=40, seed=1)
    lengths = set()
    for sc in scans:
        n = int(sc.VCtl["TRNum"])
        lengths.add(n)
        assert 20 <= n <= 40
        assert sc.n_steps() == 5 * n
        assert int(sc.VSig["SignalNum"]) == n
        assert int(sc.VObj["SpinNum"]) == 64 and sc.n_iso == 64
        amp = sc.VSeq["rfAmpLine"].reshape(n, -1)[:, 1]
        fa = np.degrees(amp * psd.GYRO * 0.5e-3)
        assert fa.min() >= 5 - 1e-9 and fa.max() <= 70 + 1e-9
        uts = sc.VSeq["utsLine"].reshape(n, -1)
        assert np.all((uts[:, -1] >= 10e-3) & (uts[:, -1] <= 15e-3))
    assert len(lengths) > 1


def test_batch_entry_fails_loudly_without_gpu():
    """The batched entry point marshals its struct arrays and, with no
    device visible, raises instead of falling back to the CPU."""
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("library not built")
    if engine.device_count() > 0:
        pytest.skip("a GPU is visible")
    scans = workloads.fingerprint_batch(n_sets=2, tr_min=5, tr_max=8)
    with pytest.raises(engine.MRSError):
        engine.DoScanBatchAtGPU([s.structs for s in scans])


@pytest.mark.parametrize("edit,match", [
    (lambda sc: sc.VSeq["rfCoilLine"].__setitem__(3, 2.0), "rfCoilLine"),
    (lambda sc: sc.VSeq["rfCoilLine"].__setitem__(3, 0.0), "rfCoilLine"),
    (lambda sc: sc.VSeq["flagsLine"].__setitem__((0, 5), 0.5), "non-negative integers"),
    (lambda sc: sc.VSeq["flagsLine"].__setitem__((4, 7), -1.0), "non-negative integers"),
    (lambda sc: sc.VVar.__setitem__("rfi", np.int32(3)), "VVar.rfi"),
    (lambda sc: sc.VVar.__setitem__("ADCi", np.int32(-1)), "VVar.ADCi"),
    (lambda sc: sc.VVar.__setitem__("utsi", np.int32(10 ** 6)), "VVar.utsi"),
])
def test_scan_rejects_lines_it_would_read_past(edit, match):
    """Sequence lines that would make the decoder index past an array (a Tx
    coil outside 1..TxCoilNum, fractional or negative flags, start indices
    past the line) are MRS_ERR_ARG before any device work; the reference
    reads out of bounds there (DoScanAtCPU.cpp:291-409)."""
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("library not built")
    sc = workloads.gre2d(n_pe=2, n_ro=8)
    edit(sc)
    with pytest.raises(engine.MRSError, match=match) as e:
        engine.DoScanAtGPU(*sc.structs)
    assert e.value.status == 1


def test_scan_rejects_more_rf_flags_than_samples():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("library not built")
    sc = workloads.gre2d(n_pe=2, n_ro=8)
    n = sc.VSeq["rfAmpLine"].size
    for k in ("rfAmpLine", "rfPhaseLine", "rfFreqLine", "rfCoilLine"):
        sc.VSeq[k] = np.asfortranarray(sc.VSeq[k][: n - 1])
    with pytest.raises(engine.MRSError, match="VVar.rfi plus the rf flags"):
        engine.DoScanAtGPU(*sc.structs)
