"""The engine's MEX gateways (mrilab_amd/mex/) under the mock MEX runtime
(tests/mex), on the CPU: every gateway builds and exports mexFunction, and
the checks it makes before any device work behave as the boundary must
(DoScanAtCPU.cpp:70-210, SURVEY's boundary row): a field of the wrong class
is refused with the field's name where the reference would reinterpret it,
a missing global or field is named, and nothing runs on the GPU."""
import os

import numpy as np
import pytest

from mrilab_amd import workloads
from tests import mexmock

ENTRIES = ["DoScanAtGPU", "DoScanAtCPU", "DoGMScanAtGPU", "DoMTScanAtGPU", "DoCESTScanAtGPU", "DoMEScanAtGPU"]
pytestmark = pytest.mark.skipif(not all(os.path.exists(mexmock.path(e)) for e in ENTRIES),
                                reason="gateways not built (make -C tests/mex)")


@pytest.mark.parametrize("entry", ENTRIES)
def test_gateway_exports_mex_function(entry):
    m = mexmock.MockMex(entry)
    assert hasattr(m.lib, "mexFunction")


@pytest.mark.parametrize("var,field,value,msg", [
    ("VObj", "T1", "double", "VObj.T1 must be single, not double"),
    ("VObj", "SpinNum", "double", "VObj.SpinNum must be int32, not double"),
    ("VSeq", "tsLine", "single", "VSeq.tsLine must be double, not single"),
    ("VVar", "utsi", "double", "VVar.utsi must be int32, not double"),
    ("VSig", "Sx", "single", "VSig.Sx must be double, not single"),
    ("VCtl", "TRNum", "double", "VCtl.TRNum must be int32, not double"),
])
def test_wrong_class_is_refused_with_its_name(var, field, value, msg):
    sc = workloads.gre2d(n_pe=2, n_ro=8)
    m = mexmock.MockMex("DoScanAtGPU")
    m.load(sc)
    a = sc.structs[("VObj", "VMag", "VCoi", "VCtl", "VSeq", "VVar", "VSig").index(var)][field]
    m.set(var, field, np.asarray(a).astype(np.float64 if value == "double" else np.float32))
    assert m.run() == ("MRiLab:dtype", msg)


def test_missing_global_and_field_are_named():
    sc = workloads.gre2d(n_pe=2, n_ro=8)
    m = mexmock.MockMex("DoScanAtCPU")
    m.load(sc)
    m.remove("VMag", "dWRnd")
    assert m.run() == ("MRiLab:field", "VMag.dWRnd is missing")
    m2 = mexmock.MockMex("DoScanAtGPU")
    m2.lib.mock_reset()
    for var, d in zip(("VObj", "VMag", "VCoi", "VCtl", "VSeq", "VVar"), sc.structs[:6]):
        for k, v in d.items():
            if isinstance(v, (np.ndarray, np.generic, str)):
                m2.set(var, k, v)
    assert m2.run() == ("MRiLab:global", "global VSig is not defined")


def test_short_array_is_refused():
    sc = workloads.gre2d(n_pe=2, n_ro=8)
    m = mexmock.MockMex("DoGMScanAtGPU")
    m.load(sc)
    m.set("VObj", "Rho", np.ones(3, dtype=np.float32))
    err = m.run()
    assert err[0] == "MRiLab:size" and err[1].startswith("VObj.Rho has 3 elements")
