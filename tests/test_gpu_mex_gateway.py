"""The engine's MEX gateways (mrilab_amd/mex/) run end to end under the mock
MEX runtime (tests/mex) on the GPU: mexFunction with no arguments reads the
globals, runs the engine through the C ABI and leaves its outputs in the
globals (VObj.M*, VSig.*, VVar), as DoScanAtGPU.cu:76 / DoScanAtCPU.cpp:70
do. The results equal the engine's own entry points called from Python on
the same structs. With Ext callbacks the gateway evaluates DoGPUFetch and
DoExtPlugin; a plugin that reallocates the globals it edits (as MATLAB does
on assignment) is followed by re-fetched pointers (DoScanAtCPU.cpp:366-405)."""
import numpy as np
import pytest

from mrilab_amd import engine, workloads
from tests import mexmock
from tests.conftest import rel_l2
from tests.test_gpu_paths import callback_scan, Recorder

pytestmark = pytest.mark.gpu


def outputs(m, sc):
    o = {}
    for k in ("Mx", "My", "Mz"):
        o["VObj." + k] = m.get("VObj", k)
    for k in ("Sx", "Sy", "Kx", "Ky", "Kz"):
        o["VSig." + k] = m.get("VSig", k)
    for k in ("utsi", "rfi", "Gzi", "Gyi", "Gxi", "ADCi", "Exti", "TRCount", "Kx", "Ky", "Kz", "rfRef", "t", "dt"):
        o["VVar." + k] = m.get("VVar", k)
    return o


def py_outputs(sc):
    o = {}
    for k in ("Mx", "My", "Mz"):
        o["VObj." + k] = sc.VObj[k].ravel(order="F")
    for k in ("Sx", "Sy", "Kx", "Ky", "Kz"):
        o["VSig." + k] = np.asarray(sc.VSig[k]).ravel()
    for k in ("utsi", "rfi", "Gzi", "Gyi", "Gxi", "ADCi", "Exti", "TRCount", "Kx", "Ky", "Kz", "rfRef", "t", "dt"):
        o["VVar." + k] = np.atleast_1d(sc.VVar[k])
    return o


def compare(a, b, tol=1e-7):
    for k in a:
        x, y = np.asarray(a[k], dtype=np.float64), np.asarray(b[k], dtype=np.float64)
        assert x.shape == y.shape, k
        if k.startswith("VVar") or k in ("VSig.Kx", "VSig.Ky", "VSig.Kz"):
            np.testing.assert_array_equal(x, y, err_msg=k)
        else:
            assert rel_l2(x, y) < tol, (k, rel_l2(x, y))


@pytest.mark.parametrize("entry", ["DoScanAtGPU", "DoScanAtCPU"])
def test_scan_through_mex_function(gpu, entry):
    sc = workloads.bssfp3d(shape=(8, 8, 4), n_ro=16, spins=2, n_coils=4, n_tr=6, tr_steps=120,
                           fov=(0.064, 0.064, 0.032))
    m = mexmock.MockMex(entry)
    m.load(sc)
    cmds = []
    m.on_eval(lambda c: cmds.append(c) or 0)
    assert m.run() is None
    assert m.lib.mock_unshare_count() == 7      # every global unshared before the in-place writes
    ref = sc.copy()
    getattr(engine, entry)(*ref.structs)
    compare(outputs(m, sc), py_outputs(ref))
    assert cmds == []                            # Ext 7 runs in the engine


def test_gm_scan_through_mex_function(gpu):
    sc = workloads.mt_cest(shape=(4, 4, 1), n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.0, sat_fa=200.0)
    m = mexmock.MockMex("DoGMScanAtGPU")
    m.load(sc)
    assert m.run() is None
    ref = sc.copy()
    engine.DoGMScanAtGPU(*ref.structs)
    compare(outputs(m, sc), py_outputs(ref))


def test_plugins_that_reallocate_the_globals(gpu):
    """Ext 8 moves the object and Ext -3 edits VVar.Kx and VObj.Mz, both by
    assigning new arrays to the globals; DoGPUFetch asks for the state at
    both. The gateway's run equals the engine's with a Python callback that
    makes the same edits in place."""
    sc = callback_scan()
    shp = sc.VObj["Mz"].shape
    mshp = sc.VMag["dWRnd"].shape
    m = mexmock.MockMex("DoScanAtGPU")
    m.load(sc)
    seen = []

    def matlab(cmd):
        ext = float(m.get("VVar", "Ext")[0])
        seen.append((cmd, ext))
        if cmd == "DoGPUFetch":
            m.set("VVar", "gpuFetch", np.array([[1.0 if ext in (8.0, -3.0) else 0.0]]))
            return 0
        if ext == 8:   # Plugin_ExecuteMotion.m: VObj.M* = circshift(...) allocates new arrays
            for var, k, s in (("VObj", "Mx", shp), ("VObj", "My", shp), ("VObj", "Mz", shp),
                              ("VObj", "Rho", shp[:3] + (1,)), ("VObj", "T1", shp[:3] + (1,)),
                              ("VObj", "T2", shp[:3] + (1,)), ("VMag", "dWRnd", mshp)):
                m.set(var, k, np.roll(m.get(var, k, s), 1, axis=0))
        if ext == -3:
            m.set("VVar", "Kx", np.array([[m.get("VVar", "Kx")[0] + 25.0]]))
            m.set("VObj", "Mz", (m.get("VObj", "Mz", shp) * np.float32(0.5)).astype(np.float32))
        return 0

    m.on_eval(matlab)
    a0 = m.address("VObj", "Mz")
    assert m.run() is None
    assert m.address("VObj", "Mz") != a0
    codes = [e for c, e in seen if c == "DoExtPlugin"]
    assert codes == [8.0, 5.0, -3.0, 8.0, 9.0]
    assert [e for c, e in seen if c == "DoGPUFetch"] == codes
    ref = sc.copy()
    engine.DoScanAtGPU(*ref.structs, ext=Recorder(ref), gpu_fetch=lambda code: code in (8, -3))
    got = outputs(m, sc)
    # the motion moved the maps: compare those too
    for var, k in (("VObj", "Rho"), ("VObj", "T1"), ("VMag", "dWRnd")):
        np.testing.assert_array_equal(m.get(var, k), ref.structs[("VObj", "VMag").index(var)][k].ravel(order="F"))
    compare(got, py_outputs(ref), tol=1e-6)


def test_plugin_error_ends_the_call(gpu):
    sc = callback_scan()
    m = mexmock.MockMex("DoScanAtCPU")
    m.load(sc)
    m.on_eval(lambda cmd: 1 if float(m.get("VVar", "Ext")[0]) == -3 else 0)
    assert m.run() == ("MRiLab:callback", "Extended process encounters ERROR!")
