"""GPU parity of the scan paths beside the plain single-Tx, single-type scan,
each against the CPU oracle on the same seeded inputs, through the C ABI:

  * multi-transmit rf (BlochKernel.h:73-121): 2 and 3 Tx coils fired
    together with their own amplitudes, phases, frequencies and B1+ maps, a
    coil switched off mid-scan, Ext 7 taking rfRef from the master coil
    (Plugin_rfRef.m:7-12); the multi-pool engine with 2 Tx coils (the
    free-pool field of BlochKernel.h's form and the mean saturation rate of
    BlochKernelGM.cuh:348-380), both schemes;
  * two species with chemical shift (TypeNum 2, BlochKernel.h:42, signal per
    type, DoScanAtCPU.cpp:485-489);
  * RunMode 1 (DoScanAtCPU.cpp:426-440,544-552): VSig.Mx/My/Mz after every
    step and VSig.Muts;
  * Ext 3/4 (Plugin_LockK.m, Plugin_ReleaseK.m);
  * Ext callbacks (DoScanAtCPU.cpp:356-409, DoScanAtGPU.cu:479-662): codes
    5, 8 and a user code, with a callback that moves the object (motion: the
    maps and the magnetization shifted in place), edits VVar.Kx and
    VObj.Mz, and records VSig at each call.
Bars as in test_gpu_scan.py: 1e-4 rel-L2 on signal and magnetization
against the fp32 oracle (multi-pool exact scheme: the fp64 oracle), decoder
outputs exactly.
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


def kspace(sc):
    return np.concatenate([sc.VSig[k] for k in ("Kx", "Ky", "Kz")])


def run_gpu(sc, entry="DoScanAtGPU", ext=None, gpu_fetch=None, **opt):
    out = sc.copy()
    getattr(engine, entry)(*out.structs, ext=ext, gpu_fetch=gpu_fetch, options=engine.make_options(**opt))
    return out


def check(g, o, what, tol=1e-4):
    e_s, e_m = rel_l2(sig(g), sig(o)), rel_l2(mag(g), mag(o))
    print(f"  {what}: signal {e_s:.3e} magnetization {e_m:.3e}", flush=True)
    assert e_s < tol and e_m < tol
    np.testing.assert_array_equal(kspace(g), kspace(o))
    for k in ("utsi", "rfi", "ADCi", "Exti", "TRCount"):
        assert int(g.VVar[k]) == int(o.VVar[k]), k
    for k in ("rfRef", "Kx", "Ky", "Kz"):
        assert float(g.VVar[k]) == float(o.VVar[k]), k


MTX = [
    ("gre_tx2", lambda: workloads.multi_transmit(workloads.gre2d(n_pe=6, n_ro=16, spins=2), [0.6, 0.5],
                                                 [0.0, 0.9], [0.0, 25.0])),
    ("gre_tx3_silent", lambda: workloads.multi_transmit(workloads.gre2d(n_pe=6, n_ro=16, spins=1), [0.5, 0.4, 0.3],
                                                        [0.0, 1.1, -2.0], [0.0, -40.0, 60.0], silent_after=(2, 300))),
    ("bssfp_tx2_master2", lambda: workloads.multi_transmit(
        workloads.bssfp3d(shape=(8, 8, 4), n_ro=16, spins=4, n_coils=4, n_tr=6, tr_steps=120,
                          fov=(0.064, 0.064, 0.032)), [0.7, 0.5], [0.0, 0.6], master=2)),
]


@pytest.mark.parametrize("free_steps", [0, 2])
@pytest.mark.parametrize("name,make", MTX, ids=[m[0] for m in MTX])
def test_multi_transmit_scan(gpu, name, make, free_steps):
    """free_steps 2 forces the deferred kernel with any Tx coil count."""
    sc = make()
    print(f"case {name}: {sc.name}, Tx {int(sc.VCoi['TxCoilNum'])}, free_steps {free_steps}", flush=True)
    o32 = oracle.scan(sc, "f32")
    g = run_gpu(sc, free_steps=free_steps)
    check(g, o32, name)
    if name.startswith("bssfp"):
        # Ext 7 took the phase of coil 2 (rf phase alternates 0/pi, coil 2 adds 0.6)
        assert float(g.VVar["rfRef"]) != 0.0


def test_two_species_chemical_shift(gpu):
    sc = workloads.two_species(workloads.gre2d(n_pe=8, n_ro=16, spins=2), cs_hz=-440.0)
    print(f"case {sc.name}", flush=True)
    o32 = oracle.scan(sc, "f32")
    g = run_gpu(sc)
    check(g, o32, "TypeNum 2")
    # the species' signals differ (a swapped type index would mix them)
    n = int(sc.VSig["SignalNum"])
    s = sig(g)
    assert rel_l2(s[:n], s[n:]) > 0.1


@pytest.mark.parametrize("name,make", [
    ("gre", lambda: workloads.gre2d(n_pe=3, n_ro=8, spins=2, tr_steps=300, te_steps=150, rf_steps=60)),
    ("gre_tx2", lambda: workloads.multi_transmit(
        workloads.gre2d(n_pe=3, n_ro=8, spins=2, tr_steps=300, te_steps=150, rf_steps=60), [0.6, 0.5], [0.0, 0.9])),
    ("gre_ty2", lambda: workloads.two_species(
        workloads.gre2d(n_pe=3, n_ro=8, spins=1, tr_steps=300, te_steps=150, rf_steps=60))),
])
@pytest.mark.parametrize("free_steps", [0, 2])
def test_run_mode1_capture(gpu, name, make, free_steps):
    """RunMode 1: the state after every step lands in VSig.Mx/My/Mz[..., utsi]
    and the elapsed time in VSig.Muts, as the oracle writes them. With
    free_steps 2 the deferred kernel applies the pending rotation before
    each capture."""
    sc = workloads.with_run_mode1(make())
    print(f"case {name} RunMode 1: {sc.name}, n_uts {sc.VSeq['utsLine'].size}, free_steps {free_steps}",
          flush=True)
    o32 = oracle.scan(sc, "f32")
    g = run_gpu(sc, free_steps=free_steps)
    check(g, o32, name)
    cap = lambda s: np.concatenate([s.VSig[k].ravel() for k in ("Mx", "My", "Mz")])  # noqa: E731
    e = rel_l2(cap(g), cap(o32))
    print(f"  captured magnetization rel-L2 {e:.3e}", flush=True)
    assert e < 1e-4
    np.testing.assert_array_equal(g.VSig["Muts"], o32.VSig["Muts"])
    assert np.count_nonzero(g.VSig["Muts"]) > 100


def test_lock_and_release_k(gpu):
    """Ext 3 at the end of TR 1 keeps k; Ext 4 in place of TR 2's k reset
    restores it, so TR 2 reads out with the kept k offset."""
    sc = workloads.gre2d(n_pe=4, n_ro=16)
    ext = sc.VSeq["ExtLine"]
    per = ext.size // 4
    i6 = [i for i in range(per) if ext[i] == 6][0]
    i1 = [i for i in range(per) if ext[i] == 1][0]
    ext[i6] = 3
    ext[per + i1] = 4
    o32 = oracle.scan(sc, "f32")
    g = run_gpu(sc)
    check(g, o32, "lock/release")
    n = int(sc.VSig["SignalNum"]) // 4
    # TR 2's readout starts away from TR 3's (reset) start by the kept k
    assert abs(g.VSig["Kx"][n] - g.VSig["Kx"][2 * n]) > 1.0


class Recorder:
    """Ext callback: code 8 shifts the object by one row (magnetization,
    Rho, T1, T2, dWRnd in place, Plugin_ExecuteMotion.m's circshift); code
    -3 adds 25 1/m to VVar.Kx and halves VObj.Mz when the state is fetched;
    every call records the code, the decoder's TRCount and the signal
    acquired so far."""

    def __init__(self, sc):
        self.sc = sc
        self.calls = []

    def __call__(self, code, vvar, state):
        self.calls.append((float(code), int(vvar.TRCount), bool(state.fetched),
                           self.sc.VSig["Sx"] + 1j * self.sc.VSig["Sy"], self.sc.VSig["Kx"].copy()))
        o, m = self.sc.VObj, self.sc.VMag
        if code == 8 and state.fetched:
            for d, k in ((o, "Mx"), (o, "My"), (o, "Mz"), (o, "Rho"), (o, "T1"), (o, "T2"), (m, "dWRnd")):
                d[k][...] = np.roll(d[k], 1, axis=0)
        if code == -3:
            vvar.Kx += 25.0
            if state.fetched:
                o["Mz"][...] *= np.float32(0.5)


def callback_scan():
    sc = workloads.gre2d(n_pe=6, n_ro=16, spins=2)
    ext = sc.VSeq["ExtLine"]
    per = ext.size // 6
    i6 = [i for i in range(per) if ext[i] == 6][0]
    for tr, code in ((1, 8), (2, 5), (3, -3), (4, 8), (5, 9)):
        ext[tr * per + i6] = code      # in place of that TR's spoiler
    return sc


def test_ext_callbacks_match_cpu_engine(gpu):
    """GPU entry point: codes 5, 8, -3 and 9 call back; code 8 fetches by
    DoGPUFetch.m's table, -3 because gpu_fetch says so; the result, the
    decoder state and what each call saw equal the oracle's (the CPU engine,
    whose plugins always see the live state)."""
    sc = callback_scan()
    ref = sc.copy()
    rec_o = Recorder(ref)
    cb, keep = engine.make_callbacks(rec_o)
    oracle.scan(ref, "f32", callbacks=cb, inplace=True)
    g = sc.copy()
    rec_g = Recorder(g)
    engine.DoScanAtGPU(*g.structs, ext=rec_g, gpu_fetch=lambda code: code in (8, -3))
    check(g, ref, "callbacks")
    assert [c[:2] for c in rec_g.calls] == [c[:2] for c in rec_o.calls]
    assert [c[2] for c in rec_g.calls] == [True, False, True, True, False]
    for a, b in zip(rec_g.calls, rec_o.calls):
        # VSig holds the samples acquired so far, and zeros after them
        assert np.count_nonzero(a[3]) == np.count_nonzero(b[3]) > 0
        assert rel_l2(a[3], b[3]) < 1e-4
        np.testing.assert_array_equal(a[4], b[4])
    # the motion moved the object: without the callbacks the result differs
    plain = run_gpu(sc)
    assert rel_l2(sig(plain), sig(g)) > 1e-2


def test_cpu_entry_point_fetches_for_every_callback(gpu):
    """mrs_DoScanAtCPU hands the live state to every callback, as the CPU
    engine does, so the Mz edit at code -3 and the motion at code 8 take
    effect with no gpu_fetch; the GPU entry point without gpu_fetch fetches
    only code 8 (DoGPUFetch.m) and the -3 edit of a stale VObj.Mz is not
    uploaded."""
    sc = callback_scan()
    ref = sc.copy()
    cb, keep = engine.make_callbacks(Recorder(ref))
    oracle.scan(ref, "f32", callbacks=cb, inplace=True)
    c = sc.copy()
    rec = Recorder(c)
    engine.DoScanAtCPU(*c.structs, ext=rec)
    assert all(x[2] for x in rec.calls)
    check(c, ref, "CPU entry, callbacks")
    gg = sc.copy()
    rec2 = Recorder(gg)
    engine.DoScanAtGPU(*gg.structs, ext=rec2)
    assert [x[2] for x in rec2.calls] == [True, False, False, True, False]
    assert rel_l2(mag(gg), mag(ref)) > 1e-3


def gm_tx2(scheme_variant_shape=(4, 4, 1)):
    sc = workloads.mt_cest(shape=scheme_variant_shape, n_offsets=2, sat_time=0.04, pulse=0.02, gap=0.0, sat_fa=200.0,
                           spins=1, n_coils=2)
    return workloads.multi_transmit(sc, [0.7, 0.5], [0.0, 1.0], [0.0, 150.0])


@pytest.mark.parametrize("scheme", [0, 1])
def test_multipool_two_tx(gpu, scheme):
    sc = gm_tx2()
    print(f"case {sc.name} scheme {scheme}", flush=True)
    ref = oracle.gm_scan(sc, "f32" if scheme == 0 else "f64", scheme)
    g = sc.copy()
    engine.DoGMScanAtGPU(*g.structs, options=engine.make_options(multipool_scheme=scheme))
    e_s, e_m = rel_l2(sig(g), sig(ref)), rel_l2(mag(g), mag(ref))
    print(f"  signal {e_s:.3e} magnetization {e_m:.3e}", flush=True)
    assert e_s < 1e-4 and e_m < 1e-4


def test_adjoint_two_losses_on_one_handle(gpu):
    """Two losses of one engine.Adjoint in one graph: each backward
    differentiates its own forward (the second one runs its forward again)."""
    import torch
    from mrilab_amd.autograd import excitation_loss
    e = workloads.excitation(shape=(8, 8), spins=4, n_steps=150, dt=4e-6, res_exc=0.04, fov=0.128, seed=3)
    dev = torch.device("cuda")
    with engine.Adjoint(e) as adj:
        adj.forward(e.rf, e.T1, e.T2, e.dB0)
        g1 = adj.backward()
        rf2 = (e.rf * 0.7).astype(np.float32)
        adj.forward(rf2, e.T1, e.T2, e.dB0)
        g2 = adj.backward()
        a = [torch.tensor(x, device=dev, requires_grad=True) for x in (e.rf, e.T1, e.T2, e.dB0)]
        b = torch.tensor(rf2, device=dev, requires_grad=True)
        loss = excitation_loss(adj, *a) + excitation_loss(adj, b, *a[1:])
        loss.backward()
        np.testing.assert_allclose(b.grad.cpu().numpy(), g2[0], rtol=1e-6, atol=1e-6 * np.abs(g2[0]).max())
        np.testing.assert_allclose(a[0].grad.cpu().numpy(), g1[0], rtol=1e-6, atol=1e-6 * np.abs(g1[0]).max())
        for t, x, y in zip(a[1:], g1[1:], g2[1:]):
            s = x + y
            np.testing.assert_allclose(t.grad.cpu().numpy(), s, rtol=1e-5, atol=1e-5 * np.abs(s).max())
