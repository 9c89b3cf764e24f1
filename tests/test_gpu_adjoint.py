"""GPU adjoint (mrs_adjoint_*) against the fp64 oracle gradient.

The forward loss must match the fp32 oracle within 1e-5 (the kernels carry
Mz - m0 instead of Mz and use a 1.5-ulp sincos, so the rounding is not the
reference's) and the fp64 one within 1e-4;
every gradient (rf, T1, T2, dB0) within 1e-3 rel-L2 of the fp64 reverse
mode, which tests/test_oracle_adjoint.py pins to central differences.
The backward pass either rebuilds each checkpoint segment backwards from its
end (inverse relaxation and rotation; the default where it is safe) or
recomputes it forwards into scratch. In the scratch mode checkpoint
intervals only change what is recomputed, never the result (1e-6); the
rebuild adds a few ulps of rounding per segment (1e-5 between the modes).
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import engine, workloads
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def make(shape=(8, 8), spins=8, n_steps=300, seed=3):
    return workloads.excitation(shape=shape, spins=spins, n_steps=n_steps, dt=4e-6, res_exc=0.04, fov=0.128,
                                seed=seed)


@pytest.mark.parametrize("shape,spins,n_steps,K", [((8, 8), 8, 300, 64), ((5, 7), 4, 130, 16), ((16, 16), 2, 64, 64)])
def test_adjoint_matches_oracle(gpu, shape, spins, n_steps, K):
    e = make(shape, spins, n_steps)
    print(f"case {e.name} K={K}", flush=True)
    params64 = [x.astype(np.float64) for x in (e.rf, e.T1, e.T2, e.dB0)]
    L64, g_rf, g1, g2, gb = oracle.exc_grad(e, *params64)
    L32, _ = oracle.exc_loss(e, *params64, precision="f32")
    with engine.Adjoint(e, checkpoint=K) as adj:
        L, mxy = adj.forward(e.rf, e.T1, e.T2, e.dB0, want_mxy=True)
        G = adj.backward()
        print(f"  loss gpu {L:.9g} f32 {L32:.9g} f64 {L64:.9g}; info {adj.info()}", flush=True)
    assert L == pytest.approx(L32, rel=1e-5)
    assert L == pytest.approx(L64, rel=1e-4)
    errs = [rel_l2(G[0], g_rf), rel_l2(G[1], g1), rel_l2(G[2], g2), rel_l2(G[3], gb)]
    print("  grad rel-L2 rf %.2e T1 %.2e T2 %.2e dB0 %.2e" % tuple(errs), flush=True)
    assert max(errs) < 1e-3


@pytest.mark.parametrize("mode,tol", [(1, 1e-6), (0, 1e-5)])
def test_checkpoint_interval_does_not_change_gradients(gpu, mode, tol):
    e = make((8, 8), 4, 200)
    outs = []
    for K in (8, 64, 200):
        with engine.Adjoint(e, checkpoint=K, segment_mode=mode) as adj:
            adj.forward(e.rf, e.T1, e.T2, e.dB0)
            outs.append(adj.backward())
            assert adj.info()["segments_rebuilt"] == (1 - mode)
    for o in outs[1:]:
        errs = [rel_l2(a, b) for a, b in zip(o, outs[0])]
        print(f"mode {mode}: rel-L2 to K=8", ["%.1e" % x for x in errs], flush=True)
        assert max(errs) < tol


def test_rebuild_matches_scratch_segments(gpu):
    e = make((16, 16), 8, 400, seed=5)
    G = []
    for mode in (0, 1):
        with engine.Adjoint(e, checkpoint=64, segment_mode=mode) as adj:
            adj.forward(e.rf, e.T1, e.T2, e.dB0)
            G.append(adj.backward())
            assert adj.info()["segments_rebuilt"] == (1 - mode)
    errs = [rel_l2(a, b) for a, b in zip(G[0], G[1])]
    print("rebuild vs scratch rel-L2 rf %.2e T1 %.2e T2 %.2e dB0 %.2e" % tuple(errs), flush=True)
    assert max(errs) < 1e-5


def test_short_t2_falls_back_to_scratch(gpu):
    """A voxel whose T2 decays by more than 8x over one checkpoint interval
    (K*dt/T2 = 64*4us/50us) must not be rebuilt backwards."""
    e = make((8, 8), 4, 200)
    e.T2[:3] = 5e-5
    params64 = [x.astype(np.float64) for x in (e.rf, e.T1, e.T2, e.dB0)]
    L64, g_rf, g1, g2, gb = oracle.exc_grad(e, *params64)
    with engine.Adjoint(e, checkpoint=64) as adj:
        L = adj.forward(e.rf, e.T1, e.T2, e.dB0)
        G = adj.backward()
        assert adj.info()["segments_rebuilt"] == 0
    assert L == pytest.approx(L64, rel=1e-4)
    errs = [rel_l2(G[0], g_rf), rel_l2(G[1], g1), rel_l2(G[2], g2), rel_l2(G[3], gb)]
    print("  short T2 grad rel-L2 rf %.2e T1 %.2e T2 %.2e dB0 %.2e" % tuple(errs), flush=True)
    assert max(errs) < 1e-3


def test_device_pointers_and_autograd(gpu):
    import torch
    from mrilab_amd.autograd import excitation_loss
    e = make((8, 8), 4, 150)
    with engine.Adjoint(e) as adj:
        L_host = adj.forward(e.rf, e.T1, e.T2, e.dB0)
        G_host = adj.backward()
        dev = torch.device("cuda")
        ts = [torch.tensor(a, device=dev, requires_grad=True) for a in (e.rf, e.T1, e.T2, e.dB0)]
        loss = excitation_loss(adj, *ts)
        loss.backward()
        assert float(loss) == pytest.approx(L_host, rel=1e-12)
        for t, g in zip(ts, G_host):
            np.testing.assert_allclose(t.grad.cpu().numpy(), g, rtol=1e-6, atol=1e-6 * np.abs(g).max())
