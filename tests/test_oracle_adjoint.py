"""The excitation oracle's reverse-mode gradients pinned by central finite
differences of its own fp64 forward model (oracle/adjoint_oracle.c; MRiLab
has no gradient code, so nothing else fixes them)."""
import numpy as np
import pytest

import oracle
from mrilab_amd import workloads


def _make(dt):
    e = workloads.excitation(shape=(4, 4), spins=4, n_steps=150, dt=dt, res_exc=0.05, fov=0.1, seed=7)
    params = [x.astype(np.float64) for x in (e.rf, e.T1, e.T2, e.dB0)]
    return e, params, oracle.exc_grad(e, *params)


@pytest.fixture(scope="module")
def problem():
    """0.6 ms excitation: rf and dB0 dominate."""
    return _make(4e-6)


@pytest.fixture(scope="module")
def slow_problem():
    """30 ms excitation: relaxation matters, so the T1/T2 derivatives are
    far above the rounding noise of the differences."""
    return _make(2e-4)


def central_fd(e, params, which, idx, h):
    out = []
    for sgn in (1, -1):
        args = [p.copy() for p in params]
        args[which].reshape(-1)[idx] += sgn * h
        out.append(oracle.exc_loss(e, *args)[0])
    return (out[0] - out[1]) / (2 * h)


def test_loss_matches_forward(problem):
    e, params, (L, *_) = problem
    assert L == pytest.approx(oracle.exc_loss(e, *params)[0], rel=1e-14)


@pytest.mark.parametrize("idx", [0, 1, 57, 150, 299])
def test_rf_gradient_vs_central_differences(problem, idx):
    e, params, (L, g_rf, *_) = problem
    fd = central_fd(e, params, 0, idx, 1e-3 * 1e-5)   # h = 1e-3 of 10 uT
    assert g_rf.reshape(-1)[idx] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(g_rf).max())


@pytest.mark.parametrize("v", [0, 6, 15])
def test_db0_gradient_vs_central_differences(problem, v):
    e, params, (L, g_rf, g1, g2, gb) = problem
    fd = central_fd(e, params, 3, v, 1e-3 * 1e-6)
    assert gb[v] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(gb).max())


@pytest.mark.parametrize("v", [0, 6, 15])
def test_relaxation_gradients_vs_central_differences(slow_problem, v):
    e, params, (L, g_rf, g1, g2, gb) = slow_problem
    for which, g in ((1, g1), (2, g2), (3, gb)):
        h = 1e-3 * (params[which][v] if which < 3 else 1e-6)
        fd = central_fd(e, params, which, v, h)
        assert g[v] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(g).max()), which


@pytest.mark.parametrize("idx", [3, 100, 250])
def test_rf_gradient_slow_problem(slow_problem, idx):
    e, params, (L, g_rf, *_) = slow_problem
    fd = central_fd(e, params, 0, idx, 1e-3 * 1e-5)
    assert g_rf.reshape(-1)[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(g_rf).max())


def test_f32_forward_tracks_f64(problem):
    e, params, _ = problem
    L32, m32 = oracle.exc_loss(e, *params, precision="f32")
    L64, m64 = oracle.exc_loss(e, *params, precision="f64")
    assert np.linalg.norm(m32 - m64) / np.linalg.norm(m64) < 1e-5
