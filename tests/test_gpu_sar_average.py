"""GPU N-gram SAR averaging (mrs_DoSARAverageAtGPU) against the oracle
(oracle/sar_average_oracle.c, pinned by tests/test_oracle_sar_average.py).

Both sum every sphere in the reference's (slice, column, row) order in
fp64 with FP contraction off, so aveSAR and avePower must agree bit for
bit: random volumes with zero-mass voxels and non-positive centres, the
one-sphere case, 2-D and single-time inputs, a body phantom whose SAR comes
from the SAR engine, and device-resident inputs.
"""
import numpy as np
import pytest

import oracle
from mrilab_amd import engine, workloads

pytestmark = pytest.mark.gpu


def _both(SAR, Power, M, d, grams, tag):
    print(f"case {tag} shape {SAR.shape} grams {grams}", flush=True)
    a = engine.DoSARAverageAtGPU(SAR, Power, M, *d, grams)
    b = oracle.sar_average(SAR, Power, M, *d, grams)
    for x, y, k in zip(a, b, ("aveSAR", "avePower")):
        assert x.shape == y.shape, f"{tag} {k} shape"
        diff = np.abs(x - y)
        assert np.array_equal(x, y), f"{tag} {k} max diff {np.nanmax(diff):.3e} at {np.argmax(diff)}"
    return a


def _random(D, seed, zero_frac=0.15):
    rng = np.random.default_rng(seed)
    M = rng.uniform(800, 1200, D[:3])
    M[rng.random(D[:3]) < zero_frac] = 0.0
    SAR = rng.uniform(-0.2, 1.0, D)
    Power = rng.uniform(0.0, 2.0, D)
    return SAR, Power, M


@pytest.mark.parametrize("grams", [0.005, 0.05, 0.3, 1.0])
def test_random_volume(gpu, grams):
    SAR, Power, M = _random((7, 6, 5, 2), 7)
    a = _both(SAR, Power, M, (2e-3, 2e-3, 2.5e-3), grams, "random-7x6x5")
    assert np.count_nonzero(a[0]) > 0


@pytest.mark.parametrize("D", [(70, 33, 9, 3), (40, 37, 1, 2), (21, 19, 17, 1)])
def test_shapes(gpu, D):
    SAR, Power, M = _random(D, sum(D))
    _both(SAR, Power, M, (2e-3, 2e-3, 2e-3), 0.5, f"random-{D}")


def test_three_dim_input(gpu):
    SAR, Power, M = _random((20, 18, 6, 1), 11)
    a = _both(SAR[..., 0], Power[..., 0], M, (2e-3, 2e-3, 2e-3), 0.4, "3-D")
    assert a[0].shape == (20, 18, 6)


def test_body_phantom_from_sar_engine(gpu):
    base = workloads.bssfp3d(shape=(32, 28, 16), n_ro=28, n_coils=1, spins=1, n_tr=16)
    sc = workloads.sar_scan(base, n_tx=2, n_samples=5, seed=3, body=True)
    engine.DoCalSARAtGPU(*sc.structs)
    ts = sc.VSig["tSample"]
    SAR, Power = engine.sar_window(sc.VObj, sc.VSig, ts[:-1], ts[1:])
    assert np.all(np.isfinite(SAR)) and SAR.max() > 0
    M = sc.VObj["MassDen"].astype(np.float64)
    d = (float(sc.VObj["XDimRes"]), float(sc.VObj["YDimRes"]), float(sc.VObj["ZDimRes"]))
    for grams in (1.0, 10.0):
        a = _both(SAR, Power, M, d, grams, "body")
        on = (M > 0) & (SAR[..., 0] > 0)
        assert np.all(a[0][on] > 0) and np.all(a[0][~on] == 0)


def test_device_tensors_match_host(gpu):
    import torch

    SAR, Power, M = _random((33, 20, 7, 2), 5)
    d = (2e-3, 2e-3, 3e-3)
    host = engine.DoSARAverageAtGPU(SAR, Power, M, *d, 0.4)

    def dev(a):  # column-major: a permuted view of a contiguous [.., slices, cols, rows] tensor
        t = torch.from_numpy(np.ascontiguousarray(a.transpose(*reversed(range(a.ndim))))).cuda()
        return t.permute(*reversed(range(a.ndim)))

    a, b = engine.DoSARAverageAtGPU(dev(SAR), dev(Power), dev(M), *d, 0.4)
    torch.cuda.synchronize()
    assert np.array_equal(a.cpu().numpy(), host[0]) and np.array_equal(b.cpu().numpy(), host[1])


def test_too_little_mass_is_an_error(gpu):
    M = np.full((3, 3, 3), 1000.0)
    SAR = np.ones((3, 3, 3, 1))
    with pytest.raises(engine.MRSError):
        engine.DoSARAverageAtGPU(SAR, SAR, M, 1e-3, 1e-3, 1e-3, 1.0)
    # the call after an error works
    _both(SAR, SAR, M, (1e-3, 1e-3, 1e-3), 0.01, "after-error")


def test_mass_knife_edge(gpu):
    # vol*MassDen = 0.125*8 = 1 exactly per voxel: the 27-voxel volume weighs
    # exactly 27 kg, and grams = 27000 gives grams*1e-3 == 27.0, where the
    # reference's search never ends; a hair less and it ends at the covering
    # sphere
    M = np.full((3, 3, 3), 8.0)
    SAR = np.linspace(0.5, 1.5, 27).reshape(3, 3, 3, 1)
    assert 27000 * 1e-3 == 27.0
    with pytest.raises(engine.MRSError):
        engine.DoSARAverageAtGPU(SAR, SAR, M, 0.5, 0.5, 0.5, 27000.0)
    _both(SAR, SAR, M, (0.5, 0.5, 0.5), 26999.999999, "knife-edge")
    # grams*1e-3 one ulp below 27: within the any-order sum's rounding bound,
    # so the entry point sums in memory order to decide
    g = 26999.999999999996
    assert g * 1e-3 == np.nextafter(27.0, 0.0)
    _both(SAR, SAR, M, (0.5, 0.5, 0.5), g, "one-ulp")
