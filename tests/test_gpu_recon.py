"""GPU Cartesian, EPI and FSE recon and k-space noise (mrs_DoCartRecon,
mrs_DoEchoTrainRecon, mrs_plan_recon, mrs_DoAddNoise) against the oracle
(oracle/recon.py, pinned by tests/test_oracle_recon.py; its echo-train
reorder by tests/test_oracle_seqgen_techniques.py).

Recon runs in fp64 on hipFFT; the bar is 1e-12 relative L2 on every VImg
array (the FFTs' rounding differs from numpy's by ulps). Noise: the Philox
words are integer work and the Box-Muller normals are compared to 1e-12
absolute (device log/sin/cos against numpy's).
"""
import numpy as np
import pytest

from mrilab_amd import engine
from oracle import recon as orc
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # rf, pe, sl, te, types, rx, zf, plane, freq_dir
    (65, 64, 1, 1, 1, 1, None, "Axial", "R/L"),
    (33, 24, 5, 2, 2, 3, None, "Axial", "A/P"),
    (17, 15, 4, 1, 3, 2, None, "Coronal", "S/I"),
    (33, 24, 2, 1, 1, 2, (64, 48, 2), "Sagittal", "A/P"),
]


def _signal(rf, pe, sl, te, types, rx, seed):
    rng = np.random.default_rng(seed)
    n = rf * pe * sl * te * types * rx
    return rng.normal(size=n), rng.normal(size=n)


def _check(out, want, tag):
    for k in ("Sx", "Sy", "Real", "Imag", "Mag"):
        e = rel_l2(out[k], want[k])
        assert e < 1e-12, f"{tag} {k}: {e:.2e}"
    # phase where the magnitude is not negligible
    m = want["Mag"] > 1e-9 * want["Mag"].max()
    d = np.angle(np.exp(1j * (out["Phase"][m] - want["Phase"][m])))
    assert np.abs(d).max() < 1e-9, tag


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}_te{c[3]}_ty{c[4]}_rx{c[5]}"
                         f"{'_zf' if c[6] else ''}_{c[7]}")
def test_recon_matches_oracle(gpu, case):
    rf, pe, sl, te, types, rx, zf, plane, fd = case
    print(f"case {case}", flush=True)
    sx, sy = _signal(rf, pe, sl, te, types, rx, seed=rf + pe)
    d = engine.recon_desc(rf, pe, sl, te, rx, types, zf=zf, scan_plane=plane, freq_dir=fd)
    out = engine.DoCartRecon(sx, sy, d)
    want = orc.cart_recon(sx, sy, rf, pe, sl, te, rx, types, zf=zf, swap=orc.swap_xy(plane, fd))
    assert out["Real"].shape == want["Real"].shape
    _check(out, want, str(case))


def test_recon_device_pointers(gpu):
    import torch
    rf, pe, sl, te, types, rx = 33, 24, 3, 2, 2, 2
    sx, sy = _signal(rf, pe, sl, te, types, rx, seed=7)
    d = engine.recon_desc(rf, pe, sl, te, rx, types)
    host = engine.DoCartRecon(sx, sy, d)
    dev = engine.DoCartRecon(torch.tensor(sx, device="cuda"), torch.tensor(sy, device="cuda"), d, device=True)
    for k in host:
        # torch views carry the column-major dims reversed
        got = dev[k].cpu().numpy().transpose(range(4, -1, -1))
        assert np.array_equal(got, host[k]), k


def test_plan_recon_from_device_resident_kspace(gpu):
    from mrilab_amd import workloads
    sc = workloads.gre2d(n_pe=32, n_ro=33, tr_steps=400, te_steps=200, rf_steps=50)
    out = sc.copy()
    with engine.Plan(*out.structs) as p:
        p.run()
        d = engine.recon_desc(33, 32, 1, 1, 1, 1)
        img = p.recon(d)
        p.fetch(VSig=out.VSig)
    want = orc.cart_recon(out.VSig["Sx"], out.VSig["Sy"], 33, 32, 1, 1, 1, 1)
    print("plan recon |img| max", float(want["Mag"].max()), flush=True)
    assert float(want["Mag"].max()) > 0
    _check(img, want, "plan")


def test_recon_rejects_mismatched_signal(gpu):
    sx, sy = _signal(9, 8, 1, 1, 1, 1, seed=1)
    with pytest.raises(engine.MRSError):
        engine.DoCartRecon(sx[:-9], sy[:-9], engine.recon_desc(9, 8, 1))


@pytest.mark.parametrize("n,seed", [(1, 0), (1000, 42), (300_001, 2 ** 40 + 17)])
def test_noise_matches_philox_oracle(gpu, n, seed):
    rng = np.random.default_rng(n)
    sx0, sy0 = rng.normal(size=n), rng.normal(size=n)
    sx, sy = sx0.copy(), sy0.copy()
    scale = 0.37
    engine.DoAddNoise(sx, sy, scale, seed)
    zx, zy = orc.noise_normals(n, seed)
    np.testing.assert_allclose((sx - sx0) / scale, zx, rtol=0, atol=1e-11)
    np.testing.assert_allclose((sy - sy0) / scale, zy, rtol=0, atol=1e-11)


def test_noise_device_pointers_and_determinism(gpu):
    import torch
    n = 4097
    a = torch.zeros(n, dtype=torch.float64, device="cuda")
    b = torch.zeros(n, dtype=torch.float64, device="cuda")
    engine.DoAddNoise(a, b, 1.0, 99)
    hx, hy = np.zeros(n), np.zeros(n)
    engine.DoAddNoise(hx, hy, 1.0, 99)
    assert np.array_equal(a.cpu().numpy(), hx)
    assert np.array_equal(b.cpu().numpy(), hy)


@pytest.mark.parametrize("epi,etl,shots,sl,rx,types", [(True, 8, 2, 2, 1, 1), (True, 6, 3, 1, 2, 2),
                                                       (False, 4, 2, 2, 3, 1), (False, 6, 1, 1, 1, 2)])
@pytest.mark.parametrize("device", [False, True])
def test_echo_train_recon_matches_oracle(gpu, epi, etl, shots, sl, rx, types, device):
    """mrs_DoEchoTrainRecon (DoEPIRecon / DoFSERecon) against the oracle's
    reorder followed by its Cartesian recon."""
    rf = 17
    print(f"case echo train epi={epi} etl={etl} shots={shots} sl={sl} rx={rx} types={types} "
          f"device={device}", flush=True)
    sx, sy = _signal(rf, etl * shots, sl, 1, types, rx, seed=etl * 7 + shots)
    d = engine.recon_desc(rf, etl * shots, sl, 1, rx, types, scan_plane="Axial", freq_dir="R/L")
    if device:
        import torch
        out = engine.DoCartRecon(torch.from_numpy(sx).cuda(), torch.from_numpy(sy).cuda(), d, device=True,
                                 echo_train=(etl, shots, epi))
        out = {k: v.cpu().numpy() if hasattr(v, "cpu") else v for k, v in out.items()}
    else:
        out = engine.DoCartRecon(sx, sy, d, echo_train=(etl, shots, epi))
    ox, oy = orc.echo_train_order(sx, sy, rf, etl, shots, sl, rx, types, epi)
    want = orc.cart_recon(ox, oy, rf, etl * shots, sl, 1, rx, types, swap=orc.swap_xy("Axial", "R/L"))
    if device:  # tensors come back with the MATLAB dims reversed
        out = {k: np.transpose(v) for k, v in out.items()}
    assert out["Real"].shape == want["Real"].shape
    _check(out, want, f"echo train {epi} {etl} {shots}")
