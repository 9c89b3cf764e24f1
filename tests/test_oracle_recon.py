"""Pins the recon/noise oracle (oracle/recon.py) and the host-side recon logic.

* Round trip: k-space built from a known image by the inverse of
  fftshift(ifftn(fftshift(.))) (ifftshift, fftn, ifftshift), split over
  types, with the extra readout sample appended, reconstructs that image
  (pins the reshape order, the type sum, the dropped sample and the permute).
* A brute-force DFT with explicit index shifts on tiny odd and even sizes
  (pins fftshift's floor(n/2) and the 1/N of ifftn).
* Zero filling puts the acquired block at DoCartRecon.m:31-37's offsets.
* The ScanPlane/FreqDir table of DoCartRecon.m:45-73.
* Philox-4x32-10 against the published known-answer vectors (Random123
  kat_vectors), Box-Muller statistics, and the DoAddNoise.m:8-16 scale.
"""
import itertools

import numpy as np
import pytest

from oracle import recon


def make_signal(img, res_freq, te, types, rx, rng):
    """VSig.Sx/Sy ([N, Rx, Ty] column-major flattened) whose recon (no
    zero filling, no swap) is img [PE, RO-1, SL, Rx, TE]."""
    pe, rf1, sl = img.shape[:3]
    N = res_freq * te * pe * sl
    sx = np.zeros((N, rx, types))
    sy = np.zeros((N, rx, types))
    for e in range(te):
        for c in range(rx):
            S = np.fft.ifftshift(np.fft.fftn(np.fft.ifftshift(img[..., c, e])))  # [PE, RO-1, SL]
            K = S.transpose(1, 0, 2)                                              # [RO-1, PE, SL]
            full = np.zeros((res_freq, pe, sl), dtype=complex)
            full[:-1] = K
            full[-1] = rng.normal(size=(pe, sl)) + 1j * rng.normal(size=(pe, sl))  # dropped sample
            parts = rng.dirichlet(np.ones(types))
            for t in range(types):
                # n = (e*RF + ro) + RF*TE*(pe + PE*sl)
                blk = np.zeros((res_freq * te, pe, sl), dtype=complex)
                blk[e * res_freq:(e + 1) * res_freq] = full * parts[t]
                flat = blk.reshape(-1, order="F")
                sx[:, c, t] += flat.real
                sy[:, c, t] += flat.imag
    return sx.reshape(-1, order="F"), sy.reshape(-1, order="F")


@pytest.mark.parametrize("rf,pe,sl,te,types,rx", [(9, 8, 1, 1, 1, 1), (8, 6, 3, 2, 2, 3), (17, 5, 4, 1, 3, 2)])
def test_round_trip_recovers_the_image(rf, pe, sl, te, types, rx):
    rng = np.random.default_rng(rf * 100 + pe)
    img = rng.normal(size=(pe, rf - 1, sl, rx, te)) + 1j * rng.normal(size=(pe, rf - 1, sl, rx, te))
    sx, sy = make_signal(img, rf, te, types, rx, rng)
    out = recon.cart_recon(sx, sy, rf, pe, sl, te, rx, types)
    assert out["Real"].shape == img.shape
    np.testing.assert_allclose(out["Real"] + 1j * out["Imag"], img, atol=1e-10)
    np.testing.assert_allclose(out["Mag"], np.abs(img), atol=1e-10)
    np.testing.assert_allclose(np.exp(1j * out["Phase"]), np.exp(1j * np.angle(img)), atol=1e-9)


def brute_recon(S):
    """fftshift(ifftn(fftshift(S))) for a 3-D array by explicit sums."""
    n = S.shape
    sh = np.zeros_like(S)
    for idx in itertools.product(*[range(k) for k in n]):
        sh[tuple((i + k // 2) % k for i, k in zip(idx, n))] = S[idx]
    out = np.zeros_like(S)
    for x in itertools.product(*[range(k) for k in n]):
        acc = 0j
        for kk in itertools.product(*[range(k) for k in n]):
            ph = sum(a * b / m for a, b, m in zip(x, kk, n))
            acc += sh[kk] * np.exp(2j * np.pi * ph)
        out[x] = acc / np.prod(n)
    res = np.zeros_like(S)
    for idx in itertools.product(*[range(k) for k in n]):
        res[tuple((i + k // 2) % k for i, k in zip(idx, n))] = out[idx]
    return res


@pytest.mark.parametrize("pe,rf,sl", [(3, 5, 2), (4, 4, 1), (5, 3, 3)])
def test_brute_force_dft(pe, rf, sl):
    rng = np.random.default_rng(pe * 10 + rf)
    N = rf * pe * sl
    sx = rng.normal(size=N)
    sy = rng.normal(size=N)
    out = recon.cart_recon(sx, sy, rf, pe, sl, 1, 1, 1)
    K = (sx + 1j * sy).reshape((rf, pe, sl), order="F")[:-1].transpose(1, 0, 2)
    want = brute_recon(K)
    np.testing.assert_allclose(out["Real"][..., 0, 0] + 1j * out["Imag"][..., 0, 0], want, atol=1e-12)


def test_zero_filling_places_the_block():
    rf, pe, sl = 9, 6, 2
    rng = np.random.default_rng(4)
    N = rf * pe * sl
    sx, sy = rng.normal(size=N), rng.normal(size=N)
    zf = (16, 10, 2)
    out = recon.cart_recon(sx, sy, rf, pe, sl, 1, 1, 1, zf=zf)
    assert out["Sx"].shape == (10, 16, 4, 1, 1)  # permuted [ZF_Ky, ZF_Kx, 2*SliceNum]
    K = sx.reshape((rf, pe, sl), order="F")[:-1]
    # 1-based MATLAB: ZF/2 - size/2 + 1; slices f*floor(S/2) - floor(S/2) + 1
    ox, oy, oz = 16 // 2 - 8 // 2, 10 // 2 - 6 // 2, 2 * 1 - 1
    blk = out["Sx"][..., 0, 0].transpose(1, 0, 2)
    np.testing.assert_array_equal(blk[ox:ox + 8, oy:oy + 6, oz:oz + 2], K)
    mask = np.ones(blk.shape, bool)
    mask[ox:ox + 8, oy:oy + 6, oz:oz + 2] = False
    assert not blk[mask].any()
    with pytest.raises(ValueError):
        recon.zf_offsets((15, 10, 2), 8, 6, 2)


def test_scan_plane_table_and_engine_mirror():
    from mrilab_amd import engine
    want = {("Axial", "A/P"): True, ("Axial", "L/R"): False, ("Sagittal", "S/I"): True,
            ("Sagittal", "A/P"): False, ("Coronal", "L/R"): False, ("Coronal", "S/I"): True}
    for (plane, fd), w in want.items():
        assert recon.swap_xy(plane, fd) == w
        assert engine.swap_xy(plane, fd) == w
    rng = np.random.default_rng(1)
    sx, sy = rng.normal(size=7 * 4), rng.normal(size=7 * 4)
    a = recon.cart_recon(sx, sy, 7, 4, 1, 1, 1, 1, swap=False)
    b = recon.cart_recon(sx, sy, 7, 4, 1, 1, 1, 1, swap=True)
    np.testing.assert_array_equal(a["Sx"][:, :, 0, 0, 0].T, b["Sx"][:, :, 0, 0, 0])
    np.testing.assert_allclose(a["Real"][:, :, 0, 0, 0].T, b["Real"][:, :, 0, 0, 0], atol=1e-14)


def test_recon_dims_match_the_oracle():
    from mrilab_amd import engine
    for args, zf, plane in [((9, 8, 1, 1, 1, 1), None, "Axial"), ((8, 6, 3, 2, 2, 3), None, "Coronal"),
                            ((9, 6, 2, 1, 1, 1), (16, 10, 2), "Axial"), ((9, 6, 2, 1, 1, 1), (16, 10, 2), "Sagittal")]:
        rf, pe, sl, te, types, rx = args
        swap = recon.swap_xy(plane, "A/P" if plane == "Axial" else "L/R")
        sx = np.zeros(rf * pe * sl * te * types * rx)
        out = recon.cart_recon(sx, sx, rf, pe, sl, te, rx, types, zf=zf, swap=swap)
        d = engine.recon_desc(rf, pe, sl, te, rx, types, zf=zf, scan_plane=plane,
                              freq_dir="A/P" if plane == "Axial" else "L/R")
        assert engine.recon_dims(d) == out["Real"].shape


def test_recon_rejects_fractional_zero_filling_offsets():
    from mrilab_amd import engine
    with pytest.raises(engine.MRSError):
        engine.recon_dims(engine.recon_desc(10, 6, 1, zf=(16, 10, 1)))  # 9 kept samples: 16/2 - 9/2


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        got = recon.philox4x32_10(np.array([ctr], dtype=np.uint32), key)[0]
        assert [int(x) for x in got] == list(want)


def test_noise_normals_statistics():
    zx, zy = recon.noise_normals(200_000, seed=12345)
    for z in (zx, zy):
        assert abs(z.mean()) < 0.01
        assert abs(z.std() - 1) < 0.01
        assert abs(np.mean(z ** 4) - 3) < 0.1  # Gaussian kurtosis
    assert abs(np.corrcoef(zx, zy)[0, 1]) < 0.01
    a, _ = recon.noise_normals(1000, seed=1)
    b, _ = recon.noise_normals(1000, seed=2)
    assert not np.allclose(a, b)


def test_noise_scale_formula():
    from mrilab_amd import engine
    # the reference point of DoAddNoise.m:8-13 gives 1
    ref = (1e3, 1.0, 1.0, 1.5, 1e-3, 1e-3, 1e-3, 100, 100, 1)
    assert recon.noise_scale(*ref) == pytest.approx(1.0)
    assert engine.noise_scale(*ref) == pytest.approx(1.0)
    # sqrt(BW), NoiseLevel, 1/sqrt(NEX), 1/B0, 1/voxel volume, 1/sqrt(samples)
    args = (4e3, 2.0, 4.0, 3.0, 2e-3, 1e-3, 1e-3, 400, 100, 1)
    want = 2 * 2 / (2 * 2 * 2 * 2)
    assert recon.noise_scale(*args) == pytest.approx(want)
    assert engine.noise_scale(*args) == pytest.approx(want)
