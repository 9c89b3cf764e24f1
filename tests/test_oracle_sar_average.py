"""Pins the N-gram SAR averaging oracle (oracle/sar_average_oracle.c,
DoSARAverageAtCPU.cpp restated).

Known answers from the mathematics of the method:
* The voxel-index sphere S_r (rado = floor(sqrt(r^2 - o^2)) columns per
  slice offset o, radm = floor(sqrt(rado^2 - m^2)) rows per column offset
  m) has 7, 25 and 83 voxels for r = 1, 2, 3 (counted by hand), and at
  r = 3 it holds the offset (slice 2, column 1, row 1) but not (1, 1, 2):
  a hot voxel at those offsets pins which array axis each loop walks.
* With uniform density rho and voxel volume v, the interpolated sphere
  holds exactly grams*1e-3/(v rho) voxels, so a hot voxel of excess a adds
  a v rho/(grams*1e-3) to every voxel whose S_(R-1) holds it, c times that
  where only S_R holds it (c the interpolation weight), nothing elsewhere.
* A uniform SAR averages to itself on any density; a linear SAR averages to
  the centre value where the sphere is not clipped (S_r is symmetric).
* Below one sphere of mass (R = 1) the result is the mean over the 7-voxel
  cross.
* Brute force: a vectorised restatement from the set definition of S_r
  (offset inequalities over a box, exact sums with math.fsum) agrees on a
  small random volume with zero-mass voxels, non-positive centres and
  clipped spheres.
"""
import math

import numpy as np
import pytest

import oracle


def sphere_offsets(r):
    """(o, m, n) offsets of S_r from its inequalities, in no particular order."""
    o, m, n = np.meshgrid(*(np.arange(-r, r + 1),) * 3, indexing="ij")
    rado2 = np.array([math.isqrt(int(x)) ** 2 if x >= 0 else -1 for x in (r * r - o * o).ravel()]).reshape(o.shape)
    keep = (r * r - o * o >= 0) & (m * m <= rado2) & (n * n <= rado2 - m * m)
    return np.stack([o[keep], m[keep], n[keep]], axis=1)


def brute_average(SAR, Power, M, dx, dy, dz, grams):
    D0, D1, D2, T = SAR.shape
    g = grams * 1e-3
    vol = dx * dy * dz
    a = np.zeros(SAR.shape)
    b = np.zeros(SAR.shape)
    cache = {}
    for s in range(D2):
        for i in range(D1):
            for j in range(D0):
                if M[j, i, s] == 0:
                    continue
                hot = SAR[j, i, s, 0] > 0
                prev = (0.0, 0, np.zeros(T), np.zeros(T))
                r = 0
                while True:
                    r += 1
                    off = cache.setdefault(r, sphere_offsets(r))
                    jj, ii, ss = j + off[:, 2], i + off[:, 1], s + off[:, 0]
                    ok = (jj >= 0) & (jj < D0) & (ii >= 0) & (ii < D1) & (ss >= 0) & (ss < D2)
                    jj, ii, ss = jj[ok], ii[ok], ss[ok]
                    w = math.fsum(vol * M[jj, ii, ss])
                    px = int(ok.sum()) if hot else 0
                    sa = np.array([math.fsum(SAR[jj, ii, ss, t]) for t in range(T)]) if hot else np.zeros(T)
                    pa = np.array([math.fsum(Power[jj, ii, ss, t]) for t in range(T)]) if hot else np.zeros(T)
                    if w > g:
                        break
                    prev = (w, px, sa, pa)
                w2, px2, sa2, pa2 = prev
                c = (g - w2) / (w - w2)
                if px > 0:
                    den = px2 + c * (px - px2)
                    a[j, i, s] = (sa2 + c * (sa - sa2)) / den
                    b[j, i, s] = (pa2 + c * (pa - pa2)) / den
    return a, b


def test_sphere_counts_by_hand():
    assert [len(sphere_offsets(r)) for r in (1, 2, 3)] == [7, 25, 83]
    s3 = {tuple(x) for x in sphere_offsets(3)}
    assert (2, 1, 1) in s3 and (1, 1, 2) not in s3 and (1, 2, 1) not in s3


def hot_volume(n, eps, a, hot):
    SAR = np.full((n, n, n, 1), eps)
    SAR[hot[2], hot[1], hot[0], 0] += a  # hot = (slice, col, row) -> [row, col, slice]
    return SAR


@pytest.mark.parametrize("r,count_prev,count", [(2, 7, 25), (3, 25, 83)])
def test_hot_voxel_spread(r, count_prev, count):
    n, c0 = 13, 6
    dx = dy = dz = 1e-3
    rho = 1000.0
    vr = dx * dy * dz * rho
    g_vox = 0.5 * (count_prev + count)  # weight c = 1/2
    grams = g_vox * vr * 1e3
    M = np.full((n, n, n), rho)
    eps, a = 0.25, 1.0
    inner = {tuple(x) for x in sphere_offsets(r - 1)}
    outer = {tuple(x) for x in sphere_offsets(r)}
    for hot in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 1, 1), (1, 1, 2), (1, 2, 1), (2, 0, 0), (3, 0, 0)]:
        SAR = hot_volume(n, eps, a, (c0 + hot[0], c0 + hot[1], c0 + hot[2]))
        ave, pw = oracle.sar_average(SAR, 2 * SAR, M, dx, dy, dz, grams)
        if hot in inner:
            want = eps + a / g_vox
        elif hot in outer:
            want = eps + 0.5 * a / g_vox
        else:
            want = eps
        got = ave[c0, c0, c0, 0]
        assert got == pytest.approx(want, rel=1e-13), hot
        assert pw[c0, c0, c0, 0] == pytest.approx(2 * want, rel=1e-13), hot


def test_uniform_sar_is_kept():
    rng = np.random.default_rng(3)
    M = rng.uniform(800, 1200, (9, 8, 6))
    M[rng.random(M.shape) < 0.1] = 0.0
    SAR = np.full((9, 8, 6, 2), 0.5)
    SAR[..., 1] = 3.0
    ave, pw = oracle.sar_average(SAR, 4 * SAR, M, 2e-3, 2e-3, 3e-3, 0.3)
    on = M != 0
    np.testing.assert_allclose(ave[on, 0], 0.5, rtol=1e-14)
    np.testing.assert_allclose(ave[on, 1], 3.0, rtol=1e-14)
    np.testing.assert_allclose(pw[on, 0], 2.0, rtol=1e-14)
    assert np.all(ave[~on] == 0) and np.all(pw[~on] == 0)


def test_linear_sar_averages_to_centre():
    n = 17
    j, i, s = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    SAR = (1.0 + 0.01 * j + 0.02 * i + 0.03 * s)[..., None]
    M = np.full((n, n, n), 1000.0)
    grams = 40 * 1e-6 * 1e3  # 40 voxels of 1 mm^3 at 1000 kg/m^3: R = 3
    ave, _ = oracle.sar_average(SAR, SAR, M, 1e-3, 1e-3, 1e-3, grams)
    core = (slice(4, n - 4),) * 3
    np.testing.assert_allclose(ave[core][..., 0], SAR[core][..., 0], rtol=1e-13)
    # clipped spheres at the faces pull the average inwards
    assert ave[0, 8, 8, 0] > SAR[0, 8, 8, 0]


def test_one_sphere_is_the_cross_mean():
    rng = np.random.default_rng(4)
    n = 7
    SAR = rng.uniform(0.1, 1.0, (n, n, n, 1))
    M = np.full((n, n, n), 1000.0)
    ave, _ = oracle.sar_average(SAR, SAR, M, 1e-3, 1e-3, 1e-3, 3e-3)  # 3 voxels of mass < 7
    c = 3
    cross = [SAR[c, c, c, 0]] + [SAR[c + d[0], c + d[1], c + d[2], 0] for d in
                                 [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]]
    assert ave[c, c, c, 0] == pytest.approx(np.mean(cross), rel=1e-14)


def test_skipped_voxels_and_first_time_point_gate():
    SAR = np.full((5, 5, 5, 2), 1.0)
    SAR[2, 2, 2, 0] = 0.0   # centre not > 0 at the first time point: both times stay 0
    SAR[1, 2, 2, 0] = -1.0
    M = np.full((5, 5, 5), 1000.0)
    M[3, 3, 3] = 0.0
    ave, pw = oracle.sar_average(SAR, SAR, M, 1e-3, 1e-3, 1e-3, 0.02)
    assert np.all(ave[2, 2, 2] == 0) and np.all(pw[2, 2, 2] == 0)
    assert np.all(ave[1, 2, 2] == 0)
    assert np.all(ave[3, 3, 3] == 0)
    assert ave[0, 0, 0, 1] > 0


def test_too_little_mass_is_an_error():
    M = np.full((3, 3, 3), 1000.0)
    SAR = np.ones((3, 3, 3, 1))
    with pytest.raises(ValueError):
        oracle.sar_average(SAR, SAR, M, 1e-3, 1e-3, 1e-3, 1.0)  # 27 mg < 1 g


def test_exact_total_mass_never_closes():
    M = np.full((3, 3, 3), 8.0)  # 0.125 m^3 voxels of 8 kg/m^3: 1 kg each, 27 kg in all
    SAR = np.ones((3, 3, 3, 1))
    with pytest.raises(ValueError):
        oracle.sar_average(SAR, SAR, M, 0.5, 0.5, 0.5, 27000.0)
    a, _ = oracle.sar_average(SAR, SAR, M, 0.5, 0.5, 0.5, 26999.999999)
    np.testing.assert_allclose(a, 1.0, rtol=1e-12)


def test_brute_force_small_volume():
    rng = np.random.default_rng(7)
    D = (7, 6, 5, 2)
    M = rng.uniform(800, 1200, D[:3])
    M[rng.random(D[:3]) < 0.15] = 0.0
    SAR = rng.uniform(-0.2, 1.0, D)
    Power = rng.uniform(0.0, 2.0, D)
    for grams in (0.05, 0.3, 1.0):
        want = brute_average(SAR, Power, M, 2e-3, 2e-3, 2.5e-3, grams)
        got = oracle.sar_average(SAR, Power, M, 2e-3, 2e-3, 2.5e-3, grams)
        for g_, w_ in zip(got, want):
            np.testing.assert_allclose(g_, w_, rtol=1e-12, atol=1e-15)
        assert np.count_nonzero(got[0]) > 0


def test_slice_range_matches_full():
    rng = np.random.default_rng(8)
    D = (6, 5, 6, 1)
    M = rng.uniform(900, 1100, D[:3])
    SAR = rng.uniform(0.1, 1.0, D)
    full = oracle.sar_average(SAR, SAR, M, 2e-3, 2e-3, 2e-3, 0.2)
    part = oracle.sar_average(SAR, SAR, M, 2e-3, 2e-3, 2e-3, 0.2, slices=(2, 4))
    np.testing.assert_array_equal(part[0][:, :, 2:4], full[0][:, :, 2:4])
    assert np.all(part[0][:, :, :2] == 0)
