"""The allocation-free SAR average (mrs_DoSARAverageAtGPU_noalloc) against
the allocating entry point (mrs_DoSARAverageAtGPU), through the C ABI.

Both run the same kernels on the same device arrays; the allocation-free one
keeps its per-voxel scratch in the library's static device array. Results
must be equal bit for bit, on two calls in a row (the second on a smaller
volume, so stale scratch of the first would show), on a caller's stream and
on the default stream; a volume above the scratch's 2^20 voxels is refused
with MRS_ERR_UNSUPPORTED before any device work.
"""
import ctypes as C

import numpy as np
import pytest

from mrilab_amd import engine

pytestmark = pytest.mark.gpu


def _volume(D, seed):
    rng = np.random.default_rng(seed)
    M = rng.uniform(800, 1200, D[:3])
    M[rng.random(D[:3]) < 0.15] = 0.0
    SAR = rng.uniform(-0.2, 1.0, D)
    Power = rng.uniform(0.0, 2.0, D)
    return SAR, Power, M


def _dev(a):
    """A CUDA float64 tensor holding ``a`` in column-major order (flat)."""
    import torch
    return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).cuda()


def _noalloc(SAR, Power, M, d, grams, stream=None):
    import torch
    lib = engine.load()
    shp = (tuple(SAR.shape) + (1, 1, 1, 1))[:4]
    s, p, m = _dev(SAR), _dev(Power), _dev(M)
    a = torch.empty_like(s)
    b = torch.empty_like(s)
    dims = (C.c_int64 * 4)(*shp)
    h = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = lib.mrs_DoSARAverageAtGPU_noalloc(C.c_void_p(s.data_ptr()), C.c_void_p(p.data_ptr()),
                                           C.c_void_p(m.data_ptr()), dims, *d, grams, C.c_void_p(a.data_ptr()),
                                           C.c_void_p(b.data_ptr()), h)
    if rc != 0:
        raise engine.MRSError(rc, lib.mrs_last_error().decode())
    torch.cuda.synchronize()
    f = lambda t: t.cpu().numpy().reshape(SAR.shape, order="F")  # noqa: E731
    return f(a), f(b)


def _check(SAR, Power, M, d, grams, tag, stream=None):
    print(f"case {tag}: shape {SAR.shape}, grams {grams}", flush=True)
    want = engine.DoSARAverageAtGPU(SAR, Power, M, *d, grams)
    got = _noalloc(SAR, Power, M, d, grams, stream)
    for x, y in zip(got, want):
        assert np.array_equal(x, y)
    assert np.count_nonzero(got[0]) > 0


def test_two_calls_in_a_row_second_smaller(gpu):
    import torch
    d = (2e-3, 2e-3, 2.5e-3)
    SAR, Power, M = _volume((48, 40, 24, 3), 1)          # 46,080 voxels
    _check(SAR, Power, M, d, 0.3, "first, larger")
    SAR2, Power2, M2 = _volume((13, 11, 7, 2), 2)         # 1,001 voxels
    _check(SAR2, Power2, M2, d, 0.05, "second, smaller")
    st = torch.cuda.Stream()
    _check(SAR2, Power2, M2, d, 0.2, "on a caller's stream", stream=st)


def test_largest_volume_and_refusal(gpu):
    d = (2e-3, 2e-3, 2e-3)
    SAR, Power, M = _volume((128, 128, 64, 1), 3)          # exactly 2^20 voxels
    _check(SAR, Power, M, d, 10.0, "2^20 voxels")
    big = _volume((129, 128, 64, 1), 4)
    with pytest.raises(engine.MRSError, match="at most 1048576 voxels"):
        _noalloc(*big, d, 10.0)
