"""GPU: cgfd_fetch_subvolume_noalloc (staging in the library's fixed device
scratch, no device allocation) returns what cgfd_fetch_subvolume returns.

Two calls in a row, the second with a smaller sub-volume, so a stale tail
of the scratch from the first call cannot pass for data; FP32 and FP64
sessions; requests beyond the scratch are refused before any device work.
"""
import numpy as np
import pytest

import cgfd
from problem import Problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", [cgfd.STORE_F32, cgfd.STORE_F64, cgfd.STORE_F16], ids=["f32", "f64", "f16"])
def test_noalloc_subvolume_equals_allocating_fetch(storage):
    P = Problem(ni=60, nj=52, nk=44, curvilinear=True, pml=(5, 5, 5, 5, 5, 0), nt_total=24, seed=41, hetero=0.03,
                f0=4.0, t0=0.15)
    s = cgfd.Session(P.ni, P.nj, P.nk, storage=storage)
    P.upload(s, dtype=np.float64 if storage == cgfd.STORE_F64 else np.float32)
    if storage == cgfd.STORE_F16:
        s.set_storage_scale(cgfd.storage_scales(storage, P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.src,
                                                P.dt))
    s.run(0, P.nt_total, P.dt)
    dt = np.float64 if storage == cgfd.STORE_F64 else np.float32
    calls = [(list(range(9)), (1, 2, 3), (50, 40, 30), (1, 1, 1)),      # large first
             ([0, 2, 5, 8], (4, 3, 7), (9, 12, 5), (5, 3, 6))]          # then smaller and strided
    for comps, start, count, stride in calls:
        want = s.fetch_subvolume(0, cgfd.LEVEL_PRE, comps, start, count, stride, dtype=dt)
        got = s.fetch_subvolume_noalloc(0, cgfd.LEVEL_PRE, comps, start, count, stride)
        assert got.dtype == dt and got.shape == want.shape
        assert np.abs(want).max() > 0
        assert np.array_equal(got, want)
    s.close()


def test_noalloc_subvolume_refuses_what_the_scratch_cannot_hold():
    P = Problem(ni=180, nj=170, nk=60, pml=(0,) * 6, source=None, nrecv=0, nt_total=1)
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    # 180 x 170 x 60 floats = 7.3 MB fits; with the wrong output type it is refused
    got = s.fetch_subvolume_noalloc(0, cgfd.LEVEL_PRE, [0], (0, 0, 0), (180, 170, 60))
    assert got.shape == (1, 60, 170, 180)
    lib = cgfd.load()
    out = np.zeros(180 * 170 * 60)
    rc = lib.cgfd_fetch_subvolume_noalloc(s.h, 0, 0, 1, 0, 0, 0, 180, 170, 60, 1, 1, 1, cgfd.DT_F64,
                                          out.ctypes.data)
    assert rc == -1 and b"model type" in lib.cgfd_last_error()
    s.close()
    big = Problem(ni=200, nj=200, nk=60, pml=(0,) * 6, source=None, nrecv=0, nt_total=1)
    s = cgfd.Session(big.ni, big.nj, big.nk)
    big.upload(s)
    with pytest.raises(cgfd.CgfdError, match="scratch"):
        s.fetch_subvolume_noalloc(0, cgfd.LEVEL_PRE, [0], (0, 0, 0), (200, 200, 60))
    s.close()
