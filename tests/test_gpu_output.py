"""GPU: the device-side sub-volume gather used for snapshots and slices."""
import numpy as np
import pytest

import cgfd
from problem import Problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", [cgfd.STORE_F32, cgfd.STORE_F16], ids=["f32", "f16"])
def test_subvolume_equals_slice_of_full_level(storage):
    P = Problem(ni=30, nj=28, nk=24, curvilinear=True, nt_total=12, seed=9)
    s = cgfd.Session(P.ni, P.nj, P.nk, storage=storage)
    P.upload(s)
    if storage == cgfd.STORE_F16:
        s.set_storage_scale(cgfd.storage_scales(storage, P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.src,
                                                P.dt))
    s.run(0, 12, P.dt)
    full = s.fetch_level(0, cgfd.LEVEL_PRE)
    comps = [0, 4, 8]
    start, count, stride = (2, 3, 1), (5, 4, 3), (3, 2, 4)
    sub = s.fetch_subvolume(0, cgfd.LEVEL_PRE, comps, start, count, stride, dtype=np.float64)
    g = 3
    want = full[comps][:, g + start[2]:g + start[2] + count[2] * stride[2]:stride[2],
                       g + start[1]:g + start[1] + count[1] * stride[1]:stride[1],
                       g + start[0]:g + start[0] + count[0] * stride[0]:stride[0]]
    assert sub.shape == want.shape
    assert np.array_equal(sub, want)
    with pytest.raises(cgfd.CgfdError):
        s.fetch_subvolume(0, cgfd.LEVEL_PRE, [0], (28, 0, 0), (2, 1, 1), (2, 1, 1))  # past the last column
    s.close()
