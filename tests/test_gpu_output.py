"""GPU: the device-side sub-volume gather used for snapshots and slices."""
import numpy as np
import pytest

import cgfd
import orc
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
    if storage == cgfd.STORE_F32:
        # and the oracle's wavefield after the same steps, sampled the same way
        wo = P.oracle_run(nt_total=12)[1]
        ref = wo[comps][:, g + start[2]:g + start[2] + count[2] * stride[2]:stride[2],
                        g + start[1]:g + start[1] + count[1] * stride[1]:stride[1],
                        g + start[0]:g + start[0] + count[0] * stride[0]:stride[0]]
        for c in range(len(comps)):
            scale = np.abs(wo[comps[c]]).max()
            assert scale > 0
            assert np.abs(sub[c] - ref[c]).max() <= 1e-5 * scale, comps[c]
    with pytest.raises(cgfd.CgfdError):
        s.fetch_subvolume(0, cgfd.LEVEL_PRE, [0], (28, 0, 0), (2, 1, 1), (2, 1, 1))  # past the last column
    s.close()


def test_peak_motion_maps_follow_the_surface_history():
    """PG maps after N steps equal the oracle's restatement of PG_calcu_gpu
    (oracle orc_peak_motion, pinned by closed forms) fed the top plane fetched
    after every step."""
    P = Problem(ni=26, nj=22, nk=20, curvilinear=True, nt_total=14, seed=17)
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    s.set_graph(False)
    s.set_peak_motion(True)
    top = [s.fetch_subvolume(0, cgfd.LEVEL_PRE, [0, 1, 2], (0, 0, P.nk - 1), (P.ni, P.nj, 1),
                             dtype=np.float64)[:, 0]]
    for it in range(P.nt_total):
        s.run(it, 1, P.dt)
        top.append(s.fetch_subvolume(0, cgfd.LEVEL_PRE, [0, 1, 2], (0, 0, P.nk - 1), (P.ni, P.nj, 1),
                                     dtype=np.float64)[:, 0])
    pg = s.read_peak_motion(0)[:, 3:-3, 3:-3]
    vtop = np.ascontiguousarray(np.stack(top).reshape(len(top), 3, -1))
    want = np.zeros(15 * P.ni * P.nj)
    orc.lib().orc_peak_motion(P.ni * P.nj, P.nt_total, orc.p(vtop), P.dt, orc.p(want))
    want = want.reshape(15, P.nj, P.ni)
    for q, name in enumerate(cgfd.PEAK_NAMES):
        scale = np.abs(want[q]).max()
        assert scale > 0, name
        assert np.abs(pg[q] - want[q]).max() <= 1e-5 * scale, name
    full = s.read_peak_motion(0)
    assert not full[:, :3].any() and not full[:, :, -3:].any()  # ghosts stay zero
    s.close()
