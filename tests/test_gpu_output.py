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


def test_peak_motion_maps_follow_the_surface_history():
    """PG maps after N steps equal maxima computed on the host from the top
    plane fetched after every step (acceleration |dV|/dt, trapezoid displacement)."""
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
    dt = np.float32(P.dt)
    want = np.zeros((15, P.nj, P.ni))
    D = np.zeros((3, P.nj, P.ni))
    for n in range(1, len(top)):
        v1, v0 = top[n], top[n - 1]
        A = np.abs((v1 - v0) / dt)
        D += 0.5 * (v1 + v0) * dt
        vals = [np.sqrt((v1 ** 2).sum(0)), np.sqrt((v1[:2] ** 2).sum(0)), *np.abs(v1),
                np.sqrt((A ** 2).sum(0)), np.sqrt((A[:2] ** 2).sum(0)), *A,
                np.sqrt((D ** 2).sum(0)), np.sqrt((D[:2] ** 2).sum(0)), *np.abs(D)]
        want = np.maximum(want, np.stack(vals))
    for q, name in enumerate(cgfd.PEAK_NAMES):
        scale = np.abs(want[q]).max()
        assert scale > 0, name
        assert np.abs(pg[q] - want[q]).max() <= 1e-4 * scale, name
    full = s.read_peak_motion(0)
    assert not full[:, :3].any() and not full[:, :, -3:].any()  # ghosts stay zero
    s.close()
