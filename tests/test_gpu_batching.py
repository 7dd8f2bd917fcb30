"""GPU: captured 8-step graphs and continuous batching of shots.

A shot's result must not depend on how it was run: eagerly or through the
captured period graphs, alone or in a batch, admitted at step 0 or later.
All sessions pin the same interior tile, so the comparisons are exact.
"""
import numpy as np
import pytest

import cgfd
from problem import Problem, rel_l2

pytestmark = pytest.mark.gpu

TILE = (2, 24)
TOL32 = 1e-5


def _model(src_ijk, nt_total):
    # identical grid, medium and receivers for every src_ijk (same seed)
    return Problem(ni=40, nj=36, nk=30, curvilinear=True, hetero=0.03, nt_total=nt_total, seed=21,
                   src_ijk=src_ijk, pml=(4, 4, 4, 4, 4, 0))


def _sess(P, nshots=1):
    s = cgfd.Session(P.ni, P.nj, P.nk, nshots=nshots)
    P.upload(s)
    s.set_tile(*TILE)
    return s


def test_graph_periods_equal_eager_launches():
    P = _model((20, 18, 15), 27)
    a = _sess(P)
    a.set_graph(False)
    a.run(0, 27, P.dt)
    b = _sess(P)
    b.run(0, 3, P.dt)      # eager 0..2
    b.run(3, 24, P.dt)     # eager 3..7, graphs 8..23, eager 24..26
    assert np.array_equal(a.read_seismo(0), b.read_seismo(0))
    assert np.array_equal(a.fetch_level(0, cgfd.LEVEL_PRE), b.fetch_level(0, cgfd.LEVEL_PRE))
    # and both are the oracle's run
    so, wo, _ = P.oracle_run()
    assert rel_l2(b.read_seismo(0), so) <= TOL32
    assert rel_l2(b.fetch_level(0, cgfd.LEVEL_PRE), wo) <= TOL32
    # a replayed period after more eager work still matches
    a.run(27, 13, P.dt)
    b.run(27, 13, P.dt)
    assert np.array_equal(a.fetch_level(0, cgfd.LEVEL_PRE), b.fetch_level(0, cgfd.LEVEL_PRE))
    a.close()
    b.close()


@pytest.mark.parametrize("reserve", [False, True], ids=["rebuild", "reserved"])
def test_shot_admitted_mid_run_matches_standalone(reserve):
    A = _model((12, 14, 20), 40)
    B = _model((25, 20, 15), 24)
    sa = _sess(A)
    sa.run(0, A.nt_total, A.dt)
    sb = _sess(B)
    sb.run(0, B.nt_total, B.dt)
    # the standalone runs are the oracle's, so the batch below is too
    for P, st in ((A, sa), (B, sb)):
        assert rel_l2(st.read_seismo(0), P.oracle_run()[0]) <= TOL32

    m = _sess(A, nshots=3)
    if reserve:
        m.reserve_receivers(len(A.recv_indx), 64)
    m.set_shot(1, False)
    m.set_shot(2, False)
    m.run(0, 16, A.dt)
    # admit B into slot 1 at step 16 while shot 0 keeps running
    m.reset_shot(1)
    B.upload_shot(m, 1)
    m.set_shot(1, True, 16)
    m.run_async(16, 24, A.dt)
    m.wait()
    assert m.poll()
    assert np.array_equal(m.read_seismo(0), sa.read_seismo(0))
    assert np.array_equal(m.read_seismo(1), sb.read_seismo(0))
    assert np.array_equal(m.fetch_level(0, cgfd.LEVEL_PRE), sa.fetch_level(0, cgfd.LEVEL_PRE))
    assert np.array_equal(m.fetch_level(1, cgfd.LEVEL_PRE), sb.fetch_level(0, cgfd.LEVEL_PRE))
    # the never-activated slot stayed untouched
    assert not np.any(m.fetch_level(2, cgfd.LEVEL_PRE))
    for s in (sa, sb, m):
        s.close()


def test_shot_start_must_align_with_the_operator_period():
    P = _model(None, 8)
    s = _sess(P, nshots=2)
    with pytest.raises(cgfd.CgfdError):
        s.set_shot(1, True, 12)
    s.close()
