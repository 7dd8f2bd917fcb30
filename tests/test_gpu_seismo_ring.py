"""GPU: the seismogram ring (device keeps `ring` samples per trace, chunks
stream to a pinned host record) records exactly what the whole-record mode
records: bit for bit, over runs cut at arbitrary steps, batched shots with
staggered starts, asynchronous runs and receivers replaced between runs."""
import numpy as np
import pytest

import cgfd
from problem import Problem

pytestmark = pytest.mark.gpu


def _pair(P, nshots=1):
    a = cgfd.Session(P.ni, P.nj, P.nk, nshots=nshots)
    b = cgfd.Session(P.ni, P.nj, P.nk, nshots=nshots)
    P.upload(a)
    P.upload(b)
    return a, b


def test_ring_equals_whole_record_over_cut_runs():
    P = Problem(ni=28, nj=24, nk=20, curvilinear=True, nt_total=90, seed=31, nrecv=7)
    whole, ring = _pair(P)
    ring.set_seismo_ring(16)
    for it0, n in ((0, 5), (5, 27), (32, 16), (48, 1), (49, 41)):
        whole.run(it0, n, P.dt)
        ring.run(it0, n, P.dt)
    a, b = whole.read_seismo(0), ring.read_seismo(0)
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)
    whole.close()
    ring.close()


def test_ring_batched_shots_with_staggered_starts_and_async_runs():
    P = Problem(ni=26, nj=22, nk=18, curvilinear=False, nt_total=70, seed=33, nrecv=5)
    whole, ring = _pair(P, nshots=3)
    ring.set_seismo_ring(32)
    for s in (whole, ring):
        s.set_shot(1, True, 8)
        s.set_shot(2, True, 40)
    # shot 2 starts at step 40, so 110 steps record all 70 of its samples
    for it0, n in ((0, 24), (24, 56), (80, 30)):
        whole.run(it0, n, P.dt)
        ring.run_async(it0, n, P.dt)
    ring.wait()
    assert ring.poll()
    for shot in range(3):
        a, b = whole.read_seismo(shot), ring.read_seismo(shot)
        assert np.abs(a).max() > 0, shot
        assert np.array_equal(a, b), shot
    # shot 1 is shot 0 delayed by 8 steps
    assert np.array_equal(whole.read_seismo(0), whole.read_seismo(1))
    whole.close()
    ring.close()


def test_ring_receivers_replaced_between_runs():
    P = Problem(ni=24, nj=20, nk=18, curvilinear=True, nt_total=40, seed=35, nrecv=4)
    whole, ring = _pair(P, nshots=2)
    ring.set_seismo_ring(16)
    for s in (whole, ring):
        s.run(0, 20, P.dt)
    # new receivers on shot 1 (fewer, longer record): shot 0 keeps its samples
    idx, frac = P.recv_indx[:2], P.recv_frac[:2]
    for s in (whole, ring):
        s.upload_receivers(1, idx, frac, 60)
        s.run(20, 20, P.dt)
    for shot in range(2):
        assert np.array_equal(whole.read_seismo(shot), ring.read_seismo(shot)), shot
    whole.close()
    ring.close()


def test_ring_argument_checks():
    P = Problem(ni=16, nj=16, nk=12, nt_total=8, seed=37, nrecv=2)
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    for bad in (-16, 8, 24):
        with pytest.raises(cgfd.CgfdError, match="multiple of 16"):
            s.set_seismo_ring(bad)
    s.set_seismo_ring(16)
    s.set_seismo_ring(0)
    s.run(0, 8, P.dt)
    assert s.read_seismo(0).shape == (len(P.recv_indx), 9, 8)
    s.close()
