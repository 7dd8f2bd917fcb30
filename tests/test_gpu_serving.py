"""GPU: the shot server (continuous batching over resident models).

Every served result must equal the same shot run alone in its own session
(relative L2 <= 1e-6 over the seismograms; the kernels are the same, only the
interior tile the autotuner picks for a batch may differ), whatever else is
in flight, and requests beyond the free slots or memory wait instead of
failing.
"""
import numpy as np
import pytest

import cgfd
from problem import Problem, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _problem(kind, src_ijk, nt):
    if kind == "a":
        return Problem(ni=40, nj=36, nk=30, curvilinear=True, hetero=0.03, nt_total=nt, seed=21, src_ijk=src_ijk,
                       pml=(4, 4, 4, 4, 4, 0))
    return Problem(ni=48, nj=40, nk=28, curvilinear=False, hetero=0.05, nt_total=nt, seed=8, src_ijk=src_ijk,
                   pml=(5, 5, 5, 5, 5, 0), source="force")


def _alone(P):
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    s.run(0, P.nt_total, P.dt)
    out = s.read_seismo(0)
    s.close()
    return out


def _add(srv, P, max_nt):
    return srv.add_model(P.ni, P.nj, P.nk, P.metric, P.lam, P.mu, P.slw, P.dt, free_top=P.free_top, mvx=P.mvx,
                         mvy=P.mvy, pml=P.pml, max_nrecv=len(P.recv_indx), max_nt=max_nt)


def _mix():
    return [("a", (12, 14, 20), 40), ("b", (20, 18, 12), 32), ("a", (25, 20, 15), 24), ("a", (18, 10, 22), 56),
            ("b", (30, 25, 14), 48), ("a", (20, 20, 10), 16)]


@pytest.mark.parametrize("budget_groups", [2.5, 1.2], ids=["both-resident", "one-at-a-time"])
def test_served_shots_match_standalone_runs(budget_groups):
    probs = [(_problem(k, ijk, nt), k) for k, ijk, nt in _mix()]
    ref = [_alone(P) for P, _ in probs]
    # the alone runs are the oracle's (FP32 tolerance), so the served ones are too
    for (P, k), r in zip(probs[:2], ref[:2]):
        assert rel_l2(r, P.oracle_run()[0]) <= 1e-5, k
    pa = next(P for P, k in probs if k == "a")
    pb = next(P for P, k in probs if k == "b")
    # memory for about budget_groups groups of two shots
    one = max(cgfd.footprint_bytes(P.ni, P.nj, P.nk, nshots=2, nrecv=len(P.recv_indx), nt_total=64)
              for P in (pa, pb))
    srv = cgfd.Server(hbm_budget=int(budget_groups * one), max_batch=2, chunk_steps=8)
    ma = _add(srv, pa, 64)
    mb = _add(srv, pb, 64)
    tickets = [srv.submit(ma if k == "a" else mb, P.nt_total, P.src, P.recv_indx, P.recv_frac) for P, k in probs]
    for t in tickets:
        assert srv.wait(t, 120.0)
    for (P, k), t, r in zip(probs, tickets, ref):
        seis, t_admit, t_done = srv.result(t)
        assert 0.0 <= t_admit <= t_done
        e = rel_l2(seis, r)
        assert e <= TOL, f"model {k} nt {P.nt_total}: served vs alone rel L2 {e:.3e}"
    st = srv.stats()
    assert st["completed"] == len(probs) and st["rejected"] == 0
    assert st["hbm_in_use"] <= st["hbm_budget"]
    srv.close()


def test_model_larger_than_budget_is_refused():
    P = _problem("a", None, 16)
    srv = cgfd.Server(hbm_budget=1 << 20, max_batch=2, chunk_steps=8)
    with pytest.raises(cgfd.CgfdError, match="insufficient HBM"):
        _add(srv, P, 16)
    srv.close()


def test_request_bounds_are_checked():
    P = _problem("a", None, 16)
    srv = cgfd.Server(max_batch=2, chunk_steps=8)
    m = _add(srv, P, 16)
    with pytest.raises(cgfd.CgfdError):
        srv.submit(m, 17, P.src, P.recv_indx, P.recv_frac)  # beyond the model's max_nt
    t = srv.submit(m, 16, P.src, P.recv_indx, P.recv_frac)
    assert srv.wait(t, 60.0)
    seis, _, _ = srv.result(t)
    assert rel_l2(seis, _alone(P)) <= TOL
    srv.close()
