"""GPU: slice files written during a run and receiver lines, against the oracle.

Slices follow io_slice_nc_put (src/forward/io_funcs.cu:667-780): one record
of the nine W^n components on a grid plane per output step.  Receiver lines
follow io_line_locate / io_line_keep (io_funcs.cu:302-416, 1669-1698): the
value at a grid node, no interpolation, recorded with the receivers.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

import cgfd
from problem import Problem, rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
NAMES = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]


def _plane(w, axis, g):
    """[9][slow][fast] plane of a full level [9][nz][ny][nx] at physical index g."""
    if axis == 0:
        return w[:, 3:-3, 3:-3, 3 + g]
    if axis == 1:
        return w[:, 3:-3, 3 + g, 3:-3]
    return w[:, 3 + g, 3:-3, 3:-3]


def test_slices_written_during_a_run_match_the_oracle(tmp_path):
    P = Problem(ni=26, nj=22, nk=20, curvilinear=True, nt_total=9, seed=21)
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    planes = [(0, 11), (1, 0), (2, P.nk - 1), (2, 9)]
    files = [cgfd.Slice(cgfd.slice_path(tmp_path, a, g), a, g, P.ni, P.nj, P.nk) for a, g in planes]
    out_steps = [3, 6, 9]
    it = 0
    for n in out_steps:
        s.run(it, n - it, P.dt)
        it = n
        for (a, g), f in zip(planes, files):
            f.append(n * P.dt, cgfd.session_slice(s, 0, a, g))
    for f in files:
        f.close()
    want = {n: P.oracle_run(nt_total=n)[1] for n in out_steps}
    for a, g in planes:
        nc = netcdf_file(cgfd.slice_path(tmp_path, a, g), "r", mmap=False)
        assert np.allclose(nc.variables["time"][:], np.float32(np.array(out_steps) * P.dt))
        for r, n in enumerate(out_steps):
            ref = _plane(want[n], a, g)
            for c, name in enumerate(NAMES):
                got = nc.variables[name][r].astype(np.float64)
                assert got.shape == ref[c].shape
                scale = np.abs(want[n][c]).max()
                assert scale > 0
                assert np.abs(got - ref[c]).max() <= TOL32 * scale, (a, g, n, name)
        nc.close()
    s.close()


def test_receiver_line_records_node_values(tmp_path):
    P = Problem(ni=24, nj=20, nk=18, curvilinear=True, nt_total=12, seed=23, nrecv=0)
    # a diagonal line that leaves the grid: only the points inside are kept
    seq, ix, fr, xyz = cgfd.line_points(P.ni, P.nj, P.nk, (2, 3, P.nk - 1), (3, 2, -2), 10,
                                        P.x.ravel(), P.y.ravel(), P.z.ravel())
    assert len(seq) == 8 and list(seq) == list(range(8))
    P.recv_indx, P.recv_frac = ix, fr
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    s.run(0, P.nt_total - 1, P.dt)
    # sample it is W^it at the node, taken before step it (drv_rk_curv_col.cu:193-203)
    flat = s.fetch_level(0, cgfd.LEVEL_PRE).reshape(9, -1)
    s.run(P.nt_total - 1, 1, P.dt)
    seis = s.read_seismo(0)
    so, wo, _ = P.oracle_run()
    assert np.abs(so).max() > 0
    assert rel_l2(seis, so) <= TOL32
    for n in range(len(seq)):
        assert np.array_equal(seis[n, :, -1], flat[:, ix[n, 0]])
    for n in range(len(seq)):
        cgfd.line_sac(tmp_path, "diag", int(seq[n]), seis[n].astype(np.float32), P.dt, *xyz[n])
    for c, name in enumerate(NAMES):
        raw = open(tmp_path / f"diag.no5.{name}.sac", "rb").read()
        assert np.array_equal(np.frombuffer(raw[632:], np.float32), seis[5, c].astype(np.float32))
    s.close()
