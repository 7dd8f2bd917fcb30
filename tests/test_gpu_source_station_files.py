"""GPU: a run set up from a source file (src_read_locate_file,
src/forward/src_t.cu:143-795) and a station file (io_recv_read_locate,
src/forward/io_funcs.cu:24-299), as the reference's driver sets one up
(main_curv_col_el_3d.cu:655, 690), against the oracle on the same tables."""
import numpy as np
import pytest

import cgfd
from problem import Problem, rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def _stations(P, path):
    st = cgfd.station_file(path, P.ni, P.nj, P.nk, P.x.ravel(), P.y.ravel(), P.z.ravel())
    kept = [s for s in st if s["inside"]]
    P.recv_indx = np.array([s["indx8"] for s in kept], dtype=np.int64)
    P.recv_frac = np.array([s["frac"] for s in kept])
    return st, kept


def _run(P):
    s = cgfd.Session(P.ni, P.nj, P.nk)
    P.upload(s)
    s.run(0, P.nt_total, P.dt)
    seis = s.read_seismo(0)
    s.close()
    so, _, _ = P.oracle_run()
    assert np.abs(so).max() > 0
    return seis, so


def test_moment_sources_and_stations_from_files(tmp_path):
    P = Problem(ni=30, nj=26, nk=22, curvilinear=True, source=None, nrecv=0, nt_total=36, seed=41)
    # two moment sources by coordinates, depth below the surface: strike,
    # dip and rake with mu from the model; a Ricker and a later Gaussian
    xc, yc = float(P.x[3 + 9, 3 + 12, 3 + 17]), float(P.y[3 + 9, 3 + 12, 3 + 17])
    coord_path = tmp_path / "ev_xy.src"
    coord_path.write_text("ev_files\n2\n0 0.18\n2 1\n1 1\n"
                          f"{float(P.x[3 + 15, 3 + 10, 3 + 12]):.3f} {float(P.y[3 + 15, 3 + 10, 3 + 12]):.3f} 450.0\n"
                          f"{xc:.3f} {yc:.3f} 750.0\n"
                          "0.0 ricker 8.0 0.12\n40 25 60 -1 0.8 2.0e6\n"
                          "0.02 gaussian 0.03 0.08\n120 70 -30 -1 0.6 1.5e6\n")
    info, src, node, _ = cgfd.source_file(coord_path, P.ni, P.nj, P.nk, 0.0, P.dt,
                                          P.x.ravel(), P.y.ravel(), P.z.ravel(), mu=P.mu)
    assert info["nsrc"] == 2 and info["by_coord"] == 1 and info["depth"] == 1
    assert np.isfinite(src["M"]).all() and np.abs(src["M"]).max() > 0
    P.src = src
    sta_path = tmp_path / "sta.list"
    sta_path.write_text("# stations\n5\n"
                        "s0 0 1 5.0 6.0 0.0\n"
                        "s1 0 0 20.5 14.25 9.75\n"
                        f"s2 1 1 {float(P.x[3, 3 + 8, 3 + 21]) + 37.0:.3f} {float(P.y[3, 3 + 8, 3 + 21]) + 21.0:.3f} 0.0\n"
                        f"s3 1 0 {float(P.x[3 + 10, 3 + 18, 3 + 7]) + 12.0:.3f} "
                        f"{float(P.y[3 + 10, 3 + 18, 3 + 7]) - 8.0:.3f} {float(P.z[3 + 10, 3 + 18, 3 + 7]) + 30.0:.3f}\n"
                        "out 0 1 400 4 0\n")
    st, kept = _stations(P, sta_path)
    assert [s["name"] for s in kept] == ["s0", "s1", "s2", "s3"] and not st[4]["inside"]
    seis, so = _run(P)
    assert seis.shape == (4, 9, P.nt_total)
    assert rel_l2(seis, so) <= TOL32


def test_force_source_with_sampled_time_function(tmp_path):
    P = Problem(ni=26, nj=24, nk=20, curvilinear=False, source=None, nrecv=0, nt_total=30, seed=43)
    # a force whose time function is given as samples every 0.01 s (half a
    # Ricker), read by piecewise cubic Lagrange at every stage
    ts = np.arange(16) * 0.01
    r = np.vectorize(lambda t: cgfd.wavelet(cgfd.WAVELET_RICKER, t, 8.0, 0.08))(ts)
    rows = "\n".join(f"{1e12 * v:.6e} {-0.4e12 * v:.6e} {0.7e12 * v:.6e}" for v in r)
    src_path = tmp_path / "f.src"
    src_path.write_text(f"ev_force\n1\n1 0.01 16\n1 0\n0 0\n11.4 12.2 7.7\n0.0\n{rows}\n")
    info, src, _, _ = cgfd.source_file(src_path, P.ni, P.nj, P.nk, 0.0, P.dt)
    assert info["stf_given"] == 1 and info["force_on"] == 1 and src["M"] is None
    P.src = src
    sta_path = tmp_path / "sta.list"
    sta_path.write_text("3\na 0 1 4 4 0\nb 0 1 20.5 18.5 3.5\nc 0 0 13 13 10\n")
    _stations(P, sta_path)
    seis, so = _run(P)
    assert rel_l2(seis, so) <= TOL32
