"""Source files (src_read_locate_file, src/forward/src_t.cu:143-795) read by
cgfd_source_file: the reference's own example files, and synthetic files that
pin each rule against what the mathematics fixes (wavelet closed forms, exact
cubic reproduction of the piecewise Lagrange rule, mu D A double couples) and
against the reference's quirks (force only for component type 3, component
values persisting from the previous source)."""
import os

import numpy as np
import pytest

import cgfd

HERE = os.path.dirname(os.path.abspath(__file__))
NG = 3
STAGE = np.array([0.0, 0.5, 0.5, 1.0])


def _grid(ni, nj, nk, h=100.0):
    return cgfd.grid_cartesian(ni, nj, nk, h, 0.0, h, 0.0, h, -(nk - 1) * h)


def _tables(src, nc):
    key = "F" if nc == 3 else "M"
    return src[key].reshape(nc, src["nsrc"], src["max_nt"], 4)


def _t(src, q, dt, t0=0.0, tstart=0.0):
    ib = src["it_begin"][q]
    t_shift = tstart - (ib * dt + t0)
    r = np.arange(src["max_nt"])[:, None]
    return r * dt + STAGE[None, :] * dt - t_shift


@pytest.mark.parametrize("fname,kind,f0,tc,length,moment,node", [
    ("test_source.src", cgfd.WAVELET_RICKER_DERIV, 2.0, 1.0, 6.0, True, (150, 100, 100)),
    ("prep_source_test_source.src", cgfd.WAVELET_RICKER, 2.0, 0.5, 4.0, False, (100, 110, 100)),
])
def test_reference_example_source_files(fname, kind, f0, tc, length, moment, node):
    ni, nj, nk, dt = 200, 160, 130, 0.005
    info, src, nd, sh = cgfd.source_file(os.path.join(HERE, "golden", fname), ni, nj, nk, 0.0, dt)
    assert info["event"] == "event_3force_ricker" and info["nsrc_file"] == 1 and info["nsrc"] == 1
    assert info["stf_given"] == 0 and info["stf_length"] == length and info["by_coord"] == 0
    assert info["max_nt"] == int(length / dt + 0.5)
    assert (info["moment_on"], info["force_on"]) == ((1, 0) if moment else (0, 1))
    assert tuple(nd[0]) == node and list(sh[0]) == [0.0, 0.0, 0.0]
    assert src["it_begin"][0] == 0 and src["it_end"][0] == info["max_nt"] - 1
    # the spreading is cgfd_source_point's for the same node
    sp = cgfd.source_point(ni, nj, nk, *node, 0, 0.0, dt, 0.0, kind, f0, tc, length)
    assert src["ext_num"][0] == sp["ext_num"] == 343
    assert np.array_equal(src["ext_indx"][:343], sp["ext_indx"]) and np.array_equal(src["ext_coef"][:343],
                                                                                      sp["ext_coef"])
    # tables: the named wavelet at t = step*dt + stage*dt times the components
    # the file gives (three values: Mxx, Myy, Mzz or Fx, Fy, Fz; a moment's
    # missing Myz, Mxz, Mxy stay zero)
    t = _t(src, 0, dt)
    stf = np.vectorize(lambda v: cgfd.wavelet(kind, v, float(np.float32(f0)), float(np.float32(tc))))(t)
    amp = float(np.float32(1.0e16))
    tab = _tables(src, 6 if moment else 3)
    for c in range(3):
        assert np.allclose(tab[c, 0], amp * stf, rtol=1e-12, atol=1e-30)
    if moment:
        assert not tab[3:].any()


def _write(tmp_path, text, name="s.src"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_sources_kept_components_persist_and_coefficients(tmp_path):
    ni, nj, nk, dt = 40, 40, 40, 0.01
    # three moment sources by index with depth; the second lies outside and
    # is dropped; the third gives only four tensor values, so Mxz and Mxy
    # keep the values of the last kept source
    p = _write(tmp_path, "# comment\nev\n3\n0 1.0\n2 0\n0 1\n"
                         "10.2 11.7 5.4\n55 3 3\n20 21 2.0\n"
                         "0.05 gaussian 0.2 0.3\n1 2 3 4 5 6\n"
                         "0.0 ricker 9 9\n7 7 7 7 7 7\n"
                         "0.10 ricker 2.5\n10 20 30 40\n")
    info, src, nd, sh = cgfd.source_file(p, ni, nj, nk, 0.0, dt)
    assert info["nsrc_file"] == 3 and info["nsrc"] == 2 and info["moment_on"] == 1 and info["max_nt"] == 100
    # index input: nearest node, signed shift; depth counted down from nk-1
    assert tuple(nd[0]) == (10, 12, nk - 1 - 5) and np.allclose(sh[0], [0.2, -0.3, -0.4], atol=2e-6)
    assert tuple(nd[1]) == (20, 21, nk - 1 - 2)
    M = _tables(src, 6)
    # source 0: gaussian(c0 = 0.2, c1 = 0.3), starting at 0.05 s = step 5
    assert src["it_begin"][0] == 5
    f = lambda v: float(np.float32(v))   # the file's values as parsed (float)
    t0v = _t(src, 0, dt, tstart=f(0.05))
    g = np.vectorize(lambda v: cgfd.wavelet(cgfd.WAVELET_GAUSSIAN, v, f(0.2), f(0.3)))(t0v)
    for c in range(6):
        assert np.allclose(M[c, 0], (c + 1) * g, rtol=1e-6)
    # source 1 (third in the file): ricker with f0 = 2.5; its second
    # coefficient is missing and keeps the 0.3 of source 0 (the outside
    # source's line is not parsed)
    assert src["it_begin"][1] == 10
    t1 = _t(src, 1, dt, tstart=f(0.10))
    r = np.vectorize(lambda v: cgfd.wavelet(cgfd.WAVELET_RICKER, v, 2.5, f(0.3)))(t1)
    for c, amp in enumerate([10, 20, 30, 40, 5, 6]):
        assert np.allclose(M[c, 1], amp * r, rtol=1e-6, atol=1e-12), c


def test_force_and_moment_activates_force_only(tmp_path):
    p = _write(tmp_path, "ev\n1\n0 0.5\n3 0\n0 0\n5 5 5\n0 ricker 2 0.2\n1 2 3 4 5 6 7 8 9\n")
    info, src, _, _ = cgfd.source_file(p, 12, 12, 12, 0.0, 0.01)
    assert info["cmp_type"] == 3 and info["force_on"] == 1 and info["moment_on"] == 0
    assert src["M"] is None
    F = _tables(src, 3)
    assert np.allclose(F[:, 0, 20, 0], [1, 2, 3])  # t = 0.2: ricker peak 1


def test_mu_d_a_double_couple(tmp_path):
    ni, nj, nk = 16, 16, 16
    mu = np.full((nk + 6, nj + 6, ni + 6), 3.0e10)
    mu[3 + 9, 3 + 9, 3 + 9] = 2.5e10   # at the second source's node, whose file mu is -1
    p = _write(tmp_path, "ev\n2\n0 0.3\n2 1\n0 0\n6 7 8\n9 9 9\n"
                         "0 gaussian 0.05 0.1\n30 60 90 3.2e10 1.5 2.0e6\n"
                         "0 gaussian 0.05 0.1\n45 30 -60 -1 0.5 1.0e6\n")
    info, src, nd, _ = cgfd.source_file(p, ni, nj, nk, 0.0, 0.01, mu=mu)
    assert info["mechanism"] == 1 and info["nsrc"] == 2
    M = _tables(src, 6)
    g = cgfd.wavelet(cgfd.WAVELET_GAUSSIAN, 0.1, 0.05, 0.1)
    for q, (sdr, m0) in enumerate([((30, 60, 90), 3.2e10 * 1.5 * 2.0e6), ((45, 30, -60), 2.5e10 * 0.5 * 1.0e6)]):
        unit = cgfd.angle2moment(*sdr)
        # a double couple: trace zero, eigenvalues -1, 0, 1
        T = np.array([[unit[0], unit[5], unit[4]], [unit[5], unit[1], unit[3]], [unit[4], unit[3], unit[2]]])
        assert np.allclose(np.linalg.eigvalsh(T), [-1, 0, 1], atol=1e-12)
        assert np.allclose(M[:, q, 10, 0], m0 * g * np.asarray(unit), rtol=1e-6), q
    assert tuple(nd[1]) == (9, 9, 9)
    p2 = _write(tmp_path, "ev\n1\n0 0.3\n2 1\n0 0\n6 7 8\n0 gaussian 0.05 0.1\n30 60 90 -1 1 1\n", "n.src")
    info2, src2, _, _ = cgfd.source_file(p2, ni, nj, nk, 0.0, 0.01)   # mu < 0 and no model: no tensor
    assert np.isnan(_tables(src2, 6)).all()


def test_sampled_time_function_reproduces_cubics(tmp_path):
    # force samples of a cubic in time every 0.02 s; the piecewise cubic
    # Lagrange rule reproduces it at every step and stage, at the ends too
    nt, sdt, dt = 11, 0.02, 0.005
    ts = np.arange(nt) * sdt
    cub = [lambda t: 1 + 2 * t - 3 * t ** 2 + 40 * t ** 3, lambda t: -5 * t ** 3 + t, lambda t: 0.5 + 0 * t]
    lines = "\n".join(" ".join(f"{np.float32(f(t)):.9g}" for f in cub) for t in ts)
    p = _write(tmp_path, f"ev\n1\n1 {sdt} {nt}\n1 0\n0 0\n4 4 4\n0.0\n{lines}\n")
    info, src, _, _ = cgfd.source_file(p, 10, 10, 10, 0.0, dt)
    assert info["stf_given"] == 1 and info["max_nt"] == int((nt - 1) * sdt / dt + 0.5) + 1
    F = _tables(src, 3)
    t = _t(src, 0, dt)
    vals = np.float64(np.float32(np.array([f(ts) for f in cub])))   # the samples as the file holds them
    for c in range(3):
        # the exact cubic through the float32 samples (they are not an exact cubic any more)
        coef = np.polyfit(ts, vals[c], 3)
        assert np.allclose(F[c, 0], np.polyval(coef, t), rtol=1e-6, atol=1e-6), c


def test_coordinates_and_depth(tmp_path):
    ni, nj, nk, h = 30, 30, 30, 100.0
    x, y, z = _grid(ni, nj, nk, h)
    p = _write(tmp_path, "ev\n2\n0 0.2\n1 0\n1 1\n1234 876 420\n5000 100 100\n0 ricker 3 0.1\n1 1 1\n"
                         "0 ricker 3 0.1\n1 1 1\n")
    info, src, nd, sh = cgfd.source_file(p, ni, nj, nk, 0.0, 0.01, x, y, z)
    assert info["by_coord"] == 1 and info["depth"] == 1 and info["nsrc"] == 1   # x = 5000 m is outside
    g = np.array([1234.0 / h, 876.0 / h, ((nk - 1) * h - 420.0) / h])
    m = np.rint(g).astype(int)
    assert tuple(nd[0]) == tuple(m)
    assert np.all(np.abs(sh[0] - (g - m)) <= 0.0100001)
    with pytest.raises(cgfd.CgfdError, match="grid coordinates"):
        cgfd.source_file(p, ni, nj, nk, 0.0, 0.01)


def test_bad_source_files(tmp_path):
    with pytest.raises(cgfd.CgfdError, match="cannot open"):
        cgfd.source_file(tmp_path / "none.src", 10, 10, 10, 0.0, 0.01)
    with pytest.raises(cgfd.CgfdError, match="positive"):
        cgfd.source_file(_write(tmp_path, "ev\n0\n0 1\n1 0\n0 0\n", "a.src"), 10, 10, 10, 0.0, 0.01)
    with pytest.raises(cgfd.CgfdError, match="not implemented"):
        cgfd.source_file(_write(tmp_path, "ev\n1\n0 1\n1 0\n0 0\n5 5 5\n0 sinc 1 1\n1 1 1\n", "b.src"),
                         10, 10, 10, 0.0, 0.01)


def test_out_of_range_values_are_rejected_or_dropped(tmp_path):
    """Counts, lengths and positions far beyond the grid or the int range
    give an error or an outside source, never a huge allocation or an
    out-of-range integer conversion."""
    def w(name, text):
        return _write(tmp_path, text, name)
    # a source count far beyond the lines in the file
    with pytest.raises(cgfd.CgfdError, match="location line 2"):
        cgfd.source_file(w("n.src", "ev\n2000000000\n0 1\n1 0\n0 0\n5 5 5\n"), 10, 10, 10, 0.0, 0.01)
    # sampled time function announcing far more samples than the file holds
    with pytest.raises(cgfd.CgfdError, match="samples of source 1"):
        cgfd.source_file(w("s.src", "ev\n1\n1 0.01 2000000000\n1 0\n0 0\n5 5 5\n0.0\n1 1 1\n"), 10, 10, 10, 0.0, 1.0)
    # no step long, or more than 2^30 steps
    for length in ("0", "-3", "1e30", "nan"):
        with pytest.raises(cgfd.CgfdError, match="at least one step"):
            cgfd.source_file(w("l.src", f"ev\n1\n0 {length}\n1 0\n0 0\n5 5 5\n0 ricker 2 0.2\n1 1 1\n"), 10, 10, 10,
                             0.0, 0.01)
    # start time beyond 2^30 steps
    with pytest.raises(cgfd.CgfdError, match="out of range"):
        cgfd.source_file(w("t.src", "ev\n1\n0 1\n1 0\n0 0\n5 5 5\n1e30 ricker 2 0.2\n1 1 1\n"), 10, 10, 10, 0.0, 0.01)
    # positions far outside (or not a number): the source is dropped
    info, src, _, _ = cgfd.source_file(w("p.src", "ev\n3\n0 1\n1 0\n0 0\n1e30 5 5\n5 -1e30 5\n5 5 nan\n"
                                                  "0 ricker 2 0.2\n1 1 1\n0 ricker 2 0.2\n1 1 1\n0 ricker 2 0.2\n1 1 1\n"),
                                       10, 10, 10, 0.0, 0.01)
    assert info["nsrc_file"] == 3 and info["nsrc"] == 0 and src["nsrc"] == 0
