"""Station files (io_recv_read_locate, src/forward/io_funcs.cu:24-299) read by
cgfd_station_file: the reference's own example station list, comments and
blank lines, stations by coordinates and depth, stations outside the grid."""
import os

import numpy as np
import pytest

import cgfd

HERE = os.path.dirname(os.path.abspath(__file__))
NG = 3


def _flat(ni, nj, nk, h=100.0):
    # Cartesian grid, x0 = y0 = 0, top surface at z = 0
    return cgfd.grid_cartesian(ni, nj, nk, h, 0.0, h, 0.0, h, -(nk - 1) * h)


def _p(ni, nj, i, j, k):
    nx, ny = ni + 2 * NG, nj + 2 * NG
    return (i + NG) + (j + NG) * nx + (k + NG) * nx * ny


def test_reference_example_station_list():
    ni, nj, nk = 160, 170, 120
    x, y, z = _flat(ni, nj, nk)
    st = cgfd.station_file(os.path.join(HERE, "golden", "station.list"), ni, nj, nk, x, y, z)
    # "recv0_x_y_z 0 1 100.0 100.0 0.0" etc.: grid index, 1-based x and y,
    # the third value a depth below the surface in grid points
    want = [("recv0_x_y_z", 100, 100, 0), ("recv1_x_y_z", 100, 150, 100), ("recv2_x_y_z", 100, 100, 50)]
    assert [s["name"] for s in st] == [w[0] for w in want]
    nx, ny = ni + 2 * NG, nj + 2 * NG
    for s, (_, ix, iy, dep) in zip(st, want):
        assert s["by_coord"] == 0 and s["depth"] == 1 and s["inside"]
        i, j, k = ix - 1, iy - 1, nk - 1 - dep
        p = _p(ni, nj, i, j, k)
        assert list(s["indx8"]) == [p, p + 1, p + nx, p + 1 + nx, p + nx * ny, p + 1 + nx * ny,
                                    p + nx + nx * ny, p + 1 + nx + nx * ny]
        assert list(s["frac"]) == [0.0, 0.0, 0.0]
        # the station's coordinates: the grid node at its lower corner
        assert list(s["xyz"]) == [100.0 * i, 100.0 * j, 100.0 * (k - (nk - 1))]


def test_comments_blank_lines_coordinates_and_outside(tmp_path):
    ni, nj, nk = 40, 36, 30
    h = 100.0
    x, y, z = _flat(ni, nj, nk, h)
    f = tmp_path / "sta.list"
    f.write_text("# a comment\n\n4\n"
                 "# between stations\n"
                 "a 1 1 1234.0 2050.0 350.0\n"      # by coordinates, depth 350 m below the surface
                 "b 1 0 777.0 888.0 -1020.0\n"      # by coordinates, z
                 "c 0 0 5.5 7.25 3.0\n"             # by grid index, fractional, 1-based z
                 "far 0 1 500.0 2.0 0.0\n")         # outside in x
    st = cgfd.station_file(f, ni, nj, nk, x, y, z)
    assert [s["name"] for s in st] == ["a", "b", "c", "far"]
    a, b, c, far = st
    # coordinate stations: the cell holding the point, fractions on the 0.01 grid
    for s, pt in ((a, (1234.0, 2050.0, -350.0)), (b, (777.0, 888.0, -1020.0))):
        assert s["inside"] and s["by_coord"] == 1
        g = np.array([pt[0] / h, pt[1] / h, (pt[2] + (nk - 1) * h) / h])
        lo = np.floor(g).astype(int)
        p = _p(ni, nj, *lo)
        assert s["indx8"][0] == p
        assert np.all(np.abs(s["frac"] - (g - lo)) <= 0.0100001)
        assert np.allclose(s["xyz"], pt)
    # grid index: 1-based, floor and fraction
    assert c["inside"] and list(c["frac"]) == [0.5, 0.25, 0.0]
    assert c["indx8"][0] == _p(ni, nj, 4, 6, 2)
    assert not far["inside"]
    # the same coordinate points through the locator directly
    idx, fr, found = cgfd.locate_points(ni, nj, nk, x, y, z, [[1234.0, 2050.0, 350.0], [777.0, 888.0, -1020.0]],
                                        depth_flag=[1, 0])
    assert found.all()
    assert a["indx8"][0] == _p(ni, nj, *idx[0]) and b["indx8"][0] == _p(ni, nj, *idx[1])
    assert np.array_equal(a["frac"], fr[0]) and np.array_equal(b["frac"], fr[1])


def test_bad_station_files(tmp_path):
    ni, nj, nk = 20, 20, 20
    x, y, z = _flat(ni, nj, nk)
    with pytest.raises(cgfd.CgfdError):
        cgfd.station_file(tmp_path / "missing.list", ni, nj, nk, x, y, z)
    f = tmp_path / "short.list"
    f.write_text("2\na 0 0 1 1 1\nb 0 0 1\n")
    with pytest.raises(cgfd.CgfdError, match="line 2"):
        cgfd.station_file(f, ni, nj, nk, x, y, z)
    g = tmp_path / "coords.list"
    g.write_text("1\na 1 0 100 100 -100\n")
    with pytest.raises(cgfd.CgfdError, match="grid coordinates"):
        cgfd.station_file(g, ni, nj, nk)


def test_far_out_stations_are_outside(tmp_path):
    ni, nj, nk = 20, 20, 20
    x, y, z = _flat(ni, nj, nk)
    f = tmp_path / "far.list"
    f.write_text("4\na 0 0 1e30 5 5\nb 0 1 5 -1e30 0\nc 1 1 1e30 100 0\nd 0 0 nan 5 5\n")
    st = cgfd.station_file(f, ni, nj, nk, x, y, z)
    assert [s["inside"] for s in st] == [False] * 4
