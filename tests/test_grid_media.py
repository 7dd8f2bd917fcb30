"""Media from a grid media file in the md3grd format (cgfd_media_grid_file), CPU.

The oracle (oracle/grid_media.py, the reference's float algorithm in numpy)
is pinned by closed forms: values linear in z between two surfaces of a
layer, the first surface's values above it and the last surface's below,
the mean of the two surfaces at a layer boundary (and within 1 m under it,
the reference's isEqual), a tilted mesh read through a window of a larger
mesh (bilinear interpolation of a plane is exact), and the equivalent-medium
weights of a flat boundary (7/9 and 2/9 of the sub-point planes), including
the sub-points on a boundary that the grid rules leave out.  The product must
equal the oracle bit for bit on the same grid and file.
"""
import importlib.util
import os

import numpy as np
import pytest

import cgfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("oracle_grid_media", os.path.join(ROOT, "oracle", "grid_media.py"))
ogm = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(ogm)


def write_md3grd(path, layers, nx=5, ny=4, minx=-1000.0, miny=-1000.0, dx=1000.0, dy=1000.0, elev=None):
    """layers: list of layers, each a list of surfaces (elevation, rho, vp, vs), top first;
    elev(s, x, y) overrides the elevation of surface s (counted over all layers) at mesh node (x, y).
    Values are written as float32 numbers so that every reader parses them alike."""
    w = lambda v: repr(float(np.float32(v)))
    with open(path, "w") as f:
        f.write("# test grid media file\nelastic_isotropic\n")
        f.write(f"{len(layers)}\n" + " ".join(str(len(L)) for L in layers) + "\n")
        f.write(f"{nx} {ny} {minx} {miny} {dx} {dy}\n")
        s = 0
        for L in layers:
            for row in L:
                for j in range(ny):
                    for i in range(nx):
                        r = list(row)
                        if elev is not None:
                            r[0] = elev(s, minx + i * dx, miny + j * dy)
                        f.write(" ".join(w(v) for v in r) + "\n")
                s += 1
    return str(path)


def _grid(ni=6, nj=5, nk=10, dz=100.0, ztop=0.0):
    return cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, dz, ztop - (nk - 1) * dz)


def _run(impl, ni, nj, nk, x, y, z, path, method="loc"):
    if impl == "oracle":
        return ogm.media(ni, nj, nk, x, y, z, path, method)
    return cgfd.media_grid_file(ni, nj, nk, x, y, z, path, method=method)


def _velocities(lam, mu, rho):
    return np.sqrt((lam + 2 * mu) / rho), np.sqrt(mu / rho), rho


LAYER_A = [(500.0, 2000.0, 4000.0, 2000.0), (-400.0, 2200.0, 4400.0, 2300.0)]
LAYER_B = [(-400.0, 2600.0, 5500.0, 3000.0), (-2000.0, 2800.0, 6000.0, 3400.0)]


def _lin(s0, s1, zz):
    t = (zz - s0[0]) / (s1[0] - s0[0])
    return [a + (b - a) * t for a, b in zip(s0[1:], s1[1:])]


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_linear_between_surfaces_and_mean_at_the_boundary(tmp_path, impl):
    ni, nj, nk = 6, 5, 10
    x, y, z = _grid(ni, nj, nk)  # z in [-900, 0], ghosts to -1200 and 300
    path = write_md3grd(tmp_path / "ab.md3grd", [LAYER_A, LAYER_B])
    lam, mu, rho = _run(impl, ni, nj, nk, x, y, z, path)
    vp, vs, r = _velocities(lam, mu, rho)
    for k in range(nk + 6):
        zz = z[k, 4, 4]
        if zz > -400.0:
            want = _lin(*LAYER_A, zz)
        elif zz == -400.0:
            want = [(a + b) / 2 for a, b in zip(LAYER_A[1][1:], LAYER_B[0][1:])]
        else:
            want = _lin(*LAYER_B, zz)
        assert r[k, 4, 4] == pytest.approx(want[0], rel=2e-6)
        assert vp[k, 4, 4] == pytest.approx(want[1], rel=2e-5)
        assert vs[k, 4, 4] == pytest.approx(want[2], rel=2e-6)
    # laterally uniform model: every column equals the checked one
    assert np.allclose(rho, rho[:, 4:5, 4:5], rtol=1e-6, atol=0)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_within_one_metre_under_a_boundary(tmp_path, impl):
    ni, nj, nk = 6, 5, 10
    path = write_md3grd(tmp_path / "ab.md3grd", [LAYER_A, LAYER_B])
    mean = [(a + b) / 2 for a, b in zip(LAYER_A[1][1:], LAYER_B[0][1:])]
    # 0.5 m under the boundary counts as on it; 0.5 m above it does not
    for ztop, want in ((-0.5, mean), (0.5, _lin(*LAYER_A, -399.5))):
        x, y, z = _grid(ni, nj, nk, ztop=ztop)
        k = int(np.argmin(np.abs(z[:, 4, 4] + 400.0)))
        assert abs(z[k, 4, 4] + 400.0) == 0.5
        vp, vs, r = _velocities(*_run(impl, ni, nj, nk, x, y, z, path))
        assert r[k, 4, 4] == pytest.approx(want[0], rel=2e-6)
        assert vs[k, 4, 4] == pytest.approx(want[2], rel=2e-6)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_above_first_and_below_last_surface(tmp_path, impl):
    ni, nj, nk = 6, 5, 10
    x, y, z = _grid(ni, nj, nk)
    surf = [(100.0, 2000.0, 4000.0, 2000.0), (-500.0, 2300.0, 4600.0, 2500.0), (-1000.0, 2500.0, 5000.0, 2800.0)]
    path = write_md3grd(tmp_path / "one.md3grd", [surf])
    vp, vs, r = _velocities(*_run(impl, ni, nj, nk, x, y, z, path))
    for k in range(nk + 6):
        zz = z[k, 2, 3]
        if zz >= 100.0:
            want = surf[0][1:]
        elif zz <= -1000.0:
            want = surf[2][1:]
        elif zz >= -500.0:
            want = _lin(surf[0], surf[1], zz)
        else:
            want = _lin(surf[1], surf[2], zz)
        assert r[k, 2, 3] == pytest.approx(want[0], rel=2e-6)
        assert vp[k, 2, 3] == pytest.approx(want[1], rel=2e-5)
        assert vs[k, 2, 3] == pytest.approx(want[2], rel=2e-6)


def _tilted(s, xx, yy):
    base = (300.0, -450.0, -450.0, -1800.0)[s]
    return base + (0.08 * xx - 0.05 * yy if s in (1, 2) else 0.0)


def test_window_of_a_larger_mesh(tmp_path):
    """A tilted boundary read from a mesh just covering the grid and from a
    larger mesh whose window starts elsewhere: the same medium, and the
    product equals the oracle on both."""
    ni, nj, nk = 8, 7, 10
    x, y, z = _grid(ni, nj, nk)
    layers = [[(0, 2000.0, 4000.0, 2000.0), (0, 2200.0, 4400.0, 2300.0)],
              [(0, 2600.0, 5500.0, 3000.0), (0, 2800.0, 6000.0, 3400.0)]]
    tight = write_md3grd(tmp_path / "tight.md3grd", layers, nx=3, ny=3, minx=-500.0, miny=-500.0, dx=800.0,
                         dy=800.0, elev=_tilted)
    # window of the larger mesh: nodes from (-450, -650); the grid (x to 1000, y to 900) stays off its last
    # row and column, where the reference's bilinear interpolation returns a node value
    big = write_md3grd(tmp_path / "big.md3grd", layers, nx=14, ny=12, minx=-5250.0, miny=-4650.0, dx=800.0,
                       dy=800.0, elev=_tilted)
    a = cgfd.media_grid_file(ni, nj, nk, x, y, z, tight)
    b = cgfd.media_grid_file(ni, nj, nk, x, y, z, big)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=2e-5, atol=0)
    assert len(np.unique(np.round(b[2], 3))) > 20  # the tilted boundary crosses the grid
    for path in (tight, big):
        for u, v in zip(ogm.media(ni, nj, nk, x, y, z, path), cgfd.media_grid_file(ni, nj, nk, x, y, z, path)):
            assert np.array_equal(u, v)


def test_errors(tmp_path):
    ni, nj, nk = 6, 5, 4
    x, y, z = _grid(ni, nj, nk)
    cases = {
        "cover": write_md3grd(tmp_path / "small.md3grd", [LAYER_A, LAYER_B], nx=2, ny=2, minx=0.0, miny=0.0,
                              dx=100.0, dy=100.0),
        "differs": write_md3grd(tmp_path / "gap.md3grd", [LAYER_A, [(-410.0,) + LAYER_B[0][1:], LAYER_B[1]]]),
        "sqrt": write_md3grd(tmp_path / "slow.md3grd", [[(500.0, 2000.0, 4000.0, 2000.0),
                                                         (-600.0, 2000.0, 4000.0, 3000.0)]]),
    }
    for what, path in cases.items():
        with pytest.raises(cgfd.CgfdError, match=what):
            cgfd.media_grid_file(ni, nj, nk, x, y, z, path)
        with pytest.raises(ValueError):
            ogm.media(ni, nj, nk, x, y, z, path)
    good = write_md3grd(tmp_path / "good.md3grd", [LAYER_A, LAYER_B])
    with pytest.raises(cgfd.CgfdError, match="method"):
        cgfd.media_grid_file(ni, nj, nk, x, y, z, good, method="geo")
    with pytest.raises(cgfd.CgfdError, match="cannot open"):
        cgfd.media_grid_file(ni, nj, nk, x, y, z, tmp_path / "missing.md3grd")


RA, PA, SA = 2500.0, 4000.0, 2000.0
RB, PB, SB = 2900.0, 6000.0, 3400.0


def _two_flat(tmp_path, z_b):
    return write_md3grd(tmp_path / "two.md3grd", [[(1000.0, RA, PA, SA), (z_b, RA, PA, SA)],
                                                  [(z_b, RB, PB, SB), (-3000.0, RB, PB, SB)]])


def _expect(method, wa, wb):
    ma, mb = SA ** 2 * RA, SB ** 2 * RB
    la, lb = PA ** 2 * RA - 2 * ma, PB ** 2 * RB - 2 * mb
    rho = wa * RA + wb * RB
    if method == "ari":
        return wa * la + wb * lb, wa * ma + wb * mb, rho
    ka, kb = la + 2 * ma / 3, lb + 2 * mb / 3
    hm, hk = 1 / (wa / ma + wb / mb), 1 / (wa / ka + wb / kb)
    return hk - 2 * hm / 3, hm, rho


@pytest.mark.parametrize("method", ["ari", "har"])
@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_equivalent_medium_flat_boundary(tmp_path, impl, method):
    """Cartesian grid, dz = 100 m, boundary at z = -330 m: the node at z = -300
    has its dual cell over [-350, -250] and 9 sub-point planes every 12.5 m,
    of which 2 (-350, -337.5) lie in the lower layer: weights 7/9 and 2/9."""
    ni, nj, nk = 6, 5, 10
    x, y, z = _grid(ni, nj, nk)
    lam, mu, rho = _run(impl, ni, nj, nk, x, y, z, _two_flat(tmp_path, -330.0), method)
    k = 3 + 6  # z = -300
    want = _expect(method, 7 / 9, 2 / 9)
    # the 729 sub-point sums accumulate in float, as the reference's: 2e-5 relative
    assert rho[k, 3, 3] == pytest.approx(want[2], rel=2e-5)
    assert mu[k, 3, 3] == pytest.approx(want[1], rel=2e-5)
    assert lam[k, 3, 3] == pytest.approx(want[0], rel=5e-5)
    # nodes away from the boundary keep their layer's values
    assert rho[3 + 5, 3, 3] == pytest.approx(RB, rel=1e-6) and rho[3 + 8, 3, 3] == pytest.approx(RA, rel=1e-6)


@pytest.mark.parametrize("method", ["ari", "har"])
@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_equivalent_medium_leaves_out_boundary_points(tmp_path, impl, method):
    """Boundary at z = -349.5 m: the node at z = -300 has a mixed dual cell,
    but its lowest sub-point plane (-350) lies within 1 m under the boundary
    and is left out; the other 8 planes are in the upper layer, so the
    average is the upper layer's medium (counting that plane would pull rho
    up by about 22)."""
    ni, nj, nk = 6, 5, 10
    x, y, z = _grid(ni, nj, nk)
    lam, mu, rho = _run(impl, ni, nj, nk, x, y, z, _two_flat(tmp_path, -349.5), method)
    k = 3 + 6
    want = _expect(method, 1.0, 0.0)
    assert rho[k, 3, 3] == pytest.approx(want[2], rel=2e-5)
    assert mu[k, 3, 3] == pytest.approx(want[1], rel=2e-5)
    assert lam[k, 3, 3] == pytest.approx(want[0], rel=5e-5)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_one_layer_equivalent_is_loc(tmp_path, impl):
    ni, nj, nk = 6, 5, 10
    x, y, z = _grid(ni, nj, nk)
    surf = [(100.0, 2000.0, 4000.0, 2000.0), (-500.0, 2300.0, 4600.0, 2500.0), (-1000.0, 2500.0, 5000.0, 2800.0)]
    path = write_md3grd(tmp_path / "one.md3grd", [surf])
    a = _run(impl, ni, nj, nk, x, y, z, path)
    for method in ("har", "ari"):
        for u, v in zip(a, _run(impl, ni, nj, nk, x, y, z, path, method)):
            assert np.array_equal(u, v)


def _topo_file(tmp_path):
    wob = lambda xx, yy: 150.0 * np.sin(xx / 900.0) * np.cos(yy / 1300.0)
    layers = [[(2000.0, 2400.0, 3000.0, 1700.0), (-1500.0, 2500.0, 3400.0, 1900.0)],
              [(-1500.0, 2600.0, 5000.0, 2900.0), (-2200.0, 2650.0, 5200.0, 3000.0), (-3000.0, 2700.0, 5600.0, 3200.0)],
              [(-3000.0, 2800.0, 6500.0, 3750.0), (-6000.0, 3000.0, 7000.0, 4000.0)]]
    base = [2000.0, -1500.0, -1500.0, -2200.0, -3000.0, -3000.0, -6000.0]
    amp = [0.0, 1.0, 1.0, 0.5, 1.0, 1.0, 0.0]
    elev = lambda s, xx, yy: base[s] + amp[s] * wob(xx, yy)
    return write_md3grd(tmp_path / "topo.md3grd", layers, nx=20, ny=18, minx=-3000.0, miny=-3000.0, dx=500.0,
                        dy=500.0, elev=elev)


@pytest.mark.parametrize("method", ["loc", "ari", "har"])
def test_product_equals_oracle_on_topography(tmp_path, method):
    ni, nj, nk = (14, 12, 10) if method == "loc" else (8, 7, 8)
    x, y, z = cgfd.grid_topography(ni, nj, nk, 250.0, 1.5, 2, (200.0, 500.0), (800.0, 1500.0), 30.0, 500.0, 4)
    path = _topo_file(tmp_path)
    a = ogm.media(ni, nj, nk, x, y, z, path, method)
    b = cgfd.media_grid_file(ni, nj, nk, x, y, z, path, method=method)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    if method == "loc":
        assert len(np.unique(b[2])) > 50  # interpolation across surfaces and layers
    else:
        loc = cgfd.media_grid_file(ni, nj, nk, x, y, z, path)
        assert (b[2] != loc[2]).sum() > 10  # the boundaries cross the grid


def test_non_finite_and_overflowing_mesh_values_are_errors(tmp_path):
    """A mesh spacing so large that the bilinear weights overflow float, and
    values that are not finite numbers, give an error instead of a layer
    index outside the file (found by fuzzing the reader)."""
    ni, nj, nk = 20, 18, 16
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -(nk - 1) * 100.0)
    surf = [(100, 2000, 3000, 1700), (-500, 2100, 3200, 1800), (-500, 2500, 5000, 2900), (-2000, 2600, 5200, 3000)]

    def grd(mesh, bad=None):
        rows = ["elastic_isotropic", "2", "2 2", mesh]
        for s_, v in enumerate(surf):
            for q in range(9):
                rows.append(" ".join(str(a) for a in v) if (s_, q) != bad else "nan 2000 3000 1700")
        p = tmp_path / "m.md3grd"
        p.write_text("\n".join(rows) + "\n")
        return p

    cgfd.media_grid_file(ni, nj, nk, x, y, z, grd("3 3 -1000 -1000 2000 2000"), "loc")   # the good file reads
    for method in ("loc", "har", "ari"):
        with pytest.raises(cgfd.CgfdError, match="not finite"):
            cgfd.media_grid_file(ni, nj, nk, x, y, z, grd("3 3 -1000 -1000 20001e30 2000"), method)
    with pytest.raises(cgfd.CgfdError, match="not a finite number"):
        cgfd.media_grid_file(ni, nj, nk, x, y, z, grd("3 3 -1000 -1000 2000 2000", bad=(1, 4)), "loc")
    with pytest.raises(cgfd.CgfdError, match="DX, DY > 0"):
        cgfd.media_grid_file(ni, nj, nk, x, y, z, grd("3 3 -1000 -1000 0 2000"), "loc")
