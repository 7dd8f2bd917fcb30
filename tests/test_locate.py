"""Points given by physical coordinates (cgfd_locate_points) on the CPU.

The oracle (oracle/locate.py, a numpy restatement of the reference's float
algorithm) is pinned by closed forms: exact fractions on Cartesian and
sheared (affine) grids, where the trilinear map is exact; depth below a flat
free surface; and the trilinear round trip on a topography grid (a point
built at a fractional index is found again within the 0.01 sampling step).
The product must then equal the oracle exactly on the same inputs.
"""
import importlib.util
import os

import numpy as np
import pytest

import cgfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("oracle_locate", os.path.join(ROOT, "oracle", "locate.py"))
oloc = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(oloc)


def _sheared(ni, nj, nk):
    k, j, i = np.meshgrid(np.arange(nk + 6) - 3.0, np.arange(nj + 6) - 3.0, np.arange(ni + 6) - 3.0, indexing="ij")
    x = 1000.0 + 100.0 * i + 20.0 * j + 10.0 * k
    y = -500.0 + 90.0 * j + 15.0 * k
    z = -4000.0 + 80.0 * k + 5.0 * i - 8.0 * j
    return x, y, z


def _at(x, y, z, ijk):
    """Physical coordinates at a fractional physical index by trilinear
    interpolation of the node coordinates (float64)."""
    i, j, k = (float(v) + 3.0 for v in ijk)
    i0, j0, k0 = int(np.floor(i)), int(np.floor(j)), int(np.floor(k))
    a, b, c = i - i0, j - j0, k - k0
    out = []
    for A in (x, y, z):
        v = 0.0
        for e in range(8):
            w = (a if e & 1 else 1 - a) * (b if e & 2 else 1 - b) * (c if e & 4 else 1 - c)
            v += w * A[k0 + (e >> 2), j0 + ((e >> 1) & 1), i0 + (e & 1)]
        out.append(v)
    return out


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_cartesian_exact_fractions(impl):
    ni, nj, nk = 14, 12, 10
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 80.0, 0.0, 50.0, -450.0)
    want = np.array([[3.37, 4.63, 2.25], [7.0, 5.5, 8.99], [10.01, 2.5, 6.5]])
    pts = np.array([[w[0] * 100.0, w[1] * 80.0, -450.0 + w[2] * 50.0] for w in want])
    f = oloc.locate if impl == "oracle" else cgfd.locate_points
    idx, frac, found = f(ni, nj, nk, x, y, z, pts)
    assert found.all()
    assert np.allclose(idx + frac, want, atol=2e-5)
    assert ((frac >= 0) & (frac < 1)).all()


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_depth_below_flat_surface(impl):
    ni, nj, nk = 12, 10, 16
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -1500.0)
    pts = np.array([[512.0, 433.0, 237.0], [250.0, 250.0, 1000.0]])
    f = oloc.locate if impl == "oracle" else cgfd.locate_points
    idx, frac, found = f(ni, nj, nk, x, y, z, pts, depth_flag=[1, 1])
    assert found.all()
    want = np.array([[5.12, 4.33, 15 - 2.37], [2.5, 2.5, 5.0]])
    assert np.allclose(idx + frac, want, atol=2e-5)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_sheared_grid_exact_inverse(impl):
    ni, nj, nk = 10, 9, 8
    x, y, z = _sheared(ni, nj, nk)
    want = np.array([[2.41, 3.07, 4.5], [6.88, 1.12, 2.03], [4.0, 5.0, 3.0]])
    pts = np.array([_at(x, y, z, w) for w in want])
    f = oloc.locate if impl == "oracle" else cgfd.locate_points
    idx, frac, found = f(ni, nj, nk, x, y, z, pts)
    assert found.all()
    assert np.allclose(idx + frac, want, atol=2e-5)


def _topo(ni, nj, nk, seed=5):
    return cgfd.grid_topography(ni, nj, nk, 100.0, 1.5, 3, (300.0, 800.0), (800.0, 2000.0), 50.0, 500.0, seed)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_topography_round_trip(impl):
    ni, nj, nk = 24, 20, 14
    x, y, z = _topo(ni, nj, nk)
    rng = np.random.default_rng(3)
    want = np.column_stack([rng.uniform(1, ni - 2, 12), rng.uniform(1, nj - 2, 12), rng.uniform(1, nk - 2, 12)])
    pts = np.array([_at(x, y, z, w) for w in want])
    f = oloc.locate if impl == "oracle" else cgfd.locate_points
    idx, frac, found = f(ni, nj, nk, x, y, z, pts)
    assert found.all()
    # the sampling step is 0.01 of a cell along each index direction
    assert np.abs(idx + frac - want).max() <= 0.011


def test_outside_points_not_found():
    ni, nj, nk = 12, 10, 8
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -700.0)
    pts = np.array([[-5000.0, 100.0, -300.0], [300.0, 300.0, 900.0], [300.0, 300.0, -300.0]])
    for f in (oloc.locate, cgfd.locate_points):
        idx, frac, found = f(ni, nj, nk, x, y, z, pts)
        assert list(found) == [False, False, True]
        assert (idx[:2] == -1000).all()


@pytest.mark.parametrize("grid", ["cartesian", "sheared", "topography"])
def test_product_equals_oracle(grid):
    ni, nj, nk = 18, 16, 12
    if grid == "cartesian":
        x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -1100.0)
    elif grid == "sheared":
        x, y, z = _sheared(ni, nj, nk)
    else:
        x, y, z = _topo(ni, nj, nk, seed=9)
    rng = np.random.default_rng(11)
    # random points inside and around the grid, half given by depth below the surface
    xs = rng.uniform(x.min() - 100, x.max() + 100, 24)
    ys = rng.uniform(y.min() - 100, y.max() + 100, 24)
    depth = rng.random(24) < 0.5
    zs = np.where(depth, rng.uniform(0, 1200, 24), rng.uniform(z.min() - 100, z.max() + 100, 24))
    pts = np.column_stack([xs, ys, zs])
    a = oloc.locate(ni, nj, nk, x, y, z, pts, depth_flag=depth.astype(int))
    b = cgfd.locate_points(ni, nj, nk, x, y, z, pts, depth_flag=depth.astype(int))
    assert a[2].any() and not a[2].all()
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])


def test_product_equals_oracle_on_a_larger_grid():
    """The product's nearest-node search prunes blocks; on a grid large enough
    for that to matter it must still pick the oracle's brute-force node."""
    ni, nj, nk = 64, 56, 40
    x, y, z = _topo(ni, nj, nk, seed=21)
    rng = np.random.default_rng(12)
    want = np.column_stack([rng.uniform(0, ni - 1, 10), rng.uniform(0, nj - 1, 10), rng.uniform(0, nk - 1, 10)])
    pts = np.array([_at(x, y, z, w) for w in want])
    a = oloc.locate(ni, nj, nk, x, y, z, pts)
    b = cgfd.locate_points(ni, nj, nk, x, y, z, pts)
    assert a[2].all() and b[2].all()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.abs(b[0] + b[1] - want).max() <= 0.011
