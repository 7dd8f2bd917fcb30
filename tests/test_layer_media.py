"""Media from a layer file in the md3lay format (cgfd_media_layer_file), CPU.

The oracle (oracle/layer_media.py, the reference's float algorithm in numpy)
is pinned by closed forms: a homogeneous layer (mu = vs^2 rho, lam = vp^2 rho
- 2 mu), a power-law gradient v0 + dz^p g at known depths, a tilted interface
(the bilinear interpolation of a plane is exact), the top layer's values above
the first interface, and the reference's on-interface averaging within 1 m.
The product must equal the oracle bit for bit on the same grid and file.
"""
import importlib.util
import os

import numpy as np
import pytest

import cgfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("oracle_layer_media", os.path.join(ROOT, "oracle", "layer_media.py"))
olm = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(olm)


def write_md3lay(path, layers, nx=5, ny=4, minx=-1000.0, miny=-1000.0, dx=1000.0, dy=1000.0, elev=None):
    """layers: list of (elevation, rho, rho_g, rho_p, vp, vp_g, vp_p, vs, vs_g, vs_p), top first;
    elev(l, x, y) overrides the elevation of interface l at mesh node (x, y)."""
    with open(path, "w") as f:
        f.write("# test interface file\n\nelastic_isotropic\n")
        f.write(f"{len(layers)}\n{nx} {ny} {minx} {miny} {dx} {dy}\n")
        for l, row in enumerate(layers):
            for j in range(ny):
                for i in range(nx):
                    r = list(row)
                    if elev is not None:
                        r[0] = elev(l, minx + i * dx, miny + j * dy)
                    f.write(" ".join(repr(float(v)) for v in r) + "\n")
    return str(path)


def _grid(ni=6, nj=5, nk=8, dz=100.0, ztop=0.0):
    return cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, dz, ztop - (nk - 1) * dz)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_homogeneous_layer(tmp_path, impl):
    ni, nj, nk = 6, 5, 8
    x, y, z = _grid(ni, nj, nk)
    path = write_md3lay(tmp_path / "h.md3lay", [(5000.0, 2670.0, 0, 1, 6000.0, 0, 1, 3464.0, 0, 1)])
    f = (lambda: olm.media(ni, nj, nk, x, y, z, path)) if impl == "oracle" else \
        (lambda: cgfd.media_layer_file(ni, nj, nk, x, y, z, path))
    lam, mu, rho = f()
    # the float bilinear interpolation of a constant field rounds: 1e-6 relative
    m = 3464.0 ** 2 * 2670.0
    assert np.allclose(mu, m, rtol=1e-6, atol=0) and np.allclose(rho, 2670.0, rtol=1e-6, atol=0)
    # lam = vp^2 rho - 2 mu cancels about 3x: float rounding shows at ~4e-6 relative
    assert np.allclose(lam, 6000.0 ** 2 * 2670.0 - 2 * m, rtol=2e-5, atol=0)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_gradient_and_two_layers(tmp_path, impl):
    ni, nj, nk = 6, 5, 21
    x, y, z = _grid(ni, nj, nk, dz=100.0, ztop=0.0)  # z from -2000 to 0
    layers = [(0.0, 2500.0, 0.0, 1.0, 4000.0, 0.5, 1.0, 2000.0, 0.0, 1.0),
              (-1000.0, 2800.0, 0.0, 1.0, 6000.0, 2e-3, 2.0, 3400.0, 0.0, 1.0)]
    path = write_md3lay(tmp_path / "g.md3lay", layers)
    f = olm.media if impl == "oracle" else cgfd.media_layer_file
    lam, mu, rho = f(ni, nj, nk, x, y, z, path)
    vs_, rho_ = np.sqrt(mu / rho), rho
    vp_ = np.sqrt((lam + 2 * mu) / rho)
    kz = lambda zz: int(round((zz + 2000.0) / 100.0)) + 3
    # upper layer at 300 m depth: vp = 4000 + 300^1 * 0.5
    assert vp_[kz(-300.0), 4, 4] == pytest.approx(4150.0, rel=1e-6)
    # lower layer, 500 m below its top: vp = 6000 + 500^2 * 2e-3
    assert vp_[kz(-1500.0), 4, 4] == pytest.approx(6500.0, rel=1e-6)
    assert rho_[kz(-1500.0), 4, 4] == pytest.approx(2800.0, rel=1e-6)
    assert vs_[kz(-1500.0), 4, 4] == pytest.approx(3400.0, rel=1e-6)
    # on the interface: the mean of the lower layer's top and the upper layer 1000 m down
    assert vp_[kz(-1000.0), 4, 4] == pytest.approx((6000.0 + 4500.0) / 2, rel=1e-6)
    # the free surface node sits on the top interface but has no layer above: upper layer top values
    assert vp_[kz(0.0), 4, 4] == pytest.approx(4000.0, rel=1e-6)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_within_one_metre_counts_as_on_the_interface(tmp_path, impl):
    """The reference's isEqual binds C's int abs: |dz| < 1 m averages the two layers."""
    ni, nj, nk = 6, 5, 3
    x, y, z = _grid(ni, nj, nk, dz=100.0, ztop=0.0)
    z = z.copy()
    z[3 + 1] = -99.4   # the middle plane, 0.4 m below the second interface
    layers = [(0.0, 2500.0, 0.0, 1.0, 4000.0, 0.0, 1.0, 2000.0, 0.0, 1.0),
              (-99.0, 2800.0, 0.0, 1.0, 6000.0, 0.0, 1.0, 3400.0, 0.0, 1.0)]
    path = write_md3lay(tmp_path / "q.md3lay", layers)
    f = olm.media if impl == "oracle" else cgfd.media_layer_file
    lam, mu, rho = f(ni, nj, nk, x, y, z, path)
    # z = -99.4 lies 0.4 m below the interface at -99: averaged
    assert rho[4, 4, 4] == pytest.approx((2500.0 + 2800.0) / 2, rel=1e-6)
    # z = -200 lies 101 m below: the lower layer alone
    assert rho[3, 4, 4] == pytest.approx(2800.0, rel=1e-6)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_tilted_interface_bilinear(tmp_path, impl):
    ni, nj, nk = 8, 5, 21
    x, y, z = _grid(ni, nj, nk, dz=50.0, ztop=0.0)   # z in [-1000, 0]
    layers = [(0.0, 2500.0, 0, 1, 4000.0, 0, 1, 2000.0, 0, 1), (0.0, 2900.0, 0, 1, 6000.0, 0, 1, 3400.0, 0, 1)]
    # second interface z = -300 - 0.5 x (exact under bilinear interpolation)
    path = write_md3lay(tmp_path / "t.md3lay", layers, elev=lambda l, xx, yy: 0.0 if l == 0 else -300.0 - 0.5 * xx)
    f = olm.media if impl == "oracle" else cgfd.media_layer_file
    lam, mu, rho = f(ni, nj, nk, x, y, z, path)
    zz, xx = z[:, 4, :], x[:, 4, :]
    iface = -300.0 - 0.5 * xx
    r = rho[:, 4, :]
    assert np.allclose(r[zz > iface + 1], 2500.0, rtol=1e-6, atol=0)
    assert np.allclose(r[zz < iface - 1], 2900.0, rtol=1e-6, atol=0)
    assert (zz > iface + 1).sum() > 20 and (zz < iface - 1).sum() > 20


def _topo_file(tmp_path):
    layers = [(2000.0, 2400.0, 0.02, 1.0, 3000.0, 0.3, 1.0, 1700.0, 0.2, 1.0),
              (-1500.0, 2600.0, 0.0, 1.0, 5000.0, 1e-4, 2.0, 2900.0, 0.0, 1.0),
              (-3000.0, 2800.0, 0.01, 1.0, 6500.0, 0.1, 1.0, 3750.0, 0.05, 1.0)]
    wob = lambda l, xx, yy: layers[l][0] + (0.0 if l == 0 else 150.0 * np.sin(xx / 900.0) * np.cos(yy / 1300.0))
    return write_md3lay(tmp_path / "topo.md3lay", layers, nx=20, ny=18, minx=-3000.0, miny=-3000.0, dx=500.0,
                        dy=500.0, elev=wob)


def test_product_equals_oracle_on_topography(tmp_path):
    ni, nj, nk = 14, 12, 10
    x, y, z = cgfd.grid_topography(ni, nj, nk, 250.0, 1.5, 2, (200.0, 500.0), (800.0, 1500.0), 30.0, 500.0, 4)
    path = _topo_file(tmp_path)
    a = olm.media(ni, nj, nk, x, y, z, path)
    b = cgfd.media_layer_file(ni, nj, nk, x, y, z, path)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert len(np.unique(b[2])) > 10  # gradients and three layers are all exercised


def test_errors(tmp_path):
    ni, nj, nk = 6, 5, 4
    x, y, z = _grid(ni, nj, nk)
    bad = write_md3lay(tmp_path / "bad.md3lay", [(5000.0, 2670.0, 0, 1, 4000.0, 0, 1, 3000.0, 0, 1)])
    with pytest.raises(cgfd.CgfdError, match="sqrt"):
        cgfd.media_layer_file(ni, nj, nk, x, y, z, bad)
    small = write_md3lay(tmp_path / "small.md3lay", [(5000.0, 2670.0, 0, 1, 6000.0, 0, 1, 3464.0, 0, 1)],
                         nx=2, ny=2, minx=0.0, miny=0.0, dx=100.0, dy=100.0)
    with pytest.raises(cgfd.CgfdError, match="outside"):
        cgfd.media_layer_file(ni, nj, nk, x, y, z, small)
    good = write_md3lay(tmp_path / "good.md3lay", [(5000.0, 2670.0, 0, 1, 6000.0, 0, 1, 3464.0, 0, 1)])
    with pytest.raises(cgfd.CgfdError, match="method"):
        cgfd.media_layer_file(ni, nj, nk, x, y, z, good, method="geo")
    with pytest.raises(cgfd.CgfdError, match="cannot open"):
        cgfd.media_layer_file(ni, nj, nk, x, y, z, tmp_path / "missing.md3lay")


def _two_flat_layers(tmp_path, z_if):
    layers = [(1000.0, 2500.0, 0.0, 1.0, 4000.0, 0.0, 1.0, 2000.0, 0.0, 1.0),
              (z_if, 2900.0, 0.0, 1.0, 6000.0, 0.0, 1.0, 3400.0, 0.0, 1.0)]
    return write_md3lay(tmp_path / "two.md3lay", layers)


@pytest.mark.parametrize("method", ["ari", "har"])
@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_equivalent_medium_flat_interface(tmp_path, impl, method):
    """Cartesian grid, dz = 100 m, interface at z = -430 m: the node at z = -400
    has its dual cell over [-450, -350] and 9 sample planes every 12.5 m, of
    which 2 (-450, -437.5) lie below the interface: weights 7/9 and 2/9."""
    ni, nj, nk = 6, 5, 10
    x, y, z = _grid(ni, nj, nk, dz=100.0, ztop=0.0)   # z in [-900, 0]
    path = _two_flat_layers(tmp_path, -430.0)
    if impl == "oracle":
        lam, mu, rho = olm.media_equivalent(ni, nj, nk, x, y, z, path, method)
    else:
        lam, mu, rho = cgfd.media_layer_file(ni, nj, nk, x, y, z, path, method=method)
    k = 3 + 5  # z = -400
    ra, rb = 2500.0, 2900.0
    ma, mb = 2000.0 ** 2 * ra, 3400.0 ** 2 * rb
    la, lb = 4000.0 ** 2 * ra - 2 * ma, 6000.0 ** 2 * rb - 2 * mb
    wa, wb = 7 / 9, 2 / 9
    # the 729 sub-point sums accumulate in float, as the reference's: 2e-5 relative
    assert rho[k, 4, 4] == pytest.approx(wa * ra + wb * rb, rel=2e-5)
    if method == "ari":
        assert mu[k, 4, 4] == pytest.approx(wa * ma + wb * mb, rel=2e-5)
        assert lam[k, 4, 4] == pytest.approx(wa * la + wb * lb, rel=5e-5)
    else:
        ka, kb = la + 2 * ma / 3, lb + 2 * mb / 3
        hm = 1 / (wa / ma + wb / mb)
        hk = 1 / (wa / ka + wb / kb)
        assert mu[k, 4, 4] == pytest.approx(hm, rel=2e-5)
        assert lam[k, 4, 4] == pytest.approx(hk - 2 * hm / 3, rel=5e-5)
    # a node far from the interface keeps its own layer's values
    assert rho[3 + 1, 4, 4] == pytest.approx(rb, rel=1e-6) and rho[3 + 8, 4, 4] == pytest.approx(ra, rel=1e-6)


@pytest.mark.parametrize("method", ["ari", "har"])
def test_equivalent_medium_product_equals_oracle(tmp_path, method):
    ni, nj, nk = 8, 7, 8
    x, y, z = cgfd.grid_topography(ni, nj, nk, 250.0, 1.5, 2, (200.0, 500.0), (800.0, 1500.0), 30.0, 500.0, 6)
    path = _topo_file(tmp_path)
    a = olm.media_equivalent(ni, nj, nk, x, y, z, path, method)
    b = cgfd.media_layer_file(ni, nj, nk, x, y, z, path, method=method)
    loc = cgfd.media_layer_file(ni, nj, nk, x, y, z, path)
    assert (b[2] != loc[2]).sum() > 10  # the interfaces cross the grid
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_header_counts_beyond_the_data_are_rejected(tmp_path):
    """An interface file whose header announces far more mesh nodes than it
    holds is refused before anything is sized by those counts."""
    f = tmp_path / "big.md3lay"
    f.write_text("elastic_isotropic\n100000\n100000 100000 0 0 1 1\n" + "1 " * 40 + "\n")
    x, y, z = cgfd.grid_cartesian(8, 8, 8, 100.0, 0.0, 100.0, 0.0, 100.0, -700.0)
    with pytest.raises(cgfd.CgfdError, match="insufficient data"):
        cgfd.media_layer_file(8, 8, 8, x, y, z, f, "loc")
