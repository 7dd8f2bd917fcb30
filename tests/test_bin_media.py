"""Media from raw binary volumes (cgfd_media_bin_file), CPU.

The oracle (oracle/bin_media.py, the reference's float algorithm in numpy) is
pinned by closed forms: trilinear interpolation reproduces fields linear in
x, y and z at any point inside the volume, whatever the file's dimension
order or the sign of the z spacing; a point outside the volume takes the
edge node's value along that axis; and the three-file layout read in any
order gives the same medium bit for bit.  The product must equal the oracle
bit for bit on the same grid and files.
"""
import importlib.util
import os

import numpy as np
import pytest

import cgfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("oracle_bin_media", os.path.join(ROOT, "oracle", "bin_media.py"))
obm = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(obm)

RHO = lambda x, y, z: 2000.0 + 0.1 * x + 0.05 * y - 0.3 * z
VP = lambda x, y, z: 4000.0 + 0.2 * x - 0.5 * z
VS = lambda x, y, z: 2000.0 + 0.1 * y - 0.2 * z


def write_bin(tmp_path, order, axes, fields=(RHO, VP, VS), tag="m"):
    """axes: {axis: (origin, spacing, n)} for 'x', 'y', 'z'; order: the axis of
    each file dimension, fastest first.  Returns (size, spacing, origin, paths)."""
    size = [axes[a][2] for a in order]
    spacing = [axes[a][1] for a in order]
    origin = [axes[a][0] for a in order]
    node = {a: axes[a][0] + axes[a][1] * np.arange(axes[a][2]) for a in "xyz"}
    # array indexed [dim3][dim2][dim1], dimension 1 fastest in the file
    g = np.meshgrid(node[order[2]], node[order[1]], node[order[0]], indexing="ij")
    coord = {order[2]: g[0], order[1]: g[1], order[0]: g[2]}
    paths = []
    for name, f in zip(("rho", "vp", "vs"), fields):
        p = tmp_path / f"{tag}_{name}_{order}.bin"
        f(coord["x"], coord["y"], coord["z"]).astype(np.float32).tofile(p)
        paths.append(str(p))
    return size, spacing, origin, paths


def _grid(ni=6, nj=5, nk=8):
    return cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -(nk - 1) * 100.0)


AXES = {"x": (-400.0, 150.0, 10), "y": (-350.0, 120.0, 10), "z": (-1100.0, 90.0, 17)}


def _run(impl, ni, nj, nk, x, y, z, order, size, spacing, origin, paths):
    f = obm.media if impl == "oracle" else cgfd.media_bin_file
    return f(ni, nj, nk, x, y, z, order, size, spacing, origin, *paths)


def _check_linear(lam, mu, rho, x, y, z, zclamp=None):
    zz = z if zclamp is None else np.clip(z, *zclamp)
    r, p, s = RHO(x, y, zz), VP(x, y, zz), VS(x, y, zz)
    assert np.allclose(rho, r, rtol=2e-6, atol=0)
    assert np.allclose(mu, s * s * r, rtol=5e-6, atol=0)
    assert np.allclose(lam, p * p * r - 2 * s * s * r, rtol=3e-5, atol=0)


@pytest.mark.parametrize("order", ["xyz", "zxy", "yzx"])
@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_linear_fields_are_reproduced(tmp_path, impl, order):
    ni, nj, nk = 6, 5, 8
    x, y, z = _grid(ni, nj, nk)  # x in [-300, 800], y in [-300, 700], z in [-1000, 300]
    size, spacing, origin, paths = write_bin(tmp_path, order, AXES)
    _check_linear(*_run(impl, ni, nj, nk, x, y, z, order, size, spacing, origin, paths), x, y, z)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_file_order_does_not_change_the_medium(tmp_path, impl):
    ni, nj, nk = 6, 5, 8
    x, y, z = _grid(ni, nj, nk)
    res = [_run(impl, ni, nj, nk, x, y, z, o, *write_bin(tmp_path, o, AXES)) for o in ("xyz", "zyx", "yxz")]
    for other in res[1:]:
        for u, v in zip(res[0], other):
            assert np.array_equal(u, v)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_decreasing_z_and_clamping_outside(tmp_path, impl):
    ni, nj, nk = 6, 5, 8
    x, y, z = _grid(ni, nj, nk)
    # z listed from 200 down to -610 (10 nodes, spacing -90): grid points below -610 or above 200
    # take the edge plane's values, still interpolated in x and y
    axes = dict(AXES, z=(200.0, -90.0, 10))
    size, spacing, origin, paths = write_bin(tmp_path, "zxy", axes)
    lam, mu, rho = _run(impl, ni, nj, nk, x, y, z, "zxy", size, spacing, origin, paths)
    _check_linear(lam, mu, rho, x, y, z, zclamp=(-610.0, 200.0))
    assert (z < -610.0).any() and (z > 200.0).any()


def test_window_of_a_larger_volume(tmp_path):
    """A volume reaching far past the grid in x and y: only the covering
    window is kept, the medium is the same and the product equals the oracle."""
    ni, nj, nk = 6, 5, 8
    x, y, z = _grid(ni, nj, nk)
    big = {"x": (-2500.0, 150.0, 40), "y": (-2630.0, 120.0, 45), "z": AXES["z"]}
    size, spacing, origin, paths = write_bin(tmp_path, "xzy", big, tag="big")
    a = cgfd.media_bin_file(ni, nj, nk, x, y, z, "xzy", size, spacing, origin, *paths)
    _check_linear(*a, x, y, z)
    for u, v in zip(obm.media(ni, nj, nk, x, y, z, "xzy", size, spacing, origin, *paths), a):
        assert np.array_equal(u, v)


def test_errors(tmp_path):
    ni, nj, nk = 6, 5, 4
    x, y, z = _grid(ni, nj, nk)
    size, spacing, origin, paths = write_bin(tmp_path, "xyz", AXES)
    cases = [
        ("order", dict(order=[0, 0, 2])),
        ("cover the grid in x", dict(origin=[5000.0, origin[1], origin[2]])),
        ("cover the grid in z", dict(origin=[origin[0], origin[1], 2000.0])),
        ("insufficient", dict(size=[size[0], size[1], size[2] + 1])),
    ]
    for what, kw in cases:
        args = dict(order=[0, 1, 2], size=size, spacing=spacing, origin=origin)
        args.update(kw)
        with pytest.raises(cgfd.CgfdError, match=what):
            cgfd.media_bin_file(ni, nj, nk, x, y, z, args["order"], args["size"], args["spacing"], args["origin"],
                                *paths)
        with pytest.raises(ValueError):
            obm.media(ni, nj, nk, x, y, z, args["order"], args["size"], args["spacing"], args["origin"], *paths)
    with pytest.raises(cgfd.CgfdError, match="cannot open"):
        cgfd.media_bin_file(ni, nj, nk, x, y, z, "xyz", size, spacing, origin, tmp_path / "missing.bin",
                            *paths[1:])


def test_product_equals_oracle_on_topography(tmp_path):
    ni, nj, nk = 14, 12, 10
    x, y, z = cgfd.grid_topography(ni, nj, nk, 250.0, 1.5, 2, (200.0, 500.0), (800.0, 1500.0), 30.0, 500.0, 4)
    rng = np.random.default_rng(8)
    axes = {"x": (float(x.min()) - 700.0, 310.0, 20), "y": (float(y.min()) - 450.0, 280.0, 19),
            "z": (float(z.max()) + 100.0, -205.0, 18)}
    n = {a: axes[a][2] for a in "xyz"}
    tables = [rng.uniform(lo, hi, (n["x"], n["y"], n["z"])) for lo, hi in ((2200, 2900), (5000, 6500), (2800, 3400))]

    def field(t):
        def f(xx, yy, zz):
            i = np.rint((xx - axes["x"][0]) / axes["x"][1]).astype(int)
            j = np.rint((yy - axes["y"][0]) / axes["y"][1]).astype(int)
            k = np.rint((zz - axes["z"][0]) / axes["z"][1]).astype(int)
            return t[i, j, k]
        return f

    size, spacing, origin, paths = write_bin(tmp_path, "zxy", axes, fields=[field(t) for t in tables])
    a = obm.media(ni, nj, nk, x, y, z, "zxy", size, spacing, origin, *paths)
    b = cgfd.media_bin_file(ni, nj, nk, x, y, z, "zxy", size, spacing, origin, *paths)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert len(np.unique(b[2])) > 1000
