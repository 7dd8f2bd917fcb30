"""The reference's example media models, written as its exporters write them.

example/prep_medium/gen_eliso_lay.m (a basin under a flat free surface, a
wavy lower interface, gradients in the top layer) and gen_eliso_grid.m
(three layers of 11, 21 and 101 surfaces, a wavy boundary, velocities
stepping with the surface index), rebuilt here on a smaller mesh with the
same formulas and written in the printf formats of md3lay_export.m /
md3grd_export.m ("%d %d %f %f %f %f" for the mesh line, "%g" for values, no
comment lines).  The product must read them and equal the oracle bit for
bit; both reject, with the reference's vp/vs check, a model whose velocities
cross vp <= sqrt(2) vs where the grid samples it (the unmodified examples
do: layer 1 of the grid example reaches vp 1610, vs 1220).
"""
import importlib.util
import os

import numpy as np
import pytest

import cgfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name):
    spec = importlib.util.spec_from_file_location("oracle_" + name, os.path.join(ROOT, "oracle", name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


olm, ogm = _load("layer_media"), _load("grid_media")

NX, NY, DH = 30, 26, 100.0
X0, Y0 = -3 * DH, -3 * DH
X1D, Y1D = X0 + DH * np.arange(NX), Y0 + DH * np.arange(NY)
LX, LY = NX * DH, NY * DH
X2, Y2 = np.meshgrid(X1D, Y1D)  # [ny][nx], as MATLAB's meshgrid


def _g(v):
    return "%g" % v


def write_lay(path, vs_grad_top=0.2):
    """gen_eliso_lay.m on a 30 x 26 mesh, written as md3lay_export.m writes it."""
    dep = [0.0, 12 * DH, 3e3, 10e3]
    vp, vpg = [1500.0, 3000.0, 5000.0, 8000.0], [0.2, 0.0, 0.0, 0.0]
    vs, vsg = [1000.0, 1500.0, 2500.0, 3000.0], [vs_grad_top, 0.0, 0.0, 0.0]
    den, deng = [1000.0, 2000.0, 3000.0, 3500.0], [0.2, 0.0, 0.0, 0.0]
    elev = np.zeros((4, NY, NX))
    elev[0] = -dep[0]
    elev[1] = elev[0]
    bx, by, r0 = X1D[12], Y1D[13], dep[1]
    r2 = np.sqrt((X2 - bx) ** 2 + (Y2 - by) ** 2)
    elev[1] = np.where(r2 <= r0, -np.sqrt(np.maximum(r0 ** 2 - r2 ** 2, 0.0)), elev[1])
    wav = np.cos(4 * X2 / LX * 2 * np.pi) * np.cos(3 * Y2 / LY * 2 * np.pi) * 0.2 + 1.0
    elev[2] = np.cos(3 * X2 / LX * 2 * np.pi) * np.cos(1 * Y2 / LY * 2 * np.pi) * 5 * DH - dep[2]
    elev[3] = -dep[3]
    with open(path, "w") as f:
        f.write("elastic_isotropic\n%d\n" % 4)
        f.write("%d %d %f %f %f %f\n" % (NX, NY, X0, Y0, DH, DH))
        for l in range(4):
            sc = wav if l == 1 else np.ones_like(wav)
            for j in range(NY):
                for i in range(NX):
                    row = [elev[l, j, i], den[l] * sc[j, i], deng[l], 1.0, vp[l] * sc[j, i], vpg[l], 1.0,
                           vs[l] * sc[j, i], vsg[l], 1.0]
                    f.write(_g(row[0]) + "".join(" " + _g(v) for v in row[1:]) + "\n")
    return str(path)


def write_grd(path, vs_step_top=20.0):
    """gen_eliso_grid.m on a 30 x 26 mesh, written as md3grd_export.m writes it."""
    npl = [11, 21, 101]
    dep = [0.0, 2e3, 5e3]
    vp, vs, den = [1500.0, 3000.0, 5000.0], [1000.0, 1500.0, 2500.0], [1000.0, 2000.0, 3000.0]
    lay = np.zeros((3, NY, NX))
    lay[1] = np.cos(3 * X2 / LX * 2 * np.pi) * np.cos(3 * Y2 / LY * 2 * np.pi) * 5 * DH - dep[1]
    lay[2] = -dep[2]
    with open(path, "w") as f:
        f.write("elastic_isotropic\n%d\n" % 3)
        f.write("".join(" %d" % n for n in npl) + "\n")
        f.write("%d %d %f %f %f %f\n" % (NX, NY, X0, Y0, DH, DH))
        for l in range(3):
            top = lay[l]
            bot = lay[l + 1] if l < 2 else top - (npl[l] - 1) * 100.0
            for ipt in range(1, npl[l] + 1):
                e = top + (bot - top) / (npl[l] - 1) * (ipt - 1)
                if l == 0:
                    p, s = vp[0] + ipt * 10, vs[0] + ipt * vs_step_top
                elif l == 1:
                    p, s = vp[1] - ipt * 10, vs[1] - ipt * 10
                else:
                    p, s = vp[2], vs[2]
                for j in range(NY):
                    for i in range(NX):
                        f.write("%s %s %s %s\n" % (_g(e[j, i]), _g(den[l]), _g(p), _g(s)))
    return str(path)


def _grid(ni=12, nj=10, nk=16):
    return cgfd.grid_cartesian(ni, nj, nk, DH, 0.0, DH, 0.0, 100.0, -(nk - 1) * 100.0)


def test_layer_example_vp_vs_check(tmp_path):
    # the basin's top layer (vp 1500 + 0.2 d, vs 1000 + 0.2 d) crosses vp = sqrt(2) vs below ~1300 m
    ni, nj, nk = 12, 10, 16
    x, y, z = _grid(ni, nj, nk)
    path = write_lay(tmp_path / "basin.md3lay")
    with pytest.raises(cgfd.CgfdError, match="sqrt"):
        cgfd.media_layer_file(ni, nj, nk, x, y, z, path)
    with pytest.raises(AssertionError):
        olm.media(ni, nj, nk, x, y, z, path)


@pytest.mark.parametrize("method", ["loc", "har"])
def test_layer_example_product_equals_oracle(tmp_path, method):
    ni, nj, nk = (12, 10, 16) if method == "loc" else (8, 7, 16)
    x, y, z = _grid(ni, nj, nk)
    path = write_lay(tmp_path / "basin.md3lay", vs_grad_top=0.05)
    a = olm.media(ni, nj, nk, x, y, z, path) if method == "loc" else \
        olm.media_equivalent(ni, nj, nk, x, y, z, path, method)
    b = cgfd.media_layer_file(ni, nj, nk, x, y, z, path, method=method)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert len(np.unique(b[2])) > 20


def test_grid_example_vp_vs_check(tmp_path):
    ni, nj, nk = 12, 10, 16
    x, y, z = _grid(ni, nj, nk)
    path = write_grd(tmp_path / "ex.md3grd")
    with pytest.raises(cgfd.CgfdError, match="sqrt"):
        cgfd.media_grid_file(ni, nj, nk, x, y, z, path)
    with pytest.raises(ValueError, match="sqrt"):
        ogm.media(ni, nj, nk, x, y, z, path)


@pytest.mark.parametrize("method", ["loc", "ari"])
def test_grid_example_product_equals_oracle(tmp_path, method):
    ni, nj, nk = (12, 10, 16) if method == "loc" else (5, 4, 24)
    x, y, z = _grid(ni, nj, nk)
    path = write_grd(tmp_path / "ex.md3grd", vs_step_top=5.0)
    a = ogm.media(ni, nj, nk, x, y, z, path, method)
    b = cgfd.media_grid_file(ni, nj, nk, x, y, z, path, method=method)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    if method != "loc":
        loc = cgfd.media_grid_file(ni, nj, nk, x, y, z, path)
        assert (b[2] != loc[2]).sum() > 5  # the layer 1/2 boundary crosses the grid
        return
    # 1000 m down at the mesh node (600, 600), in layer 1: its 11 surfaces split [bot, 0] evenly and
    # surface n (1-based) carries vp 1500 + 10 n, so vp = 1500 + 10 (1 + 1000 / h), h = |bot| / 10
    vp = np.sqrt((b[0] + 2 * b[1]) / b[2])
    k = 3 + (nk - 1) - 10
    assert (x[k, 9, 9], y[k, 9, 9], z[k, 9, 9]) == (600.0, 600.0, -1000.0)
    bot = np.cos(3 * 600.0 / LX * 2 * np.pi) * np.cos(3 * 600.0 / LY * 2 * np.pi) * 5 * DH - 2e3
    assert vp[k, 9, 9] == pytest.approx(1500.0 + 10 * (1 + 1000.0 / (abs(bot) / 10)), rel=2e-5)
