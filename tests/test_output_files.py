"""Output formats: SAC traces and NetCDF-3 snapshots, read back independently.

The SAC header is decoded from the published v6 word layout; snapshots are
read with scipy's NetCDF-3 reader.  Host only.
"""
import struct

import numpy as np
import pytest
from scipy.io import netcdf_file

import cgfd


def _sac(path):
    raw = open(path, "rb").read()
    hf = np.frombuffer(raw[:280], dtype=np.float32)
    hi = np.frombuffer(raw[280:440], dtype=np.int32)
    hk = raw[440:632]
    data = np.frombuffer(raw[632:], dtype=np.float32)
    return hf, hi, hk, data


def test_sac_header_and_data(tmp_path):
    x = np.sin(np.linspace(0, 3, 50)).astype(np.float32)
    p = tmp_path / "a.sac"
    cgfd.sac_write(p, x, 0.005, b=0.25, ev=(1, 2, 3, 4), st=(5, 6, 7, 8), kstnm="STA1", kcmpnm="Vz")
    hf, hi, hk, data = _sac(p)
    assert len(open(p, "rb").read()) == 632 + 4 * 50
    assert np.array_equal(data, x)
    assert hf[0] == np.float32(0.005) and hf[5] == np.float32(0.25)
    assert hf[6] == np.float32(0.25) + np.float32(0.005) * np.float32(49)
    assert list(hf[35:39]) == [1, 2, 3, 4] and list(hf[31:35]) == [5, 6, 7, 8]
    assert hf[57] == 0 and hf[58] == 0 and hf[1] == -12345.0  # cmpaz, cmpinc set; depmin undefined
    assert hi[6] == 6 and hi[9] == 50 and hi[15] == 1 and hi[35] == 1 and hi[0] == -12345
    assert hk[0:4] == b"STA1" and hk[160:162] == b"Vz"
    assert hk[24:32] == b"-12345  "  # khole undefined


def test_station_files_names_quirk_and_strain(tmp_path):
    nt = 40
    rng = np.random.default_rng(0)
    seis = rng.normal(size=(9, nt)).astype(np.float32)
    lam, mu = 3.2e10, 3.0e10
    cgfd.station_sac(tmp_path, "r1", seis, 0.005, 100.0, 200.0, -50.0, strain=True, lam=lam, mu=mu)
    names = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    for c, n in enumerate(names):
        hf, hi, _, data = _sac(tmp_path / f"r1.{n}.sac")
        assert np.array_equal(data, seis[c])
        assert hf[5] == 0.0 and hf[34] == np.float32(0.005)  # b = 0; station depth carries dt
        assert list(hf[31:34]) == [100.0, 200.0, -50.0]
    # strain from the isotropic compliance: E = (1/2mu) T - lam/(2mu(3lam+2mu)) tr(T) I
    T = {n: seis[c].astype(np.float64) for c, n in enumerate(names)}
    tr = T["Txx"] + T["Tyy"] + T["Tzz"]
    want = {"Exx": T["Txx"] / (2 * mu) - lam / (2 * mu * (3 * lam + 2 * mu)) * tr,
            "Eyy": T["Tyy"] / (2 * mu) - lam / (2 * mu * (3 * lam + 2 * mu)) * tr,
            "Ezz": T["Tzz"] / (2 * mu) - lam / (2 * mu * (3 * lam + 2 * mu)) * tr,
            "Eyz": T["Tyz"] / (2 * mu), "Exz": T["Txz"] / (2 * mu), "Exy": T["Txy"] / (2 * mu)}
    for n, w in want.items():
        _, _, _, data = _sac(tmp_path / f"r1.{n}.sac")
        assert np.allclose(data, w, rtol=1e-5, atol=1e-6 * np.abs(w).max())


def test_snapshot_netcdf_layout(tmp_path):
    ni, nj, nk = 7, 5, 4
    rng = np.random.default_rng(1)
    p = tmp_path / "snap_px0_py0_pz0.nc"
    sn = cgfd.Snapshot(p, ni, nj, nk, vel=True, stress=True, strain=True, first_global=(2, 3, 4),
                       first_local=(5, 6, 7), stride=(2, 1, 3))
    lam = rng.uniform(1e10, 3e10, size=(nk, nj, ni)).astype(np.float32)
    mu = rng.uniform(1e10, 3e10, size=(nk, nj, ni)).astype(np.float32)
    frames = [rng.normal(size=(9, nk, nj, ni)).astype(np.float32) for _ in range(3)]
    for t, w in enumerate(frames):
        sn.append(0.5 * t, w, lam, mu)
    sn.close()
    f = netcdf_file(str(p), "r", mmap=False)
    assert f.dimensions["time"] is None and f.dimensions["i"] == ni and f.dimensions["k"] == nk
    assert list(f._attributes["first_index_to_snapshot_output"]) == [2, 3, 4]
    assert list(f._attributes["first_index_in_this_thread_with_ghosts"]) == [5, 6, 7]
    assert list(f._attributes["index_stride_in_this_thread"]) == [2, 1, 3]
    assert list(f._attributes["coords_of_mpi_topo"]) == [0, 0, 0]
    assert np.array_equal(f.variables["time"][:], [0.0, 0.5, 1.0])
    order = {"Vx": 0, "Vy": 1, "Vz": 2, "Txx": 3, "Tyy": 4, "Tzz": 5, "Tyz": 6, "Txz": 7, "Txy": 8}
    for n, c in order.items():
        v = f.variables[n]
        assert v.dimensions == ("time", "k", "j", "i")
        for t, w in enumerate(frames):
            assert np.array_equal(v[t], w[c])
    w = frames[1].astype(np.float64)
    l, m = lam.astype(np.float64), mu.astype(np.float64)
    tr = w[3] + w[4] + w[5]
    k = l / (2 * m * (3 * l + 2 * m))
    assert np.allclose(f.variables["Exx"][1], w[3] / (2 * m) - k * tr, rtol=1e-4, atol=1e-14)
    assert np.allclose(f.variables["Eyz"][1], w[6] / (2 * m), rtol=1e-5)
    assert np.allclose(f.variables["Exz"][1], w[7] / (2 * m), rtol=1e-5)
    f.close()


def test_snapshot_velocity_only(tmp_path):
    p = tmp_path / "v.nc"
    sn = cgfd.Snapshot(p, 3, 2, 2, vel=True, stress=False, strain=False)
    w = np.arange(9 * 12, dtype=np.float32).reshape(9, 2, 2, 3)
    sn.append(0.0, w)
    sn.close()
    f = netcdf_file(str(p), "r", mmap=False)
    assert set(f.variables) == {"time", "Vx", "Vy", "Vz"}
    assert np.array_equal(f.variables["Vz"][0], w[2])
    f.close()


def test_peak_motion_file(tmp_path):
    rng = np.random.default_rng(2)
    pg = rng.random((15, 9, 11)).astype(np.float32)
    p = tmp_path / "PG_V_A_D_px0_py0.nc"
    cgfd.peak_motion_write(p, pg, first_physical=(3, 3), count_physical=(5, 3), topo=(0, 0, 0))
    f = netcdf_file(str(p), "r", mmap=False)
    assert f.dimensions["j"] == 9 and f.dimensions["i"] == 11
    assert list(f._attributes["global_index_of_first_physical_points"]) == [3, 3]
    assert list(f._attributes["count_index_of_physical_points"]) == [5, 3]
    for q, n in enumerate(cgfd.PEAK_NAMES):
        assert f.variables[n].dimensions == ("j", "i")
        assert np.array_equal(f.variables[n][:], pg[q])
    f.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_slice_netcdf_layout(tmp_path, axis):
    ni, nj, nk = 7, 5, 4
    gindex = 2
    path = cgfd.slice_path(tmp_path, axis, gindex)
    # file name carries the 1-based plane number (io_funcs.cu:466-468)
    assert path.endswith(f"slice{'xyz'[axis]}_{'ijk'[axis]}3_px0_py0_pz0.nc")
    sl = cgfd.Slice(path, axis, gindex, ni, nj, nk, topo=(1, 0, 2))
    rng = np.random.default_rng(10 + axis)
    frames = [rng.normal(size=(9,) + sl.shape).astype(np.float32) for _ in range(3)]
    for t, w in enumerate(frames):
        sl.append(0.25 * (t + 1), w)
    sl.close()
    f = netcdf_file(path, "r", mmap=False)
    dims = [("k", "j"), ("k", "i"), ("j", "i")][axis]
    sizes = {"i": ni, "j": nj, "k": nk}
    assert f.dimensions["time"] is None
    assert {d: f.dimensions[d] for d in dims} == {d: sizes[d] for d in dims}
    assert "ijk"[axis] not in f.dimensions
    attr = f"{'ijk'[axis]}_index_with_ghosts_in_this_thread"
    assert list(np.atleast_1d(f._attributes[attr])) == [gindex + 3]
    assert list(f._attributes["coords_of_mpi_topo"]) == [1, 0, 2]
    assert np.array_equal(f.variables["time"][:], np.float32([0.25, 0.5, 0.75]))
    names = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    assert list(f.variables) == ["time"] + names
    for c, n in enumerate(names):
        v = f.variables[n]
        assert v.dimensions == ("time",) + dims and v.typecode() == "f"
        for t, w in enumerate(frames):
            assert np.array_equal(v[t], w[c])
    f.close()


def test_slice_rejects_bad_plane(tmp_path):
    with pytest.raises(ValueError):
        cgfd.Slice(tmp_path / "s.nc", 3, 0, 4, 4, 4)
    with pytest.raises(ValueError):
        cgfd.Slice(tmp_path / "s.nc", 1, 4, 4, 4, 4)
    with pytest.raises(cgfd.CgfdError):
        cgfd.Slice(tmp_path / "missing_dir" / "s.nc", 2, 0, 4, 4, 4)


def test_line_points_follow_the_reference_walk():
    # io_line_locate (io_funcs.cu:302-416): point ipt at start + ipt*incre,
    # kept when inside the physical grid; seq is ipt, the value is the node's
    ni, nj, nk = 9, 6, 5
    g = (ni + 6, nj + 6, nk + 6)
    x, y, z = np.meshgrid(np.arange(g[0]) * 10.0, np.arange(g[1]) * 20.0, np.arange(g[2]) * 30.0, indexing="ij")
    x, y, z = (np.ascontiguousarray(a.transpose(2, 1, 0)).ravel() for a in (x, y, z))
    seq, ix, fr, xyz = cgfd.line_points(ni, nj, nk, (-2, 1, 4), (2, 1, -1), 8, x, y, z)
    want = [ipt for ipt in range(8)
            if 0 <= -2 + 2 * ipt < ni and 0 <= 1 + ipt < nj and 0 <= 4 - ipt < nk]
    assert list(seq) == want == [1, 2, 3, 4]
    sy, sz = g[0], g[0] * g[1]
    for n, ipt in enumerate(seq):
        i, j, k = -2 + 2 * ipt + 3, 1 + ipt + 3, 4 - ipt + 3
        p = i + j * sy + k * sz
        assert ix[n, 0] == p and list(ix[n]) == [p, p + 1, p + sy, p + 1 + sy, p + sz, p + 1 + sz,
                                                 p + sy + sz, p + 1 + sy + sz]
        assert list(fr[n]) == [0.0, 0.0, 0.0]
        assert list(xyz[n]) == [10.0 * i, 20.0 * j, 30.0 * k]
    s2, _, _, _ = cgfd.line_points(ni, nj, nk, (ni, 0, 0), (1, 0, 0), 3)
    assert s2.size == 0


def test_line_sac_names(tmp_path):
    nt = 12
    seis = np.arange(9 * nt, dtype=np.float32).reshape(9, nt)
    cgfd.line_sac(tmp_path, "line_a", 7, seis, 0.01, 1.0, 2.0, 3.0)
    names = ["Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"]
    for c, n in enumerate(names):
        # io_line_output_sac (io_funcs.cu:1966-1968): {dir}/{line}.no{seq}.{cmp}.sac
        hf, hi, _, data = _sac(tmp_path / f"line_a.no7.{n}.sac")
        assert np.array_equal(data, seis[c]) and hi[9] == nt
        assert list(hf[31:35]) == [1.0, 2.0, 3.0, np.float32(0.01)]
