"""Host preparation of the product (FP64, include/cgfd_setup.h) against the oracle.

The oracle restates the reference's FP32 setup; the product evaluates the
same formulas in FP64, so agreement is to FP32 rounding (plus the float
promotions the reference applies), and exact where the math is exact.
These run on the CPU only.
"""
import ctypes as C
import math

import numpy as np
import pytest

import cgfd
import orc
from problem import Problem

L = orc.lib()


def test_cartesian_grid_matches_oracle():
    ni, nj, nk = 12, 10, 9
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -800.0)
    n = x.size
    xo, yo, zo = np.zeros(n), np.zeros(n), np.zeros(n)
    L.orc_gen_cart(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -800.0, orc.p(xo), orc.p(yo), orc.p(zo))
    assert np.array_equal(x.ravel(), xo) and np.array_equal(y.ravel(), yo) and np.array_equal(z.ravel(), zo)


@pytest.mark.parametrize("curv", [False, True])
def test_metric_fp64_matches_oracle(curv):
    P = Problem(ni=18, nj=16, nk=14, curvilinear=curv, source=None, nrecv=0, metric_precision=64)
    m = cgfd.metric_calc(P.ni, P.nj, P.nk, P.x, P.y, P.z)
    assert np.allclose(m, P.metric, rtol=1e-12, atol=1e-15)
    P32 = Problem(ni=18, nj=16, nk=14, curvilinear=curv, source=None, nrecv=0, metric_precision=32)
    for c in range(10):
        # zero components carry FP32 noise: scale by the component group (jac, or all metrics)
        scale = np.abs(P32.metric[0]).max() if c == 0 else np.abs(P32.metric[1:]).max()
        assert np.abs(m[c] - P32.metric[c]).max() <= 2e-5 * scale


def test_topography_grid_properties():
    ni, nj, nk, dh = 60, 50, 40, 100.0
    x, y, z = cgfd.grid_topography(ni, nj, nk, dh, 1.5, 3, (300, 800), (800, 1500), 50.0, 1000.0, 7)
    top = z[nk + 2, 3:-3, 3:-3]
    # vertical spacing grows from ~dh at the top to ~1.5 dh at the bottom (flat-surface column)
    dz = np.diff(z[3:3 + nk, :, :], axis=0)
    assert np.all(dz > 0)
    col = dz[:, 3, 3]
    assert 1.3 < col[0] / col[-1] < 1.7
    assert np.all(np.abs(z[3, 3:-3, 3:-3] - z[3, 3, 3]) < 1e-6)   # flat bottom
    assert top.max() - top.min() > 250                              # hills present
    assert np.array_equal(x[0, 0], (np.arange(ni + 6) - 3) * dh)
    # metrics stay well conditioned and the grid is not folded
    m = cgfd.metric_calc(ni, nj, nk, x, y, z)
    assert m[0].min() > 0
    # deterministic for a seed
    x2, y2, z2 = cgfd.grid_topography(ni, nj, nk, dh, 1.5, 3, (300, 800), (800, 1500), 50.0, 1000.0, 7)
    assert np.array_equal(z, z2)


def test_dt_estimate_matches_oracle():
    P = Problem(ni=14, nj=12, nk=10, curvilinear=True, source=None, nrecv=0, hetero=0.1, seed=3)
    dt_p = cgfd.dt_estimate(P.ni, P.nj, P.nk, P.x, P.y, P.z, P.lam, P.mu, P.rho, 1.3)
    dt_o = L.orc_dt_esti(P.ni, P.nj, P.nk, orc.p(P.x), orc.p(P.y), orc.p(P.z), orc.p(P.lam), orc.p(P.mu),
                         orc.p(P.rho), 1.3)
    assert abs(dt_p - dt_o) < 1e-6 * dt_o


def test_pml_profile_matches_oracle():
    P = Problem(ni=24, nj=20, nk=18, curvilinear=True, source=None, nrecv=0, pml=(6, 6, 6, 6, 6, 0))
    for f, (nl, A, B, D) in P.pml.items():
        L0, dh = cgfd.pml_len_dh(P.ni, P.nj, P.nk, P.x, P.y, P.z, f, nl)
        Ap, Bp, Dp = cgfd.pml_profile(nl, f, L0, dh, 6000.0, 3.0, 2.5)
        assert np.allclose(Ap, A, rtol=2e-6) and np.allclose(Bp, B, rtol=2e-6) and np.allclose(Dp, D, rtol=2e-6)


def test_free_surface_matrices_match_oracle():
    P = Problem(ni=16, nj=14, nk=10, curvilinear=True, source=None, nrecv=0, hetero=0.1, seed=8,
                metric_precision=64)
    mx, my = cgfd.free_surface_matrices(P.ni, P.nj, P.nk, P.metric, P.lam, P.mu)
    mxo, myo = np.zeros_like(mx), np.zeros_like(my)
    L.orc_free_matrices(64, P.ni, P.nj, P.nk, orc.p(P.metric), orc.p(P.lam), orc.p(P.mu), orc.p(mxo), orc.p(myo))
    assert np.allclose(mx, mxo, rtol=1e-10, atol=1e-12)
    assert np.allclose(my, myo, rtol=1e-10, atol=1e-12)


def test_wavelets_and_moment_match_oracle():
    for t in np.linspace(0, 1.5, 31):
        assert abs(cgfd.wavelet(0, t, 1.5, 0.8) - L.orc_ricker(t, 1.5, 0.8)) < 2e-6
        assert abs(cgfd.wavelet(2, t, 1.5, 0.8) - L.orc_ricker_deriv(t, 1.5, 0.8)) < 2e-5 * 10
        assert abs(cgfd.wavelet(1, t, 0.2, 0.8) - L.orc_gauss(t, 0.2, 0.8)) < 2e-5
    m = np.zeros(6)
    for s, d, r in [(30, 60, 90), (123, 45, -30), (250, 80, 170)]:
        L.orc_angle2moment(s, d, r, orc.p(m))
        assert np.allclose(cgfd.angle2moment(s, d, r), m, atol=2e-6)


@pytest.mark.parametrize("depth_flag,sz", [(0, 9.3), (1, 4.0)])
def test_source_point_matches_oracle(depth_flag, sz):
    ni, nj, nk, dt = 30, 28, 24, 0.005
    s = cgfd.source_point(ni, nj, nk, 14.4, 12.6, sz, depth_flag, 0.0, dt, 0.01, 0, 2.0, 0.3, 0.4)
    ei = np.zeros(343, dtype=np.int32)
    ec = np.zeros(343)
    itb, ite, en, mnt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    stf = np.zeros(4 * 200)
    L.orc_src_point(ni, nj, nk, 14.4, 12.6, sz, depth_flag, 0.0, dt, 0.01, 0, 2.0, 0.3, 0.4, C.byref(itb),
                    C.byref(ite), C.byref(en), orc.p(ei), orc.p(ec), C.byref(mnt), orc.p(stf))
    assert (s["it_begin"], s["it_end"], s["ext_num"], s["max_nt"]) == (itb.value, ite.value, en.value, mnt.value)
    assert np.array_equal(s["ext_indx"][:en.value], ei[:en.value])
    assert np.allclose(s["ext_coef"][:en.value], ec[:en.value], rtol=1e-5)
    assert np.allclose(s["stf"].ravel(), stf[:4 * mnt.value], atol=5e-6)


def test_receiver_index_matches_oracle():
    ni, nj, nk = 30, 25, 20
    for args in [(10.0, 12.0, 0.0, 1), (10.5, 12.25, 3.0, 0), (1.0, 1.0, 2.0, 1), (30.0, 25.0, 0.0, 1)]:
        ix, fr = cgfd.receiver_index(ni, nj, nk, *args)
        io, fo = np.zeros(8, dtype=np.int64), np.zeros(3)
        assert L.orc_recv_index(ni, nj, nk, *args, orc.p(io), orc.p(fo)) == 0
        assert np.array_equal(ix, io) and np.allclose(fr, fo, atol=1e-6)
    assert cgfd.receiver_index(ni, nj, nk, 31.0, 5.0, 0.0, 1)[0] is None


def test_layered_media_gardner_and_perturbation():
    ni, nj, nk = 40, 36, 30
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, 100.0, 0.0, 100.0, 0.0, 100.0, -2900.0)
    zi = np.array([-800.0, -2000.0])
    vp = np.array([3000.0, 5000.0, 6500.0])
    lam, mu, rho = cgfd.media_layered(ni, nj, nk, x, y, z, zi, vp, np.full(3, 1.73), np.zeros(3), 0.025, 0.05,
                                      1000.0, 5)
    Vp = np.sqrt((lam + 2 * mu) / rho)
    Vs = np.sqrt(mu / rho)
    assert np.allclose(Vp / Vs, 1.73)
    assert np.allclose(rho, 310.0 * Vp ** 0.25)                    # Gardner
    for l, (lo, hi) in enumerate([(-800, 1e9), (-2000, -800), (-1e9, -2000)]):
        sel = (z < hi) & (z >= lo)
        rel = Vp[sel] / vp[l] - 1
        assert np.abs(rel).max() <= 0.05 + 1e-12
        assert 0.01 < rel.std() < 0.04
    # correlated: neighbouring points differ much less than the std
    d = Vp[:, :, 1:] - Vp[:, :, :-1]
    sel = (z[:, :, 1:] < -2100) & (z[:, :, 1:] > -2800)
    assert np.abs(d[sel]).mean() < 0.3 * (Vp[(z < -2100)].std())
