"""Pins of the CPU oracle against closed forms, invariants and literal tables.

The reference ships no tests or golden outputs (its only fixtures are example
input files), so each oracle function is pinned here by mathematics chosen so
that a plausible slip -- a dropped term, a wrong sign or index, a transposed
operand, a reversed profile -- fails at least one check.
"""
import ctypes as C
import math

import numpy as np
import pytest

import orc
from problem import Problem, rel_l2

L = orc.lib()


def op(layer, d):
    idx = np.zeros(5, dtype=np.int32)
    coef = np.zeros(5)
    n = L.orc_fd_op(layer, d, orc.p(idx), orc.p(coef))
    return idx[:n], coef[:n]


# ---------------------------------------------------------------- operators
@pytest.mark.parametrize("layer", [1, 2, 3])
@pytest.mark.parametrize("d", [0, 1])
def test_operator_consistency(layer, d):
    idx, c = op(layer, d)
    assert abs(c.sum()) < 1e-8                      # kills constants
    assert abs((c * idx).sum() - 1.0) < 2e-9       # exact slope (DRP: 0.999999999)
    if d == 0:
        assert idx.max() <= 1 and idx.min() < 0    # backward biased
    else:
        assert idx.min() >= -1 and idx.max() > 0   # forward biased


@pytest.mark.parametrize("layer", [1, 2, 3])
def test_forward_is_mirrored_backward(layer):
    ib, cb = op(layer, 0)
    jf, cf = op(layer, 1)
    fb = dict(zip(ib.tolist(), cb.tolist()))
    for j, c in zip(jf.tolist(), cf.tolist()):
        assert abs(c + fb[-j]) < 1e-12


def test_drp_average_is_centered_metric_operator():
    ib, cb = op(3, 0)
    jf, cf = op(3, 1)
    avg = np.zeros(7)
    for i, c in zip(ib, cb):
        avg[i + 3] += 0.5 * c
    for j, c in zip(jf, cf):
        avg[j + 3] += 0.5 * c
    fdc = np.zeros(7)
    L.orc_fdc(orc.p(fdc))
    assert np.allclose(avg, fdc, atol=1e-4)
    assert np.allclose(fdc, -fdc[::-1])


def test_drp_dispersion_is_better_than_plain_4th_order():
    """The DRP centered average resolves kh = 1.0 closer than the 4th-order stencil."""
    fdc = np.zeros(7)
    L.orc_fdc(orc.p(fdc))
    kh = 1.0
    drp = sum(fdc[n + 3] * math.sin(n * kh) for n in range(-3, 4))
    fourth = (8 * math.sin(kh) - math.sin(2 * kh)) / 6
    assert abs(drp - kh) < abs(fourth - kh)


def test_pair_table_literal_and_alternation():
    # FD_Flags literal (src/forward/fd_t.cu:184-194)
    flags = [(0, 0, 0), (1, 1, 0), (1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0), (1, 0, 1), (0, 1, 1)]
    got = []
    for ip in range(8):
        d = np.zeros(3, dtype=np.int32)
        L.orc_pair_dirs(ip, 0, orc.p(d))
        got.append(tuple(d))
        for st in range(1, 4):
            e = np.zeros(3, dtype=np.int32)
            L.orc_pair_dirs(ip, st, orc.p(e))
            assert tuple(e) == tuple((np.array(d) + st) % 2)
    assert got == flags
    assert len(set(got)) == 8


def test_rk4_coefficients():
    a, b, t = np.zeros(3), np.zeros(4), np.zeros(4)
    L.orc_rk(orc.p(a), orc.p(b), orc.p(t))
    assert np.allclose(b.sum(), 1.0)
    assert np.allclose(a, [0.5, 0.5, 1.0]) and np.allclose(t, [0, 0.5, 0.5, 1.0])
    assert np.allclose(b, [1 / 6, 1 / 3, 1 / 3, 1 / 6])


# ---------------------------------------------------------------- grid metrics
def test_cartesian_metric_closed_form():
    ni, nj, nk = 8, 7, 6
    n = (ni + 6) * (nj + 6) * (nk + 6)
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    L.orc_gen_cart(ni, nj, nk, 50.0, 0.0, 80.0, 0.0, 100.0, -500.0, orc.p(x), orc.p(y), orc.p(z))
    met = np.zeros(10 * n)
    L.orc_metric_cal(64, ni, nj, nk, orc.p(x), orc.p(y), orc.p(z), orc.p(met))
    met = met.reshape(10, -1)
    fdc = np.zeros(7)
    L.orc_fdc(orc.p(fdc))
    g = sum(fdc[q] * (q - 3) for q in range(7))  # slope of the centered table, 1.00004
    assert abs(g - 1.00004) < 1e-12
    assert np.allclose(met[0], g ** 3 * 50 * 80 * 100, rtol=1e-6)
    assert np.allclose(met[1], 1 / (g * 50), rtol=1e-6)
    assert np.allclose(met[5], 1 / (g * 80), rtol=1e-6)
    assert np.allclose(met[9], 1 / (g * 100), rtol=1e-6)
    for c in (2, 3, 4, 6, 7, 8):
        assert np.abs(met[c]).max() < 1e-9


def test_affine_grid_metric_is_inverse_jacobian():
    """x = A xi for a sheared A: metrics are A^-1/g everywhere, ghosts included."""
    ni, nj, nk = 7, 6, 6
    nx, ny, nz = ni + 6, nj + 6, nk + 6
    A = np.array([[90.0, 12.0, -5.0], [8.0, 110.0, 15.0], [-6.0, 9.0, 95.0]])
    I, J, K = np.meshgrid(np.arange(nx) - 3.0, np.arange(ny) - 3.0, np.arange(nz) - 3.0, indexing="ij")
    xi = np.stack([I, J, K]).transpose(0, 3, 2, 1)  # [3][k][j][i]
    X = np.einsum("ab,bkji->akji", A, xi)
    x, y, z = [np.ascontiguousarray(X[q]) for q in range(3)]
    met = np.zeros(10 * nx * ny * nz)
    L.orc_metric_cal(64, ni, nj, nk, orc.p(x), orc.p(y), orc.p(z), orc.p(met))
    met = met.reshape(10, -1)
    g = 1.00004
    Ainv = np.linalg.inv(A) / g
    assert np.allclose(met[0], np.linalg.det(A) * g ** 3, rtol=1e-9)
    for r in range(3):
        for c in range(3):
            assert np.allclose(met[1 + 3 * r + c], Ainv[r, c], rtol=1e-8, atol=1e-14)


# ---------------------------------------------------------------- inner RHS
def _linear_stage(P, G, S, ipair=0, istage=0):
    nz, ny, nx = P.nz, P.ny, P.nx
    w = np.zeros((9, nz, ny, nx))
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([P.x, P.y, P.z])
    phys = (i >= 3) & (i < nx - 3) & (j >= 3) & (j < ny - 3) & (k >= 3) & (k < nz - 3)
    for a in range(3):
        w[a] = np.einsum("b,bkji->kji", G[a], X)
    Tmap = {(0, 0): 3, (1, 1): 4, (2, 2): 5, (1, 2): 6, (0, 2): 7, (0, 1): 8}
    for (a, b), c in Tmap.items():
        w[c] = np.einsum("d,dkji->kji", S[a, b], X)
    w *= phys
    rhs, _ = P.oracle_stage(w, ipair, istage, 0, precision=64)
    inner = (i >= 6) & (i < nx - 6) & (j >= 6) & (j < ny - 6) & (k >= 6) & (k < nz - 6)
    return rhs, inner


def test_inner_rhs_linear_velocity_gives_hooke():
    P = Problem(ni=16, nj=16, nk=16, free_top=False, pml=(0,) * 6, source=None, nrecv=0, metric_precision=64)
    rng = np.random.default_rng(0)
    G = rng.normal(size=(3, 3)) * 1e-3
    rhs, inner = _linear_stage(P, G, np.zeros((3, 3, 3)))
    lam, mu = P.lam[8, 8, 8], P.mu[8, 8, 8]
    f = 0.999999999 / 1.00004   # DRP slope x centered-metric slope
    Ge = G * f
    sig = lam * np.trace(Ge) * np.eye(3) + mu * (Ge + Ge.T)
    for c, (a, b) in zip(range(3, 9), [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]):
        # the DRP table sums to -1e-9, which leaves x/h * 1e-9 relative residue
        assert np.allclose(rhs[c][inner], sig[a, b], rtol=0, atol=1e-7 * np.abs(sig).max())
    for c in range(3):
        assert np.abs(rhs[c][inner]).max() < 1e-12


@pytest.mark.parametrize("ipair,istage", [(0, 0), (2, 1), (5, 3)])
def test_inner_rhs_linear_stress_gives_divergence(ipair, istage):
    P = Problem(ni=16, nj=16, nk=16, free_top=False, pml=(0,) * 6, source=None, nrecv=0, curvilinear=False,
                metric_precision=64)
    rng = np.random.default_rng(1)
    S = rng.normal(size=(3, 3, 3)) * 1e3
    S = 0.5 * (S + S.transpose(1, 0, 2))
    rhs, inner = _linear_stage(P, np.zeros((3, 3)), S, ipair, istage)
    f = 0.999999999 / 1.00004
    div = np.array([sum(S[a, b, b] for b in range(3)) for a in range(3)]) * f
    slw = P.slw[8, 8, 8]
    for a in range(3):
        assert np.allclose(rhs[a][inner], slw * div[a], rtol=0, atol=1e-7 * slw * np.abs(div).max())


# ---------------------------------------------------------------- free surface
def test_free_surface_matrices_flat_closed_form():
    P = Problem(ni=10, nj=9, nk=8, free_top=True, pml=(0,) * 6, source=None, nrecv=0)
    lam, mu = P.lam[0, 0, 0], P.mu[0, 0, 0]
    j, i = 5, 6
    xix, ety, ztz = P.metric[1, -4, j, i], P.metric[5, -4, j, i], P.metric[9, -4, j, i]
    mx = P.mvx[j, i].reshape(3, 3)
    my = P.mvy[j, i].reshape(3, 3)
    ex = np.zeros((3, 3))
    ex[0, 2] = -xix / ztz
    ex[2, 0] = -lam / (lam + 2 * mu) * xix / ztz
    ey = np.zeros((3, 3))
    ey[1, 2] = -ety / ztz
    ey[2, 1] = -lam / (lam + 2 * mu) * ety / ztz
    # matrices are evaluated in FP32: entries carry float-level residue
    assert np.allclose(mx, ex, rtol=1e-6, atol=1e-6)
    assert np.allclose(my, ey, rtol=1e-6, atol=1e-6)


def test_free_surface_matrices_zero_traction_on_topography():
    """Velocity gradient built with the matrices makes the surface traction rate vanish."""
    P = Problem(ni=14, nj=12, nk=10, free_top=True, curvilinear=True, pml=(0,) * 6, source=None, nrecv=0,
                hetero=0.1, seed=9)
    rng = np.random.default_rng(2)
    k = P.nz - 4
    for (j, i) in [(4, 5), (8, 9), (10, 7)]:
        xi = P.metric[1:4, k, j, i]
        et = P.metric[4:7, k, j, i]
        zt = P.metric[7:10, k, j, i]
        lam, mu = P.lam[k, j, i], P.mu[k, j, i]
        DxV = rng.normal(size=3)
        DyV = rng.normal(size=3)
        DzV = P.mvx[j, i].reshape(3, 3) @ DxV + P.mvy[j, i].reshape(3, 3) @ DyV
        Lg = np.outer(DxV, xi) + np.outer(DyV, et) + np.outer(DzV, zt)   # dV_a/dx_b
        sig = lam * np.trace(Lg) * np.eye(3) + mu * (Lg + Lg.T)
        tr = sig @ zt
        assert np.linalg.norm(tr) < 2e-5 * np.linalg.norm(sig) * np.linalg.norm(zt)


# ---------------------------------------------------------------- PML
def _pml(nl, iside, L0=1000.0, dh=100.0, vel=6000.0, amax=3.0, bmax=2.5):
    A, B, D = np.zeros(nl + 1), np.zeros(nl + 1), np.zeros(nl + 1)
    L.orc_pml_coefs(nl, iside, L0, dh, vel, amax, bmax, orc.p(A), orc.p(B), orc.p(D))
    return A, B, D


def test_pml_profile_shape_and_orientation():
    A, B, D = _pml(10, 1)
    # high side: index 0 is the inner (inert) point, the last is the outer edge
    assert D[0] == 0.0 and B[0] == 1.0 and abs(A[0] - 3.0) < 1e-6
    assert np.all(np.diff(D) > 0) and np.all(np.diff(B) < 0)
    A1, B1, D1 = _pml(10, 0)
    assert np.allclose(A1, A[::-1]) and np.allclose(B1, B[::-1]) and np.allclose(D1, D[::-1])


def test_pml_profile_values_cfs():
    """Zhang & Shen CFS-PML: d0 = -(p+1) V ln R / (2L), R = 10^-((log10 N - 1)/log10 2 + 4)."""
    nl, L0, dh, vel, amax, bmax = 10, 1000.0, 100.0, 6000.0, 3.0, 2.5
    A, B, D = _pml(nl, 1, L0, dh, vel, amax, bmax)
    R = 10 ** (-((math.log10(nl) - 1) / math.log10(2) + 4))
    d0 = -(2 + 1) * vel * math.log(R) / (2 * L0)
    for n in range(nl + 1):
        r = n * dh / L0
        d, a, b = d0 * r * r, amax * (1 - r), 1 + (bmax - 1) * r * r
        assert abs(B[n] - 1 / b) < 1e-6
        assert abs(D[n] - d / b) < 1e-4 * max(1, d0)
        assert abs(A[n] - (a + d / b)) < 1e-4 * max(1, d0)


def test_pml_identity_coefficients_leave_rhs_unchanged():
    base = Problem(ni=20, nj=18, nk=16, free_top=True, pml=(0,) * 6, init_random=True, seed=4)
    withp = Problem(ni=20, nj=18, nk=16, free_top=True, pml=(4, 4, 4, 4, 4, 0), init_random=True, seed=4)
    for f in list(withp.pml):
        nl = withp.pml[f][0]
        withp.pml[f] = (nl, np.zeros(nl + 1), np.ones(nl + 1), np.zeros(nl + 1))
    r0, _ = base.oracle_stage(base.init, 3, 1, 0)
    sizes = withp.face_sizes()
    aux0 = np.random.default_rng(5).normal(size=9 * int(sizes.sum()))
    # B = 1 makes the memory term -Phi; zero memory must reproduce the plain RHS exactly
    r1, a1 = withp.oracle_stage(withp.init, 3, 1, 0, aux=np.zeros_like(aux0))
    assert np.array_equal(r0, r1)
    assert np.abs(a1).max() == 0.0


def test_rk4_amplification_on_pml_memory_ode():
    """D=0, A=alpha, B=1: Phi' = -alpha Phi decouples; one RK4 step multiplies by the Taylor polynomial."""
    P = Problem(ni=16, nj=14, nk=12, free_top=False, pml=(3, 0, 0, 0, 0, 0), source=None, nrecv=0, dt=0.01)
    alpha = 37.0
    nl = P.pml[0][0]
    P.pml[0] = (nl, np.full(nl + 1, alpha), np.ones(nl + 1), np.zeros(nl + 1))
    sizes = P.face_sizes()
    phi0 = np.random.default_rng(7).normal(size=9 * int(sizes.sum()))
    nsteps = 3
    _, _, aux = P.oracle_run(64, nt_total=nsteps, want_aux=True, init_aux=phi0)
    z = -alpha * P.dt
    amp = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.allclose(aux, phi0 * amp ** nsteps, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------- source
def test_ricker_closed_form():
    f0, t0 = 2.0, 0.6
    for t in np.linspace(0, 1.2, 41):
        tau = t - t0
        exact = (1 - 2 * (math.pi * f0 * tau) ** 2) * math.exp(-(math.pi * f0 * tau) ** 2)
        assert abs(L.orc_ricker(t, f0, t0) - exact) < 2e-6
    assert abs(L.orc_ricker(t0, f0, t0) - 1.0) < 1e-7
    tz = t0 + 1 / (math.sqrt(2) * math.pi * f0)
    assert abs(L.orc_ricker(tz, f0, t0)) < 1e-5


def test_gaussian_spreading_weights():
    for off in [(0, 0, 0), (0.3, -0.2, 0.45)]:
        d = np.zeros(343)
        assert L.orc_norm_delt3d(off[0], off[1], off[2], 1.5, 1.5, 1.5, 3, orc.p(d)) == 0
        d = d.reshape(7, 7, 7)  # [k][j][i]
        assert abs(d.sum() - 1) < 1e-6
        g = np.arange(-3, 4)
        cx = (d.sum(axis=(0, 1)) * g).sum()
        cz = (d.sum(axis=(1, 2)) * g).sum()
        assert abs(cx - off[0]) < 0.02 and abs(cz - off[2]) < 0.02
        if off == (0, 0, 0):
            assert np.allclose(d, d[::-1, ::-1, ::-1], rtol=1e-6)


def test_angle2moment_known_mechanisms():
    m = np.zeros(6)
    L.orc_angle2moment(0.0, 90.0, 0.0, orc.p(m))     # vertical strike-slip, strike north
    assert np.allclose(m, [0, 0, 0, 0, 0, 1], atol=1e-6)
    L.orc_angle2moment(0.0, 45.0, 90.0, orc.p(m))    # 45 degree thrust striking north
    assert np.allclose(m, [-1, 0, 1, 0, 0, 0], atol=1e-6)
    rng = np.random.default_rng(3)
    for _ in range(20):
        s, d, r = rng.uniform(0, 360), rng.uniform(1, 89), rng.uniform(-180, 180)
        L.orc_angle2moment(s, d, r, orc.p(m))
        M = np.array([[m[0], m[5], m[4]], [m[5], m[1], m[3]], [m[4], m[3], m[2]]])
        ev = np.sort(np.linalg.eigvalsh(M))
        assert np.allclose(ev, [-1, 0, 1], atol=2e-6)   # double couple


def test_source_stage_time_table():
    ni, nj, nk, dt, f0, tc = 30, 30, 30, 0.005, 2.0, 0.3
    ext_i = np.zeros(343, dtype=np.int32)
    ext_c = np.zeros(343)
    itb, ite, en, mnt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    stf = np.zeros(4 * 200)
    rc = L.orc_src_point(ni, nj, nk, 15.0, 15.0, 5.0, 1, 0.0, dt, 0.0123, 0, f0, tc, 0.5,
                         C.byref(itb), C.byref(ite), C.byref(en), orc.p(ext_i), orc.p(ext_c), C.byref(mnt),
                         orc.p(stf))
    assert rc == 0 and mnt.value == 100 and itb.value == 2 and ite.value == 101 and en.value == 343
    # depth 5 from the surface: k = (nk-1) - 5 = 24, local index 27
    nx, ny = ni + 6, nj + 6
    centre = ext_i[171]
    assert centre == 18 + 18 * nx + 27 * nx * ny
    t_shift = 0.0123 - itb.value * dt
    for it in range(0, 100, 7):
        for s, frac in enumerate([0, 0.5, 0.5, 1.0]):
            t = it * dt + frac * dt - t_shift
            tau = t - tc
            exact = (1 - 2 * (math.pi * f0 * tau) ** 2) * math.exp(-(math.pi * f0 * tau) ** 2)
            assert abs(stf[it * 4 + s] - exact) < 5e-6


def test_source_near_boundary_is_clipped_but_normalised_before():
    ni, nj, nk = 20, 20, 20
    ext_i = np.zeros(343, dtype=np.int32)
    ext_c = np.zeros(343)
    itb, ite, en, mnt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    L.orc_src_point(ni, nj, nk, 1.0, 10.0, 10.0, 0, 0.0, 0.005, 0.0, 0, 2.0, 0.3, 0.1,
                    C.byref(itb), C.byref(ite), C.byref(en), orc.p(ext_i), orc.p(ext_c), C.byref(mnt), None)
    assert en.value == 5 * 49          # i = -2, -1 fall outside
    assert 0.8 < ext_c[:en.value].sum() < 1.0


# ---------------------------------------------------------------- receivers, dt
def test_receiver_index_conventions():
    ni, nj, nk = 30, 25, 20
    ix = np.zeros(8, dtype=np.int64)
    fr = np.zeros(3)
    # station.list style: x,y 1-based; z given as depth below the surface
    assert L.orc_recv_index(ni, nj, nk, 10.0, 12.0, 0.0, 1, orc.p(ix), orc.p(fr)) == 0
    nx, ny = ni + 6, nj + 6
    assert ix[0] == (9 + 3) + (11 + 3) * nx + (nk - 1 + 3) * nx * ny
    assert ix[7] == ix[0] + 1 + nx + nx * ny
    assert np.all(fr == 0)
    assert L.orc_recv_index(ni, nj, nk, 10.5, 12.25, 3.0, 0, orc.p(ix), orc.p(fr)) == 0
    assert np.allclose(fr, [0.5, 0.25, 0.0])
    assert ix[0] == (9 + 3) + (11 + 3) * nx + (2 + 3) * nx * ny
    assert L.orc_recv_index(ni, nj, nk, 31.0, 12.0, 0.0, 1, orc.p(ix), orc.p(fr)) != 0


def test_dt_estimate_cube_closed_form():
    ni = nj = nk = 8
    n = (ni + 6) ** 3
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    L.orc_gen_cart(ni, nj, nk, 100.0, 0, 100.0, 0, 100.0, -700.0, orc.p(x), orc.p(y), orc.p(z))
    lam = np.full(n, 3.2044e10)
    mu = np.full(n, 3.2038e10)
    rho = np.full(n, 2670.0)
    dt = L.orc_dt_esti(ni, nj, nk, orc.p(x), orc.p(y), orc.p(z), orc.p(lam), orc.p(mu), orc.p(rho), 1.3)
    vp = math.sqrt((3.2044e10 + 2 * 3.2038e10) / 2670.0)
    assert abs(dt - 1.3 * 100 / math.sqrt(3) / vp) < 1e-6


# ---------------------------------------------------------------- whole loop
def test_oracle_fp32_tracks_fp64():
    P = Problem(ni=24, nj=20, nk=16, curvilinear=True, nt_total=40, seed=0)
    s32, w32, _ = P.oracle_run(32)
    s64, w64, _ = P.oracle_run(64)
    assert np.abs(s64).max() > 0
    assert rel_l2(s32, s64) < 1e-5 and rel_l2(w32, w64) < 1e-5


def test_homogeneous_p_arrival_time():
    """A buried explosion reaches a receiver 10 cells away at r/Vp."""
    dt, vp = 0.002, 6000.0
    P = Problem(ni=40, nj=40, nk=40, free_top=False, pml=(6, 6, 6, 6, 6, 6), source="moment", nrecv=0,
                src_ijk=(14, 20, 20), nt_total=120, dt=dt, f0=12.0, t0=0.1)
    P.src["M"][:] = P.src["M"][0:1]        # isotropic: Mxx=Myy=Mzz
    P.src["M"][3:] = 0
    ix = np.zeros(8, dtype=np.int64)
    fr = np.zeros(3)
    L.orc_recv_index(40, 40, 40, 25.0, 21.0, 21.0, 0, orc.p(ix), orc.p(fr))   # global (24, 20, 20)
    P.recv_indx = ix[None, :]
    P.recv_frac = fr[None, :]
    seis, _, _ = P.oracle_run(64)
    vx = seis[0, 0]
    t_peak = np.argmax(np.abs(vx)) * dt
    r = 10 * 100.0
    # velocity of an explosion peaks near arrival + Ricker centre; allow half a period
    assert abs(t_peak - (0.1 + r / vp)) < 0.5 / 12.0
