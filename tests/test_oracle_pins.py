"""Pins of the oracle's free-surface rows, receiver interpolation and source injection.

Each check is fixed by mathematics, not by the oracle's own formulas:

* traction image (sv_curv_col_el_gpu.cu:19-221): on a grid whose top surface
  is a grid plane, a linear stress field whose traction on that plane vanishes
  is its own odd extension, so the imaged rows must return (1/rho) div T
  exactly; and the rows must not depend on any stress value on or above the
  surface that the image replaces.
* vlow rows (sv_curv_col_el_iso_gpu.cu:462-603): a linear velocity field whose
  Hooke stress has zero traction on the surface satisfies the matrix relation
  of row nk2 exactly and is differentiated exactly by the low-order z
  operators of rows nk2-1 and nk2-2, so those rows must return Hooke's law.
* receivers (io_funcs.cu:1640-1660): trilinear weights reproduce a linear
  field exactly, plus the reference's corner-5 term (b5 - b4) Lx1 Ly1 Lz2.
* source injection (sv_curv_col_el_gpu.cu:227-301): the velocities of an
  explosion, a general moment tensor and a single force in a homogeneous
  whole space against Aki & Richards (4.29)/(4.23), superposed over the
  Gaussian-spread source points, so polarity, 1/J scaling, the component
  mapping and the moment/force units are all fixed.
"""
import ctypes as C

import numpy as np
import pytest

import orc
from analytic import Ricker, force_velocity, moment_velocity
from problem import Problem, rel_l2

L = orc.lib()
DRP_SLOPE = 0.999999999     # sum of c*idx of the DRP tables (fd_t.cu:90-93, 111-114)
MET_SLOPE = 1.00004         # slope of the centred 7-point metric table (fd_t.cu:164-170)
TMAP = {(0, 0): 3, (1, 1): 4, (2, 2): 5, (1, 2): 6, (0, 2): 7, (0, 1): 8}

SHEAR = np.array([[90.0, 12.0, -5.0], [8.0, 110.0, 15.0], [-6.0, 9.0, 95.0]])
ANISO = np.diag([90.0, 110.0, 80.0])


def affine(P, A, metric_precision=64):
    """Replace P's grid by x = A (i, j, k) (ghost-inclusive indices shifted by 3) and redo metrics/matrices."""
    nx, ny, nz = P.nx, P.ny, P.nz
    K, J, I = np.meshgrid(np.arange(nz) - 3.0, np.arange(ny) - 3.0, np.arange(nx) - 3.0, indexing="ij")
    X = np.einsum("ab,bkji->akji", A, np.stack([I, J, K]))
    P.x, P.y, P.z = [np.ascontiguousarray(X[q]) for q in range(3)]
    met = np.zeros(10 * nx * ny * nz)
    L.orc_metric_cal(metric_precision, P.ni, P.nj, P.nk, orc.p(P.x), orc.p(P.y), orc.p(P.z), orc.p(met))
    P.metric = met.reshape(10, nz, ny, nx)
    if P.free_top:
        P.mvx = np.zeros((ny, nx, 9))
        P.mvy = np.zeros((ny, nx, 9))
        L.orc_free_matrices(64, P.ni, P.nj, P.nk, orc.p(P.metric), orc.p(P.lam), orc.p(P.mu),
                            orc.p(P.mvx), orc.p(P.mvy))
    return np.stack([I, J, K]), X


def stage_dirs(ipair, istage):
    d = np.zeros(3, dtype=np.int32)
    L.orc_pair_dirs(ipair, istage, orc.p(d))
    return tuple(int(v) for v in d)


# every (x, y, z) direction combination occurs among these (pair, stage) choices
STAGES = [(p, p % 4) for p in range(8)]


def test_stage_choice_covers_both_z_directions():
    assert {stage_dirs(p, s)[2] for p, s in STAGES} == {0, 1}
    assert len({stage_dirs(p, s) for p, s in STAGES}) == 8


# ---------------------------------------------------------------- traction image
@pytest.mark.parametrize("grid", ["aniso", "shear"])
@pytest.mark.parametrize("ipair,istage", STAGES)
def test_timg_linear_traction_free_stress_gives_divergence(grid, ipair, istage):
    P = Problem(ni=14, nj=13, nk=12, free_top=True, pml=(0,) * 6, source=None, nrecv=0)
    A = ANISO if grid == "aniso" else SHEAR
    xi, X = affine(P, A)
    nk2 = P.nz - 4
    n = np.linalg.inv(A)[2]                      # grad(zeta): surface normal direction
    Pn = np.eye(3) - np.outer(n, n) / (n @ n)
    rng = np.random.default_rng(10 + ipair)
    sym = lambda: (lambda S: S + S.T)(rng.normal(size=(3, 3)))  # noqa: E731
    Sd = [Pn @ sym() @ Pn for _ in range(3)]     # tangential: zero traction on every zeta plane
    U = sym()
    T = np.einsum("dab,dkji->abkji", np.array(Sd), X) * 1e3
    T += U[:, :, None, None, None] * (xi[2] - (nk2 - 3))[None, None] * 1e5   # odd in (k - nk2)
    w = np.zeros((9, P.nz, P.ny, P.nx))
    for (a, b), c in TMAP.items():
        w[c] = T[a, b]
    w[:, nk2 + 1:] = rng.normal(size=w[:, nk2 + 1:].shape) * 1e9   # above the surface: garbage
    rhs, _ = P.oracle_stage(w, ipair, istage, 0, precision=64)
    div = np.einsum("dab,db->a", np.array(Sd), np.eye(3)) * 1e3 + U @ n * 1e5
    want = P.slw[0, 0, 0] * DRP_SLOPE / MET_SLOPE * div
    phys = (slice(3, nk2 + 1), slice(3, P.ny - 3), slice(3, P.nx - 3))
    for c in range(3):
        got = rhs[c][phys]
        assert np.allclose(got, want[c], rtol=0, atol=1e-6 * np.abs(want).max()), (c, np.abs(got - want[c]).max())
    for c in range(3, 9):
        assert np.abs(rhs[c][phys]).max() == 0.0


@pytest.mark.parametrize("ipair,istage", STAGES)
def test_timg_ignores_stress_on_and_above_the_surface(ipair, istage):
    """The image sets the zeta-flux to 0 at the surface and mirrors it above: those values are never read.

    On this Cartesian grid Tzz enters only the zeta-flux; Txz and Tyz on the surface row still enter
    the x/y derivatives of that row, so only Tzz is changed there (and every stress above it).
    """
    P = Problem(ni=14, nj=13, nk=12, free_top=True, pml=(0,) * 6, source=None, nrecv=0)
    affine(P, ANISO)
    nk2 = P.nz - 4
    rng = np.random.default_rng(3 + istage)
    w = np.zeros((9, P.nz, P.ny, P.nx))
    w[3:, 3:nk2 + 1, 3:-3, 3:-3] = rng.normal(size=(6, nk2 - 2, P.nj, P.ni)) * 1e6
    r0, _ = P.oracle_stage(w, ipair, istage, 0, precision=64)
    w2 = w.copy()
    w2[3:, nk2 + 1:] = rng.normal(size=w2[3:, nk2 + 1:].shape) * 1e6        # all stresses above
    w2[5, nk2] = rng.normal(size=w2[5, nk2].shape) * 1e6                   # Tzz on it (zeta-flux only)
    r1, _ = P.oracle_stage(w2, ipair, istage, 0, precision=64)
    scale = np.abs(r0[:3]).max()
    assert scale > 0
    assert np.abs(r1[:3] - r0[:3]).max() <= 1e-12 * scale
    # and a stress value the image does read (Txx on the surface row) does change it
    w3 = w.copy()
    w3[3, nk2] += rng.normal(size=w3[3, nk2].shape) * 1e6
    r3, _ = P.oracle_stage(w3, ipair, istage, 0, precision=64)
    assert np.abs(r3[0, nk2] - r0[0, nk2]).max() > 1e-3 * scale


# ---------------------------------------------------------------- vlow rows
def _hooke(G, lam, mu):
    return lam * np.trace(G) * np.eye(3) + mu * (G + G.T)


def _traction_free_gradient(n, lam, mu, rng):
    """A random velocity gradient G (dV_a/dx_b) whose Hooke stress has zero traction on the plane n."""
    Cm = np.zeros((3, 9))
    for q in range(9):
        E = np.zeros(9)
        E[q] = 1
        Cm[:, q] = _hooke(E.reshape(3, 3), lam, mu) @ n
    g0 = rng.normal(size=9)
    g = g0 - np.linalg.pinv(Cm) @ (Cm @ g0)
    return g.reshape(3, 3)


@pytest.mark.parametrize("grid", ["aniso", "shear"])
@pytest.mark.parametrize("ipair,istage", STAGES)
def test_vlow_traction_free_linear_velocity_gives_hooke(grid, ipair, istage):
    P = Problem(ni=14, nj=13, nk=12, free_top=True, pml=(0,) * 6, source=None, nrecv=0)
    A = ANISO if grid == "aniso" else SHEAR
    _, X = affine(P, A)
    nk2 = P.nz - 4
    lam, mu = P.lam[0, 0, 0], P.mu[0, 0, 0]
    n = np.linalg.inv(A)[2]
    rng = np.random.default_rng(20 + ipair)
    G = _traction_free_gradient(n, lam, mu, rng) * 1e-3
    assert np.abs(_hooke(G, lam, mu) @ n).max() < 1e-9 * np.abs(_hooke(G, lam, mu)).max()
    w = np.zeros((9, P.nz, P.ny, P.nx))
    w[:3] = np.einsum("ab,bkji->akji", G, X)
    w[:, nk2 + 1:] = rng.normal(size=w[:, nk2 + 1:].shape) * 1e3            # above the surface: garbage
    rhs, _ = P.oracle_stage(w, ipair, istage, 0, precision=64)
    sig = _hooke(G * DRP_SLOPE / MET_SLOPE, lam, mu)
    tol = 1e-6 * np.abs(sig).max()
    for row in (nk2, nk2 - 1, nk2 - 2, nk2 - 3, 8):
        for (a, b), c in TMAP.items():
            got = rhs[c, row, 3:-3, 3:-3]
            assert np.allclose(got, sig[a, b], rtol=0, atol=tol), (row, c, np.abs(got - sig[a, b]).max())
    assert np.abs(rhs[:3, 3:nk2 + 1]).max() == 0.0


def test_vlow_surface_row_uses_the_zero_traction_matrices():
    """A gradient with non-zero surface traction is NOT reproduced on row nk2 (the matrices replace Dz)."""
    P = Problem(ni=14, nj=13, nk=12, free_top=True, pml=(0,) * 6, source=None, nrecv=0)
    _, X = affine(P, ANISO)
    nk2 = P.nz - 4
    lam, mu = P.lam[0, 0, 0], P.mu[0, 0, 0]
    G = np.diag([0.0, 0.0, 1e-3])                 # pure dVz/dz: traction Tzz = (lam + 2 mu) G33
    w = np.zeros((9, P.nz, P.ny, P.nx))
    w[:3] = np.einsum("ab,bkji->akji", G, X)
    rhs, _ = P.oracle_stage(w, 0, 0, 0, precision=64)
    # row nk2: Dz V from the matrices of zero x/y gradients is 0, so hTzz = 0; one row down it is exact
    assert np.abs(rhs[5, nk2, 3:-3, 3:-3]).max() < 1e-9 * (lam + 2 * mu) * 1e-3
    assert np.allclose(rhs[5, nk2 - 1, 3:-3, 3:-3], (lam + 2 * mu) * 1e-3 / MET_SLOPE, rtol=1e-6)


# ---------------------------------------------------------------- receivers
def test_receiver_trilinear_with_corner5_term():
    ni, nj, nk = 20, 18, 16
    P = Problem(ni=ni, nj=nj, nk=nk, free_top=False, pml=(0,) * 6, source=None, nrecv=0)
    K, J, I = np.meshgrid(np.arange(P.nz), np.arange(P.ny), np.arange(P.nx), indexing="ij")
    rng = np.random.default_rng(5)
    a, bx, by, bz = rng.normal(size=(4, 9)) * np.array([1, 0.1, 0.2, 0.3])[:, None]
    w = a[:, None, None, None] + bx[:, None, None, None] * I + by[:, None, None, None] * J \
        + bz[:, None, None, None] * K
    P.init = w
    stations = [(5.0, 6.0, 3.0, 0), (5.25, 6.5, 3.75, 0), (10.6, 4.3, 8.2, 0), (12.9, 11.1, 2.5, 1),
                (3.5, 15.75, 6.125, 1), (7.0, 7.0, 5.5, 0)]
    rows_i, rows_f, want = [], [], []
    for rx, ry, rz, dflag in stations:
        ix = np.zeros(8, dtype=np.int64)
        fr = np.zeros(3)
        assert L.orc_recv_index(ni, nj, nk, rx, ry, rz, dflag, orc.p(ix), orc.p(fr)) == 0
        rows_i.append(ix)
        rows_f.append(fr)
        k0, rem = divmod(int(ix[0]), P.nx * P.ny)
        j0, i0 = divmod(rem, P.nx)
        fx, fy, fz = [float(np.float32(v)) for v in fr]
        exact = a + bx * (i0 + fx) + by * (j0 + fy) + bz * (k0 + fz)
        want.append(exact + bx * (1 - fx) * (1 - fy) * fz)     # corner 5 counted for corner 4
    P.recv_indx = np.array(rows_i)
    P.recv_frac = np.array(rows_f)
    seis, _, _ = P.oracle_run(64, nt_total=1)
    got = seis[:, :, 0]
    want = np.array(want)
    assert np.allclose(got, want, rtol=0, atol=1e-9 * np.abs(want).max())
    # the corner-5 term is really there (fz != 0 stations differ from plain trilinear by it)
    assert np.abs(got[1] - (want[1] - bx * (1 - 0.25) * (1 - 0.5) * 0.75)).max() > 1e-3


# ---------------------------------------------------------------- source injection vs Aki & Richards
def _whole_space_run(kind, M=None, F=None, stations=(), f0=1.5, tc=0.8, nt=230, dt=0.01, n=64, nlay=10):
    src = (n // 2 - 1, n // 2 - 1, n // 2 - 1)
    P = Problem(ni=n, nj=n, nk=n, free_top=False, pml=(nlay,) * 6, source=kind, nrecv=0, src_ijk=src,
                nt_total=nt, dt=dt, f0=f0, t0=tc, stf_len=nt * dt)
    stf = np.float32(P.src["M"][0] / np.float32(1.0e16)) if kind == "moment" else \
        np.float32(P.src["F"][0] / np.float32(1.0e12))
    if kind == "moment":
        m6 = np.array([M[0, 0], M[1, 1], M[2, 2], M[1, 2], M[0, 2], M[0, 1]])
        P.src["M"] = np.stack([stf * m for m in m6])
    else:
        P.src["F"] = np.stack([stf * f for f in F])
    rows_i, rows_f = [], []
    for (i, j, k) in stations:     # 0-based global physical indices -> 1-based station-file indices
        ix = np.zeros(8, dtype=np.int64)
        fr = np.zeros(3)
        assert L.orc_recv_index(n, n, n, i + 1.0, j + 1.0, k + 1.0, 0, orc.p(ix), orc.p(fr)) == 0
        rows_i.append(ix)
        rows_f.append(fr)
    P.recv_indx = np.array(rows_i)
    P.recv_frac = np.array(rows_f)
    seis, _, _ = P.oracle_run(64)
    return P, seis


def _analytic(P, kind, M, F, stations, f0, tc, nt, dt):
    rho = P.rho[0, 0, 0]
    alpha = np.sqrt((P.lam[0, 0, 0] + 2 * P.mu[0, 0, 0]) / rho)
    beta = np.sqrt(P.mu[0, 0, 0] / rho)
    stf = Ricker(f0, tc)
    t = np.arange(nt) * dt
    s = P.src
    pts = []
    for e in range(int(s["ext_num"][0])):
        q = int(s["ext_indx"][e])
        k, rem = divmod(q, P.nx * P.ny)
        j, i = divmod(rem, P.nx)
        pts.append(((P.x[k, j, i], P.y[k, j, i], P.z[k, j, i]), float(s["ext_coef"][e])))
    out = []
    for (i, j, k) in stations:
        r = (P.x[k + 3, j + 3, i + 3], P.y[k + 3, j + 3, i + 3], P.z[k + 3, j + 3, i + 3])
        v = np.zeros((3, nt))
        for xs, wgt in pts:
            v += wgt * (moment_velocity(M, xs, r, t, alpha, beta, rho, stf) if kind == "moment"
                        else force_velocity(F, xs, r, t, alpha, beta, rho, stf))
        out.append(v)
    return np.array(out)


def test_ricker_antiderivatives():
    """The closed-form integrals used by the whole-space solutions, against quadrature."""
    stf = Ricker(1.5, 0.8)
    s = np.linspace(-2.0, 4.0, 600001)
    h = s[1] - s[0]
    cum = np.concatenate([[0], np.cumsum(0.5 * (stf.R(s[1:]) + stf.R(s[:-1])) * h)])
    assert np.allclose(cum, stf.Q(s) - stf.Q(s[0]), atol=1e-9)
    for t in (0.5, 1.2, 2.0):
        ta, tb = 0.3, 0.55
        tau = np.linspace(ta, tb, 20001)
        assert abs(np.trapezoid(tau * stf.R(t - tau), tau) - stf.near_R(t, ta, tb)) < 1e-9
        assert abs(np.trapezoid(tau * stf.dR(t - tau), tau) - stf.near_dR(t, ta, tb)) < 1e-7
    # R' is the derivative of R
    assert np.allclose((stf.R(s[2:]) - stf.R(s[:-2])) / (2 * h), stf.dR(s[1:-1]), atol=1e-6)


@pytest.fixture(scope="module")
def explosion():
    M = np.eye(3) * 1e16
    st = [(51, 31, 31), (31, 31, 11), (45, 45, 31)]          # 20 cells along +x, -z, and a diagonal
    P, seis = _whole_space_run("moment", M=M, stations=st)
    want = _analytic(P, "moment", M, None, st, 1.5, 0.8, 230, 0.01)
    return P, seis, want, st


def test_explosion_matches_aki_richards(explosion):
    P, seis, want, st = explosion
    got = seis[:, :3, :]
    for r in range(len(st)):
        e = rel_l2(got[r], want[r])
        assert e < 0.04, (st[r], e)         # measured 0.1% on the axes, 2.5% on the diagonal
        # radial motion: first the outward push (positive along the source-receiver direction)
        d = np.array(st[r], float) - 31
        rad_g = (d / np.linalg.norm(d)) @ got[r]
        rad_w = (d / np.linalg.norm(d)) @ want[r]
        ig = int(np.argmax(np.abs(rad_w)))
        assert np.sign(rad_g[ig]) == np.sign(rad_w[ig])
        assert 0.97 < np.abs(rad_g).max() / np.abs(rad_w).max() < 1.03


def test_explosion_polarity_and_scale_slips_fail(explosion):
    """The check above is sharp: a flipped sign, a dropped 1/J or a wrong density power fails it."""
    P, seis, want, st = explosion
    got = seis[:, :3, :]
    J = P.metric[0, 10, 10, 10]
    for bad in (-got, got * J, got / 2670.0):
        assert rel_l2(bad[0], want[0]) > 0.5


def test_moment_tensor_components_match_aki_richards():
    rng = np.random.default_rng(7)
    M = rng.normal(size=(3, 3))
    M = (M + M.T) * 0.5e16
    st = [(51, 31, 31), (31, 47, 17), (17, 19, 44), (45, 17, 45)]
    P, seis = _whole_space_run("moment", M=M, stations=st)
    want = _analytic(P, "moment", M, None, st, 1.5, 0.8, 230, 0.01)
    for r in range(len(st)):
        assert rel_l2(seis[r, :3], want[r]) < 0.03, (st[r], rel_l2(seis[r, :3], want[r]))   # measured <= 1.3%
    # swapping the Myz and Mxz slots of the tables is caught
    Msw = M.copy()
    Msw[1, 2], Msw[0, 2] = M[0, 2], M[1, 2]
    Msw[2, 1], Msw[2, 0] = Msw[1, 2], Msw[0, 2]
    bad = _analytic(P, "moment", Msw, None, st, 1.5, 0.8, 230, 0.01)
    assert max(rel_l2(seis[r, :3], bad[r]) for r in range(len(st))) > 0.2


def test_single_force_matches_aki_richards():
    F = np.array([1.0e12, -0.5e12, 2.0e12])
    st = [(51, 31, 31), (31, 31, 11), (17, 45, 40)]
    P, seis = _whole_space_run("force", F=F, stations=st)
    want = _analytic(P, "force", None, F, st, 1.5, 0.8, 230, 0.01)
    for r in range(len(st)):
        e = rel_l2(seis[r, :3], want[r])
        assert e < 0.01, (st[r], e)         # measured <= 0.22%
        assert rel_l2(-seis[r, :3], want[r]) > 1.5


# ---------------------------------------------------------------- peak ground motion
def _peak(vtop, dt):
    nt = vtop.shape[0] - 1
    npts = vtop.shape[2]
    pg = np.zeros(15 * npts)
    L.orc_peak_motion(npts, nt, orc.p(np.ascontiguousarray(vtop, dtype=np.float64)), dt, orc.p(pg))
    return pg.reshape(15, npts)


def test_peak_motion_linear_velocity_closed_form():
    """V(t) = a + b t on every point: |dV|/dt = |b| exactly, the trapezoid rule
    integrates a linear history exactly (D = a t + b t^2 / 2), and the maxima
    are those of the sampled values (wav_t.cu:281-360)."""
    rng = np.random.default_rng(13)
    npts, nt, dt = 7, 40, 0.01
    a = rng.normal(size=(3, npts))
    b = rng.normal(size=(3, npts)) * 20
    b[:, 0] = 0.0                                   # a constant velocity: zero acceleration
    t = np.arange(nt + 1) * np.float32(dt)
    vtop = np.float32(a[None] + b[None] * t[:, None, None])
    pg = _peak(vtop, dt)
    V = np.float64(vtop[1:])                         # maxima over V^1 .. V^nt (after each step)
    D = a[None] * t[1:, None, None] + 0.5 * b[None] * t[1:, None, None] ** 2
    want = [np.sqrt((V ** 2).sum(1)).max(0), np.sqrt((V[:, :2] ** 2).sum(1)).max(0), *np.abs(V).max(0),
            np.full(npts, 0.0) + np.sqrt((b ** 2).sum(0)), np.sqrt((b[:2] ** 2).sum(0)), *np.abs(b),
            np.sqrt((D ** 2).sum(1)).max(0), np.sqrt((D[:, :2] ** 2).sum(1)).max(0), *np.abs(D).max(0)]
    want = np.array(want)
    # float arithmetic: differences of V in float give |dV|/dt to ~1e-5 relative
    assert np.allclose(pg, want, rtol=2e-4, atol=1e-5)
    assert np.all(pg[5:10, 0] < 1e-4)                # the constant column has no acceleration


def test_peak_motion_keeps_running_maxima():
    """A velocity pulse that rises and falls: the maps keep the peak, not the last value."""
    npts, nt, dt = 3, 60, 0.005
    t = np.arange(nt + 1) * dt
    pulse = np.exp(-((t - 0.15) / 0.03) ** 2)
    vtop = np.zeros((nt + 1, 3, npts))
    vtop[:, 2, :] = pulse[:, None] * np.array([1.0, -2.0, 0.5])
    pg = _peak(vtop, dt)
    assert np.allclose(pg[4], np.abs(vtop[1:, 2]).max(0), rtol=1e-6)   # PGVz
    assert np.allclose(pg[0], pg[4], rtol=1e-6)                        # |V| = |Vz|
    assert np.all(pg[1] == 0) and np.all(pg[2] == 0)                   # no horizontal motion
    # the displacement of a pulse that ends near zero velocity: its integral, the area of the Gaussian
    area = 0.03 * np.sqrt(np.pi) * np.array([1.0, 2.0, 0.5])
    assert np.allclose(pg[14], area, rtol=2e-3)
