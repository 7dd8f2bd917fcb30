"""Synthetic el_iso workloads, prepared with the library's FP64 host setup.

small    : 100x100x60 Cartesian, 100 m, flat free surface, 10-layer CFS-PML on
           the other faces, homogeneous Vp 6000 / Vs 3464 / rho 2670, explosive
           source (Mxx=Myy=Mzz=1e16 N m) at the centre, 3 km deep, Ricker 2 Hz
           t0 0.6 s, dt 0.005 s, 500 steps, 20 seeded surface receivers.
topo     : 400x400x200 curvilinear: surface = 3 Gaussian hills (300-800 m,
           sigma 2-5 km, seeded) + Fourier roughness (rms 50 m, wavelengths
           > 1 km), vertical spacing 100 m stretched to 150 m at depth;
           3 flat layers Vp 3000/5000/6500, Vs = Vp/1.73, Gardner density,
           each with a Gaussian-correlated +-5% perturbation (1 km); seeded
           double couple, Ricker 1.5 Hz; 1000 steps; 400 receivers.
Every model array is returned in the reference host layout (nx = ni+6).
"""
import numpy as np

import cgfd

PML = dict(nlay=10, amax=3.0, bmax=2.5)


def _pml_faces(m, ni, nj, nk, x, y, z, vel, faces=(0, 1, 2, 3, 4)):
    out = {}
    for f in faces:
        L0, dh = cgfd.pml_len_dh(ni, nj, nk, x, y, z, f, PML["nlay"])
        out[f] = (PML["nlay"],) + cgfd.pml_profile(PML["nlay"], f, L0, dh, vel, PML["amax"], PML["bmax"])
    return out


def _surface_receivers(ni, nj, nk, n, seed, margin):
    rng = np.random.default_rng(seed)
    cells = set()
    while len(cells) < n:
        cells.add((int(rng.integers(margin, ni - margin)), int(rng.integers(margin, nj - margin))))
    idx, frac = [], []
    for (i, j) in sorted(cells):
        ix, fr = cgfd.receiver_index(ni, nj, nk, i + 1.0, j + 1.0, 0.0, 1)
        idx.append(ix)
        frac.append(fr)
    return np.array(idx), np.array(frac)


def _source(ni, nj, nk, sx, sy, depth_idx, dt, f0, t0, m6, nt_total, wavelet_tstart=0.0):
    s = cgfd.source_point(ni, nj, nk, sx, sy, depth_idx, 1, 0.0, dt, wavelet_tstart, 0, f0, t0, nt_total * dt)
    mnt = s["max_nt"]
    M = np.stack([s["stf"] * m for m in m6]).reshape(6, 1, mnt, 4)
    return dict(nsrc=1, force_on=0, moment_on=1, max_nt=mnt, max_ext=343,
                it_begin=np.array([s["it_begin"]]), it_end=np.array([s["it_end"]]),
                ext_num=np.array([s["ext_num"]]), ext_indx=s["ext_indx"], ext_coef=s["ext_coef"], F=None, M=M)


def small(seed=1, nt=500, ni=100, nj=100, nk=60):
    """The 100x100x60 explosion model; other sizes keep its physics, with the
    source at the horizontal centre and 3 km deep."""
    h = 100.0
    x, y, z = cgfd.grid_cartesian(ni, nj, nk, h, 0.0, h, 0.0, h, -(nk - 1) * h)
    metric = cgfd.metric_calc(ni, nj, nk, x, y, z)
    lam, mu, rho = cgfd.media_homogeneous(ni, nj, nk, 6000.0, 3464.0, 2670.0)
    m = dict(name="small" if (ni, nj, nk) == (100, 100, 60) else f"explosion{ni}x{nj}x{nk}", ni=ni, nj=nj, nk=nk,
             dt=0.005, nt_total=nt + 1, x=x, y=y, z=z, metric=metric,
             lam=lam, mu=mu, rho=rho)
    m["dtmax"] = cgfd.dt_estimate(ni, nj, nk, x, y, z, lam, mu, rho)
    m["slw"] = cgfd.rho_to_slow(rho)
    m["mvx"], m["mvy"] = cgfd.free_surface_matrices(ni, nj, nk, metric, lam, mu)
    m["pml"] = _pml_faces(m, ni, nj, nk, x, y, z, 6000.0)
    m["src_params"] = dict(sx=ni / 2, sy=nj / 2, depth=30.0, f0=2.0, t0=0.6, m6=[1e16, 1e16, 1e16, 0, 0, 0],
                           pml_vel=6000.0)
    m["src"] = _source(ni, nj, nk, ni / 2, nj / 2, 30.0, m["dt"], 2.0, 0.6, [1e16, 1e16, 1e16, 0, 0, 0],
                       m["nt_total"])
    m["recv_indx"], m["recv_frac"] = _surface_receivers(ni, nj, nk, 20, seed, PML["nlay"] + 1)
    return m


def topo(seed=2, nt=1000, ni=400, nj=400, nk=200):
    h = 100.0
    x, y, z = cgfd.grid_topography(ni, nj, nk, h, 1.5, 3, (300.0, 800.0), (2000.0, 5000.0), 50.0, 1000.0, seed)
    metric = cgfd.metric_calc(ni, nj, nk, x, y, z)
    vp = np.array([3000.0, 5000.0, 6500.0])
    lam, mu, rho = cgfd.media_layered(ni, nj, nk, x, y, z, np.array([-2000.0, -8000.0]), vp, np.full(3, 1.73),
                                      np.zeros(3), 0.025, 0.05, 1000.0, seed + 100)
    m = dict(name="topo", ni=ni, nj=nj, nk=nk, dt=0.005, nt_total=nt + 1, x=x, y=y, z=z, metric=metric,
             lam=lam, mu=mu, rho=rho)
    m["dtmax"] = cgfd.dt_estimate(ni, nj, nk, x, y, z, lam, mu, rho)
    m["slw"] = cgfd.rho_to_slow(rho)
    m["mvx"], m["mvy"] = cgfd.free_surface_matrices(ni, nj, nk, metric, lam, mu)
    m["pml"] = _pml_faces(m, ni, nj, nk, x, y, z, float(vp.max() * 1.05))
    rng = np.random.default_rng(seed)
    strike, dip, rake = rng.uniform(0, 360), rng.uniform(20, 80), rng.uniform(-180, 180)
    m["mech"] = (strike, dip, rake)
    m6 = cgfd.angle2moment(strike, dip, rake) * 1e16
    m["src_params"] = dict(sx=ni / 2, sy=nj / 2, depth=30.0, f0=1.5, t0=0.8, m6=list(m6),
                           pml_vel=float(vp.max() * 1.05))
    m["src"] = _source(ni, nj, nk, ni / 2, nj / 2, 30.0, m["dt"], 1.5, 0.8, m6, m["nt_total"])
    m["recv_indx"], m["recv_frac"] = _surface_receivers(ni, nj, nk, 400, seed, PML["nlay"] + 1)
    return m


def upload(sess, m, shots=None, nt_total=None):
    """Upload a workload into a session (model once, source and receivers per shot)."""
    f = np.float64 if sess.storage == cgfd.STORE_F64 else np.float32
    sess.upload_metric(np.asarray(m["metric"], f))
    sess.upload_medium(np.asarray(m["lam"], f), np.asarray(m["mu"], f), np.asarray(m["slw"], f))
    sess.upload_free_surface(True, np.asarray(m["mvx"], f), np.asarray(m["mvy"], f))
    for face, (nl, A, B, D) in m["pml"].items():
        sess.upload_pml(face, nl, np.asarray(A, f), np.asarray(B, f), np.asarray(D, f))
    nt = m["nt_total"] if nt_total is None else nt_total
    for shot in (range(sess.nshots) if shots is None else shots):
        sess.upload_source(shot, m["src"])
        sess.upload_receivers(shot, m["recv_indx"], np.asarray(m["recv_frac"], f), nt)


def cells(m):
    return m["ni"] * m["nj"] * m["nk"]


def interior_cells(m, free_top=True):
    """Points of the box inside every PML face box (the sweep's set when the
    per-point kernel takes the PML slabs)."""
    lay = PML["nlay"] + 1
    return (m["ni"] - 2 * lay) * (m["nj"] - 2 * lay) * (m["nk"] - lay - (4 if free_top else 0))


def sweep_cells(m, free_top=True):
    """Points of the default z-marching sweep: whole planes from the bottom to
    the last plane below the 4 free-surface rows."""
    return m["ni"] * m["nj"] * (m["nk"] - (4 if free_top else 0))


def sweep_pml_pairs(m, free_top=True):
    """(point, PML face) pairs inside the default sweep: each carries nine
    memory variables (faces x1, x2, y1, y2 and z1 of nlay+1 points each)."""
    lay = PML["nlay"] + 1
    nz = m["nk"] - (4 if free_top else 0)
    return 2 * lay * m["nj"] * nz + 2 * lay * m["ni"] * nz + lay * m["ni"] * m["nj"]


def shot_requests(m, n, seed, nt_range=(500, 2000), f0_range=(1.0, 4.0), depth_m=(1000.0, 10000.0)):
    """Seeded shot requests on a model: random epicentre inside the PML frame,
    depth in depth_m clamped to [1 km, bottom - PML - 3 cells] (index units of
    the top spacing), random moment tensor of scale 1e16 N m, Ricker f0 in
    f0_range with t0 = 1.2 / f0, and a step count in nt_range."""
    rng = np.random.default_rng(seed)
    ni, nj, nk = m["ni"], m["nj"], m["nk"]
    lay = PML["nlay"] + 1
    z = m["z"]
    dz = float(z[nk + 2, 3, 3] - z[nk + 1, 3, 3])  # top spacing
    kmax = nk - 1 - lay - 3
    out = []
    for _ in range(n):
        sx = float(rng.uniform(lay + 3, ni - lay - 4))
        sy = float(rng.uniform(lay + 3, nj - lay - 4))
        d = float(np.clip(rng.uniform(*depth_m) / dz, 1000.0 / dz, kmax))
        f0 = float(rng.uniform(*f0_range))
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        m6 = rng.normal(size=6) * 1e16
        src = _source(ni, nj, nk, sx, sy, d, m["dt"], f0, 1.2 / f0, m6, nt)
        out.append(dict(src=src, nt=nt, f0=f0, depth_idx=d, epicentre=(sx, sy)))
    return out


def model_arrays(m, dtype=np.float32):
    """Keyword arguments of cgfd.Server.add_model for a workload."""
    f = dtype
    return dict(metric=np.asarray(m["metric"], f), lam=np.asarray(m["lam"], f), mu=np.asarray(m["mu"], f),
                slw=np.asarray(m["slw"], f), free_top=True, mvx=np.asarray(m["mvx"], f),
                mvy=np.asarray(m["mvy"], f), pml={k: (nl, np.asarray(A, f), np.asarray(B, f), np.asarray(D, f))
                                                  for k, (nl, A, B, D) in m["pml"].items()})
