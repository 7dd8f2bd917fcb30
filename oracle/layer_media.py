"""layer_media.py -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's layered medium for elastic_isotropic,
"loc" method, the checker of cgfd_media_layer_file (the product's C++ version
in gpu-cgfd3d-elastic_amd/csrc/cgfd_setup.cpp).  Only tests/ import it.

Follows, in float32 with the reference's promotions:
  src/media/media_read_file.cpp:22-260      interface file (elastic_isotropic rows)
  src/media/media_layer2model.cpp:401-533   AssignLayerMediaPara2Point, CalPointValue_layer
  src/media/media_layer2model.cpp:1313-1362 parametrization_layer_el_iso_loc
  src/media/media_utility.cpp:25-180        isEqual, findLast/FirstGreaterEqualIndex,
                                            BilinearInterpolation
isEqual is `abs(a - b) < FLT_EPSILON` with C's int abs, i.e. |a - b| < 1.
"""
import numpy as np

f32 = np.float32
FLT_MAX = float(np.finfo(np.float32).max)


def read(path):
    toks = []
    for line in open(path):
        if line.startswith("#") or line == "\n":
            continue
        toks += line.split()
    assert toks[0] == "elastic_isotropic"
    ni, nx, ny = int(toks[1]), int(toks[2]), int(toks[3])
    minx, miny, dx, dy = (f32(v) for v in toks[4:8])
    vals = np.array(toks[8:8 + ni * nx * ny * 10], dtype=np.float32).reshape(ni, ny, nx, 10)
    return dict(ni=ni, nx=nx, ny=ny, minx=minx, miny=miny, dx=dx, dy=dy, v=vals)


def _bilinear(L, xv, yv, f, xq, yq):
    ix = int(f32(xq - xv[0]) / f32(xv[1] - xv[0]))
    iy = int(f32(yq - yv[0]) / f32(yv[1] - yv[0]))
    if ix < 0 or iy < 0:
        return f32(FLT_MAX)
    if ix == L["nx"] - 1 or iy == L["ny"] - 1:
        return f[iy, ix]
    area = f32(f32(xv[ix + 1] - xv[ix]) * f32(yv[iy + 1] - yv[iy]))
    a = f32(f32(f[iy, ix] * f32(xv[ix + 1] - xq)) * f32(yv[iy + 1] - yq))
    b = f32(f32(f[iy, ix + 1] * f32(xq - xv[ix])) * f32(yv[iy + 1] - yq))
    c = f32(f32(f[iy + 1, ix] * f32(xv[ix + 1] - xq)) * f32(yq - yv[iy]))
    d = f32(f32(f[iy + 1, ix + 1] * f32(xq - xv[ix])) * f32(yq - yv[iy]))
    return f32(f32(f32(f32(a + b) + c) + d) / area)


def _last_ge(value, x):
    if value > x[0]:
        return -1
    dist, idx = f32(FLT_MAX), -1
    for i, xi in enumerate(x):
        nd = f32(xi - value)
        if nd < 0:
            break
        if nd <= dist:
            dist, idx = nd, i
    return idx


def _first_ge(value, x):
    dist, idx = f32(FLT_MAX), -1
    for i, xi in enumerate(x):
        nd = f32(xi - value)
        if nd >= 0 and nd <= dist:
            if nd <= 1e-6:
                return i
            dist, idx = nd, i
    return idx


def media(ni, nj, nk, x, y, z, path):
    """lam, mu, rho (float64 arrays of the grid's shape) as cgfd.media_layer_file."""
    L = read(path)
    xv = [f32(L["minx"] + L["dx"] * i) for i in range(L["nx"])]
    yv = [f32(L["miny"] + L["dy"] * j) for j in range(L["ny"])]
    V = L["v"]
    X, Y, Z = (np.asarray(a, np.float64).astype(f32) for a in (x, y, z))
    lam, mu, rho = (np.zeros(X.shape) for _ in range(3))
    col = dict(elev=0, rho=1, rho_g=2, rho_p=3, vp=4, vp_g=5, vp_p=6, vs=7, vs_g=8, vs_p=9)
    maxx, maxy = f32(L["minx"] + (L["nx"] - 1) * L["dx"]), f32(L["miny"] + (L["ny"] - 1) * L["dy"])
    for p in np.ndindex(X.shape):
        ax, ay, az = X[p], Y[p], Z[p]
        if ax < L["minx"] or ax > maxx or ay < L["miny"] or ay > maxy:
            raise ValueError("grid point outside the interface mesh")
        fld = lambda name, l: _bilinear(L, xv, yv, V[l, :, :, col[name]], ax, ay)
        elev = [fld("elev", l) for l in range(L["ni"])]
        mi = _last_ge(az, elev)

        def at(l, dz):
            out = []
            for nm in ("rho", "vp", "vs"):
                out.append(f32(float(fld(nm, l)) + float(dz) ** float(fld(nm + "_p", l)) * float(fld(nm + "_g", l))))
            return out

        if mi == -1:
            mi = _last_ge(elev[0], elev)
            r, vp, vs = fld("rho", mi), fld("vp", mi), fld("vs", mi)
        else:
            dz = f32(elev[mi] - az)
            r, vp, vs = at(mi, dz)
            assert vp > float(vs) * np.sqrt(2.0)
            if abs(int(dz)) < 1:  # the reference's isEqual(dz, 0)
                up = _first_ge(az, elev) - 1
                if up >= 0:
                    r2, p2, s2 = at(up, f32(elev[up] - az))
                    r = f32(float(f32(r + r2)) / 2.0)
                    vp = f32(float(f32(vp + p2)) / 2.0)
                    vs = f32(float(f32(vs + s2)) / 2.0)
        m = f32(f32(vs * vs) * r)
        mu[p] = m
        lam[p] = f32(float(f32(f32(vp * vp) * r)) - 2.0 * float(m))
        rho[p] = r
    return lam, mu, rho


# ---- "har" / "ari" (media_layer2model.cpp:1963-2174; media_utility.cpp:404-600;
# MarkInterfaceNumber, media_layer2model.cpp:1494-1580), vectorized over points

def _bilinear_v(L, xv, yv, f, xq, yq):
    xv0, xv1, yv0, yv1 = xv[0], xv[1], yv[0], yv[1]
    ix = np.trunc((xq - xv0).astype(f32) / f32(xv1 - xv0)).astype(np.int64)
    iy = np.trunc((yq - yv0).astype(f32) / f32(yv1 - yv0)).astype(np.int64)
    nx, ny = L["nx"], L["ny"]
    XV, YV = np.array(xv, f32), np.array(yv, f32)
    edge = (ix == nx - 1) | (iy == ny - 1)
    i0, j0 = np.clip(ix, 0, nx - 2), np.clip(iy, 0, ny - 2)
    x0, x1, y0, y1 = XV[i0], XV[i0 + 1], YV[j0], YV[j0 + 1]
    area = ((x1 - x0) * (y1 - y0)).astype(f32)
    a = (f[j0, i0] * (x1 - xq)).astype(f32) * (y1 - yq).astype(f32)
    b = (f[j0, i0 + 1] * (xq - x0)).astype(f32) * (y1 - yq).astype(f32)
    c = (f[j0 + 1, i0] * (x1 - xq)).astype(f32) * (yq - y0).astype(f32)
    d = (f[j0 + 1, i0 + 1] * (xq - x0)).astype(f32) * (yq - y0).astype(f32)
    v = ((((a + b).astype(f32) + c).astype(f32) + d).astype(f32) / area).astype(f32)
    v = np.where(edge, f[np.minimum(iy, ny - 1), np.minimum(ix, nx - 1)], v)
    return np.where((ix < 0) | (iy < 0), f32(FLT_MAX), v).astype(f32)


def _last_ge_v(value, E):
    n = value.shape[0]
    idx = np.full(n, -1)
    dist = np.full(n, FLT_MAX, f32)
    live = ~(value > E[0])
    for i in range(E.shape[0]):
        nd = (E[i] - value).astype(f32)
        live &= ~(nd < 0)
        upd = live & (nd <= dist)
        dist = np.where(upd, nd, dist)
        idx = np.where(upd, i, idx)
    return idx


def _first_ge_v(value, E):
    n = value.shape[0]
    idx = np.full(n, -1)
    dist = np.full(n, FLT_MAX, f32)
    done = np.zeros(n, bool)
    for i in range(E.shape[0]):
        nd = (E[i] - value).astype(f32)
        ok = ~done & (nd >= 0) & (nd <= dist)
        hit = ok & (nd <= 1e-6)
        idx = np.where(ok, i, idx)
        dist = np.where(ok & ~hit, nd, dist)
        done |= hit
    return idx


def _points(L, xv, yv, px, py, pz):
    """rho, vp, vs of the layered model at arrays of points (float32)."""
    V = L["v"]
    col = dict(elev=0, rho=1, rho_g=2, rho_p=3, vp=4, vp_g=5, vp_p=6, vs=7, vs_g=8, vs_p=9)
    fl = lambda nm, l: _bilinear_v(L, xv, yv, V[l, :, :, col[nm]], px, py)
    E = np.array([fl("elev", l) for l in range(L["ni"])], f32)
    mi = _last_ge_v(pz, E)
    top = mi == -1
    mtop = _last_ge_v(E[0], E)
    out = {}
    for nm in ("rho", "vp", "vs"):
        F = np.array([fl(nm, l) for l in range(L["ni"])], f32)
        G = np.array([fl(nm + "_g", l) for l in range(L["ni"])], f32)
        P = np.array([fl(nm + "_p", l) for l in range(L["ni"])], f32)
        out[nm] = (F, G, P)
    cols = np.arange(pz.shape[0])

    def at(l, dz):
        res = []
        for nm in ("rho", "vp", "vs"):
            F, G, P = out[nm]
            li = np.clip(l, 0, L["ni"] - 1)
            res.append((F[li, cols].astype(np.float64)
                        + np.power(dz.astype(np.float64), P[li, cols].astype(np.float64)) * G[li, cols]).astype(f32))
        return res

    mc = np.clip(mi, 0, L["ni"] - 1)
    dz = (E[mc, cols] - pz).astype(f32)
    r, vp, vs = at(mi, dz)
    assert np.all(top | (vp > vs.astype(np.float64) * np.sqrt(2.0)))
    onif = ~top & (np.abs(np.trunc(dz)) < 1)
    up = _first_ge_v(pz, E) - 1
    av = onif & (up >= 0)
    uc = np.clip(up, 0, L["ni"] - 1)
    r2, p2, s2 = at(up, (E[uc, cols] - pz).astype(f32))
    r = np.where(av, ((r + r2).astype(f32).astype(np.float64) / 2.0).astype(f32), r)
    vp = np.where(av, ((vp + p2).astype(f32).astype(np.float64) / 2.0).astype(f32), vp)
    vs = np.where(av, ((vs + s2).astype(f32).astype(np.float64) / 2.0).astype(f32), vs)
    tc = np.clip(mtop, 0, L["ni"] - 1)
    r = np.where(top, out["rho"][0][tc, cols], r)
    vp = np.where(top, out["vp"][0][tc, cols], vp)
    vs = np.where(top, out["vs"][0][tc, cols], vs)
    return r.astype(f32), vp.astype(f32), vs.astype(f32)


def media_equivalent(ni, nj, nk, x, y, z, path, method):
    """cgfd.media_layer_file with method "har" or "ari"."""
    lam, mu, rho = media(ni, nj, nk, x, y, z, path)
    L = read(path)
    xv = [f32(L["minx"] + L["dx"] * i) for i in range(L["nx"])]
    yv = [f32(L["miny"] + L["dy"] * j) for j in range(L["ny"])]
    X, Y, Z = (np.asarray(a, np.float64).astype(f32) for a in (x, y, z))
    nz, ny, nx = X.shape
    H = []
    for A in (X, Y, Z):
        h = A.copy()
        s = A[:-1, :-1, :-1].copy()
        for dk, dj, di in ((0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)):
            s = (s + A[dk:nz - 1 + dk, dj:ny - 1 + dj, di:nx - 1 + di]).astype(f32)
        h[:-1, :-1, :-1] = (s.astype(np.float64) / 8.0).astype(f32)
        H.append(h)
    Hx, Hy, Hz = H
    fl = lambda l, px, py: _bilinear_v(L, xv, yv, L["v"][l, :, :, 0], px, py)
    E = np.array([fl(l, Hx.ravel(), Hy.ravel()) for l in range(L["ni"])], f32)
    mat = _last_ge_v(Hz.ravel(), E)
    mat = np.where(mat == -1, _last_ge_v(E[0], E), mat).reshape(X.shape)
    ng = f32(8)
    for k in range(1, nz - 1):
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                cn = [(k - 1, j - 1, i - 1), (k - 1, j - 1, i), (k - 1, j, i), (k - 1, j, i - 1),
                      (k, j - 1, i - 1), (k, j - 1, i), (k, j, i), (k, j, i - 1)]
                v = [mat[c] for c in cn]
                if -1 in v or len(set(v)) <= 1:
                    continue
                Vc = np.array([[Hx[c], Hy[c], Hz[c]] for c in cn], f32)
                t = np.arange(9, dtype=f32)
                ii, jj, kk = t[None, None, :], t[None, :, None], t[:, None, None]
                lerp = lambda A, B, s: (A + ((B - A) / ng).astype(f32) * s).astype(f32)
                pts = []
                for d in range(3):
                    lx00 = lerp(Vc[0, d], Vc[1, d], ii)
                    lx10 = lerp(Vc[3, d], Vc[2, d], ii)
                    lx01 = lerp(Vc[4, d], Vc[5, d], ii)
                    lx11 = lerp(Vc[7, d], Vc[6, d], ii)
                    lxy0 = lerp(lx00, lx10, jj)
                    lxy1 = lerp(lx01, lx11, jj)
                    pts.append(np.broadcast_to(lerp(lxy0, lxy1, kk), (9, 9, 9)).ravel())
                r, vp, vs = _points(L, xv, yv, *pts)
                m = ((vs * vs).astype(f32) * r).astype(f32)
                lm = (((vp * vp).astype(f32) * r).astype(f32).astype(np.float64) - 2.0 * m.astype(np.float64)).astype(f32)
                a_rho = f32(0)
                a_lam = f32(0)
                a_mu = f32(0)
                h_k = f32(0)
                h_m = f32(0)
                for q in range(729):  # sequential float sums, as there
                    a_rho = f32(a_rho + r[q])
                    if method == "har":
                        h_k = f32(float(h_k) + 1.0 / (float(lm[q]) + 2.0 / 3.0 * float(m[q])))
                        h_m = f32(float(h_m) + 1.0 / float(m[q]))
                    else:
                        a_lam = f32(a_lam + lm[q])
                        a_mu = f32(a_mu + m[q])
                if method == "har":
                    hm, hk = f32(729.0 / float(h_m)), f32(729.0 / float(h_k))
                    lam[k, j, i] = float(f32(float(hk) - 2.0 / 3.0 * float(hm)))
                    mu[k, j, i] = float(hm)
                else:
                    lam[k, j, i] = float(f32(float(a_lam) / 729.0))
                    mu[k, j, i] = float(f32(float(a_mu) / 729.0))
                rho[k, j, i] = float(f32(float(a_rho) / 729.0))
    return lam, mu, rho
