"""locate.py -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's location of points given by physical
coordinates, the checker of cgfd_locate_points (the product's C++ version in
gpu-cgfd3d-elastic_amd/csrc/cgfd_setup.cpp).  Only tests/ import it.

Follows, in float32 with the reference's operation order:
  src/forward/gd_t.cu:1395-1486   gd_curv_depth_to_axis (tile order, cell
                                  bounds, quad test, inverse-distance z)
  src/lib/fdlib_math.cu:168-250   fdlib_math_isPoint2InQuad, fdlib_math_rdinterp_2d
  src/forward/gd_t.cu:1241-1392   gd_curv_coord_to_local_indx (nearest node,
                                  first of the 8 cells holding the point)
  src/forward/gd_t.cu:1574-1650   gd_curv_coord2index_sample (101^3 sampling)
  src/forward/gd_t.cu:1767-1800   isPointInHexahedron_c
  src/forward/io_funcs.cu:156-170 negative fractions move to the lower cell
One deliberate difference: a point exactly on quad corner 0 takes that
corner's z (the reference reads an unset weight there).
"""
import numpy as np

NG = 3
f32 = np.float32


def _nearest(X, Y, Z, sx, sy, sz):
    d = np.sqrt((sx - X) * (sx - X) + (sy - Y) * (sy - Y) + (sz - Z) * (sz - Z))
    return np.unravel_index(int(np.argmin(d)), X.shape)  # first minimum in k, j, i order


def _in_hex(p, v):
    faces = ((0, 4, 6), (7, 5, 1), (5, 4, 0), (2, 6, 7), (4, 5, 7), (3, 1, 0))
    for a, b, c in faces:
        p2f = v[a] - p
        A, B = v[b] - v[a], v[c] - v[a]
        nrm = np.array([A[1] * B[2] - A[2] * B[1], A[2] * B[0] - A[0] * B[2], A[0] * B[1] - A[1] * B[0]], f32)
        ln = np.sqrt(f32(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]))
        nu = nrm / ln
        sign = f32(p2f[0] * nu[0] + p2f[1] * nu[1] + p2f[2] * nu[2])
        with np.errstate(invalid="ignore", divide="ignore"):  # a corner point gives 0/0: NaN passes, as there
            sign = sign / np.sqrt(f32(p2f[0] * p2f[0] + p2f[1] * p2f[1] + p2f[2] * p2f[2]))
        if sign < 0.0:
            return False
    return True


def _sample(p, v):
    L1 = (np.arange(101, dtype=f32) / f32(100)).astype(f32)
    L0 = (1.0 - L1.astype(np.float64)).astype(f32)
    L = (L0, L1)
    xp = np.zeros((101, 101, 101), f32)
    yp = np.zeros_like(xp)
    zp = np.zeros_like(xp)
    for c in range(8):
        w = (L[c & 1][None, None, :] * L[(c >> 1) & 1][None, :, None]) * L[c >> 2][:, None, None]
        xp += w * v[c, 0]
        yp += w * v[c, 1]
        zp += w * v[c, 2]
    d = np.sqrt((p[0] - xp) * (p[0] - xp) + (p[1] - yp) * (p[1] - yp) + (p[2] - zp) * (p[2] - zp))
    n3, n2, n1 = np.unravel_index(int(np.argmin(d)), d.shape)
    return n1, n2, n3


def _in_quad(px, py, qx, qy):
    vx, vy = (qx[0], qx[1], qx[3], qx[2]), (qy[0], qy[1], qy[3], qy[2])
    for i in range(4):
        n = (i + 1) % 4
        ex, ey = f32(vx[n] - vx[i]), f32(vy[n] - vy[i])
        sign = f32(ex * f32(py - vy[i]) - ey * f32(px - vx[i]))
        sign = sign / np.sqrt(f32(ex * ex + ey * ey))
        if sign < -1e-6:
            return False
    return True


def _idw(px, py, qx, qy, qv):
    w = np.zeros(4, f32)
    tot = f32(0)
    at = -1
    for c in range(4):
        d = np.sqrt(f32(f32(px - qx[c]) * f32(px - qx[c]) + f32(py - qy[c]) * f32(py - qy[c])))
        if d < 1e-9:
            at = c
        else:
            w[c] = f32(1.0 / float(d))
            tot = f32(tot + w[c])
    if at >= 0:
        w[:] = 0
        w[at] = 1
        tot = f32(1)
    v = f32(0)
    for c in range(4):
        w[c] = f32(float(w[c]) * (1.0 / float(tot)))
        v = f32(v + w[c] * qv[c])
    return v


def _tiles(n):
    out, start = [], NG
    for q in range(4):
        ln = n // 4 + (1 if q < n % 4 else 0)
        out.append((start, start + ln - 1))
        start += ln
    return out


def _surface_z(X, Y, Z, ni, nj, nk, sx, sy):
    k = NG + nk - 1
    for j0, j1 in _tiles(nj):
        for i0, i1 in _tiles(ni):
            for j in range(j0, j1 + 1):
                for i in range(i0, i1 + 1):
                    cx, cy = X[k:k + 2, j:j + 2, i:i + 2], Y[k:k + 2, j:j + 2, i:i + 2]
                    if sx < cx.min() or sx > cx.max() or sy < cy.min() or sy > cy.max():
                        continue
                    qx = [X[k, j + (c >> 1), i + (c & 1)] for c in range(4)]
                    qy = [Y[k, j + (c >> 1), i + (c & 1)] for c in range(4)]
                    if _in_quad(sx, sy, qx, qy):
                        qz = [Z[k, j + (c >> 1), i + (c & 1)] for c in range(4)]
                        return _idw(sx, sy, qx, qy, qz)
    return None


def locate(ni, nj, nk, x, y, z, pts, depth_flag=None):
    """Same contract as cgfd.locate_points; x, y, z are [nk+6][nj+6][ni+6]."""
    shp = (nk + 2 * NG, nj + 2 * NG, ni + 2 * NG)
    X, Y, Z = (np.asarray(a, dtype=np.float64).reshape(shp).astype(f32) for a in (x, y, z))
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    n = pts.shape[0]
    idx = np.full((n, 3), -1000, np.int32)
    frac = np.zeros((n, 3))
    found = np.zeros(n, bool)
    xmin, xmax, ymin, ymax, zmin, zmax = X.min(), X.max(), Y.min(), Y.max(), Z.min(), Z.max()
    for q in range(n):
        sx, sy, sz = (f32(v) for v in pts[q])
        if depth_flag is not None and np.broadcast_to(depth_flag, (n,))[q]:
            if xmin <= sx <= xmax and ymin <= sy <= ymax:
                zs = _surface_z(X, Y, Z, ni, nj, nk, sx, sy)
                if zs is not None:
                    sz = f32(zs - sz)
        if not (xmin <= sx <= xmax and ymin <= sy <= ymax and zmin <= sz <= zmax):
            continue
        mk, mj, mi = _nearest(X, Y, Z, sx, sy, sz)
        if not (NG <= mi <= NG + ni - 1 and NG <= mj <= NG + nj - 1 and NG <= mk <= NG + nk - 1):
            continue
        found[q] = True
        inc = [f32(0), f32(0), f32(0)]
        p = np.array([sx, sy, sz], f32)
        for c in range(8):
            ci, cj, ck = mi - 1 + (c & 1), mj - 1 + ((c >> 1) & 1), mk - 1 + (c >> 2)
            v = np.array([[X[ck + (e >> 2), cj + ((e >> 1) & 1), ci + (e & 1)],
                           Y[ck + (e >> 2), cj + ((e >> 1) & 1), ci + (e & 1)],
                           Z[ck + (e >> 2), cj + ((e >> 1) & 1), ci + (e & 1)]] for e in range(8)], f32)
            if not _in_hex(p, v):
                continue
            s1, s2, s3 = _sample(p, v)
            inc = [f32(f32(f32(ci) + f32(s1) / f32(100)) - f32(mi)),
                   f32(f32(f32(cj) + f32(s2) / f32(100)) - f32(mj)),
                   f32(f32(f32(ck) + f32(s3) / f32(100)) - f32(mk))]
            break
        for d, m in enumerate((mi - NG, mj - NG, mk - NG)):
            f = inc[d]
            if f < 0:
                f = f32(1.0 + float(f))
                m -= 1
            idx[q, d] = m
            frac[q, d] = float(f)
    return idx, frac, found
