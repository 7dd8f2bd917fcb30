"""bin_media.py -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's binary-volume medium for
elastic_isotropic, the checker of cgfd_media_bin_file (the product's C++
version in gpu-cgfd3d-elastic_amd/csrc/cgfd_setup.cpp).  Only tests/ import it.

Follows, in float32 with the reference's operation order:
  src/media/media_bin2model.cpp:27-230     media_bin2model_el_iso (window of the
                                           file covering the grid, z coincidence
                                           check), parameterization_bin_el_iso_loc
  src/media/media_read_file.cpp:1274-1320  read_bin_file (dimension 1 fastest)
  src/media/media_utility.cpp:189-284      TrilinearInterpolation (clamped
                                           floor/ceil nodes; x, then y, then z)
Vectorized over grid points.
"""
import numpy as np

f32 = np.float32


def media(ni, nj, nk, x, y, z, order, size, spacing, origin, rho_path, vp_path, vs_path):
    """lam, mu, rho (float64 arrays of the grid's shape) as cgfd.media_bin_file."""
    if isinstance(order, str):
        order = ["xyz".index(c) for c in order]
    if sorted(order) != [0, 1, 2]:
        raise ValueError("bad order")
    dim = [list(order).index(a) for a in range(3)]
    X, Y, Z = (np.asarray(a, np.float64).astype(f32) for a in (x, y, z))
    o = [f32(origin[dim[a]]) for a in range(3)]
    h = [f32(spacing[dim[a]]) for a in range(3)]
    n = [int(size[dim[a]]) for a in range(3)]
    lo, hi = [0, 0, 0], [0, 0, n[2] - 1]
    for a, (amin, amax) in enumerate(((X.min(), X.max()), (Y.min(), Y.max()))):
        lo[a] = 0 if amin < o[a] else int(f32(f32(amin - o[a]) / h[a]))
        if lo[a] >= n[a] or amax < o[a]:
            raise ValueError("does not cover")
        hi[a] = min(int(np.ceil(f32(f32(amax - o[a]) / h[a]))), n[a] - 1)
        if hi[a] <= lo[a]:
            raise ValueError("one-node window")
    zn = f32(o[2] + f32(h[2] * f32(n[2] - 1)))
    if Z.min() >= max(zn, o[2]) or Z.max() <= min(o[2], zn):
        raise ValueError("does not cover in z")
    vec = [np.array([f32(o[a] + f32(f32(lo[a] + q) * h[a])) for q in range(hi[a] - lo[a] + 1)], f32)
           for a in range(3)]
    fields = []
    for path in (rho_path, vp_path, vs_path):
        raw = np.fromfile(path, dtype=np.float32)
        if raw.size < size[0] * size[1] * size[2]:
            raise ValueError("insufficient data")
        F = raw[:size[0] * size[1] * size[2]].reshape(size[2], size[1], size[0])  # [dim3][dim2][dim1]
        F = np.transpose(F, [2 - dim[2], 2 - dim[1], 2 - dim[0]])  # [z][y][x]
        fields.append(F[:, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1])
    q = [X.ravel(), Y.ravel(), Z.ravel()]
    i0, i2 = [], []
    for a in range(3):
        d = f32(vec[a][1] - vec[a][0])
        t = ((q[a] - vec[a][0]).astype(f32) / d).astype(f32)
        nb = vec[a].shape[0]
        i0.append(np.clip(np.floor(t).astype(np.int64), 0, nb - 1))
        i2.append(np.clip(np.ceil(t).astype(np.int64), 0, nb - 1))
    out = []
    # w is 0/0 where a node coincides (i0 == i2); np.where then takes the node value
    with np.errstate(divide="ignore", invalid="ignore"):
        wx = ((q[0] - vec[0][i0[0]]).astype(f32) / (vec[0][i2[0]] - vec[0][i0[0]]).astype(f32)).astype(f32)
        wy = ((q[1] - vec[1][i0[1]]).astype(f32) / (vec[1][i2[1]] - vec[1][i0[1]]).astype(f32)).astype(f32)
        wz = ((q[2] - vec[2][i0[2]]).astype(f32) / (vec[2][i2[2]] - vec[2][i0[2]]).astype(f32)).astype(f32)
        lerp = lambda a, b, w: (a + (w * (b - a).astype(f32)).astype(f32)).astype(f32)
        for F in fields:
            at = lambda a, b, c: F[c, b, a]
            v000, v002 = at(i0[0], i0[1], i0[2]), at(i0[0], i0[1], i2[2])
            v200, v202 = at(i2[0], i0[1], i0[2]), at(i2[0], i0[1], i2[2])
            v020, v022 = at(i0[0], i2[1], i0[2]), at(i0[0], i2[1], i2[2])
            v220, v222 = at(i2[0], i2[1], i0[2]), at(i2[0], i2[1], i2[2])
            sx, sy, sz = i0[0] == i2[0], i0[1] == i2[1], i0[2] == i2[2]
            v100 = np.where(sx, v000, lerp(v000, v200, wx))
            v102 = np.where(sx, v002, lerp(v002, v202, wx))
            v122 = np.where(sx, v022, lerp(v022, v222, wx))
            v120 = np.where(sx, v020, lerp(v020, v220, wx))
            v110 = np.where(sy, v100, lerp(v100, v120, wy))
            v112 = np.where(sy, v102, lerp(v102, v122, wy))
            v111 = np.where(sz, v110, lerp(v110, v112, wz))
            out.append(np.where(sx & sy & sz, v000, v111).astype(f32))
    r, vp, vs = out
    m = ((vs * vs).astype(f32) * r).astype(f32)
    lm = (((vp * vp).astype(f32) * r).astype(f32).astype(np.float64) - 2.0 * m.astype(np.float64)).astype(f32)
    return tuple(a.astype(np.float64).reshape(X.shape) for a in (lm, m, r))
