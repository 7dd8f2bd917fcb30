"""The oracle's own preparation of a workload, and oracle runs of it (test infrastructure).

`oracle_setup(m)` rebuilds everything a workload dict of
gpu-cgfd3d-elastic_amd/workloads.py derives from its grid and medium with
the oracle's restatements of the reference's setup code instead of the
product's FP64 host setup: metrics (gd_curv_metric_cal, gd_t.cu:209-314),
1/rho (md_rho_to_slow, md_t.cu:622-637), the free-surface matrices
(sv_curv_col_el_iso_dvh2dvz, sv_curv_col_el_iso_gpu.cu:1152-1261), the
CFS-PML profiles (bdry_pml_set, bdry_t.cu:85-451), the source tables
(src_read_locate_file, src_t.cu:143-790) and the receiver indices
(io_recv_read_locate, io_funcs.cu:24-299).  A HIP run of the product setup
compared with an oracle run of this one checks the whole drop-in chain
(setup -> device path), not only the time stepping.
"""
import ctypes as C

import numpy as np

import orc

L = orc.lib()


def oracle_setup(m, metric_precision=32):
    ni, nj, nk = m["ni"], m["nj"], m["nk"]
    x, y, z = [np.ascontiguousarray(m[k], dtype=np.float64) for k in ("x", "y", "z")]
    n = x.size
    out = dict(ni=ni, nj=nj, nk=nk, dt=m["dt"], nt_total=m["nt_total"], x=x, y=y, z=z)
    met = np.zeros(10 * n)
    L.orc_metric_cal(metric_precision, ni, nj, nk, orc.p(x), orc.p(y), orc.p(z), orc.p(met))
    out["metric"] = met.reshape((10,) + x.shape)
    out["lam"] = np.ascontiguousarray(m["lam"], dtype=np.float64)
    out["mu"] = np.ascontiguousarray(m["mu"], dtype=np.float64)
    rho = np.ascontiguousarray(m["rho"], dtype=np.float64)
    slw = np.zeros(n)
    L.orc_rho_to_slow(n, orc.p(rho), orc.p(slw))
    out["slw"] = slw.reshape(x.shape)
    ny, nx = nj + 6, ni + 6
    mvx, mvy = np.zeros((ny, nx, 9)), np.zeros((ny, nx, 9))
    L.orc_free_matrices(32, ni, nj, nk, orc.p(out["metric"]), orc.p(out["lam"]), orc.p(out["mu"]), orc.p(mvx),
                        orc.p(mvy))
    out["mvx"], out["mvy"] = mvx, mvy
    sp = m["src_params"]
    pml = {}
    for f, (nl, _, _, _) in m["pml"].items():
        i1, i2, j1, j2, k1, k2 = 3, 3 + ni - 1, 3, 3 + nj - 1, 3, 3 + nk - 1
        if f == 0: i2 = i1 + nl
        if f == 1: i1 = i2 - nl
        if f == 2: j2 = j1 + nl
        if f == 3: j1 = j2 - nl
        if f == 4: k2 = k1 + nl
        if f == 5: k1 = k2 - nl
        L0, dh = C.c_double(), C.c_double()
        L.orc_abl_len_dh(ni, nj, nk, orc.p(x), orc.p(y), orc.p(z), i1, i2, j1, j2, k1, k2, f // 2, C.byref(L0),
                         C.byref(dh))
        A, B, D = np.zeros(nl + 1), np.zeros(nl + 1), np.zeros(nl + 1)
        L.orc_pml_coefs(nl, f % 2, L0.value, dh.value, sp["pml_vel"], 3.0, 2.5, orc.p(A), orc.p(B), orc.p(D))
        pml[f] = (nl, A, B, D)
    out["pml"] = pml
    # source: by grid index, third coordinate a depth below the surface (depth flag 1)
    ext_i, ext_c = np.zeros(343, dtype=np.int32), np.zeros(343)
    itb, ite, en, mnt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    args = (ni, nj, nk, sp["sx"], sp["sy"], sp["depth"], 1, 0.0, m["dt"], 0.0, 0, sp["f0"], sp["t0"],
            m["nt_total"] * m["dt"])
    L.orc_src_point(*args, C.byref(itb), C.byref(ite), C.byref(en), orc.p(ext_i), orc.p(ext_c), C.byref(mnt), None)
    stf = np.zeros(mnt.value * 4)
    L.orc_src_point(*args, C.byref(itb), C.byref(ite), C.byref(en), orc.p(ext_i), orc.p(ext_c), C.byref(mnt),
                    orc.p(stf))
    stf32 = np.float32(stf).reshape(1, mnt.value, 4)
    M = np.float64(np.stack([stf32 * np.float32(v) for v in sp["m6"]]))
    out["src"] = dict(nsrc=1, force_on=0, moment_on=1, max_nt=mnt.value, max_ext=343,
                      it_begin=np.array([itb.value], np.int32), it_end=np.array([ite.value], np.int32),
                      ext_num=np.array([en.value], np.int32), ext_indx=np.int64(ext_i),
                      ext_coef=np.float64(np.float32(ext_c)), F=None, M=M)
    # receivers: the same stations (surface, by grid index), located by the oracle
    idx, frac = [], []
    for ix8 in np.asarray(m["recv_indx"]):
        k, rem = divmod(int(ix8[0]), nx * ny)
        j, i = divmod(rem, nx)
        ix, fr = np.zeros(8, dtype=np.int64), np.zeros(3)
        assert L.orc_recv_index(ni, nj, nk, i - 3 + 1.0, j - 3 + 1.0, float(nk - 1 - (k - 3)), 1, orc.p(ix),
                                orc.p(fr)) == 0
        idx.append(ix)
        frac.append(fr)
    out["recv_indx"], out["recv_frac"] = np.array(idx), np.array(frac)
    return out


def oracle_run(m, precision, nt_total):
    """Seismograms [nrecv][9][nt_total] and the final wavefield [9][nz][ny][nx] of the oracle."""
    P = orc.problem_struct(m)
    nr = len(m["recv_indx"])
    seis = np.zeros((nr, 9, nt_total))
    wave = np.zeros((9,) + np.asarray(m["x"]).shape)
    rc = L.orc_run(C.byref(P), precision, m["dt"], nt_total, None, None, orc.p(seis), orc.p(wave), None)
    assert rc == 0
    return seis, wave
