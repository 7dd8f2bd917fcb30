"""Small seeded el_iso problems built with the oracle's setup restatements.

The same arrays feed the HIP session (through the C ABI) and the CPU oracle,
so a parity test compares the two time-stepping implementations on identical
inputs.  Arrays use the reference host layout: [comp][k][j][i] with 3 ghost
points per side.
"""
import ctypes as C

import numpy as np

import orc

FACES = ("x1", "x2", "y1", "y2", "z1", "z2")


class Problem:
    def __init__(self, ni=24, nj=20, nk=16, dh=100.0, curvilinear=False, free_top=True,
                 pml=(3, 3, 3, 3, 3, 0), vp=6000.0, vs=3464.0, rho=2670.0, hetero=0.0,
                 seed=0, source="moment", src_ijk=None, nrecv=6, nt_total=20, dt=0.005,
                 init_random=False, stf_len=None, f0=2.0, t0=0.1, pml_amax=3.0, pml_bmax=2.5,
                 metric_precision=32):
        L = orc.lib()
        rng = np.random.default_rng(seed)
        self.ni, self.nj, self.nk = ni, nj, nk
        nx, ny, nz = ni + 6, nj + 6, nk + 6
        self.nx, self.ny, self.nz = nx, ny, nz
        self.free_top = bool(free_top)
        self.dt = dt
        self.nt_total = nt_total
        n = nx * ny * nz
        # ---- grid
        x = np.zeros(n)
        y = np.zeros(n)
        z = np.zeros(n)
        L.orc_gen_cart(ni, nj, nk, dh, 0.0, dh, 0.0, dh, -(nk - 1) * dh, orc.p(x), orc.p(y), orc.p(z))
        x, y, z = x.reshape(nz, ny, nx), y.reshape(nz, ny, nx), z.reshape(nz, ny, nx)
        if curvilinear:
            # smooth topography and lateral shear, evaluated on ghosts too
            I = np.arange(nx) - 3.0
            J = np.arange(ny) - 3.0
            K = np.arange(nz) - 3.0
            Xg, Yg = np.meshgrid(I * dh, J * dh)
            topo = 0.08 * (nk * dh) * np.exp(-((Xg - 0.5 * ni * dh) ** 2 + (Yg - 0.45 * nj * dh) ** 2)
                                            / (2 * (0.3 * ni * dh) ** 2))
            zb = -(nk - 1) * dh
            frac = (K / (nk - 1))[:, None, None]
            z = zb + (0.0 + topo[None, :, :] - zb) * frac
            x = np.broadcast_to((I * dh)[None, None, :] + 0.05 * dh * np.sin(J * 0.7)[None, :, None]
                                * (1 + 0 * K[:, None, None]), (nz, ny, nx)).copy()
            y = np.broadcast_to((J * dh)[None, :, None] + 0.04 * dh * np.cos(I * 0.5)[None, None, :]
                                + 0 * K[:, None, None], (nz, ny, nx)).copy()
            # round through float32 as the reference keeps coordinates in float
            x, y, z = [np.float64(np.float32(a)) for a in (x, y, z)]
        self.x, self.y, self.z = [np.ascontiguousarray(a) for a in (x, y, z)]
        met = np.zeros(10 * n)
        L.orc_metric_cal(metric_precision, ni, nj, nk, orc.p(self.x), orc.p(self.y), orc.p(self.z), orc.p(met))
        self.metric = met.reshape(10, nz, ny, nx)
        # ---- medium (on every point, ghosts included, as media discretization does)
        pert = 1.0 + hetero * (rng.random((nz, ny, nx)) - 0.5) * 2
        Vp = np.float32(vp * pert)
        Vs = np.float32(vs * pert)
        Rho = np.float32(rho * np.ones((nz, ny, nx)))
        mu = np.float32(Rho * Vs * Vs)
        lam = np.float32(Rho * Vp * Vp - np.float32(2.0) * mu)
        self.lam = np.float64(lam)
        self.mu = np.float64(mu)
        self.rho = np.float64(Rho)
        slw = np.zeros(n)
        L.orc_rho_to_slow(n, orc.p(np.ascontiguousarray(self.rho.ravel())), orc.p(slw))
        self.slw = slw.reshape(nz, ny, nx)
        # ---- free surface
        self.mvx = np.zeros((ny, nx, 9))
        self.mvy = np.zeros((ny, nx, 9))
        if self.free_top:
            L.orc_free_matrices(32, ni, nj, nk, orc.p(self.metric), orc.p(self.lam), orc.p(self.mu),
                                orc.p(self.mvx), orc.p(self.mvy))
        # ---- PML (bdry_pml_set): face boxes in local indices
        self.pml = {}
        for f, nl in enumerate(pml):
            if nl <= 0 or (f == 5 and self.free_top):
                continue
            i1, i2, j1, j2, k1, k2 = 3, 3 + ni - 1, 3, 3 + nj - 1, 3, 3 + nk - 1
            if f == 0: i2 = i1 + nl
            if f == 1: i1 = i2 - nl
            if f == 2: j2 = j1 + nl
            if f == 3: j1 = j2 - nl
            if f == 4: k2 = k1 + nl
            if f == 5: k1 = k2 - nl
            L0 = C.c_double()
            dhh = C.c_double()
            L.orc_abl_len_dh(ni, nj, nk, orc.p(self.x), orc.p(self.y), orc.p(self.z), i1, i2, j1, j2, k1, k2,
                             f // 2, C.byref(L0), C.byref(dhh))
            A = np.zeros(nl + 1)
            B = np.zeros(nl + 1)
            D = np.zeros(nl + 1)
            L.orc_pml_coefs(nl, f % 2, L0.value, dhh.value, vp, pml_amax, pml_bmax, orc.p(A), orc.p(B), orc.p(D))
            self.pml[f] = (nl, A, B, D)
        # ---- source
        self.src = None
        if source is not None:
            si, sj, sk = src_ijk if src_ijk is not None else (ni // 2 + 0.3, nj // 2 - 0.2, nk // 2 + 0.1)
            stf_len = stf_len if stf_len is not None else nt_total * dt
            ext_indx = np.zeros(343, dtype=np.int32)
            ext_coef = np.zeros(343)
            itb = C.c_int()
            ite = C.c_int()
            en = C.c_int()
            mnt = C.c_int()
            L.orc_src_point(ni, nj, nk, si, sj, sk, 0, 0.0, dt, 0.0, 0, f0, t0, stf_len,
                            C.byref(itb), C.byref(ite), C.byref(en), orc.p(ext_indx), orc.p(ext_coef),
                            C.byref(mnt), None)
            max_nt = mnt.value
            stf = np.zeros(max_nt * 4)
            L.orc_src_point(ni, nj, nk, si, sj, sk, 0, 0.0, dt, 0.0, 0, f0, t0, stf_len,
                            C.byref(itb), C.byref(ite), C.byref(en), orc.p(ext_indx), orc.p(ext_coef),
                            C.byref(mnt), orc.p(stf))
            stf32 = np.float32(stf).reshape(1, max_nt, 4)
            F = M = None
            force_on = moment_on = 0
            if source == "moment":
                m6 = np.float32([1.0e16, 0.7e16, 1.2e16, 0.3e16, -0.2e16, 0.5e16])
                M = np.float64(np.stack([stf32 * m for m in m6]))
                moment_on = 1
            else:
                f3 = np.float32([1.0e12, -0.5e12, 2.0e12])
                F = np.float64(np.stack([stf32 * f for f in f3]))
                force_on = 1
            self.src = dict(nsrc=1, force_on=force_on, moment_on=moment_on, max_nt=max_nt, max_ext=343,
                            it_begin=np.array([itb.value], np.int32), it_end=np.array([ite.value], np.int32),
                            ext_num=np.array([en.value], np.int32), ext_indx=np.int64(ext_indx),
                            ext_coef=np.float64(np.float32(ext_coef)), F=F, M=M)
        # ---- receivers: surface and buried, some fractional
        self.recv_indx = np.zeros((0, 8), dtype=np.int64)
        self.recv_frac = np.zeros((0, 3))
        rows_i, rows_f = [], []
        for r in range(nrecv):
            rx = 1 + rng.integers(1, ni - 2) + (0.0 if r % 2 == 0 else rng.random() * 0.9)
            ry = 1 + rng.integers(1, nj - 2) + (0.0 if r % 3 == 0 else rng.random() * 0.9)
            rz = float(rng.integers(0, nk // 2)) + (0.0 if r % 2 == 0 else rng.random() * 0.9)
            ix = np.zeros(8, dtype=np.int64)
            fr = np.zeros(3)
            if L.orc_recv_index(ni, nj, nk, rx, ry, rz, 1, orc.p(ix), orc.p(fr)) == 0:
                rows_i.append(ix)
                rows_f.append(fr)
        if rows_i:
            self.recv_indx = np.array(rows_i, dtype=np.int64)
            self.recv_frac = np.array(rows_f)
        # ---- initial condition
        self.init = None
        if init_random:
            w = np.zeros((9, nz, ny, nx))
            scale = np.array([1.0, 1.0, 1.0, 1e7, 1e7, 1e7, 1e7, 1e7, 1e7])
            w[:, 3:3 + nk, 3:3 + nj, 3:3 + ni] = (rng.random((9, nk, nj, ni)) - 0.5) * scale[:, None, None, None]
            self.init = np.float64(np.float32(w))

    # ---------------------------------------------------------------- oracle
    def orc_struct(self):
        orc.check_problem(self.ni, self.nj, self.nk, self.metric, self.lam, self.mu, self.slw, self.mvx, self.mvy,
                          self.pml, self.src, self.recv_indx, self.recv_frac)
        keep = []

        def k(a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return orc.p(a)

        P = orc.OrcProblem()
        P.ni, P.nj, P.nk = self.ni, self.nj, self.nk
        P.metric = k(self.metric)
        P.lam, P.mu, P.slw = k(self.lam), k(self.mu), k(self.slw)
        P.free_top = int(self.free_top)
        P.matVx2Vz, P.matVy2Vz = k(self.mvx), k(self.mvy)
        for f in range(6):
            if f in self.pml:
                nl, A, B, D = self.pml[f]
                P.pml_on[f] = 1
                P.pml_nlay[f] = nl
                P.pml_A[f], P.pml_B[f], P.pml_D[f] = k(A), k(B), k(D)
        if self.src is not None:
            s = self.src
            P.nsrc, P.force_on, P.moment_on = s["nsrc"], s["force_on"], s["moment_on"]
            P.max_nt, P.max_ext = s["max_nt"], s["max_ext"]
            P.src_it_begin = k(s["it_begin"], np.int32)
            P.src_it_end = k(s["it_end"], np.int32)
            P.src_ext_num = k(s["ext_num"], np.int32)
            P.src_ext_indx = k(s["ext_indx"], np.int32)
            P.src_ext_coef = k(s["ext_coef"])
            P.src_F = k(s["F"]) if s["F"] is not None else None
            P.src_M = k(s["M"]) if s["M"] is not None else None
        P.nrecv = len(self.recv_indx)
        if P.nrecv:
            P.recv_indx = k(self.recv_indx, np.int64)
            P.recv_frac = k(self.recv_frac)
        P._keep = keep   # arrays stay alive as long as the struct does
        return P

    def face_sizes(self):
        P = self.orc_struct()
        out = np.zeros(6, dtype=np.int64)
        orc.lib().orc_face_sizes(C.byref(P), orc.p(out))
        return out

    def oracle_run(self, precision=32, nt_total=None, want_aux=False, init_aux=None):
        nt = self.nt_total if nt_total is None else nt_total
        P = self.orc_struct()
        lev = 9 * self.nx * self.ny * self.nz
        if self.init is not None and np.asarray(self.init).size != lev:
            raise ValueError("initial wavefield size does not match the grid")
        if init_aux is not None and np.asarray(init_aux).size != 9 * int(self.face_sizes().sum()):
            raise ValueError("initial PML memory size does not match the face boxes")
        seis = np.zeros((max(len(self.recv_indx), 1), 9, nt))
        wave = np.zeros((9, self.nz, self.ny, self.nx))
        aux = np.zeros(9 * int(self.face_sizes().sum())) if want_aux else None
        init = np.ascontiguousarray(self.init) if self.init is not None else None
        ia = np.ascontiguousarray(init_aux, dtype=np.float64) if init_aux is not None else None
        rc = orc.lib().orc_run(C.byref(P), precision, self.dt, nt, orc.p(init), orc.p(ia), orc.p(seis),
                               orc.p(wave), orc.p(aux))
        assert rc == 0
        return seis[:len(self.recv_indx)], wave, aux

    def oracle_stage(self, w, ipair, istage, it, aux=None, precision=32):
        P = self.orc_struct()
        if np.asarray(w).size != 9 * self.nx * self.ny * self.nz:
            raise ValueError("stage input size does not match the grid")
        if aux is not None and np.asarray(aux).size != 9 * int(self.face_sizes().sum()):
            raise ValueError("PML memory size does not match the face boxes")
        if not (0 <= ipair < 8 and 0 <= istage < 4):
            raise ValueError("pair 0..7 and stage 0..3")
        rhs = np.zeros((9, self.nz, self.ny, self.nx))
        aux_rhs = np.zeros(9 * int(self.face_sizes().sum()))
        w = np.ascontiguousarray(w, dtype=np.float64)
        a = np.ascontiguousarray(aux, dtype=np.float64) if aux is not None else None
        rc = orc.lib().orc_onestage(C.byref(P), precision, ipair, istage, it, orc.p(w), orc.p(a), orc.p(rhs),
                                    orc.p(aux_rhs))
        assert rc == 0
        return rhs, aux_rhs

    # ---------------------------------------------------------------- device
    def upload(self, sess, shots=None, dtype=np.float32):
        sess.upload_metric(np.asarray(self.metric, dtype=dtype))
        sess.upload_medium(np.asarray(self.lam, dtype), np.asarray(self.mu, dtype), np.asarray(self.slw, dtype))
        if self.free_top:
            sess.upload_free_surface(True, np.asarray(self.mvx, dtype), np.asarray(self.mvy, dtype))
        else:
            sess.upload_free_surface(False)
        for f in range(6):
            if f in self.pml:
                nl, A, B, D = self.pml[f]
                sess.upload_pml(f, nl, np.asarray(A, dtype), np.asarray(B, dtype), np.asarray(D, dtype))
        for shot in (range(sess.nshots) if shots is None else shots):
            self.upload_shot(sess, shot, dtype)

    def upload_shot(self, sess, shot, dtype=np.float32):
        """This problem's source, receivers and initial state into one shot slot."""
        if self.src is not None:
            s = dict(self.src)
            s["ext_coef"] = np.asarray(s["ext_coef"], dtype)
            if s["F"] is not None:
                s["F"] = np.asarray(s["F"], dtype)
            if s["M"] is not None:
                s["M"] = np.asarray(s["M"], dtype)
            sess.upload_source(shot, s)
        if len(self.recv_indx):
            sess.upload_receivers(shot, self.recv_indx, np.asarray(self.recv_frac, dtype), self.nt_total)
        if self.init is not None:
            sess.set_level(shot, 0, np.asarray(self.init, dtype))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
