// el_iso_point.hip -- per-point stage kernel of the el_iso RK4 path.
//
// One thread updates one grid point of one shot: 27 one-sided DRP
// derivatives, the momentum/Hooke right-hand side (inner form, or the
// traction-image / vlow forms on the free-surface rows), the CFS-PML ADE
// corrections of every face box containing the point, then the fused RK
// update of the wavefield and of the PML memory (stage-value form, see Mode
// in el_iso_device.h).  It covers any list of boxes,
// so it serves both as the boundary-shell kernel next to the z-marching
// interior kernel and as the whole-domain kernel of the plain method.
//
// Semantics follow, in order, sv_curv_col_el_iso_rhs_inner_gpu
// (src/forward/sv_curv_col_el_iso_gpu.cu:291-452),
// sv_curv_col_el_iso_rhs_timg_z2_gpu (src/forward/sv_curv_col_el_gpu.cu:19-221),
// sv_curv_col_el_iso_rhs_vlow_z2_gpu (sv_curv_col_el_iso_gpu.cu:462-603),
// sv_curv_col_el_iso_rhs_cfspml_gpu (sv_curv_col_el_iso_gpu.cu:680-1147) and the
// RK axpys wav_update / wav_update_end (src/forward/wav_t.cu:362-378) as driven by
// drv_rk_curv_col.cu:339-491.
#include "el_iso_device.h"
#include "el_iso_launch.h"
#include "el_iso_pml.h"

namespace cgfd {

template <int DIR, typename TC, typename F>
__device__ __forceinline__ TC drp5(F f, long s) {
  constexpr int lo = Drp<TC, DIR>::lo;
  return Drp<TC, DIR>::c(0) * f((lo + 0) * s) + Drp<TC, DIR>::c(1) * f((lo + 1) * s) +
         Drp<TC, DIR>::c(2) * f((lo + 2) * s) + Drp<TC, DIR>::c(3) * f((lo + 3) * s) +
         Drp<TC, DIR>::c(4) * f((lo + 4) * s);
}

// RK update of the wavefield at one point from its right-hand side h (MODE:
// the RK mode, a compile-time constant, so each mode's loads are
// unconditional and issue together)
template <int MODE, typename TS, typename TC>
__device__ __forceinline__ void rk_store(const StageArgs<TS, TC> &a, long so, int o, const TC *h) {
  const long cs = a.L.cs;
  const unsigned qb = (unsigned)o * (unsigned)sizeof(TS);
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    TS *wout = a.w_out + so + c * cs, *wend = a.w_end + so + c * cs;
    const TS *wpre = a.w_pre + so + c * cs, *wu2 = a.w_u2 + so + c * cs;
    const TS *win = a.w_in + so + c * cs;
    const TC sc = St<TS>::scaled ? (TC)a.sc[c] : TC(1);
    const TC is = St<TS>::scaled ? (TC)a.isc[c] : TC(1);
    if constexpr (MODE == MODE_RHS) {
      St<TS>::st(&at_b(wout, qb), h[c] * is);
    } else if constexpr (MODE == MODE_S0) {
      const TC v = (TC)St<TS>::ld(&at_b(win, qb)) * sc;
      St<TS>::st(&at_b(wout, qb), (v + a.ca * h[c]) * is);
    } else if constexpr (MODE == MODE_S12) {
      const TC pre = (TC)St<TS>::ld(&at_b(wpre, qb)) * sc;
      St<TS>::st(&at_b(wout, qb), (pre + a.ca * h[c]) * is);
    } else {
      const TC pre = (TC)St<TS>::ld(&at_b(wpre, qb)) * sc;
      const TC u3 = (TC)St<TS>::ld(&at_b(win, qb)) * sc;
      const TC part = rk_partial(a.rk_w, pre, (TC)St<TS>::ld(&at_b(wend, qb)) * sc, (TC)St<TS>::ld(&at_b(wu2, qb)) * sc);
      St<TS>::st(&at_b(wend, qb), rk_final(pre, part, a.rk_w[2], u3, a.cb * h[c]) * is);
    }
  }
}

// the point of a box list that global thread gid maps to; false past the end
template <typename TS, typename TC>
__device__ __forceinline__ bool box_point(const StageArgs<TS, TC> &a, long gid, int &i, int &j, int &k) {
  int b = 0;
  while (b < a.nbox && gid >= a.box_end[b]) b++;
  if (b >= a.nbox) return false;
  const long loc = gid - (b ? a.box_end[b - 1] : 0);
  const int bx = a.box[b][1] - a.box[b][0] + 1, by = a.box[b][3] - a.box[b][2] + 1;
  i = a.box[b][0] + (int)(loc % bx);
  const long t = loc / bx;
  j = a.box[b][2] + (int)(t % by);
  k = a.box[b][4] + (int)(t / by);
  return true;
}

template <int DX, int DY, int DZ, int MODE, typename TS, typename TC>
__device__ __forceinline__ void point_update(const StageArgs<TS, TC> &a, int shot, int i, int j, int k) {
  const Layout &L = a.L;
  // element offsets inside one component stay below 2^31: 32-bit offsets from
  // uniform component bases let the compiler use scalar-base addressing
  const int o = (int)L.off(i, j, k);
  const long so = (long)shot * a.shot_stride;
  const TS *win = a.w_in + so;
  const long cs = L.cs;
  auto W = [&](int c, long d) -> TC {
    return (TC)St<TS>::ld(&at_b(win + c * cs, (unsigned)(o + (int)d) * (unsigned)sizeof(TS)));
  };
  auto S = [&](int c) -> TC { return St<TS>::scaled ? (TC)a.sc[c] : TC(1); };

  TC Dx[kNC], Dy[kNC], Dz[kNC];
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    auto f = [&](long d) { return W(c, d); };
    Dx[c] = drp5<DX, TC>(f, 1);
    Dy[c] = drp5<DY, TC>(f, L.px);
    Dz[c] = drp5<DZ, TC>(f, L.pz);
    if (St<TS>::scaled) { Dx[c] *= S(c); Dy[c] *= S(c); Dz[c] *= S(c); }
  }
  const TC *met = a.met;
  auto Mt = [&](int m, long d) -> TC {
    return at_b(met + m * cs, (unsigned)(o + (int)d) * (unsigned)sizeof(TC));
  };
  const TC xix = Mt(XIX, 0), xiy = Mt(XIY, 0), xiz = Mt(XIZ, 0);
  const TC etx = Mt(ETX, 0), ety = Mt(ETY, 0), etz = Mt(ETZ, 0);
  const TC ztx = Mt(ZTX, 0), zty = Mt(ZTY, 0), ztz = Mt(ZTZ, 0);
  const unsigned ob = (unsigned)o * (unsigned)sizeof(TC);
  const TC lam = at_b(a.med, ob), mu = at_b(a.med + cs, ob), slw = at_b(a.med + 2 * cs, ob);
  // lam + 2.0*mu: the double-promoted sum of the reference rounds to the same
  // float as the float sum whenever lam and 2mu are within 2^29 of each other.
  const TC lam2mu = lam + TC(2) * mu;

  TC h[kNC];
  const bool top = a.free_top != 0;
  const int ktop = L.nk - 1;
  constexpr int zlast = (DZ == 0) ? 1 : 3;
  if (top && k >= ktop - zlast) {
    // traction image: conservative form with the odd extension of J*zeta.T
    constexpr int zlo = (DZ == 0) ? -3 : -1;
    const int n_free = ktop - k - zlo;
    const TC slwjac = slw / Mt(JAC, 0);
    constexpr int T1[3] = {TXX, TXY, TXZ}, T2[3] = {TXY, TYY, TYZ}, T3[3] = {TXZ, TYZ, TZZ};
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      auto vec = [&](long d, int e1, int e2, int e3) -> TC {
        return Mt(JAC, d) * (Mt(e1, d) * (W(T1[cc], d) * S(T1[cc])) + Mt(e2, d) * (W(T2[cc], d) * S(T2[cc])) +
                             Mt(e3, d) * (W(T3[cc], d) * S(T3[cc])));
      };
      constexpr int xlo = (DX == 0) ? -3 : -1, ylo = (DY == 0) ? -3 : -1;
      TC DxTx = TC(0), DyTy = TC(0), DzTz = TC(0);
#pragma unroll
      for (int n = 0; n < 5; n++) {
        DxTx += Drp<TC, DX>::c(n) * vec((long)(xlo + n), XIX, XIY, XIZ);
        DyTy += Drp<TC, DY>::c(n) * vec((long)(ylo + n) * L.px, ETX, ETY, ETZ);
      }
#pragma unroll
      for (int n = 0; n < 5; n++) {
        TC v;
        if (n < n_free) v = vec((long)(zlo + n) * L.pz, ZTX, ZTY, ZTZ);
        else if (n == n_free) v = TC(0);
        else v = -vec((long)(zlo + n - 2 * (n - n_free)) * L.pz, ZTX, ZTY, ZTZ);
        DzTz += Drp<TC, DZ>::c(n) * v;
      }
      h[VX + cc] = (DxTx + DyTy + DzTz) * slwjac;
    }
  } else {
    h[VX] = slw * (xix * Dx[TXX] + xiy * Dx[TXY] + xiz * Dx[TXZ] + etx * Dy[TXX] + ety * Dy[TXY] + etz * Dy[TXZ] +
                   ztx * Dz[TXX] + zty * Dz[TXY] + ztz * Dz[TXZ]);
    h[VY] = slw * (xix * Dx[TXY] + xiy * Dx[TYY] + xiz * Dx[TYZ] + etx * Dy[TXY] + ety * Dy[TYY] + etz * Dy[TYZ] +
                   ztx * Dz[TXY] + zty * Dz[TYY] + ztz * Dz[TYZ]);
    h[VZ] = slw * (xix * Dx[TXZ] + xiy * Dx[TYZ] + xiz * Dx[TZZ] + etx * Dy[TXZ] + ety * Dy[TYZ] + etz * Dy[TZZ] +
                   ztx * Dz[TXZ] + zty * Dz[TYZ] + ztz * Dz[TZZ]);
  }

  TC DzVx = Dz[VX], DzVy = Dz[VY], DzVz = Dz[VZ];
  if (top && k >= ktop - 2) {
    const int n = ktop - k;
    if (n == 0) {
      const long pl = (long)L.ni * L.nj, q = (long)j * L.ni + i;
      const TC *A = a.mat + q, *B = a.mat + 9 * pl + q;
      DzVx = A[0] * Dx[VX] + A[pl] * Dx[VY] + A[2 * pl] * Dx[VZ] + B[0] * Dy[VX] + B[pl] * Dy[VY] + B[2 * pl] * Dy[VZ];
      DzVy = A[3 * pl] * Dx[VX] + A[4 * pl] * Dx[VY] + A[5 * pl] * Dx[VZ] + B[3 * pl] * Dy[VX] + B[4 * pl] * Dy[VY] +
             B[5 * pl] * Dy[VZ];
      DzVz = A[6 * pl] * Dx[VX] + A[7 * pl] * Dx[VY] + A[8 * pl] * Dx[VZ] + B[6 * pl] * Dy[VX] + B[7 * pl] * Dy[VY] +
             B[8 * pl] * Dy[VZ];
    } else if (n == 1) {
      // 2nd-order op, layer 1 (fd_t.cu:82-85, 103-106)
      const long z = L.pz;
      if (DZ == 0) {
        DzVx = -W(VX, -z) + W(VX, 0); DzVy = -W(VY, -z) + W(VY, 0); DzVz = -W(VZ, -z) + W(VZ, 0);
      } else {
        DzVx = -W(VX, 0) + W(VX, z); DzVy = -W(VY, 0) + W(VY, z); DzVz = -W(VZ, 0) + W(VZ, z);
      }
      DzVx *= S(VX); DzVy *= S(VY); DzVz *= S(VZ);
    } else {
      // 4th-order op, layer 2 (fd_t.cu:86-89, 107-110)
      const long z = L.pz;
      const TC c0 = (DZ == 0) ? TC(1.0 / 6.0) : TC(-7.0 / 6.0);
      const TC c1 = (DZ == 0) ? TC(-8.0 / 6.0) : TC(8.0 / 6.0);
      const TC c2 = (DZ == 0) ? TC(7.0 / 6.0) : TC(-1.0 / 6.0);
      const long s0 = (DZ == 0) ? -2 * z : 0, s1 = (DZ == 0) ? -z : z, s2 = (DZ == 0) ? 0 : 2 * z;
      DzVx = (c0 * W(VX, s0) + c1 * W(VX, s1) + c2 * W(VX, s2)) * S(VX);
      DzVy = (c0 * W(VY, s0) + c1 * W(VY, s1) + c2 * W(VY, s2)) * S(VY);
      DzVz = (c0 * W(VZ, s0) + c1 * W(VZ, s1) + c2 * W(VZ, s2)) * S(VZ);
    }
  }
  h[TXX] = lam2mu * (xix * Dx[VX] + etx * Dy[VX] + ztx * DzVx) +
           lam * (xiy * Dx[VY] + ety * Dy[VY] + zty * DzVy + xiz * Dx[VZ] + etz * Dy[VZ] + ztz * DzVz);
  h[TYY] = lam2mu * (xiy * Dx[VY] + ety * Dy[VY] + zty * DzVy) +
           lam * (xix * Dx[VX] + etx * Dy[VX] + ztx * DzVx + xiz * Dx[VZ] + etz * Dy[VZ] + ztz * DzVz);
  h[TZZ] = lam2mu * (xiz * Dx[VZ] + etz * Dy[VZ] + ztz * DzVz) +
           lam * (xix * Dx[VX] + etx * Dy[VX] + ztx * DzVx + xiy * Dx[VY] + ety * Dy[VY] + zty * DzVy);
  h[TXY] = mu * (xiy * Dx[VX] + xix * Dx[VY] + ety * Dy[VX] + etx * Dy[VY] + zty * DzVx + ztx * DzVy);
  h[TXZ] = mu * (xiz * Dx[VX] + xix * Dx[VZ] + etz * Dy[VX] + etx * Dy[VZ] + ztz * DzVx + ztx * DzVz);
  h[TYZ] = mu * (xiz * Dx[VY] + xiy * Dx[VZ] + etz * Dy[VY] + ety * Dy[VZ] + ztz * DzVy + zty * DzVz);

  // ---- CFS-PML faces, in the reference's face order ----
  pml_apply<MODE, TS, TC>(a, shot, i, j, k, Dx, Dy, Dz, xix, xiy, xiz, etx, ety, etz, ztx, zty, ztz, lam, mu, lam2mu,
                          slw, h);

  rk_store<MODE, TS, TC>(a, so, o, h);
}

template <int DX, int DY, int DZ, int MODE, typename TS, typename TC>
__global__ __launch_bounds__(256) void k_point(StageArgs<TS, TC> a) {
  const int shot = blockIdx.y;
  if (a.shot_active && !a.shot_active[shot]) return;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_point(a, gid, i, j, k)) return;
  point_update<DX, DY, DZ, MODE, TS, TC>(a, shot, i, j, k);
}

// ---- source injection (sv_curv_col_el_iso_rhs_src_gpu, sv_curv_col_el_gpu.cu:227-301) ----
// One thread per distinct grid point; contributions of the shot's sources are
// summed in source order and applied to the RK outputs of this stage.
template <typename TS, typename TC>
__global__ __launch_bounds__(256) void k_source(SrcArgs<TS, TC> a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npt) return;
  const int shot = a.pt_shot[p];
  if (a.only_shot >= 0 && shot != a.only_shot) return;
  if (a.only_shot < 0 && !a.shot_active[shot]) return;
  const int it = a.only_shot >= 0 ? a.fixed_it : *a.clock - a.shot_start[shot];
  const long o = a.pt_off[p];
  const long cs = a.L.cs;
  const TC jac = a.met[JAC * cs + o], slw = a.med[2 * cs + o];
  TC v[kNC];
#pragma unroll
  for (int c = 0; c < kNC; c++) v[c] = TC(0);
  bool any = false;
  for (int q = a.pt_beg[p]; q < a.pt_beg[p + 1]; q++) {
    const int s = a.ct_src[q];
    if (it < a.s_itb[s] || it > a.s_ite[s]) continue;
    any = true;
    const long ist = a.s_tab[s] + (long)(it - a.s_itb[s]) * 4 + a.istage;
    const TC coef = a.ct_coef[q];
    if (a.s_kind[s] & 1) {
      const TC V = coef * slw / jac;
      v[VX] += a.F[ist] * V;
      v[VY] += a.F[a.ntab + ist] * V;
      v[VZ] += a.F[2 * a.ntab + ist] * V;
    }
    if (a.s_kind[s] & 2) {
      const TC rjac = coef / jac;
      v[TXX] += -a.M[ist] * rjac;
      v[TYY] += -a.M[a.ntab + ist] * rjac;
      v[TZZ] += -a.M[2 * a.ntab + ist] * rjac;
      v[TYZ] += -a.M[3 * a.ntab + ist] * rjac;
      v[TXZ] += -a.M[4 * a.ntab + ist] * rjac;
      v[TXY] += -a.M[5 * a.ntab + ist] * rjac;
    }
  }
  if (!any) return;
  const long so = (long)shot * a.shot_stride;
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    if (v[c] == TC(0)) continue;
    const long q = so + c * cs + o;
    const TC sc = St<TS>::scaled ? TC(1) / (TC)a.isc[c] : TC(1);
    const TC is = St<TS>::scaled ? (TC)a.isc[c] : TC(1);
    // stage values take a_i dt of the source (U_{i+1} = W + a_i dt (K_i + S_i)),
    // W^{n+1} takes b3 dt of stage 3's; those of stages 0..2 reach it through U1..U3
    if (a.mode == MODE_RHS) {
      St<TS>::st(a.w_out + q, ((TC)St<TS>::ld(a.w_out + q) * sc + v[c]) * is);
    } else if (a.mode == MODE_S0 || a.mode == MODE_S12) {
      St<TS>::st(a.w_out + q, ((TC)St<TS>::ld(a.w_out + q) * sc + a.ca * v[c]) * is);
    } else {
      St<TS>::st(a.w_end + q, ((TC)St<TS>::ld(a.w_end + q) * sc + a.cb * v[c]) * is);
    }
  }
}

// ---- receivers (io_recv_keep + io_recv_line_interp_pack_buff, io_funcs.cu:1624-1666, 1747-1763) ----
// Corner 5 carries both (x1,y1,z2) and (x2,y1,z2) weights and corner 4 is
// unused, as in the reference; with on-node receivers the quirk vanishes.
template <typename TS, typename TC>
__global__ __launch_bounds__(256) void k_recv(RecvArgs<TS, TC> a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nrec * kNC) return;
  const int r = t / kNC, c = t % kNC;
  const int shot = a.r_shot[r];
  if (a.shot_active && !a.shot_active[shot]) return;
  const int it = *a.clock - a.shot_start[shot];
  if (it < 0 || it >= a.r_nt[r]) return;
  const TS *w = a.w_pre + (long)shot * a.shot_stride + c * a.cs;
  const long *ix = a.r_off + (long)r * 8;
  const TC sc = St<TS>::scaled ? (TC)a.sc[c] : TC(1);
  TC b[8];
#pragma unroll
  for (int q = 0; q < 8; q++) b[q] = (TC)St<TS>::ld(w + ix[q]) * sc;
  const TC Lx2 = a.r_frac[r * 3 + 0], Ly2 = a.r_frac[r * 3 + 1], Lz2 = a.r_frac[r * 3 + 2];
  const TC Lx1 = TC(1) - Lx2, Ly1 = TC(1) - Ly2, Lz1 = TC(1) - Lz2;
  const TC s = b[0] * Lx1 * Ly1 * Lz1 + b[1] * Lx2 * Ly1 * Lz1 + b[2] * Lx1 * Ly2 * Lz1 + b[3] * Lx2 * Ly2 * Lz1 +
               b[5] * Lx1 * Ly1 * Lz2 + b[5] * Lx2 * Ly1 * Lz2 + b[6] * Lx1 * Ly2 * Lz2 + b[7] * Lx2 * Ly2 * Lz2;
  const long pitch = a.ring > 0 ? a.ring : a.r_nt[r];
  a.seismo[a.r_out[r] + (long)c * pitch + (a.ring > 0 ? it % a.ring : it)] = s;
}

#if !defined(CGFD_TYPE) || CGFD_TYPE == 0
__global__ void k_clock_advance(int *clock) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *clock += 1;
}
#endif

template <typename T> __global__ __launch_bounds__(256) void k_fill(T *p, long n, T v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---- peak ground motion on the free surface (PG_calcu_gpu, src/forward/wav_t.cu:281-360) ----
// After each step, from W^{n+1} (w_new) and W^n (w_old) on the top plane:
// running maxima of |V|, |V_h|, |V_x|, |V_y|, |V_z|, the same for the
// acceleration |dV|/dt and for the displacement integrated with the trapezoid
// rule (accumulated in double, as the reference's 0.5 literal promotes it).
// Maps per shot: [15 maps + 3 displacements][nj][ni].
template <typename TS, typename TC>
__global__ __launch_bounds__(256) void k_peak(const TS *w_new, const TS *w_old, long shot_stride, Layout L, TC *pg,
                                              TC dt, float sx, float sy, float sz, const int *shot_active) {
  const int shot = blockIdx.y;
  if (shot_active && !shot_active[shot]) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = L.ni * L.nj;
  if (t >= n) return;
  const int i = t % L.ni, j = t / L.ni;
  const long o = L.off(i, j, L.nk - 1) + (long)shot * shot_stride;
  const long cs = L.cs;
  const TC vx1 = (TC)St<TS>::ld(w_new + o) * (TC)sx, vy1 = (TC)St<TS>::ld(w_new + cs + o) * (TC)sy;
  const TC vz1 = (TC)St<TS>::ld(w_new + 2 * cs + o) * (TC)sz;
  const TC vx0 = (TC)St<TS>::ld(w_old + o) * (TC)sx, vy0 = (TC)St<TS>::ld(w_old + cs + o) * (TC)sy;
  const TC vz0 = (TC)St<TS>::ld(w_old + 2 * cs + o) * (TC)sz;
  TC *m = pg + (long)shot * 18 * n + t;
  const TC Ax = fabs((vx1 - vx0) / dt), Ay = fabs((vy1 - vy0) / dt), Az = fabs((vz1 - vz0) / dt);
  TC Dx = m[15 * n], Dy = m[16 * n], Dz = m[17 * n];
  Dx = (TC)((double)Dx + 0.5 * (double)(vx1 + vx0) * (double)dt);
  Dy = (TC)((double)Dy + 0.5 * (double)(vy1 + vy0) * (double)dt);
  Dz = (TC)((double)Dz + 0.5 * (double)(vz1 + vz0) * (double)dt);
  m[15 * n] = Dx; m[16 * n] = Dy; m[17 * n] = Dz;
  const TC V = sqrt(vx1 * vx1 + vy1 * vy1 + vz1 * vz1), Vh = sqrt(vx1 * vx1 + vy1 * vy1);
  const TC A = sqrt(Ax * Ax + Ay * Ay + Az * Az), Ah = sqrt(Ax * Ax + Ay * Ay);
  const TC D = sqrt(Dx * Dx + Dy * Dy + Dz * Dz), Dh = sqrt(Dx * Dx + Dy * Dy);
  const TC val[15] = {V, Vh, fabs(vx1), fabs(vy1), fabs(vz1), A, Ah, Ax, Ay, Az, D, Dh, fabs(Dx), fabs(Dy), fabs(Dz)};
#pragma unroll
  for (int q = 0; q < 15; q++)
    if (m[q * n] < val[q]) m[q * n] = val[q];
}

// ---- strided sub-volume gather (io_snap_pack_buff, io_funcs.cu:1568-1596) ----
template <typename TS, typename TC>
__global__ __launch_bounds__(256) void k_gather(const TS *w, Layout L, int i1, int j1, int k1, int ni, int nj, int nk,
                                                int di, int dj, int dk, TC scale, TC *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)ni * nj * nk;
  if (t >= n) return;
  const int i = (int)(t % ni);
  const long r = t / ni;
  const int j = (int)(r % nj), k = (int)(r / nj);
  out[t] = (TC)St<TS>::ld(w + L.off(i1 + i * di, j1 + j * dj, k1 + k * dk)) * scale;
}

// ---- launchers ------------------------------------------------------------------
template <typename TS, typename TC>
hipError_t launch_peak(const TS *w_new, const TS *w_old, long shot_stride, const Layout &L, TC *pg, TC dt,
                       const float *sc, const int *shot_active, int nshots, hipStream_t st) {
  const int n = L.ni * L.nj;
  hipLaunchKernelGGL((k_peak<TS, TC>), dim3((unsigned)((n + 255) / 256), (unsigned)nshots), dim3(256), 0, st, w_new,
                     w_old, shot_stride, L, pg, dt, sc[0], sc[1], sc[2], shot_active);
  return hipGetLastError();
}

template <typename TS, typename TC>
hipError_t launch_gather(const TS *w, const Layout &L, int i1, int j1, int k1, int ni, int nj, int nk, int di,
                         int dj, int dk, TC scale, TC *out, hipStream_t st) {
  const long n = (long)ni * nj * nk;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_gather<TS, TC>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, L, i1, j1, k1, ni,
                     nj, nk, di, dj, dk, scale, out);
  return hipGetLastError();
}

// one RK mode's k_point for the operator directions sel = 4 dx + 2 dy + dz
template <int MODE, typename TS, typename TC>
static void launch_point_m(const StageArgs<TS, TC> &a, dim3 grid, int sel, hipStream_t st) {
  const dim3 block(256);
#define CGFD_L(X, Y, Z) hipLaunchKernelGGL((k_point<X, Y, Z, MODE, TS, TC>), grid, block, 0, st, a)
  switch (sel) {
    case 0: CGFD_L(0, 0, 0); break;
    case 1: CGFD_L(0, 0, 1); break;
    case 2: CGFD_L(0, 1, 0); break;
    case 3: CGFD_L(0, 1, 1); break;
    case 4: CGFD_L(1, 0, 0); break;
    case 5: CGFD_L(1, 0, 1); break;
    case 6: CGFD_L(1, 1, 0); break;
    default: CGFD_L(1, 1, 1); break;
  }
#undef CGFD_L
}

template <typename TS, typename TC>
hipError_t launch_point(const StageArgs<TS, TC> &a, int dx, int dy, int dz, int nshots, hipStream_t st) {
  const long n = a.nbox ? a.box_end[a.nbox - 1] : 0;
  if (n == 0) return hipSuccess;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)nshots);
  const int sel = dx * 4 + dy * 2 + dz;
  switch (a.mode) {
    case MODE_S0: launch_point_m<MODE_S0>(a, grid, sel, st); break;
    case MODE_S12: launch_point_m<MODE_S12>(a, grid, sel, st); break;
    case MODE_S3: launch_point_m<MODE_S3>(a, grid, sel, st); break;
    default: launch_point_m<MODE_RHS>(a, grid, sel, st); break;
  }
  return hipGetLastError();
}

template <typename TS, typename TC> hipError_t launch_source(const SrcArgs<TS, TC> &a, hipStream_t st) {
  if (a.npt == 0) return hipSuccess;
  hipLaunchKernelGGL((k_source<TS, TC>), dim3((a.npt + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename TS, typename TC> hipError_t launch_recv(const RecvArgs<TS, TC> &a, hipStream_t st) {
  if (a.nrec == 0) return hipSuccess;
  hipLaunchKernelGGL((k_recv<TS, TC>), dim3((a.nrec * kNC + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

#if !defined(CGFD_TYPE) || CGFD_TYPE == 0
hipError_t launch_clock_advance(int *clock, hipStream_t st) {
  hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, st, clock);
  return hipGetLastError();
}
#endif

template <typename T> hipError_t launch_fill(T *p, long n, T v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((k_fill<T>), dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

#define CGFD_INST(TS, TC)                                                                                   \
  template hipError_t launch_point<TS, TC>(const StageArgs<TS, TC> &, int, int, int, int, hipStream_t);    \
  template hipError_t launch_peak<TS, TC>(const TS *, const TS *, long, const Layout &, TC *, TC, const float *, \
                                          const int *, int, hipStream_t);                                      \
  template hipError_t launch_gather<TS, TC>(const TS *, const Layout &, int, int, int, int, int, int, int, int,   \
                                            int, TC, TC *, hipStream_t);                                        \
  template hipError_t launch_source<TS, TC>(const SrcArgs<TS, TC> &, hipStream_t);                          \
  template hipError_t launch_recv<TS, TC>(const RecvArgs<TS, TC> &, hipStream_t);
// CGFD_TYPE selects one storage type per object so the four compile in
// parallel; without it every type is instantiated here.
#if !defined(CGFD_TYPE) || CGFD_TYPE == 0
CGFD_INST(float, float)
#endif
#if !defined(CGFD_TYPE) || CGFD_TYPE == 1
CGFD_INST(double, double)
#endif
#if !defined(CGFD_TYPE) || CGFD_TYPE == 2
CGFD_INST(__half, float)
#endif
#if !defined(CGFD_TYPE) || CGFD_TYPE == 3
CGFD_INST(__hip_bfloat16, float)
#endif
#undef CGFD_INST
#if !defined(CGFD_TYPE) || CGFD_TYPE == 0
template hipError_t launch_fill<float>(float *, long, float, hipStream_t);
template hipError_t launch_fill<double>(double *, long, double, hipStream_t);
template hipError_t launch_fill<uint16_t>(uint16_t *, long, uint16_t, hipStream_t);
template hipError_t launch_fill<int>(int *, long, int, hipStream_t);
#endif

}  // namespace cgfd
