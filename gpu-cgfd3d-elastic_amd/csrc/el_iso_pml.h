// el_iso_pml.h -- CFS-PML ADE correction and memory-variable RK update of one
// grid point, shared by the per-point kernel and the PML variant of the
// z-marching kernel.
//
// For each face box containing the point, in the reference's face order
// x1,x2,y1,y2,z1,z2 (sv_curv_col_el_iso_rhs_cfspml, src/forward/
// sv_curv_col_el_iso_gpu.cu:613-1147): R = the part of the right-hand side
// along the face's axis, rhs += (B-1) R - B Phi, Phi' = D R - A Phi, plus the
// free-surface conversion term on the top row of x and y faces (:852-915,
// :1000-1060).  The memory variables are then advanced with the same RK
// step as the wavefield (drv_rk_curv_col.cu:357-477), in the stage-value
// form of Mode (el_iso_device.h).
#pragma once
#include "el_iso_device.h"

namespace cgfd {

// One face fi (its box must contain the point): Dd are the raw derivatives
// along the face's axis and e1..e3 the matching metric row; MODE is the RK
// mode (compile-time: each mode's loads are unconditional).
template <int MODE, typename TS, typename TC>
__device__ __forceinline__ void pml_face(const StageArgs<TS, TC> &a, int fi, int shot, int i, int j, int k,
                                         const TC *Dd, TC e1, TC e2, TC e3, TC ztx, TC zty, TC ztz, TC lam, TC mu,
                                         TC lam2mu, TC slw, TC *h) {
  const Layout &L = a.L;
  const bool top = a.free_top != 0;
  const int ktop = L.nk - 1;
  auto S = [&](int c) -> TC { return St<TS>::scaled ? (TC)a.sc[c] : TC(1); };
  const FaceArgs<TS, TC> &F = a.f[fi];
  const int dim = fi >> 1;
  const int ic = (dim == 0) ? i - F.i1 : (dim == 1) ? j - F.j1 : k - F.k1;
  const unsigned ia = (unsigned)(((k - F.k1) * F.bj + (j - F.j1)) * F.bi + (i - F.i1));
  const TC cD = F.D[ic], cA = F.A[ic], cB = F.B[ic];
  const TC Bm1 = cB - TC(1);
  TC R[kNC];
  R[VX] = slw * (e1 * Dd[TXX] + e2 * Dd[TXY] + e3 * Dd[TXZ]);
  R[VY] = slw * (e1 * Dd[TXY] + e2 * Dd[TYY] + e3 * Dd[TYZ]);
  R[VZ] = slw * (e1 * Dd[TXZ] + e2 * Dd[TYZ] + e3 * Dd[TZZ]);
  R[TXX] = lam2mu * e1 * Dd[VX] + lam * e2 * Dd[VY] + lam * e3 * Dd[VZ];
  R[TYY] = lam * e1 * Dd[VX] + lam2mu * e2 * Dd[VY] + lam * e3 * Dd[VZ];
  R[TZZ] = lam * e1 * Dd[VX] + lam * e2 * Dd[VY] + lam2mu * e3 * Dd[VZ];
  R[TXY] = mu * (e2 * Dd[VX] + e1 * Dd[VY]);
  R[TXZ] = mu * (e3 * Dd[VX] + e1 * Dd[VZ]);
  R[TYZ] = mu * (e3 * Dd[VY] + e2 * Dd[VZ]);
  const long fso = (long)shot * F.shot_stride;
  const unsigned fs = (unsigned)F.fsize;
  TC P[kNC], Ph[kNC], pre_a[kNC], u1_a[kNC], u2_a[kNC];
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    const unsigned qb = (c * fs + ia) * (unsigned)sizeof(TS);
    P[c] = (TC)St<TS>::ld(&at_b(F.a_in + fso, qb)) * S(c);
    if constexpr (MODE == MODE_S12 || MODE == MODE_S3) pre_a[c] = (TC)St<TS>::ld(&at_b(F.a_pre + fso, qb)) * S(c);
    if constexpr (MODE == MODE_S3) {
      u1_a[c] = (TC)St<TS>::ld(&at_b(F.a_end + fso, qb)) * S(c);
      u2_a[c] = (TC)St<TS>::ld(&at_b(F.a_u2 + fso, qb)) * S(c);
    }
  }
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    h[c] += Bm1 * R[c] - cB * P[c];
    Ph[c] = cD * R[c] - cA * P[c];
  }
  if (top && dim < 2 && k == ktop) {
    const long pl = (long)L.ni * L.nj, q = (long)j * L.ni + i;
    const TC *A = a.mat + (dim == 0 ? 0 : 9 * pl) + q;
    const TC dvx = A[0] * Dd[VX] + A[pl] * Dd[VY] + A[2 * pl] * Dd[VZ];
    const TC dvy = A[3 * pl] * Dd[VX] + A[4 * pl] * Dd[VY] + A[5 * pl] * Dd[VZ];
    const TC dvz = A[6 * pl] * Dd[VX] + A[7 * pl] * Dd[VY] + A[8 * pl] * Dd[VZ];
    TC t[kNC];
    t[TXX] = lam2mu * (ztx * dvx) + lam * (zty * dvy + ztz * dvz);
    t[TYY] = lam2mu * (zty * dvy) + lam * (ztx * dvx + ztz * dvz);
    t[TZZ] = lam2mu * (ztz * dvz) + lam * (ztx * dvx + zty * dvy);
    t[TXY] = mu * (zty * dvx + ztx * dvy);
    t[TXZ] = mu * (ztz * dvx + ztx * dvz);
    t[TYZ] = mu * (ztz * dvy + zty * dvz);
    // (coef_B - 1.0) * rhs accumulated in double (sv_curv_col_el_iso_gpu.cu:900-905)
    const double bm1 = (double)cB - 1.0;
#pragma unroll
    for (int c = TXX; c < kNC; c++) {
      h[c] = (TC)((double)h[c] + bm1 * (double)t[c]);
      Ph[c] += cD * t[c];
    }
  }
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    const unsigned qb = (c * fs + ia) * (unsigned)sizeof(TS);
    const TC is = St<TS>::scaled ? (TC)a.isc[c] : TC(1);
    if constexpr (MODE == MODE_RHS) {
      St<TS>::st(&at_b(F.a_out + fso, qb), Ph[c] * is);
    } else if constexpr (MODE == MODE_S0) {
      St<TS>::st(&at_b(F.a_out + fso, qb), (P[c] + a.ca * Ph[c]) * is);
    } else if constexpr (MODE == MODE_S12) {
      St<TS>::st(&at_b(F.a_out + fso, qb), (pre_a[c] + a.ca * Ph[c]) * is);
    } else {
      const TC part = rk_partial(a.rk_w, pre_a[c], u1_a[c], u2_a[c]);
      St<TS>::st(&at_b(F.a_end + fso, qb), rk_final(pre_a[c], part, a.rk_w[2], P[c], a.cb * Ph[c]) * is);
    }
  }
}

template <typename TS, typename TC>
__device__ __forceinline__ bool pml_in_face(const StageArgs<TS, TC> &a, int fi, int i, int j, int k) {
  const FaceArgs<TS, TC> &F = a.f[fi];
  return F.on && i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2;
}

// pml_face for a point off the free-surface row (no surface term), the form
// the z-marching sweep uses: the face axis DIM and the RK mode MODE are
// compile-time constants (straight-line component loops whose loads are
// unconditional), every face argument is read once, and on z faces the
// profile index is uniform across the wave, so the three profile values come
// through scalar loads.
template <int DIM, int MODE, typename TS, typename TC>
__device__ __forceinline__ void pml_face_march(const StageArgs<TS, TC> &a, int fi, int shot, int i, int j, int k,
                                               const TC *Dd, TC e1, TC e2, TC e3, TC lam, TC mu, TC lam2mu, TC slw,
                                               TC *h) {
  const FaceArgs<TS, TC> &F = a.f[fi];
  const int i1 = F.i1, j1 = F.j1, k1 = F.k1, bi = F.bi, bj = F.bj;
  const long fso = (long)shot * F.shot_stride;
  const TS *pin = F.a_in + fso, *ppre = F.a_pre + fso, *pu2 = F.a_u2 + fso;
  TS *pout = F.a_out + fso, *pend = F.a_end + fso;
  const unsigned fs = (unsigned)F.fsize;
  const unsigned ia = (unsigned)(((k - k1) * bj + (j - j1)) * bi + (i - i1));
  auto qb = [&](int c) { return (c * fs + ia) * (unsigned)sizeof(TS); };
  auto S = [&](int c) -> TC { return St<TS>::scaled ? (TC)a.sc[c] : TC(1); };
  auto IS = [&](int c) -> TC { return St<TS>::scaled ? (TC)a.isc[c] : TC(1); };
  TC R[kNC];
  R[VX] = slw * (e1 * Dd[TXX] + e2 * Dd[TXY] + e3 * Dd[TXZ]);
  R[VY] = slw * (e1 * Dd[TXY] + e2 * Dd[TYY] + e3 * Dd[TYZ]);
  R[VZ] = slw * (e1 * Dd[TXZ] + e2 * Dd[TYZ] + e3 * Dd[TZZ]);
  R[TXX] = lam2mu * e1 * Dd[VX] + lam * e2 * Dd[VY] + lam * e3 * Dd[VZ];
  R[TYY] = lam * e1 * Dd[VX] + lam2mu * e2 * Dd[VY] + lam * e3 * Dd[VZ];
  R[TZZ] = lam * e1 * Dd[VX] + lam * e2 * Dd[VY] + lam2mu * e3 * Dd[VZ];
  R[TXY] = mu * (e2 * Dd[VX] + e1 * Dd[VY]);
  R[TXZ] = mu * (e3 * Dd[VX] + e1 * Dd[VZ]);
  R[TYZ] = mu * (e3 * Dd[VY] + e2 * Dd[VZ]);
  TC P[kNC], pre[kNC], part[kNC];
#pragma unroll
  for (int c = 0; c < kNC; c++) P[c] = (TC)St<TS>::ld(&at_b(pin, qb(c))) * S(c);
  if constexpr (MODE == MODE_S12 || MODE == MODE_S3) {
#pragma unroll
    for (int c = 0; c < kNC; c++) pre[c] = (TC)St<TS>::ld(&at_b(ppre, qb(c))) * S(c);
  }
  if constexpr (MODE == MODE_S3) {
#pragma unroll
    for (int c = 0; c < kNC; c++)
      part[c] = rk_partial(a.rk_w, pre[c], (TC)St<TS>::ld(&at_b(pend, qb(c))) * S(c),
                           (TC)St<TS>::ld(&at_b(pu2, qb(c))) * S(c));
  }
  TC cD, cA, cB;
  if constexpr (DIM == 2) {
    typedef const __attribute__((address_space(4))) TC *cptr;
    const int ic = k - k1;
    cD = ((cptr)F.D)[ic];
    cA = ((cptr)F.A)[ic];
    cB = ((cptr)F.B)[ic];
  } else {
    const int ic = DIM == 0 ? i - i1 : j - j1;
    cD = F.D[ic];
    cA = F.A[ic];
    cB = F.B[ic];
  }
  const TC Bm1 = cB - TC(1);
  TC Ph[kNC];
#pragma unroll
  for (int c = 0; c < kNC; c++) {
    h[c] += Bm1 * R[c] - cB * P[c];
    Ph[c] = cD * R[c] - cA * P[c];
  }
  if constexpr (MODE == MODE_S12) {
#pragma unroll
    for (int c = 0; c < kNC; c++) St<TS>::st(&at_b(pout, qb(c)), (pre[c] + a.ca * Ph[c]) * IS(c));
  } else if constexpr (MODE == MODE_S0) {
#pragma unroll
    for (int c = 0; c < kNC; c++) St<TS>::st(&at_b(pout, qb(c)), (P[c] + a.ca * Ph[c]) * IS(c));
  } else if constexpr (MODE == MODE_S3) {
#pragma unroll
    for (int c = 0; c < kNC; c++)
      St<TS>::st(&at_b(pend, qb(c)), rk_final(pre[c], part[c], a.rk_w[2], P[c], a.cb * Ph[c]) * IS(c));
  } else {
#pragma unroll
    for (int c = 0; c < kNC; c++) St<TS>::st(&at_b(pout, qb(c)), Ph[c] * IS(c));
  }
}

// Every face box containing the point, in the reference's face order.
template <int MODE, typename TS, typename TC>
__device__ __forceinline__ void pml_apply(const StageArgs<TS, TC> &a, int shot, int i, int j, int k, const TC *Dx,
                                          const TC *Dy, const TC *Dz, TC xix, TC xiy, TC xiz, TC etx, TC ety,
                                          TC etz, TC ztx, TC zty, TC ztz, TC lam, TC mu, TC lam2mu, TC slw,
                                          TC *h) {
#pragma unroll
  for (int fi = 0; fi < 6; fi++) {
    const FaceArgs<TS, TC> &F = a.f[fi];
    if (!F.on) continue;
    if (i < F.i1 || i > F.i2 || j < F.j1 || j > F.j2 || k < F.k1 || k > F.k2) continue;
    const int dim = fi >> 1;
    const TC *Dd = (dim == 0) ? Dx : (dim == 1) ? Dy : Dz;
    const TC e1 = (dim == 0) ? xix : (dim == 1) ? etx : ztx;
    const TC e2 = (dim == 0) ? xiy : (dim == 1) ? ety : zty;
    const TC e3 = (dim == 0) ? xiz : (dim == 1) ? etz : ztz;
    pml_face<MODE, TS, TC>(a, fi, shot, i, j, k, Dd, e1, e2, e3, ztx, zty, ztz, lam, mu, lam2mu, slw, h);
  }
}

}  // namespace cgfd
