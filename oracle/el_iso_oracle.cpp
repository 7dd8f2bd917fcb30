/*
 * el_iso_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of CGFD3D's elastic-isotropic RK4 time-stepping path
 * (curvilinear collocated MacCormack/DRP RHS, traction-image and
 * velocity-gradient free surface, CFS-PML ADE, moment/force source,
 * trilinear receivers) written from the formulas in the reference files
 * cited below.  It keeps the reference's *unfused* stage order (RHS over
 * all points, then the RK axpys over whole levels) and its float/double
 * promotions, so that the FP32 instance is "the reference's FP32 algorithm"
 * and the FP64 instance is the same scheme in double precision.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load this library, and only as the checker (or as the timed CPU baseline).
 * The product (the HIP library) never links or calls it.
 *
 * All array arguments are double for ctypes convenience; the instance type R
 * (float or double) is what the arithmetic is carried out in.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/forward/fd_t.cu:18-310           operator tables, RK coefficients
 *   src/forward/drv_rk_curv_col.cu:193-542  time loop, stage order, swaps
 *   src/forward/sv_curv_col_el_iso_gpu.cu:291-452   inner RHS
 *   src/forward/sv_curv_col_el_gpu.cu:19-221        traction image (timg)
 *   src/forward/sv_curv_col_el_iso_gpu.cu:462-603   vlow free-surface rows
 *   src/forward/sv_curv_col_el_iso_gpu.cu:680-1147  CFS-PML
 *   src/forward/sv_curv_col_el_gpu.cu:227-301       source injection
 *   src/forward/sv_curv_col_el_iso_gpu.cu:1152-1261 free-surface matrices
 *   src/lib/fdlib_math.cu:5-95                      3x3 inverse, products
 *   src/forward/wav_t.cu:362-378                    RK axpys
 *   src/forward/bdry_t.cu:85-451                    PML profiles
 *   src/forward/gd_t.cu:209-314,404-498,599-630     metrics, ghosts, cart grid
 *   src/forward/src_t.cu:143-950                    source tables
 *   src/forward/io_funcs.cu:24-299,1624-1666        receivers
 *   src/forward/blk_t.cu:673-763                    dt estimate
 *   src/forward/md_t.cu:622-637                     rho -> 1/rho
 *   src/forward/wav_t.cu:281-360                    peak ground motion maps
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int NG = 3;      // ghost points per side (fd_t.cu:64-66)
constexpr int NCMP = 9;    // Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy (wav_t.cu:100-140)
enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };
// metric component order (gd_t.cu metric init: jac, xi, eta, zeta)
enum { JAC = 0, XIX, XIY, XIZ, ETX, ETY, ETZ, ZTX, ZTY, ZTZ, NMET };

// ---------------------------------------------------------------------------
// operator tables (fd_t.cu:76-148, 164-170, 184-194); layer 3 = DRP normal op
// ---------------------------------------------------------------------------
const int kIdx[4][2][5] = {
    {{-1, 0, 0, 0, 0}, {0, 1, 0, 0, 0}},
    {{-1, 0, 0, 0, 0}, {0, 1, 0, 0, 0}},
    {{-2, -1, 0, 0, 0}, {0, 1, 2, 0, 0}},
    {{-3, -2, -1, 0, 1}, {-1, 0, 1, 2, 3}}};
const double kCoef[4][2][5] = {
    {{-1.0, 1.0, 0, 0, 0}, {-1.0, 1.0, 0, 0, 0}},
    {{-1.0, 1.0, 0, 0, 0}, {-1.0, 1.0, 0, 0, 0}},
    {{1.0 / 6.0, -8.0 / 6.0, 7.0 / 6.0, 0, 0}, {-7.0 / 6.0, 8.0 / 6.0, -1.0 / 6.0, 0, 0}},
    {{-0.041686285, 0.333411808, -1.250117713, 0.666745142, 0.291647047},
     {-0.291647047, -0.666745142, 1.250117713, -0.333411808, 0.041686285}}};
const int kLen[4] = {2, 2, 3, 5};
const int kFlags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                          {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
const double kFdc[7] = {-0.02084, 0.1667, -0.7709, 0.0, 0.7709, -0.1667, 0.02084};
const double kRkA[3] = {0.5, 0.5, 1.0};
const double kRkB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
const double kRkT[4] = {0.0, 0.5, 0.5, 1.0};

struct Grid {
  int ni, nj, nk, nx, ny, nz;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  size_t siz_iy, siz_iz, siz_icmp;
  explicit Grid(int ni_, int nj_, int nk_) {
    ni = ni_; nj = nj_; nk = nk_;
    nx = ni + 2 * NG; ny = nj + 2 * NG; nz = nk + 2 * NG;
    ni1 = NG; ni2 = NG + ni - 1;
    nj1 = NG; nj2 = NG + nj - 1;
    nk1 = NG; nk2 = NG + nk - 1;
    siz_iy = nx; siz_iz = (size_t)nx * ny; siz_icmp = siz_iz * nz;
  }
};

// One operator instance in arithmetic type R.
template <typename R> struct Op {
  int len;
  int idx[5];
  R coef[5];
  long shift[5];
};

template <typename R> Op<R> make_op(int layer, int dir, long stride) {
  Op<R> op;
  op.len = kLen[layer];
  for (int n = 0; n < 5; n++) {
    op.idx[n] = kIdx[layer][dir][n];
    op.coef[n] = (R)kCoef[layer][dir][n];
    op.shift[n] = (long)op.idx[n] * stride;
  }
  return op;
}

// M_FD_SHIFT_PTR_MACDRP (fd_t.h:24-29): c0*f0 + c1*f1 + ... left to right
template <typename R> inline R d5(const R *p, const Op<R> &o) {
  return o.coef[0] * p[o.shift[0]] + o.coef[1] * p[o.shift[1]] + o.coef[2] * p[o.shift[2]] +
         o.coef[3] * p[o.shift[3]] + o.coef[4] * p[o.shift[4]];
}
// M_FD_SHIFT (fd_t.h:9-13): deriv = c0*f0; deriv += cn*fn
template <typename R> inline R dn(const R *p, const Op<R> &o) {
  R d = o.coef[0] * p[o.shift[0]];
  for (int n = 1; n < o.len; n++) d += o.coef[n] * p[o.shift[n]];
  return d;
}

// lam + 2.0*mu evaluated in double then rounded (sv_curv_col_el_iso_gpu.cu:408)
template <typename R> inline R lam2mu_of(R lam, R mu) { return (R)((double)lam + 2.0 * (double)mu); }

// ---------------------------------------------------------------------------
// problem description (C ABI for ctypes)
// ---------------------------------------------------------------------------
}  // namespace

extern "C" {

typedef struct {
  int ni, nj, nk;
  const double *metric;            /* NMET x siz_icmp, order jac,xi*,eta*,zeta* */
  const double *lam, *mu, *slw;    /* siz_icmp each, slw = 1/rho */
  int free_top;
  const double *matVx2Vz;          /* nx*ny*9 [j][i][row][col] */
  const double *matVy2Vz;
  int pml_on[6];                   /* x1,x2,y1,y2,z1,z2 */
  int pml_nlay[6];
  const double *pml_A[6], *pml_B[6], *pml_D[6]; /* nlay+1 each */
  int nsrc, force_on, moment_on, max_nt, max_ext;
  const int *src_it_begin, *src_it_end, *src_ext_num, *src_ext_indx;
  const double *src_ext_coef;      /* nsrc*max_ext */
  const double *src_F;             /* 3 x (nsrc*max_nt*4) */
  const double *src_M;             /* 6 x (nsrc*max_nt*4): Mxx,Myy,Mzz,Myz,Mxz,Mxy */
  int nrecv;
  const long long *recv_indx;      /* nrecv x 8 flat indices into one component */
  const double *recv_frac;         /* nrecv x 3: di, dj, dk */
} orc_problem;

}  // extern "C"

namespace {

// PML face box in local (ghost-inclusive) indices (bdry_t.cu:127-150)
struct Face {
  bool on;
  int idim, iside, nlay;
  int i1, i2, j1, j2, k1, k2;
  int ni, nj, nk;
  size_t size;  // points in the face box
};

inline void face_boxes(const Grid &g, const orc_problem *p, Face f[6]) {
  for (int idim = 0; idim < 3; idim++)
    for (int iside = 0; iside < 2; iside++) {
      Face &F = f[idim * 2 + iside];
      F.on = p->pml_on[idim * 2 + iside] != 0;
      F.idim = idim; F.iside = iside;
      F.nlay = F.on ? p->pml_nlay[idim * 2 + iside] : 0;
      F.i1 = g.ni1; F.i2 = g.ni2; F.j1 = g.nj1; F.j2 = g.nj2; F.k1 = g.nk1; F.k2 = g.nk2;
      if (idim == 0 && iside == 0) F.i2 = g.ni1 + F.nlay;
      if (idim == 0 && iside == 1) F.i1 = g.ni2 - F.nlay;
      if (idim == 1 && iside == 0) F.j2 = g.nj1 + F.nlay;
      if (idim == 1 && iside == 1) F.j1 = g.nj2 - F.nlay;
      if (idim == 2 && iside == 0) F.k2 = g.nk1 + F.nlay;
      if (idim == 2 && iside == 1) F.k1 = g.nk2 - F.nlay;
      F.ni = F.i2 - F.i1 + 1; F.nj = F.j2 - F.j1 + 1; F.nk = F.k2 - F.k1 + 1;
      F.size = F.on ? (size_t)F.ni * F.nj * F.nk : 0;
    }
}

template <typename R> struct Model {
  Grid g;
  std::vector<R> met, lam, mu, slw, mvx, mvy;
  int free_top;
  Face face[6];
  std::vector<R> pA[6], pB[6], pD[6];
  // source
  int nsrc, force_on, moment_on, max_nt, max_ext;
  std::vector<int> it_begin, it_end, ext_num, ext_indx;
  std::vector<R> ext_coef, F[3], M[6];
  // receivers
  int nrecv;
  std::vector<long long> rindx;
  std::vector<R> rfrac;

  explicit Model(const orc_problem *p) : g(p->ni, p->nj, p->nk) {
    size_t n = g.siz_icmp;
    met.resize(NMET * n);
    for (size_t i = 0; i < NMET * n; i++) met[i] = (R)p->metric[i];
    lam.resize(n); mu.resize(n); slw.resize(n);
    for (size_t i = 0; i < n; i++) { lam[i] = (R)p->lam[i]; mu[i] = (R)p->mu[i]; slw[i] = (R)p->slw[i]; }
    free_top = p->free_top;
    size_t nm = g.siz_iz * 9;
    mvx.assign(nm, R(0)); mvy.assign(nm, R(0));
    if (free_top && p->matVx2Vz && p->matVy2Vz)
      for (size_t i = 0; i < nm; i++) { mvx[i] = (R)p->matVx2Vz[i]; mvy[i] = (R)p->matVy2Vz[i]; }
    face_boxes(g, p, face);
    for (int f = 0; f < 6; f++) {
      if (!face[f].on) continue;
      int np = face[f].nlay + 1;
      pA[f].resize(np); pB[f].resize(np); pD[f].resize(np);
      for (int i = 0; i < np; i++) { pA[f][i] = (R)p->pml_A[f][i]; pB[f][i] = (R)p->pml_B[f][i]; pD[f][i] = (R)p->pml_D[f][i]; }
    }
    nsrc = p->nsrc; force_on = p->force_on; moment_on = p->moment_on;
    max_nt = p->max_nt; max_ext = p->max_ext;
    if (nsrc > 0) {
      it_begin.assign(p->src_it_begin, p->src_it_begin + nsrc);
      it_end.assign(p->src_it_end, p->src_it_end + nsrc);
      ext_num.assign(p->src_ext_num, p->src_ext_num + nsrc);
      ext_indx.assign(p->src_ext_indx, p->src_ext_indx + (size_t)nsrc * max_ext);
      ext_coef.resize((size_t)nsrc * max_ext);
      for (size_t i = 0; i < ext_coef.size(); i++) ext_coef[i] = (R)p->src_ext_coef[i];
      size_t nt4 = (size_t)nsrc * max_nt * 4;
      if (force_on)
        for (int c = 0; c < 3; c++) {
          F[c].resize(nt4);
          for (size_t i = 0; i < nt4; i++) F[c][i] = (R)p->src_F[c * nt4 + i];
        }
      if (moment_on)
        for (int c = 0; c < 6; c++) {
          M[c].resize(nt4);
          for (size_t i = 0; i < nt4; i++) M[c][i] = (R)p->src_M[c * nt4 + i];
        }
    }
    nrecv = p->nrecv;
    if (nrecv > 0) {
      rindx.assign(p->recv_indx, p->recv_indx + (size_t)nrecv * 8);
      rfrac.resize((size_t)nrecv * 3);
      for (size_t i = 0; i < rfrac.size(); i++) rfrac[i] = (R)p->recv_frac[i];
    }
  }
  const R *m(int c) const { return met.data() + (size_t)c * g.siz_icmp; }
};

// ---------------------------------------------------------------------------
// one RK stage right-hand side, reference order (sv_curv_col_el_iso_gpu.cu:23-285)
// ---------------------------------------------------------------------------
template <typename R>
void onestage(const Model<R> &M, const R *w, R *rhs, R *const aux_cur[6], R *const aux_rhs[6],
              int ipair, int istage, int it) {
  const Grid &g = M.g;
  const size_t sc = g.siz_icmp;
  const int idir = (kFlags[ipair][0] + istage) % 2;
  const int jdir = (kFlags[ipair][1] + istage) % 2;
  const int kdir = (kFlags[ipair][2] + istage) % 2;
  const Op<R> ox = make_op<R>(3, idir, 1);
  const Op<R> oy = make_op<R>(3, jdir, (long)g.siz_iy);
  const Op<R> oz = make_op<R>(3, kdir, (long)g.siz_iz);

  const R *Vx = w + VX * sc, *Vy = w + VY * sc, *Vz = w + VZ * sc;
  const R *Txx = w + TXX * sc, *Tyy = w + TYY * sc, *Tzz = w + TZZ * sc;
  const R *Tyz = w + TYZ * sc, *Txz = w + TXZ * sc, *Txy = w + TXY * sc;
  R *hVx = rhs + VX * sc, *hVy = rhs + VY * sc, *hVz = rhs + VZ * sc;
  R *hTxx = rhs + TXX * sc, *hTyy = rhs + TYY * sc, *hTzz = rhs + TZZ * sc;
  R *hTyz = rhs + TYZ * sc, *hTxz = rhs + TXZ * sc, *hTxy = rhs + TXY * sc;
  const R *xi_x = M.m(XIX), *xi_y = M.m(XIY), *xi_z = M.m(XIZ);
  const R *et_x = M.m(ETX), *et_y = M.m(ETY), *et_z = M.m(ETZ);
  const R *zt_x = M.m(ZTX), *zt_y = M.m(ZTY), *zt_z = M.m(ZTZ);
  const R *jac3d = M.m(JAC);
  const R *lam3d = M.lam.data(), *mu3d = M.mu.data(), *slw3d = M.slw.data();

  // ---- inner (sv_curv_col_el_iso_gpu.cu:291-452) ----
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = g.nk1; k <= g.nk2; k++)
    for (int j = g.nj1; j <= g.nj2; j++)
      for (int i = g.ni1; i <= g.ni2; i++) {
        size_t iptr = i + j * g.siz_iy + k * g.siz_iz;
        R DxVx = d5(Vx + iptr, ox), DyVx = d5(Vx + iptr, oy), DzVx = d5(Vx + iptr, oz);
        R DxVy = d5(Vy + iptr, ox), DyVy = d5(Vy + iptr, oy), DzVy = d5(Vy + iptr, oz);
        R DxVz = d5(Vz + iptr, ox), DyVz = d5(Vz + iptr, oy), DzVz = d5(Vz + iptr, oz);
        R DxTxx = d5(Txx + iptr, ox), DyTxx = d5(Txx + iptr, oy), DzTxx = d5(Txx + iptr, oz);
        R DxTyy = d5(Tyy + iptr, ox), DyTyy = d5(Tyy + iptr, oy), DzTyy = d5(Tyy + iptr, oz);
        R DxTzz = d5(Tzz + iptr, ox), DyTzz = d5(Tzz + iptr, oy), DzTzz = d5(Tzz + iptr, oz);
        R DxTxz = d5(Txz + iptr, ox), DyTxz = d5(Txz + iptr, oy), DzTxz = d5(Txz + iptr, oz);
        R DxTyz = d5(Tyz + iptr, ox), DyTyz = d5(Tyz + iptr, oy), DzTyz = d5(Tyz + iptr, oz);
        R DxTxy = d5(Txy + iptr, ox), DyTxy = d5(Txy + iptr, oy), DzTxy = d5(Txy + iptr, oz);
        R xix = xi_x[iptr], xiy = xi_y[iptr], xiz = xi_z[iptr];
        R etx = et_x[iptr], ety = et_y[iptr], etz = et_z[iptr];
        R ztx = zt_x[iptr], zty = zt_y[iptr], ztz = zt_z[iptr];
        R lam = lam3d[iptr], mu = mu3d[iptr], slw = slw3d[iptr];
        R lam2mu = lam2mu_of(lam, mu);
        hVx[iptr] = slw * (xix * DxTxx + xiy * DxTxy + xiz * DxTxz + etx * DyTxx + ety * DyTxy +
                           etz * DyTxz + ztx * DzTxx + zty * DzTxy + ztz * DzTxz);
        hVy[iptr] = slw * (xix * DxTxy + xiy * DxTyy + xiz * DxTyz + etx * DyTxy + ety * DyTyy +
                           etz * DyTyz + ztx * DzTxy + zty * DzTyy + ztz * DzTyz);
        hVz[iptr] = slw * (xix * DxTxz + xiy * DxTyz + xiz * DxTzz + etx * DyTxz + ety * DyTyz +
                           etz * DyTzz + ztx * DzTxz + zty * DzTyz + ztz * DzTzz);
        hTxx[iptr] = lam2mu * (xix * DxVx + etx * DyVx + ztx * DzVx) +
                     lam * (xiy * DxVy + ety * DyVy + zty * DzVy + xiz * DxVz + etz * DyVz + ztz * DzVz);
        hTyy[iptr] = lam2mu * (xiy * DxVy + ety * DyVy + zty * DzVy) +
                     lam * (xix * DxVx + etx * DyVx + ztx * DzVx + xiz * DxVz + etz * DyVz + ztz * DzVz);
        hTzz[iptr] = lam2mu * (xiz * DxVz + etz * DyVz + ztz * DzVz) +
                     lam * (xix * DxVx + etx * DyVx + ztx * DzVx + xiy * DxVy + ety * DyVy + zty * DzVy);
        hTxy[iptr] = mu * (xiy * DxVx + xix * DxVy + ety * DyVx + etx * DyVy + zty * DzVx + ztx * DzVy);
        hTxz[iptr] = mu * (xiz * DxVx + xix * DxVz + etz * DyVx + etx * DyVz + ztz * DzVx + ztx * DzVz);
        hTyz[iptr] = mu * (xiz * DxVy + xiy * DxVz + etz * DyVy + ety * DyVz + ztz * DzVy + zty * DzVz);
      }

  if (M.free_top) {
    // ---- traction image (sv_curv_col_el_gpu.cu:19-221) ----
    const int *fdz_indx = oz.idx;
    const int k_min = g.nk2 - fdz_indx[oz.len - 1];
    for (int k = k_min; k <= g.nk2; k++) {
      const int n_free = g.nk2 - k - fdz_indx[0];
#pragma omp parallel for schedule(static)
      for (int j = g.nj1; j <= g.nj2; j++)
        for (int i = g.ni1; i <= g.ni2; i++) {
          size_t iptr = i + j * g.siz_iy + k * g.siz_iz;
          R slwjac = slw3d[iptr] / jac3d[iptr];
          const R *T1s[3] = {Txx, Txy, Txz};
          const R *T2s[3] = {Txy, Tyy, Tyz};
          const R *T3s[3] = {Txz, Tyz, Tzz};
          R *outs[3] = {hVx, hVy, hVz};
          for (int c = 0; c < 3; c++) {
            const R *T1 = T1s[c], *T2 = T2s[c], *T3 = T3s[c];
            R vecxi[5] = {0, 0, 0, 0, 0}, vecet[5] = {0, 0, 0, 0, 0}, veczt[5] = {0, 0, 0, 0, 0};
            for (int n = 0; n < ox.len; n++) {
              size_t q = iptr + ox.idx[n];
              vecxi[n] = jac3d[q] * (xi_x[q] * T1[q] + xi_y[q] * T2[q] + xi_z[q] * T3[q]);
            }
            for (int n = 0; n < oy.len; n++) {
              size_t q = iptr + (long)oy.idx[n] * (long)g.siz_iy;
              vecet[n] = jac3d[q] * (et_x[q] * T1[q] + et_y[q] * T2[q] + et_z[q] * T3[q]);
            }
            int n;
            for (n = 0; n < n_free; n++) {
              size_t q = iptr + (long)fdz_indx[n] * (long)g.siz_iz;
              veczt[n] = jac3d[q] * (zt_x[q] * T1[q] + zt_y[q] * T2[q] + zt_z[q] * T3[q]);
            }
            veczt[n_free] = 0;
            for (n = n_free + 1; n < oz.len; n++) {
              int n_img = fdz_indx[n] - 2 * (n - n_free);
              size_t q = iptr + (long)n_img * (long)g.siz_iz;
              veczt[n] = -jac3d[q] * (zt_x[q] * T1[q] + zt_y[q] * T2[q] + zt_z[q] * T3[q]);
            }
            R DxTx = ox.coef[0] * vecxi[0];
            for (int m = 1; m < ox.len; m++) DxTx += ox.coef[m] * vecxi[m];
            R DyTy = oy.coef[0] * vecet[0];
            for (int m = 1; m < oy.len; m++) DyTy += oy.coef[m] * vecet[m];
            R DzTz = oz.coef[0] * veczt[0];
            for (int m = 1; m < oz.len; m++) DzTz += oz.coef[m] * veczt[m];
            outs[c][iptr] = (DxTx + DyTy + DzTz) * slwjac;
          }
        }
    }

    // ---- vlow (sv_curv_col_el_iso_gpu.cu:462-603) ----
    for (int n = 0; n < 3; n++) {
      const int k = g.nk2 - n;
      const Op<R> ozl = make_op<R>(n, kdir, (long)g.siz_iz);
#pragma omp parallel for schedule(static)
      for (int j = g.nj1; j <= g.nj2; j++)
        for (int i = g.ni1; i <= g.ni2; i++) {
          size_t iptr = i + j * g.siz_iy + k * g.siz_iz;
          R xix = xi_x[iptr], xiy = xi_y[iptr], xiz = xi_z[iptr];
          R etx = et_x[iptr], ety = et_y[iptr], etz = et_z[iptr];
          R ztx = zt_x[iptr], zty = zt_y[iptr], ztz = zt_z[iptr];
          R lam = lam3d[iptr], mu = mu3d[iptr];
          R lam2mu = lam2mu_of(lam, mu);
          R DxVx = dn(Vx + iptr, ox), DyVx = dn(Vx + iptr, oy);
          R DxVy = dn(Vy + iptr, ox), DyVy = dn(Vy + iptr, oy);
          R DxVz = dn(Vz + iptr, ox), DyVz = dn(Vz + iptr, oy);
          R DzVx, DzVy, DzVz;
          if (k == g.nk2) {
            size_t ij = ((size_t)i + (size_t)j * g.siz_iy) * 9;
            const R *A = M.mvx.data() + ij, *B = M.mvy.data() + ij;
            DzVx = A[0] * DxVx + A[1] * DxVy + A[2] * DxVz + B[0] * DyVx + B[1] * DyVy + B[2] * DyVz;
            DzVy = A[3] * DxVx + A[4] * DxVy + A[5] * DxVz + B[3] * DyVx + B[4] * DyVy + B[5] * DyVz;
            DzVz = A[6] * DxVx + A[7] * DxVy + A[8] * DxVz + B[6] * DyVx + B[7] * DyVy + B[8] * DyVz;
          } else {
            DzVx = dn(Vx + iptr, ozl);
            DzVy = dn(Vy + iptr, ozl);
            DzVz = dn(Vz + iptr, ozl);
          }
          hTxx[iptr] = lam2mu * (xix * DxVx + etx * DyVx + ztx * DzVx) +
                       lam * (xiy * DxVy + ety * DyVy + zty * DzVy + xiz * DxVz + etz * DyVz + ztz * DzVz);
          hTyy[iptr] = lam2mu * (xiy * DxVy + ety * DyVy + zty * DzVy) +
                       lam * (xix * DxVx + etx * DyVx + ztx * DzVx + xiz * DxVz + etz * DyVz + ztz * DzVz);
          hTzz[iptr] = lam2mu * (xiz * DxVz + etz * DyVz + ztz * DzVz) +
                       lam * (xix * DxVx + etx * DyVx + ztx * DzVx + xiy * DxVy + ety * DyVy + zty * DzVy);
          hTxy[iptr] = mu * (xiy * DxVx + xix * DxVy + ety * DyVx + etx * DyVy + zty * DzVx + ztx * DzVy);
          hTxz[iptr] = mu * (xiz * DxVx + xix * DxVz + etz * DyVx + etx * DyVz + ztz * DzVx + ztx * DzVz);
          hTyz[iptr] = mu * (xiz * DxVy + xiy * DxVz + etz * DyVy + ety * DyVz + ztz * DzVy + zty * DzVz);
        }
    }
  }

  // ---- CFS-PML, faces in order x1,x2,y1,y2,z1,z2 (sv_curv_col_el_iso_gpu.cu:613-1147) ----
  for (int f = 0; f < 6; f++) {
    const Face &F = M.face[f];
    if (!F.on) continue;
    const R *cA = M.pA[f].data(), *cB = M.pB[f].data(), *cD = M.pD[f].data();
    const R *P = aux_cur[f];
    R *Ph = aux_rhs[f];
    const size_t fs = F.size;
    const Op<R> &od = (F.idim == 0) ? ox : (F.idim == 1) ? oy : oz;
    const R *m1 = (F.idim == 0) ? xi_x : (F.idim == 1) ? et_x : zt_x;
    const R *m2 = (F.idim == 0) ? xi_y : (F.idim == 1) ? et_y : zt_y;
    const R *m3 = (F.idim == 0) ? xi_z : (F.idim == 1) ? et_z : zt_z;
    const R *mat = (F.idim == 0) ? M.mvx.data() : M.mvy.data();
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < F.nk; iz++)
      for (int iy = 0; iy < F.nj; iy++)
        for (int ix = 0; ix < F.ni; ix++) {
          size_t ia = (size_t)iz * (F.nj * F.ni) + (size_t)iy * F.ni + ix;
          size_t iptr = (size_t)(ix + F.i1) + (size_t)(iy + F.j1) * g.siz_iy + (size_t)(iz + F.k1) * g.siz_iz;
          int ic = (F.idim == 0) ? ix : (F.idim == 1) ? iy : iz;
          R cD_ = cD[ic], cA_ = cA[ic], cB_ = cB[ic];
          R Bm1 = (R)((double)cB_ - 1.0);
          R e1 = m1[iptr], e2 = m2[iptr], e3 = m3[iptr];
          R lam = lam3d[iptr], mu = mu3d[iptr], slw = slw3d[iptr];
          R lam2mu = lam2mu_of(lam, mu);
          R DVx = d5(Vx + iptr, od), DVy = d5(Vy + iptr, od), DVz = d5(Vz + iptr, od);
          R DTxx = d5(Txx + iptr, od), DTyy = d5(Tyy + iptr, od), DTzz = d5(Tzz + iptr, od);
          R DTxz = d5(Txz + iptr, od), DTyz = d5(Tyz + iptr, od), DTxy = d5(Txy + iptr, od);
          R hVx_r = slw * (e1 * DTxx + e2 * DTxy + e3 * DTxz);
          R hVy_r = slw * (e1 * DTxy + e2 * DTyy + e3 * DTyz);
          R hVz_r = slw * (e1 * DTxz + e2 * DTyz + e3 * DTzz);
          R hTxx_r = lam2mu * e1 * DVx + lam * e2 * DVy + lam * e3 * DVz;
          R hTyy_r = lam * e1 * DVx + lam2mu * e2 * DVy + lam * e3 * DVz;
          R hTzz_r = lam * e1 * DVx + lam * e2 * DVy + lam2mu * e3 * DVz;
          R hTxy_r = mu * (e2 * DVx + e1 * DVy);
          R hTxz_r = mu * (e3 * DVx + e1 * DVz);
          R hTyz_r = mu * (e3 * DVy + e2 * DVz);
          hVx[iptr] += Bm1 * hVx_r - cB_ * P[VX * fs + ia];
          hVy[iptr] += Bm1 * hVy_r - cB_ * P[VY * fs + ia];
          hVz[iptr] += Bm1 * hVz_r - cB_ * P[VZ * fs + ia];
          hTxx[iptr] += Bm1 * hTxx_r - cB_ * P[TXX * fs + ia];
          hTyy[iptr] += Bm1 * hTyy_r - cB_ * P[TYY * fs + ia];
          hTzz[iptr] += Bm1 * hTzz_r - cB_ * P[TZZ * fs + ia];
          hTxz[iptr] += Bm1 * hTxz_r - cB_ * P[TXZ * fs + ia];
          hTyz[iptr] += Bm1 * hTyz_r - cB_ * P[TYZ * fs + ia];
          hTxy[iptr] += Bm1 * hTxy_r - cB_ * P[TXY * fs + ia];
          Ph[VX * fs + ia] = cD_ * hVx_r - cA_ * P[VX * fs + ia];
          Ph[VY * fs + ia] = cD_ * hVy_r - cA_ * P[VY * fs + ia];
          Ph[VZ * fs + ia] = cD_ * hVz_r - cA_ * P[VZ * fs + ia];
          Ph[TXX * fs + ia] = cD_ * hTxx_r - cA_ * P[TXX * fs + ia];
          Ph[TYY * fs + ia] = cD_ * hTyy_r - cA_ * P[TYY * fs + ia];
          Ph[TZZ * fs + ia] = cD_ * hTzz_r - cA_ * P[TZZ * fs + ia];
          Ph[TXZ * fs + ia] = cD_ * hTxz_r - cA_ * P[TXZ * fs + ia];
          Ph[TYZ * fs + ia] = cD_ * hTyz_r - cA_ * P[TYZ * fs + ia];
          Ph[TXY * fs + ia] = cD_ * hTxy_r - cA_ * P[TXY * fs + ia];
          if (M.free_top && F.idim < 2 && (iz + F.k1) == g.nk2) {
            size_t ij = ((size_t)(ix + F.i1) + (size_t)(iy + F.j1) * g.siz_iy) * 9;
            const R *A = mat + ij;
            R D_DzVx = A[0] * DVx + A[1] * DVy + A[2] * DVz;
            R D_DzVy = A[3] * DVx + A[4] * DVy + A[5] * DVz;
            R D_DzVz = A[6] * DVx + A[7] * DVy + A[8] * DVz;
            R ztx = zt_x[iptr], zty = zt_y[iptr], ztz = zt_z[iptr];
            R tTxx = lam2mu * (ztx * D_DzVx) + lam * (zty * D_DzVy + ztz * D_DzVz);
            R tTyy = lam2mu * (zty * D_DzVy) + lam * (ztx * D_DzVx + ztz * D_DzVz);
            R tTzz = lam2mu * (ztz * D_DzVz) + lam * (ztx * D_DzVx + zty * D_DzVy);
            R tTxy = mu * (zty * D_DzVx + ztx * D_DzVy);
            R tTxz = mu * (ztz * D_DzVx + ztx * D_DzVz);
            R tTyz = mu * (ztz * D_DzVy + zty * D_DzVz);
            // (coef_B - 1.0) * rhs accumulated in double (sv_curv_col_el_iso_gpu.cu:900-905)
            double bm1 = (double)cB_ - 1.0;
            hTxx[iptr] = (R)((double)hTxx[iptr] + bm1 * (double)tTxx);
            hTyy[iptr] = (R)((double)hTyy[iptr] + bm1 * (double)tTyy);
            hTzz[iptr] = (R)((double)hTzz[iptr] + bm1 * (double)tTzz);
            hTxz[iptr] = (R)((double)hTxz[iptr] + bm1 * (double)tTxz);
            hTyz[iptr] = (R)((double)hTyz[iptr] + bm1 * (double)tTyz);
            hTxy[iptr] = (R)((double)hTxy[iptr] + bm1 * (double)tTxy);
            Ph[TXX * fs + ia] += cD_ * tTxx;
            Ph[TYY * fs + ia] += cD_ * tTyy;
            Ph[TZZ * fs + ia] += cD_ * tTzz;
            Ph[TXZ * fs + ia] += cD_ * tTxz;
            Ph[TYZ * fs + ia] += cD_ * tTyz;
            Ph[TXY * fs + ia] += cD_ * tTxy;
          }
        }
  }

  // ---- source (sv_curv_col_el_gpu.cu:227-301), sources in order ----
  for (int is = 0; is < M.nsrc; is++) {
    if (it < M.it_begin[is] || it > M.it_end[is]) continue;
    size_t ist = (size_t)is * M.max_nt * 4 + (size_t)(it - M.it_begin[is]) * 4 + istage;
    for (int e = 0; e < M.ext_num[is]; e++) {
      size_t iptr = (size_t)M.ext_indx[(size_t)is * M.max_ext + e];
      R coef = M.ext_coef[(size_t)is * M.max_ext + e];
      if (M.force_on) {
        R V = coef * slw3d[iptr] / jac3d[iptr];
        hVx[iptr] += M.F[0][ist] * V;
        hVy[iptr] += M.F[1][ist] * V;
        hVz[iptr] += M.F[2][ist] * V;
      }
      if (M.moment_on) {
        R rjac = coef / jac3d[iptr];
        hTxx[iptr] += -M.M[0][ist] * rjac;
        hTyy[iptr] += -M.M[1][ist] * rjac;
        hTzz[iptr] += -M.M[2][ist] * rjac;
        hTxz[iptr] += -M.M[4][ist] * rjac;
        hTyz[iptr] += -M.M[3][ist] * rjac;
        hTxy[iptr] += -M.M[5][ist] * rjac;
      }
    }
  }
}

// wav_update / wav_update_end (wav_t.cu:362-378)
template <typename R> void axpy_set(size_t n, R coef, R *out, const R *a, const R *b) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) out[i] = a[i] + coef * b[i];
}
template <typename R> void axpy_acc(size_t n, R coef, R *out, const R *b) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) out[i] = out[i] + coef * b[i];
}

// receivers: io_recv_keep (io_funcs.cu:1624-1666), with its corner-5 quirk
template <typename R> void recv_keep(const Model<R> &M, const R *w, int it, int nt_total, double *seismo) {
  const size_t sc = M.g.siz_icmp;
  for (int r = 0; r < M.nrecv; r++) {
    const long long *ix = M.rindx.data() + (size_t)r * 8;
    R Lx2 = M.rfrac[r * 3 + 0], Ly2 = M.rfrac[r * 3 + 1], Lz2 = M.rfrac[r * 3 + 2];
    R Lx1 = (R)(1.0 - (double)Lx2), Ly1 = (R)(1.0 - (double)Ly2), Lz1 = (R)(1.0 - (double)Lz2);
    for (int c = 0; c < NCMP; c++) {
      const R *v = w + c * sc;
      R b[8];
      for (int q = 0; q < 8; q++) b[q] = v[ix[q]];
      R s = b[0] * Lx1 * Ly1 * Lz1 + b[1] * Lx2 * Ly1 * Lz1 + b[2] * Lx1 * Ly2 * Lz1 + b[3] * Lx2 * Ly2 * Lz1 +
            b[5] * Lx1 * Ly1 * Lz2 + b[5] * Lx2 * Ly1 * Lz2 + b[6] * Lx1 * Ly2 * Lz2 + b[7] * Lx2 * Ly2 * Lz2;
      seismo[((size_t)r * NCMP + c) * nt_total + it] = (double)s;
    }
  }
}

double g_loop_seconds = 0.0;

template <typename R>
int run(const orc_problem *p, double dt_in, int nt_total, const double *init_wave, const double *init_aux,
        double *seismo, double *final_wave, double *final_aux) {
  Model<R> M(p);
  const Grid &g = M.g;
  const size_t lev = NCMP * g.siz_icmp;
  std::vector<R> wpre(lev, R(0)), wtmp(lev, R(0)), wrhs(lev, R(0)), wend(lev, R(0));
  if (init_wave)
    for (size_t i = 0; i < lev; i++) wpre[i] = (R)init_wave[i];
  std::vector<R> apre[6], atmp[6], arhs[6], aend[6];
  for (int f = 0; f < 6; f++) {
    size_t n = NCMP * M.face[f].size;
    apre[f].assign(n, R(0)); atmp[f].assign(n, R(0)); arhs[f].assign(n, R(0)); aend[f].assign(n, R(0));
  }
  if (init_aux) {
    size_t off = 0;
    for (int f = 0; f < 6; f++) {
      size_t n = NCMP * M.face[f].size;
      for (size_t i = 0; i < n; i++) apre[f][i] = (R)init_aux[off + i];
      off += n;
    }
  }
  R *pre = wpre.data(), *tmp = wtmp.data(), *rhs = wrhs.data(), *end = wend.data();
  R *a_pre[6], *a_tmp[6], *a_rhs[6], *a_end[6], *a_cur[6];
  for (int f = 0; f < 6; f++) { a_pre[f] = apre[f].data(); a_tmp[f] = atmp[f].data(); a_rhs[f] = arhs[f].data(); a_end[f] = aend[f].data(); }
  const R dt = (R)dt_in;
  const auto t_loop0 = std::chrono::steady_clock::now();
  for (int it = 0; it < nt_total; it++) {
    if (seismo) recv_keep(M, pre, it, nt_total, seismo);
    const int ipair = it % 8;
    for (int istage = 0; istage < 4; istage++) {
      const R *cur = (istage == 0) ? pre : tmp;
      for (int f = 0; f < 6; f++) a_cur[f] = (istage == 0) ? a_pre[f] : a_tmp[f];
      onestage<R>(M, cur, rhs, a_cur, a_rhs, ipair, istage, it);
      if (istage == 0) {
        R ca = (R)kRkA[0] * dt, cb = (R)kRkB[0] * dt;
        axpy_set(lev, ca, tmp, pre, rhs);
        for (int f = 0; f < 6; f++) axpy_set(NCMP * M.face[f].size, ca, a_tmp[f], a_pre[f], a_rhs[f]);
        axpy_set(lev, cb, end, pre, rhs);
        for (int f = 0; f < 6; f++) axpy_set(NCMP * M.face[f].size, cb, a_end[f], a_pre[f], a_rhs[f]);
      } else if (istage < 3) {
        R ca = (R)kRkA[istage] * dt, cb = (R)kRkB[istage] * dt;
        axpy_set(lev, ca, tmp, pre, rhs);
        for (int f = 0; f < 6; f++) axpy_set(NCMP * M.face[f].size, ca, a_tmp[f], a_pre[f], a_rhs[f]);
        axpy_acc(lev, cb, end, rhs);
        for (int f = 0; f < 6; f++) axpy_acc(NCMP * M.face[f].size, cb, a_end[f], a_rhs[f]);
      } else {
        R cb = (R)kRkB[3] * dt;
        axpy_acc(lev, cb, end, rhs);
        for (int f = 0; f < 6; f++) axpy_acc(NCMP * M.face[f].size, cb, a_end[f], a_rhs[f]);
      }
    }
    std::swap(pre, end);
    for (int f = 0; f < 6; f++) std::swap(a_pre[f], a_end[f]);
  }
  g_loop_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop0).count();
  if (final_wave)
    for (size_t i = 0; i < lev; i++) final_wave[i] = (double)pre[i];
  if (final_aux) {
    size_t off = 0;
    for (int f = 0; f < 6; f++) {
      size_t n = NCMP * M.face[f].size;
      for (size_t i = 0; i < n; i++) final_aux[off + i] = (double)a_pre[f][i];
      off += n;
    }
  }
  return 0;
}

template <typename R>
int stage_once(const orc_problem *p, int ipair, int istage, int it, const double *w_in, const double *aux_in,
               double *rhs_out, double *aux_rhs_out) {
  Model<R> M(p);
  const size_t lev = NCMP * M.g.siz_icmp;
  std::vector<R> w(lev), rhs(lev, R(0));
  for (size_t i = 0; i < lev; i++) w[i] = (R)w_in[i];
  std::vector<R> ac[6], ar[6];
  R *pc[6], *pr[6];
  size_t off = 0;
  for (int f = 0; f < 6; f++) {
    size_t n = NCMP * M.face[f].size;
    ac[f].assign(n, R(0)); ar[f].assign(n, R(0));
    if (aux_in)
      for (size_t i = 0; i < n; i++) ac[f][i] = (R)aux_in[off + i];
    off += n;
    pc[f] = ac[f].data(); pr[f] = ar[f].data();
  }
  onestage<R>(M, w.data(), rhs.data(), pc, pr, ipair, istage, it);
  for (size_t i = 0; i < lev; i++) rhs_out[i] = (double)rhs[i];
  if (aux_rhs_out) {
    off = 0;
    for (int f = 0; f < 6; f++) {
      size_t n = NCMP * M.face[f].size;
      for (size_t i = 0; i < n; i++) aux_rhs_out[off + i] = (double)ar[f][i];
      off += n;
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// setup restatements
// ---------------------------------------------------------------------------

// gd_curv_metric_cal + geometric_symmetry (gd_t.cu:209-314, 404-498)
template <typename R> void metric_cal(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                                      double *out) {
  Grid g(ni, nj, nk);
  const size_t n = g.siz_icmp;
  std::vector<R> x3(n), y3(n), z3(n), v(NMET * n, R(0));
  for (size_t i = 0; i < n; i++) { x3[i] = (R)x[i]; y3[i] = (R)y[i]; z3[i] = (R)z[i]; }
  R c[7];
  for (int q = 0; q < 7; q++) c[q] = (R)kFdc[q];
  const long sx = 1, sy = (long)g.siz_iy, sz = (long)g.siz_iz;
  auto der = [&](const R *f, size_t iptr, long s) {
    R d = c[0] * f[iptr - 3 * s];
    for (int q = 1; q < 7; q++) d += c[q] * f[iptr + (q - 3) * s];
    return d;
  };
  for (int k = g.nk1; k <= g.nk2; k++)
    for (int j = g.nj1; j <= g.nj2; j++)
      for (int i = g.ni1; i <= g.ni2; i++) {
        size_t iptr = i + j * g.siz_iy + k * g.siz_iz;
        R v1[3] = {der(x3.data(), iptr, sx), der(y3.data(), iptr, sx), der(z3.data(), iptr, sx)};
        R v2[3] = {der(x3.data(), iptr, sy), der(y3.data(), iptr, sy), der(z3.data(), iptr, sy)};
        R v3[3] = {der(x3.data(), iptr, sz), der(y3.data(), iptr, sz), der(z3.data(), iptr, sz)};
        auto cross = [](const R *a, const R *b, R *o) {
          o[0] = a[1] * b[2] - a[2] * b[1];
          o[1] = a[2] * b[0] - a[0] * b[2];
          o[2] = a[0] * b[1] - a[1] * b[0];
        };
        R vg[3];
        cross(v1, v2, vg);
        R jac = R(0);
        for (int q = 0; q < 3; q++) jac += vg[q] * v3[q];
        v[JAC * n + iptr] = jac;
        cross(v2, v3, vg);
        v[XIX * n + iptr] = vg[0] / jac; v[XIY * n + iptr] = vg[1] / jac; v[XIZ * n + iptr] = vg[2] / jac;
        cross(v3, v1, vg);
        v[ETX * n + iptr] = vg[0] / jac; v[ETY * n + iptr] = vg[1] / jac; v[ETZ * n + iptr] = vg[2] / jac;
        cross(v1, v2, vg);
        v[ZTX * n + iptr] = vg[0] / jac; v[ZTY * n + iptr] = vg[1] / jac; v[ZTZ * n + iptr] = vg[2] / jac;
      }
  for (int cmp = 0; cmp < NMET; cmp++) {
    R *a = v.data() + cmp * n;
    auto at = [&](int i, int j, int k) -> R & { return a[i + j * g.siz_iy + k * g.siz_iz]; };
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.ni1; i++) at(i, j, k) = 2 * at(g.ni1, j, k) - at(2 * g.ni1 - i, j, k);
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = g.ni2 + 1; i < g.nx; i++) at(i, j, k) = 2 * at(g.ni2, j, k) - at(2 * g.ni2 - i, j, k);
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.nj1; j++)
        for (int i = 0; i < g.nx; i++) at(i, j, k) = 2 * at(i, g.nj1, k) - at(i, 2 * g.nj1 - j, k);
    for (int k = 0; k < g.nz; k++)
      for (int j = g.nj2 + 1; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) at(i, j, k) = 2 * at(i, g.nj2, k) - at(i, 2 * g.nj2 - j, k);
    for (int k = 0; k < g.nk1; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) at(i, j, k) = 2 * at(i, j, g.nk1) - at(i, j, 2 * g.nk1 - k);
    for (int k = g.nk2 + 1; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) at(i, j, k) = 2 * at(i, j, g.nk2) - at(i, j, 2 * g.nk2 - k);
  }
  for (size_t i = 0; i < NMET * n; i++) out[i] = (double)v[i];
}

// fdlib_math_invert3x3 (fdlib_math.cu:5-36): cofactors in float, det in double
void invert3x3_f(float m[3][3]) {
  double inv[3][3];
  inv[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  inv[0][1] = m[2][1] * m[0][2] - m[0][1] * m[2][2];
  inv[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  inv[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  inv[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
  inv[1][2] = m[1][0] * m[0][2] - m[0][0] * m[1][2];
  inv[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  inv[2][1] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
  inv[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  double det = inv[0][0] * m[0][0] + inv[0][1] * m[1][0] + inv[0][2] * m[2][0];
  det = 1.0f / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = (float)(inv[i][j] * det);
}
void invert3x3_d(double m[3][3]) {
  double inv[3][3];
  inv[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  inv[0][1] = m[2][1] * m[0][2] - m[0][1] * m[2][2];
  inv[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  inv[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  inv[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
  inv[1][2] = m[1][0] * m[0][2] - m[0][0] * m[1][2];
  inv[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  inv[2][1] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
  inv[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  double det = inv[0][0] * m[0][0] + inv[0][1] * m[1][0] + inv[0][2] * m[2][0];
  det = 1.0 / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = inv[i][j] * det;
}
template <typename R> void invert3x3(R m[3][3]);
template <> void invert3x3<float>(float m[3][3]) { invert3x3_f(m); }
template <> void invert3x3<double>(double m[3][3]) { invert3x3_d(m); }

// sv_curv_col_el_iso_dvh2dvz (sv_curv_col_el_iso_gpu.cu:1152-1261)
template <typename R> void free_matrices(int ni, int nj, int nk, const double *metric, const double *lam3,
                                         const double *mu3, double *mvx, double *mvy) {
  Grid g(ni, nj, nk);
  const size_t n = g.siz_icmp;
  const int k = g.nk2;
  for (size_t q = 0; q < g.siz_iz * 9; q++) { mvx[q] = 0; mvy[q] = 0; }
  for (int j = g.nj1; j <= g.nj2; j++)
    for (int i = g.ni1; i <= g.ni2; i++) {
      size_t iptr = i + j * g.siz_iy + k * g.siz_iz;
      R e11 = (R)metric[XIX * n + iptr], e12 = (R)metric[XIY * n + iptr], e13 = (R)metric[XIZ * n + iptr];
      R e21 = (R)metric[ETX * n + iptr], e22 = (R)metric[ETY * n + iptr], e23 = (R)metric[ETZ * n + iptr];
      R e31 = (R)metric[ZTX * n + iptr], e32 = (R)metric[ZTY * n + iptr], e33 = (R)metric[ZTZ * n + iptr];
      R lam = (R)lam3[iptr], mu = (R)mu3[iptr];
      R lam2mu = lam + R(2) * mu;
      R A[3][3], B[3][3], C[3][3], AB[3][3], AC[3][3];
      A[0][0] = lam2mu * e31 * e31 + mu * (e32 * e32 + e33 * e33);
      A[0][1] = lam * e31 * e32 + mu * e32 * e31;
      A[0][2] = lam * e31 * e33 + mu * e33 * e31;
      A[1][0] = lam * e32 * e31 + mu * e31 * e32;
      A[1][1] = lam2mu * e32 * e32 + mu * (e31 * e31 + e33 * e33);
      A[1][2] = lam * e32 * e33 + mu * e33 * e32;
      A[2][0] = lam * e33 * e31 + mu * e31 * e33;
      A[2][1] = lam * e33 * e32 + mu * e32 * e33;
      A[2][2] = lam2mu * e33 * e33 + mu * (e31 * e31 + e32 * e32);
      invert3x3<R>(A);
      B[0][0] = -lam2mu * e31 * e11 - mu * (e32 * e12 + e33 * e13);
      B[0][1] = -lam * e31 * e12 - mu * e32 * e11;
      B[0][2] = -lam * e31 * e13 - mu * e33 * e11;
      B[1][0] = -lam * e32 * e11 - mu * e31 * e12;
      B[1][1] = -lam2mu * e32 * e12 - mu * (e31 * e11 + e33 * e13);
      B[1][2] = -lam * e32 * e13 - mu * e33 * e12;
      B[2][0] = -lam * e33 * e11 - mu * e31 * e13;
      B[2][1] = -lam * e33 * e12 - mu * e32 * e13;
      B[2][2] = -lam2mu * e33 * e13 - mu * (e31 * e11 + e32 * e12);
      C[0][0] = -lam2mu * e31 * e21 - mu * (e32 * e22 + e33 * e23);
      C[0][1] = -lam * e31 * e22 - mu * e32 * e21;
      C[0][2] = -lam * e31 * e23 - mu * e33 * e21;
      C[1][0] = -lam * e32 * e21 - mu * e31 * e22;
      C[1][1] = -lam2mu * e32 * e22 - mu * (e31 * e21 + e33 * e23);
      C[1][2] = -lam * e32 * e23 - mu * e33 * e22;
      C[2][0] = -lam * e33 * e21 - mu * e31 * e23;
      C[2][1] = -lam * e33 * e22 - mu * e32 * e23;
      C[2][2] = -lam2mu * e33 * e23 - mu * (e31 * e21 + e32 * e22);
      for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) {
          AB[r][cc] = 0; AC[r][cc] = 0;
          for (int q = 0; q < 3; q++) { AB[r][cc] += A[r][q] * B[q][cc]; AC[r][cc] += A[r][q] * C[q][cc]; }
        }
      size_t ij = (j * g.siz_iy + i) * 9;
      for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) { mvx[ij + r * 3 + cc] = (double)AB[r][cc]; mvy[ij + r * 3 + cc] = (double)AC[r][cc]; }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int orc_version(void) { return 1; }

/* wall seconds of the time loop (setup excluded) of the last orc_run */
double orc_last_loop_seconds(void) { return g_loop_seconds; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* face box sizes (points) for the problem, x1,x2,y1,y2,z1,z2 */
void orc_face_sizes(const orc_problem *p, long long out[6]) {
  Grid g(p->ni, p->nj, p->nk);
  Face f[6];
  face_boxes(g, p, f);
  for (int i = 0; i < 6; i++) out[i] = (long long)f[i].size;
}

/* run nt_total steps of the reference loop; precision 32 or 64 */
int orc_run(const orc_problem *p, int precision, double dt, int nt_total, const double *init_wave,
            const double *init_aux, double *seismo, double *final_wave, double *final_aux) {
  if (precision == 32) return run<float>(p, dt, nt_total, init_wave, init_aux, seismo, final_wave, final_aux);
  if (precision == 64) return run<double>(p, dt, nt_total, init_wave, init_aux, seismo, final_wave, final_aux);
  return -1;
}

/* operator table of fd_set_macdrp (fd_t.cu:76-148): layer 0..3, dir 0 backward / 1 forward */
int orc_fd_op(int layer, int dir, int *idx, double *coef) {
  if (layer < 0 || layer > 3 || dir < 0 || dir > 1) return -1;
  for (int n = 0; n < 5; n++) { idx[n] = kIdx[layer][dir][n]; coef[n] = kCoef[layer][dir][n]; }
  return kLen[layer];
}

/* centered 7-point metric operator (fd_t.cu:164-170) */
void orc_fdc(double *coef) { for (int n = 0; n < 7; n++) coef[n] = kFdc[n]; }

/* operator directions (0 backward, 1 forward) of x,y,z at (pair, stage) (fd_t.cu:184-194, 233) */
void orc_pair_dirs(int ipair, int istage, int *dirs) {
  for (int d = 0; d < 3; d++) dirs[d] = (kFlags[ipair][d] + istage) % 2;
}

/* RK4 a, b and stage times (fd_t.cu:36-49) */
void orc_rk(double *a3, double *b4, double *t4) {
  for (int i = 0; i < 3; i++) a3[i] = kRkA[i];
  for (int i = 0; i < 4; i++) { b4[i] = kRkB[i]; t4[i] = kRkT[i]; }
}

/* one RHS evaluation (sv_curv_col_el_iso_onestage) */
int orc_onestage(const orc_problem *p, int precision, int ipair, int istage, int it, const double *w_in,
                 const double *aux_in, double *rhs_out, double *aux_rhs_out) {
  if (precision == 32) return stage_once<float>(p, ipair, istage, it, w_in, aux_in, rhs_out, aux_rhs_out);
  if (precision == 64) return stage_once<double>(p, ipair, istage, it, w_in, aux_in, rhs_out, aux_rhs_out);
  return -1;
}

int orc_metric_cal(int precision, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                   double *out) {
  if (precision == 32) metric_cal<float>(ni, nj, nk, x, y, z, out);
  else if (precision == 64) metric_cal<double>(ni, nj, nk, x, y, z, out);
  else return -1;
  return 0;
}

int orc_free_matrices(int precision, int ni, int nj, int nk, const double *metric, const double *lam,
                      const double *mu, double *mvx, double *mvy) {
  if (precision == 32) free_matrices<float>(ni, nj, nk, metric, lam, mu, mvx, mvy);
  else if (precision == 64) free_matrices<double>(ni, nj, nk, metric, lam, mu, mvx, mvy);
  else return -1;
  return 0;
}

/* gd_curv_gen_cart (gd_t.cu:599-630), single rank: x0 = x0_glob - 3*dx */
void orc_gen_cart(int ni, int nj, int nk, double dx, double x0g, double dy, double y0g, double dz, double z0g,
                  double *x, double *y, double *z) {
  Grid g(ni, nj, nk);
  float fx = (float)dx, fy = (float)dy, fz = (float)dz;
  float x0 = (float)x0g + (float)(0 - NG) * fx;
  float y0 = (float)y0g + (float)(0 - NG) * fy;
  float z0 = (float)z0g + (float)(0 - NG) * fz;
  size_t iptr = 0;
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        x[iptr] = (double)(x0 + (float)i * fx);
        y[iptr] = (double)(y0 + (float)j * fy);
        z[iptr] = (double)(z0 + (float)k * fz);
        iptr++;
      }
}

/* bdry_cal_abl_len_dh (bdry_t.cu:334-451) for one face box given in local indices */
void orc_abl_len_dh(int ni, int nj, int nk, const double *x, const double *y, const double *z, int i1, int i2,
                    int j1, int j2, int k1, int k2, int idim, double *L0, double *dh) {
  Grid g(ni, nj, nk);
  double L = 0.0;
  int num = 0;
  auto P = [&](int i, int j, int k) { return (size_t)i + (size_t)j * g.siz_iy + (size_t)k * g.siz_iz; };
  auto dist = [&](size_t a, size_t b) {
    double x0 = (float)x[a], y0 = (float)y[a], z0 = (float)z[a];
    double x1 = (float)x[b], y1 = (float)y[b], z1 = (float)z[b];
    return std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0) + (z1 - z0) * (z1 - z0));
  };
  if (idim == 0) {
    for (int k = k1; k <= k2; k++)
      for (int j = j1; j <= j2; j++)
        for (int i = i1 + 1; i <= i2; i++) { L += dist(P(i - 1, j, k), P(i, j, k)); num++; }
  } else if (idim == 1) {
    for (int k = k1; k <= k2; k++)
      for (int i = i1; i <= i2; i++)
        for (int j = j1 + 1; j <= j2; j++) { L += dist(P(i, j - 1, k), P(i, j, k)); num++; }
  } else {
    for (int j = j1; j <= j2; j++)
      for (int i = i1; i <= i2; i++)
        for (int k = k1 + 1; k <= k2; k++) { L += dist(P(i, j, k - 1), P(i, j, k)); num++; }
  }
  float fdh = (float)(L / num);
  int span = (idim == 0) ? (i2 - i1) : (idim == 1) ? (j2 - j1) : (k2 - k1);
  float fL = fdh * (float)span;
  *dh = fdh;
  *L0 = fL;
}

/* bdry_pml_set profile loop + bdry_pml_cal_{R,dmax,d,a,b} (bdry_t.cu:193-241, 297-330), FP32 host
   semantics of the reference: C++ overloads make log/pow of floats single precision. */
void orc_pml_coefs(int nlay, int iside, double L0_, double dh_, double vel_, double amax_, double bmax_,
                   double *A, double *B, double *D) {
  const float L0 = (float)L0_, dh = (float)dh_, vel = (float)vel_, amax = (float)amax_, bmax = (float)bmax_;
  const int npoints = nlay + 1;
  const float num_lay = (float)(npoints - 1);
  const float Rpp = (float)(std::pow(10.0, -((std::log10((double)num_lay) - 1.0) / std::log10(2.0) + 4.0)));
  const float dmax = (float)(-vel / (2.0 * L0) * (double)logf(Rpp) * (2.0f + 1.0));
  for (int n = 0; n < npoints; n++) {
    float L = (float)n * dh;
    int i = (iside == 0) ? npoints - 1 - n : n;
    float r = L / L0;
    float d = (L < 0) ? 0.0f : (float)(dmax * powf(r, 2.0f));
    float a = (L < 0) ? 0.0f : (float)(amax * (1.0 - (double)powf(r, 1.0f)));
    float b = (L < 0) ? 1.0f : (float)(1.0 + (bmax - 1.0) * (double)powf(r, 2.0f));
    d /= b;
    a += d;
    b = (float)(1.0 / b);
    A[i] = a; B[i] = b; D[i] = d;
  }
}

/* fun_ricker (src_t.cu:834-841) */
double orc_ricker(double t_, double fc_, double t0_) {
  float t = (float)t_, fc = (float)fc_, t0 = (float)t0_;
  float u = (float)((double)(t - t0) * 2.0 * M_PI * fc);
  float v = (float)((1 - (double)(u * u) / 2.0) * std::exp((double)(-u * u) / 4.0));
  return v;
}
double orc_ricker_deriv(double t_, double fc_, double t0_) {
  float t = (float)t_, fc = (float)fc_, t0 = (float)t0_;
  float u = (float)((double)(t - t0) * 2.0 * M_PI * fc);
  float v = (float)((double)(u * (-3 + 1.0 / 2.0 * u * u)) * (double)expf(-u * u / 4) * M_PI * fc);
  return v;
}
/* fun_gauss (src_t.cu:853-859): float arithmetic */
double orc_gauss(double t_, double a_, double t0_) {
  float t = (float)t_, a = (float)a_, t0 = (float)t0_;
  float f = expf(-(t - t0) * (t - t0) / (a * a)) / (sqrtf((float)M_PI) * a);
  return f;
}

/* src_cal_norm_delt3d (src_t.cu:796-827) */
int orc_norm_delt3d(double x0, double y0, double z0, double rx0, double ry0, double rz0, int LenDelt, double *delt) {
  int n = 2 * LenDelt + 1;
  std::vector<float> d((size_t)n * n * n);
  float SUM = 0.0f;
  int iptr = 0;
  for (int k = -LenDelt; k <= LenDelt; k++)
    for (int j = -LenDelt; j <= LenDelt; j++)
      for (int i = -LenDelt; i <= LenDelt; i++) {
        float D1 = (float)orc_gauss((float)i - (float)x0, (float)rx0, 0.0);
        float D2 = (float)orc_gauss((float)j - (float)y0, (float)ry0, 0.0);
        float D3 = (float)orc_gauss((float)k - (float)z0, (float)rz0, 0.0);
        d[iptr] = D1 * D2 * D3;
        SUM += d[iptr];
        iptr++;
      }
  if (SUM < 1e-20f) return -1;
  for (size_t q = 0; q < d.size(); q++) delt[q] = (double)(d[q] / SUM);
  return 0;
}

/* angle2moment (src_t.cu:910-950) */
void orc_angle2moment(double strike_, double dip_, double rake_, double *m6) {
  float strike = (float)strike_, dip = (float)dip_, rake = (float)rake_;
  float dip_pi = (float)(dip / 180.0 * M_PI);
  float strike_pi = (float)(strike / 180.0 * M_PI);
  float rake_pi = (float)(rake / 180.0 * M_PI);
  // sin/cos of float arguments are float overloads; products with 2.0 promote to double
  float M11 = (float)(-((double)(sinf(dip_pi) * cosf(rake_pi)) * std::sin(2.0 * strike_pi) +
                        std::sin(2.0 * dip_pi) * sinf(rake_pi) * sinf(strike_pi) * sinf(strike_pi)));
  float M22 = (float)((double)(sinf(dip_pi) * cosf(rake_pi)) * std::sin(2.0 * strike_pi) -
                      std::sin(2.0 * dip_pi) * sinf(rake_pi) * cosf(strike_pi) * cosf(strike_pi));
  float M33 = -(M11 + M22);
  float M12 = (float)((double)(sinf(dip_pi) * cosf(rake_pi)) * std::cos(2.0 * strike_pi) +
                      0.5 * std::sin(2.0 * dip_pi) * sinf(rake_pi) * std::sin(2.0 * strike_pi));
  float M13 = (float)(-((double)(cosf(dip_pi) * cosf(rake_pi) * cosf(strike_pi)) +
                        std::cos(2.0 * dip_pi) * sinf(rake_pi) * sinf(strike_pi)));
  float M23 = (float)(-((double)(cosf(dip_pi) * cosf(rake_pi) * sinf(strike_pi)) -
                        std::cos(2.0 * dip_pi) * sinf(rake_pi) * cosf(strike_pi)));
  m6[0] = M22;   // Mxx
  m6[1] = M11;   // Myy
  m6[2] = M33;   // Mzz
  m6[3] = -M13;  // Myz
  m6[4] = -M23;  // Mxz
  m6[5] = M12;   // Mxy
}

/*
 * Source tables for one point source given by grid index (src_t.cu:143-790, index branch,
 * wavelet-by-name branch).  Coordinates are 0-based global physical indices; depth_flag=1
 * means the third coordinate is counted down from the surface (k = nk-1 - z).
 * wavelet: 0 ricker, 1 gaussian, 2 ricker_deriv, 3 gaussian_deriv.
 * Outputs: it_begin, it_end, ext_num; ext_indx[343], ext_coef[343]; stf[max_nt*4].
 */
int orc_src_point(int ni, int nj, int nk, double sx_, double sy_, double sz_, int depth_flag, double t0_, double dt_,
                  double wavelet_tstart_, int wavelet, double c0, double c1, double stf_length_, int *it_begin,
                  int *it_end, int *ext_num, int *ext_indx, double *ext_coef, int *max_nt_out, double *stf) {
  Grid g(ni, nj, nk);
  float sx = (float)sx_, sy = (float)sy_, sz = (float)sz_;
  const float t0 = (float)t0_, dt = (float)dt_, wavelet_tstart = (float)wavelet_tstart_;
  const float stf_length = (float)stf_length_;
  const int gnk2 = nk - 1;
  if (depth_flag == 1) sz = (float)gnk2 - sz;
  int si_glob = (int)((double)sx + 0.5), sj_glob = (int)((double)sy + 0.5), sk_glob = (int)((double)sz + 0.5);
  float sx_inc = sx - (float)si_glob, sy_inc = sy - (float)sj_glob, sz_inc = sz - (float)sk_glob;
  const int half = 3, len_ext = 7;
  int si = si_glob + NG, sj = sj_glob + NG, sk = sk_glob + NG;
  double delt[343];
  float wid = (float)half / 2.0f;
  if (orc_norm_delt3d(sx_inc, sy_inc, sz_inc, wid, wid, wid, half, delt) != 0) return -1;
  int ne = 0;
  for (int k = sk - half; k <= sk + half; k++)
    for (int j = sj - half; j <= sj + half; j++)
      for (int i = si - half; i <= si + half; i++) {
        if (i >= g.ni1 && i <= g.ni2 && j >= g.nj1 && j <= g.nj2 && k >= g.nk1 && k <= g.nk2) {
          int ig = i + j * (int)g.siz_iy + k * (int)g.siz_iz;
          int ic = (i - (si - half)) + len_ext * (j - (sj - half)) + len_ext * len_ext * (k - (sk - half));
          ext_indx[ne] = ig;
          ext_coef[ne] = delt[ic];
          ne++;
        }
      }
  *ext_num = ne;
  int max_nt = (int)((double)(stf_length / dt) + 0.5);
  *max_nt_out = max_nt;
  int ib = (int)((wavelet_tstart - t0) / dt);
  *it_begin = ib;
  *it_end = ib + max_nt - 1;
  if (stf) {
    float t_shift = wavelet_tstart - ((float)ib * dt + t0);
    for (int it = ib; it <= ib + max_nt - 1; it++) {
      int itr = it - ib;
      for (int s = 0; s < 4; s++) {
        float t = (float)itr * dt + (float)kRkT[s] * dt - t_shift;
        double v = 0;
        if (wavelet == 0) v = orc_ricker(t, c0, c1);
        else if (wavelet == 1) v = orc_gauss(t, c0, c1);
        else if (wavelet == 2) v = orc_ricker_deriv(t, c0, c1);
        else if (wavelet == 3) {
          float tt = t, a = (float)c0, tc = (float)c1;
          v = (float)(expf(-(tt - tc) * (tt - tc) / (a * a)) / (sqrtf((float)M_PI) * a) * (-2 * (tt - tc) / (a * a)));
        }
        stf[(size_t)itr * 4 + s] = v;
      }
    }
  }
  return 0;
}

/* receiver by grid index (io_funcs.cu:162-251): x,y 1-based; z by depth (flag 1) or 1-based */
int orc_recv_index(int ni, int nj, int nk, double rx_, double ry_, double rz_, int depth_flag, long long indx[8],
                   double frac[3]) {
  Grid g(ni, nj, nk);
  float rx = (float)rx_ - 1, ry = (float)ry_ - 1, rz = (float)rz_;
  if (depth_flag == 1) rz = (float)(nk - 1) - rz;
  else rz -= 1;
  int ix = (int)rx, iy = (int)ry, iz = (int)rz;
  float dx = rx - (float)ix, dy = ry - (float)iy, dz = rz - (float)iz;
  if (ix < 0 || ix > ni - 1 || iy < 0 || iy > nj - 1 || iz < 0 || iz > nk - 1) return -1;
  long long i = ix + NG, j = iy + NG, k = iz + NG;
  long long sy = (long long)g.siz_iy, sz = (long long)g.siz_iz;
  indx[0] = i + j * sy + k * sz;
  indx[1] = i + 1 + j * sy + k * sz;
  indx[2] = i + (j + 1) * sy + k * sz;
  indx[3] = i + 1 + (j + 1) * sy + k * sz;
  indx[4] = i + j * sy + (k + 1) * sz;
  indx[5] = i + 1 + j * sy + (k + 1) * sz;
  indx[6] = i + (j + 1) * sy + (k + 1) * sz;
  indx[7] = i + 1 + (j + 1) * sy + (k + 1) * sz;
  frac[0] = dx; frac[1] = dy; frac[2] = dz;
  return 0;
}

/* blk_dt_esti_curv (blk_t.cu:673-763), elastic iso, rho not yet inverted */
double orc_dt_esti(int ni, int nj, int nk, const double *x, const double *y, const double *z, const double *lam,
                   const double *mu, const double *rho, double CFL) {
  Grid g(ni, nj, nk);
  float dtmax = 1.0e10f;
  auto dist = [](const float *x0, const float *x1, const float *x2, const float *x3) {
    float a[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
    float b[3] = {x3[0] - x1[0], x3[1] - x1[1], x3[2] - x1[2]};
    float p[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    auto dot = [](const float *u, const float *v) { float r = 0; for (int q = 0; q < 3; q++) r += u[q] * v[q]; return r; };
    float d = dot(p, x1);
    float L = (float)std::fabs(dot(p, x0) - d);
    return L / sqrtf(dot(p, p));
  };
  for (int k = g.nk1; k < g.nk2; k++)
    for (int j = g.nj1; j < g.nj2; j++)
      for (int i = g.ni1; i < g.ni2; i++) {
        size_t ip = i + j * g.siz_iy + k * g.siz_iz;
        float Vp = (float)std::sqrt(((double)(float)lam[ip] + 2.0 * (double)(float)mu[ip]) / (double)(float)rho[ip]);
        float dtLe = 1.0e20f;
        float p0[3] = {(float)x[ip], (float)y[ip], (float)z[ip]};
        for (int kk = -1; kk <= 1; kk++)
          for (int jj = -1; jj <= 1; jj++)
            for (int ii = -1; ii <= 1; ii++) {
              if (ii != 0 && jj != 0 && kk != 0) {
                size_t a = ip - ii, b = ip - jj * (long)g.siz_iy, c = ip - kk * (long)g.siz_iz;
                float p1[3] = {(float)x[a], (float)y[a], (float)z[a]};
                float p2[3] = {(float)x[b], (float)y[b], (float)z[b]};
                float p3[3] = {(float)x[c], (float)y[c], (float)z[c]};
                float L = dist(p0, p1, p2, p3);
                if (dtLe > L) dtLe = L;
              }
            }
        float dtp = (float)CFL / Vp * dtLe;
        if (dtp < dtmax) dtmax = dtp;
      }
  return dtmax;
}

/* md_rho_to_slow (md_t.cu:622-637) in FP32 */
void orc_rho_to_slow(size_t n, const double *rho, double *slw) {
  for (size_t i = 0; i < n; i++) {
    float r = (float)rho[i];
    slw[i] = ((double)r > 1e-10) ? (double)(float)(1.0 / r) : 0.0;
  }
}

/*
 * Peak ground motion maps (PG_calcu_gpu, src/forward/wav_t.cu:281-360, called
 * after every step at drv_rk_curv_col.cu:519-527) from the surface velocity
 * history vtop[nt+1][3][npts] (W^0 .. W^nt on the top plane): per step n the
 * acceleration |V^{n+1} - V^n| / dt, the displacement accumulated as
 * D += 0.5 (V^{n+1} + V^n) dt (the 0.5 literal makes it a double sum, stored
 * to float), and the running maxima of |V|, |V_h|, |V_x..z|, the same for A
 * and D.  pg[15][npts]: PGV PGVh PGVx PGVy PGVz PGA PGAh PGAx PGAy PGAz PGD
 * PGDh PGDx PGDy PGDz, float arithmetic as the reference.
 */
void orc_peak_motion(int npts, int nt, const double *vtop, double dt_, double *pg) {
  const float dt = (float)dt_;
  std::vector<float> P((size_t)15 * npts, 0.0f), D((size_t)3 * npts, 0.0f);
  for (int n = 0; n < nt; n++) {
    const double *v0 = vtop + (size_t)n * 3 * npts, *v1 = vtop + (size_t)(n + 1) * 3 * npts;
    for (int q = 0; q < npts; q++) {
      const float x1 = (float)v1[q], y1 = (float)v1[npts + q], z1 = (float)v1[2 * npts + q];
      const float x0 = (float)v0[q], y0 = (float)v0[npts + q], z0 = (float)v0[2 * npts + q];
      const float Ax = std::fabs((x1 - x0) / dt), Ay = std::fabs((y1 - y0) / dt), Az = std::fabs((z1 - z0) / dt);
      float &Dx = D[q], &Dy = D[npts + q], &Dz = D[2 * (size_t)npts + q];
      Dx = (float)((double)Dx + 0.5 * (double)(x1 + x0) * (double)dt);
      Dy = (float)((double)Dy + 0.5 * (double)(y1 + y0) * (double)dt);
      Dz = (float)((double)Dz + 0.5 * (double)(z1 + z0) * (double)dt);
      const float V = std::sqrt(x1 * x1 + y1 * y1 + z1 * z1), Vh = std::sqrt(x1 * x1 + y1 * y1);
      const float A = std::sqrt(Ax * Ax + Ay * Ay + Az * Az), Ah = std::sqrt(Ax * Ax + Ay * Ay);
      const float Dm = std::sqrt(Dx * Dx + Dy * Dy + Dz * Dz), Dh = std::sqrt(Dx * Dx + Dy * Dy);
      const float val[15] = {V, Vh, std::fabs(x1), std::fabs(y1), std::fabs(z1), A, Ah, Ax, Ay, Az,
                             Dm, Dh, std::fabs(Dx), std::fabs(Dy), std::fabs(Dz)};
      for (int m = 0; m < 15; m++)
        if (P[(size_t)m * npts + q] < val[m]) P[(size_t)m * npts + q] = val[m];
    }
  }
  for (size_t q = 0; q < P.size(); q++) pg[q] = P[q];
}

}  // extern "C"
