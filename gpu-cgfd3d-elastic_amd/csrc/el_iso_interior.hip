// el_iso_interior.hip -- z-marching interior kernel of the el_iso RK4 stage.
//
// A workgroup owns a TX x TY column tile and sweeps a chunk of z planes.
// Each thread keeps, per wavefield component, a 5-plane register queue
// holding exactly the planes of the one-sided z operator of this stage
// (backward {-3..1} or forward {-1..3}), so every wavefield value is fetched
// from HBM once per sweep.  The centre plane of the queue is staged in LDS
// together with an asymmetric x/y halo (3 points on the operator's long side,
// 1 on the short side) for the x and y derivatives.  The right-hand side is
// the inner form of sv_curv_col_el_iso_rhs_inner_gpu
// (src/forward/sv_curv_col_el_iso_gpu.cu:291-452); the RK update
// (src/forward/wav_t.cu:362-378, drv_rk_curv_col.cu:339-491) is fused into
// the epilogue so the RHS never goes to HBM.
//
// Only points that are interior in the reference's sense are handed to this
// kernel: outside every PML face box and at least 4 rows below a free
// surface.  The boundary shell goes to the per-point kernel.
#include "el_iso_device.h"
#include "el_iso_launch.h"

namespace cgfd {

template <int DX, int DY, int DZ, typename TS, typename TC, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void k_interior(StageArgs<TS, TC> a, InteriorBox b, int ntx, int nty,
                                                     int nzc) {
  constexpr int NT = TX * TY;
  constexpr int HXL = (DX == 0) ? 3 : 1, HXR = 4 - HXL;
  constexpr int HYL = (DY == 0) ? 3 : 1, HYR = 4 - HYL;
  constexpr int LW = TX + 4, LH = TY + 4;
  constexpr int zlo = (DZ == 0) ? -3 : -1;
  constexpr int zc0 = -zlo;  // queue slot of the centre plane
  __shared__ TC lds[kNC][LH][LW];

  // XCD-aware remap: consecutive hardware workgroup ids land on different
  // XCDs; give each XCD a contiguous range of tiles so halos hit its L2.
  const int nwg = gridDim.x;
  const int hw = blockIdx.x;
  const int q8 = nwg / 8, r8 = nwg % 8;
  const int xcd = hw % 8, m = hw / 8;
  const int wg = xcd * q8 + (xcd < r8 ? xcd : r8) + m;
  const int tx_t = wg % ntx;
  const int ty_t = (wg / ntx) % nty;
  const int zc = (wg / (ntx * nty)) % nzc;
  const int shot = wg / (ntx * nty * nzc);
  if (a.shot_active && !a.shot_active[shot]) return;

  const Layout &L = a.L;
  const long cs = L.cs, px = L.px, pz = L.pz;
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = tid / TX;
  const int i0 = b.i0 + tx_t * TX, j0 = b.j0 + ty_t * TY;
  const int i = i0 + lx, j = j0 + ly;
  const int kbeg = b.k0 + zc * b.zchunk;
  int kend = kbeg + b.zchunk - 1;
  if (kend > b.k1) kend = b.k1;
  const bool owner = (i <= b.i1) && (j <= b.j1);
  const bool inmem = (i <= L.ni + 2) && (j <= L.nj + 2);
  const long so = (long)shot * a.shot_stride;
  const TS *win = a.w_in + so;
  const long col = (long)(j + kNG) * px + (long)(i + L.xa);

  auto sc = [&](int c) -> TC { return St<TS>::scaled ? (TC)a.sc[c] : TC(1); };
  auto gld = [&](int c, long off) -> TC { return (TC)St<TS>::ld(win + c * cs + off) * sc(c); };

  // register queues: planes k+zlo .. k+zlo+4
  TC q[kNC][5];
#pragma unroll
  for (int c = 0; c < kNC; c++)
#pragma unroll
    for (int s = 0; s < 4; s++)
      q[c][s + 1] = inmem ? gld(c, (long)(kbeg + zlo + s + kNG) * pz + col) : TC(0);

  for (int k = kbeg; k <= kend; k++) {
    const long pk = (long)(k + kNG) * pz;
#pragma unroll
    for (int c = 0; c < kNC; c++) {
#pragma unroll
      for (int s = 0; s < 4; s++) q[c][s] = q[c][s + 1];
      q[c][4] = inmem ? gld(c, (long)(k + zlo + 4 + kNG) * pz + col) : TC(0);
      lds[c][ly + HYL][lx + HXL] = q[c][zc0];
    }
    // halo of the centre plane: x columns for the tile rows, y rows for the tile columns
    constexpr int NXH = 4 * TY, NYH = 4 * TX, NH = NXH + NYH;
    for (int e = tid; e < kNC * NH; e += NT) {
      const int c = e / NH, r = e % NH;
      int hi, hj, li, lj;
      if (r < NXH) {
        const int h = r % 4, row = r / 4;
        li = (h < HXL) ? h : TX + h;
        hi = i0 - HXL + li;
        lj = row + HYL;
        hj = j0 + row;
      } else {
        const int rr = r - NXH;
        const int h = rr / TX, cc = rr % TX;
        lj = (h < HYL) ? h : TY + h;
        hj = j0 - HYL + lj;
        li = cc + HXL;
        hi = i0 + cc;
      }
      TC v = TC(0);
      if (hi <= L.ni + 2 && hj <= L.nj + 2) v = gld(c, pk + (long)(hj + kNG) * px + (hi + L.xa));
      lds[c][lj][li] = v;
    }
    __syncthreads();

    TC Dx[kNC], Dy[kNC], Dz[kNC];
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      constexpr int xl = HXL + ((DX == 0) ? -3 : -1), yl = HYL + ((DY == 0) ? -3 : -1);
      TC dx = TC(0), dy = TC(0), dz = TC(0);
#pragma unroll
      for (int n = 0; n < 5; n++) {
        dx += Drp<TC, DX>::c(n) * lds[c][ly + HYL][lx + xl + n];
        dy += Drp<TC, DY>::c(n) * lds[c][ly + yl + n][lx + HXL];
        dz += Drp<TC, DZ>::c(n) * q[c][n];
      }
      Dx[c] = dx; Dy[c] = dy; Dz[c] = dz;
    }

    if (owner) {
      const long o = pk + col;
      const TC *met = a.met;
      const TC xix = met[XIX * cs + o], xiy = met[XIY * cs + o], xiz = met[XIZ * cs + o];
      const TC etx = met[ETX * cs + o], ety = met[ETY * cs + o], etz = met[ETZ * cs + o];
      const TC ztx = met[ZTX * cs + o], zty = met[ZTY * cs + o], ztz = met[ZTZ * cs + o];
      const TC lam = a.med[o], mu = a.med[cs + o], slw = a.med[2 * cs + o];
      const TC lam2mu = lam + TC(2) * mu;
      TC h[kNC];
      h[VX] = slw * (xix * Dx[TXX] + xiy * Dx[TXY] + xiz * Dx[TXZ] + etx * Dy[TXX] + ety * Dy[TXY] +
                     etz * Dy[TXZ] + ztx * Dz[TXX] + zty * Dz[TXY] + ztz * Dz[TXZ]);
      h[VY] = slw * (xix * Dx[TXY] + xiy * Dx[TYY] + xiz * Dx[TYZ] + etx * Dy[TXY] + ety * Dy[TYY] +
                     etz * Dy[TYZ] + ztx * Dz[TXY] + zty * Dz[TYY] + ztz * Dz[TYZ]);
      h[VZ] = slw * (xix * Dx[TXZ] + xiy * Dx[TYZ] + xiz * Dx[TZZ] + etx * Dy[TXZ] + ety * Dy[TYZ] +
                     etz * Dy[TZZ] + ztx * Dz[TXZ] + zty * Dz[TYZ] + ztz * Dz[TZZ]);
      h[TXX] = lam2mu * (xix * Dx[VX] + etx * Dy[VX] + ztx * Dz[VX]) +
               lam * (xiy * Dx[VY] + ety * Dy[VY] + zty * Dz[VY] + xiz * Dx[VZ] + etz * Dy[VZ] + ztz * Dz[VZ]);
      h[TYY] = lam2mu * (xiy * Dx[VY] + ety * Dy[VY] + zty * Dz[VY]) +
               lam * (xix * Dx[VX] + etx * Dy[VX] + ztx * Dz[VX] + xiz * Dx[VZ] + etz * Dy[VZ] + ztz * Dz[VZ]);
      h[TZZ] = lam2mu * (xiz * Dx[VZ] + etz * Dy[VZ] + ztz * Dz[VZ]) +
               lam * (xix * Dx[VX] + etx * Dy[VX] + ztx * Dz[VX] + xiy * Dx[VY] + ety * Dy[VY] + zty * Dz[VY]);
      h[TXY] = mu * (xiy * Dx[VX] + xix * Dx[VY] + ety * Dy[VX] + etx * Dy[VY] + zty * Dz[VX] + ztx * Dz[VY]);
      h[TXZ] = mu * (xiz * Dx[VX] + xix * Dx[VZ] + etz * Dy[VX] + etx * Dy[VZ] + ztz * Dz[VX] + ztx * Dz[VZ]);
      h[TYZ] = mu * (xiz * Dx[VY] + xiy * Dx[VZ] + etz * Dy[VY] + ety * Dy[VZ] + ztz * Dz[VY] + zty * Dz[VZ]);

      TS *wout = a.w_out + so, *wend = a.w_end + so;
      const TS *wpre = a.w_pre + so;
#pragma unroll
      for (int c = 0; c < kNC; c++) {
        const long qo = c * cs + o;
        const TC is = St<TS>::scaled ? (TC)a.isc[c] : TC(1);
        if (a.mode == MODE_S12) {
          const TC pre = (TC)St<TS>::ld(wpre + qo) * sc(c);
          const TC en = (TC)St<TS>::ld(wend + qo) * sc(c);
          St<TS>::st(wout + qo, (pre + a.ca * h[c]) * is);
          St<TS>::st(wend + qo, (en + a.cb * h[c]) * is);
        } else if (a.mode == MODE_S0) {
          const TC v = q[c][zc0];
          St<TS>::st(wout + qo, (v + a.ca * h[c]) * is);
          St<TS>::st(wend + qo, (v + a.cb * h[c]) * is);
        } else if (a.mode == MODE_S3) {
          const TC en = (TC)St<TS>::ld(wend + qo) * sc(c);
          St<TS>::st(wend + qo, (en + a.cb * h[c]) * is);
        } else {
          St<TS>::st(wout + qo, h[c] * is);
        }
      }
    }
    __syncthreads();
  }
}

template <typename TS, typename TC, int TX, int TY>
static hipError_t launch_interior_t(const StageArgs<TS, TC> &a, const InteriorBox &b, int dx, int dy, int dz,
                                    int nshots, hipStream_t st) {
  const int ntx = (b.i1 - b.i0 + TX) / TX;
  const int nty = (b.j1 - b.j0 + TY) / TY;
  const int nzc = (b.k1 - b.k0 + b.zchunk) / b.zchunk;
  const long nwg = (long)ntx * nty * nzc * nshots;
  dim3 grid((unsigned)nwg), block(TX * TY);
  const int sel = dx * 4 + dy * 2 + dz;
#define CGFD_L(X, Y, Z) \
  hipLaunchKernelGGL((k_interior<X, Y, Z, TS, TC, TX, TY>), grid, block, 0, st, a, b, ntx, nty, nzc)
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
  return hipGetLastError();
}

template <typename TS, typename TC>
hipError_t launch_interior(const StageArgs<TS, TC> &a, const InteriorBox &b, int dx, int dy, int dz, int nshots,
                           hipStream_t st) {
  if (b.i1 < b.i0 || b.j1 < b.j0 || b.k1 < b.k0) return hipSuccess;
  return launch_interior_t<TS, TC, 64, 4>(a, b, dx, dy, dz, nshots, st);
}

template hipError_t launch_interior<float, float>(const StageArgs<float, float> &, const InteriorBox &, int, int, int,
                                                  int, hipStream_t);
template hipError_t launch_interior<double, double>(const StageArgs<double, double> &, const InteriorBox &, int, int,
                                                    int, int, hipStream_t);
template hipError_t launch_interior<__half, float>(const StageArgs<__half, float> &, const InteriorBox &, int, int,
                                                   int, int, hipStream_t);
template hipError_t launch_interior<__hip_bfloat16, float>(const StageArgs<__hip_bfloat16, float> &,
                                                           const InteriorBox &, int, int, int, int, hipStream_t);

}  // namespace cgfd
