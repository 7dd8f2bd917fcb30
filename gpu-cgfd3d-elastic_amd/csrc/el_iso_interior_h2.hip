// el_iso_interior_h2.hip -- the interior z-marching sweep for 16-bit storage
// (FP16 / BF16, FP32 arithmetic) with two adjacent columns per lane.
//
// Same scheme and arithmetic per point as k_march in el_iso_interior.hip
// (inner right-hand side of sv_curv_col_el_iso_rhs_inner_gpu,
// src/forward/sv_curv_col_el_iso_gpu.cu:291-452, fused RK update).  With
// 2-byte storage the one-column sweep issues as many memory instructions as
// the FP32 one and is bound by them, not by HBM; here every wavefield access
// moves a pair of columns in one 32-bit load or store (tiles start at an even
// column so each pair is 4-byte aligned), the coefficients move as float2,
// and the z register queue holds the pairs packed.
#include "el_iso_device.h"
#include "el_iso_launch.h"

namespace cgfd {

// a pair of adjacent 16-bit values <-> two floats
template <typename TS> struct Pair;
template <> struct Pair<__half> {
  __device__ __forceinline__ static void unpack(unsigned u, float &lo, float &hi) {
    lo = __half2float(__ushort_as_half((unsigned short)(u & 0xffffu)));
    hi = __half2float(__ushort_as_half((unsigned short)(u >> 16)));
  }
  __device__ __forceinline__ static unsigned pack(float lo, float hi) {
    return (unsigned)__half_as_ushort(__float2half(lo)) | ((unsigned)__half_as_ushort(__float2half(hi)) << 16);
  }
  __device__ __forceinline__ static float one(unsigned short u) { return __half2float(__ushort_as_half(u)); }
};
template <> struct Pair<__hip_bfloat16> {
  __device__ __forceinline__ static void unpack(unsigned u, float &lo, float &hi) {
    lo = __uint_as_float(u << 16);
    hi = __uint_as_float(u & 0xffff0000u);
  }
  __device__ __forceinline__ static unsigned bits(float v) {
    const __hip_bfloat16 b = __float2bfloat16(v);
    unsigned short s;
    __builtin_memcpy(&s, &b, 2);
    return s;
  }
  __device__ __forceinline__ static unsigned pack(float lo, float hi) { return bits(lo) | (bits(hi) << 16); }
  __device__ __forceinline__ static float one(unsigned short u) { return __uint_as_float((unsigned)u << 16); }
};

template <int DX, int DY, int DZ, typename TS, int TXP, int TY, int MINB, bool CART>
__global__ __launch_bounds__(TXP *TY, MINB) void k_march_h2(StageArgs<TS, float> a, MarchSet b, MarchGrid g) {
  using TC = float;
  constexpr int TX = 2 * TXP;                     // points per tile row
  constexpr int NT = TXP * TY;
  constexpr int HXL = (DX == 0) ? 3 : 1;
  constexpr int HYL = (DY == 0) ? 3 : 1;
  constexpr int LW = TX + 4, LH = TY + 4;
  constexpr int zlo = (DZ == 0) ? -3 : -1;
  constexpr int zc0 = -zlo;
  constexpr int NYP = 4 * TXP;                    // y-halo pairs per component
  constexpr int NXH = 4 * TY;                     // x-halo single cells per component
  static_assert(NYP + NXH <= NT, "one halo item per thread and component");
  __shared__ TC lds[2][kNC][LH][LW];

  // XCD-aware remap: consecutive hardware workgroup ids land on different
  // XCDs (round robin).  Each XCD takes a contiguous eighth of every shot's
  // tiles, shot after shot: halos of neighbouring tiles meet in its L2, and
  // any subset of active shots keeps all eight XCDs busy.  The grid holds
  // 8 x part x nshots workgroups; the few past a shot's last tile exit.
  const int per_shot = g.wg_beg[b.nbox];
  const int part = (per_shot + 7) / 8;
  const int hw = blockIdx.x;
  const int xcd = hw % 8, m = hw / 8;
  const int shot = m / part;
  const int rw = xcd * part + (m - shot * part);
  if (rw >= per_shot) return;
  int bx = 0;
  while (bx < b.nbox - 1 && rw >= g.wg_beg[bx + 1]) bx++;
  const int lw = rw - g.wg_beg[bx];
  const int ntx = g.ntx[bx], nty = g.nty[bx];
  const int tx_t = lw % ntx;
  const int ty_t = (lw / ntx) % nty;
  const int zc = lw / (ntx * nty);
  const int bi0 = b.box[bx][0], bi1 = b.box[bx][1], bj0 = b.box[bx][2], bj1 = b.box[bx][3];
  const int bk0 = b.box[bx][4], bk1 = b.box[bx][5];
  if (a.shot_active && !a.shot_active[shot]) return;

  const Layout &L = a.L;
  const long cs = L.cs;
  const int tid = threadIdx.x;
  const int lx = tid % TXP, ly = tid / TXP;
  const int i0 = (bi0 & ~1) + tx_t * TX, j0 = bj0 + ty_t * TY;  // even first column
  const int i = i0 + 2 * lx, j = j0 + ly;
  const int kbeg = bk0 + zc * b.zchunk;
  int kend = kbeg + b.zchunk - 1;
  if (kend > bk1) kend = bk1;
  const bool own0 = i >= bi0 && i <= bi1 && j <= bj1;
  const bool own1 = i + 1 >= bi0 && i + 1 <= bi1 && j <= bj1;
  // lanes past the last ghost column/row only feed LDS cells no owner reads;
  // clamping keeps every load in bounds (the last even start whose pair ends
  // at or before the last ghost column)
  const int imax = (L.ni + 1) & ~1;
  const int ic = i <= imax ? i : imax, jc = j <= L.nj + 2 ? j : L.nj + 2;
  const long so = (long)shot * a.shot_stride;
  const unsigned upz = (unsigned)L.pz, upx = (unsigned)L.px;
  const unsigned col = (unsigned)(jc + kNG) * upx + (unsigned)(ic + L.xa);
  auto sc = [&](int c) -> TC { return (TC)a.sc[c]; };
  // 32-bit pair at element offset off of component c of a level
  auto ldp = [&](const TS *base, int c, unsigned off) -> unsigned {
    return *reinterpret_cast<const unsigned *>(reinterpret_cast<const char *>(base + so + c * cs) +
                                               off * (unsigned)sizeof(TS));
  };
  auto ld1 = [&](const TS *base, int c, unsigned off) -> unsigned short {
    return *reinterpret_cast<const unsigned short *>(reinterpret_cast<const char *>(base + so + c * cs) +
                                                     off * (unsigned)sizeof(TS));
  };
  auto stp = [&](TS *base, int c, unsigned off, unsigned v) {
    *reinterpret_cast<unsigned *>(reinterpret_cast<char *>(base + so + c * cs) + off * (unsigned)sizeof(TS)) = v;
  };

  // this thread's halo item (the same for every component): a y-halo pair
  // (threads 0..NYP-1) or an x-halo cell (the next NXH threads)
  int h_kind = 0, h_lds = -1;
  unsigned h_off = 0;
  if (tid < NYP) {
    const int h = tid / TXP, cp = tid % TXP;
    const int lj = (h < HYL) ? h : TY + h;
    int hj = j0 - HYL + lj;
    hj = hj <= L.nj + 2 ? hj : L.nj + 2;
    int hi = i0 + 2 * cp;
    hi = hi <= imax ? hi : imax;
    h_kind = 1;
    h_lds = lj * LW + HXL + 2 * cp;
    h_off = (unsigned)(hj + kNG) * upx + (unsigned)(hi + L.xa);
  } else if (tid < NYP + NXH) {
    const int r = tid - NYP;
    const int h = r % 4, row = r / 4;
    const int li = (h < HXL) ? h : TX + h;
    int hi = i0 - HXL + li;
    hi = hi <= L.ni + 2 ? hi : L.ni + 2;
    int hj = j0 + row;
    hj = hj <= L.nj + 2 ? hj : L.nj + 2;
    h_kind = 2;
    h_lds = (row + HYL) * LW + li;
    h_off = (unsigned)(hj + kNG) * upx + (unsigned)(hi + L.xa);
  }
  unsigned q[kNC][5];  // packed column pairs, planes k+zlo .. k+zlo+4
#pragma unroll
  for (int c = 0; c < kNC; c++)
#pragma unroll
    for (int s = 0; s < 4; s++) q[c][s + 1] = ldp(a.w_in, c, (unsigned)(kbeg + zlo + s + kNG) * upz + col);

  const int mode = a.mode;
  // S0 takes its base from the queue's centre plane (the stage input is W^n)
  const bool need_end = mode == MODE_S12 || mode == MODE_S3, need_pre = mode == MODE_S12;
  // wavefield and coefficients of the next plane are fetched one plane ahead;
  // the RK levels, needed only in the epilogue, are fetched for the current
  // plane before that prefetch
  unsigned n_lead[kNC], n_h[kNC] = {0};
  float2 n_met[12];
  auto issue = [&](int kk) {
    const unsigned pk = (unsigned)(kk + kNG) * upz;
    const unsigned o = pk + col;
#pragma unroll
    for (int c = 0; c < kNC; c++) n_lead[c] = ldp(a.w_in, c, (unsigned)(kk + zlo + 4 + kNG) * upz + col);
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      if (h_kind == 1) n_h[c] = ldp(a.w_in, c, pk + h_off);
      else if (h_kind == 2) n_h[c] = ld1(a.w_in, c, pk + h_off);
    }
    const unsigned ob = o * (unsigned)sizeof(TC);
#pragma unroll
    for (int mm = 0; mm < 9; mm++)
      n_met[mm] = CART ? make_float2(a.cmet[mm], a.cmet[mm])
                       : *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(a.met + (XIX + mm) * cs) + ob);
#pragma unroll
    for (int mm = 0; mm < 3; mm++)
      n_met[9 + mm] = *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(a.med + mm * cs) + ob);

  };
  issue(kbeg);

  for (int k = kbeg; k <= kend; k++) {
    const unsigned o = (unsigned)(k + kNG) * upz + col;
    unsigned hv[kNC], pre_v[kNC], end_v[kNC];
    float2 mt[12];
#pragma unroll
    for (int c = 0; c < kNC; c++) {
#pragma unroll
      for (int s = 0; s < 4; s++) q[c][s] = q[c][s + 1];
      q[c][4] = n_lead[c];
      hv[c] = n_h[c];
    }
#pragma unroll
    for (int mm = 0; mm < 12; mm++) mt[mm] = n_met[mm];
    // this plane's RK levels go out before the next plane's prefetch: memory
    // counters complete in order, so a load issued after the prefetch could
    // only be waited for together with all of it
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      pre_v[c] = need_pre ? ldp(a.w_pre, c, o) : 0u;
      end_v[c] = need_end ? ldp(a.w_end, c, o) : 0u;
    }
    issue(k < kend ? k + 1 : kend);
    const int buf = (k - kbeg) & 1;
    TC *lb = &lds[buf][0][0][0];
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      float v0, v1;
      Pair<TS>::unpack(q[c][zc0], v0, v1);
      TC *row = lb + (c * LH + ly + HYL) * LW + HXL + 2 * lx;
      row[0] = v0 * sc(c);
      row[1] = v1 * sc(c);
      if (h_kind == 1) {
        float h0, h1;
        Pair<TS>::unpack(hv[c], h0, h1);
        lb[c * LH * LW + h_lds] = h0 * sc(c);
        lb[c * LH * LW + h_lds + 1] = h1 * sc(c);
      } else if (h_kind == 2) {
        lb[c * LH * LW + h_lds] = Pair<TS>::one((unsigned short)hv[c]) * sc(c);
      }
    }
    __syncthreads();

    // derivatives of both columns together (one unpack of each queue pair)
    TC Lg[2][3][3];
    TC dv[2][3] = {{TC(0), TC(0), TC(0)}, {TC(0), TC(0), TC(0)}};
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      TC dx[2] = {TC(0), TC(0)}, dy[2] = {TC(0), TC(0)}, dz[2] = {TC(0), TC(0)};
#pragma unroll
      for (int n = 0; n < 5; n++) {
        float z0, z1;
        Pair<TS>::unpack(q[c][n], z0, z1);
        const TC cz = Drp<TC, DZ>::c(n), cx = Drp<TC, DX>::c(n), cy = Drp<TC, DY>::c(n);
        dz[0] += cz * (z0 * sc(c));
        dz[1] += cz * (z1 * sc(c));
#pragma unroll
        for (int p = 0; p < 2; p++) {
          dx[p] += cx * lb[(c * LH + ly + HYL) * LW + 2 * lx + p + n];
          dy[p] += cy * lb[(c * LH + ly + n) * LW + 2 * lx + p + HXL];
        }
      }
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const TC xix = p ? mt[0].y : mt[0].x, xiy = p ? mt[1].y : mt[1].x, xiz = p ? mt[2].y : mt[2].x;
        const TC etx = p ? mt[3].y : mt[3].x, ety = p ? mt[4].y : mt[4].x, etz = p ? mt[5].y : mt[5].x;
        const TC ztx = p ? mt[6].y : mt[6].x, zty = p ? mt[7].y : mt[7].x, ztz = p ? mt[8].y : mt[8].x;
        const TC g0 = xix * dx[p] + etx * dy[p] + ztx * dz[p];
        const TC g1 = xiy * dx[p] + ety * dy[p] + zty * dz[p];
        const TC g2 = xiz * dx[p] + etz * dy[p] + ztz * dz[p];
        if (c < 3) {
          Lg[p][c][0] = g0; Lg[p][c][1] = g1; Lg[p][c][2] = g2;
        } else if (c == TXX) {
          dv[p][0] += g0;
        } else if (c == TYY) {
          dv[p][1] += g1;
        } else if (c == TZZ) {
          dv[p][2] += g2;
        } else if (c == TYZ) {
          dv[p][1] += g2; dv[p][2] += g1;
        } else if (c == TXZ) {
          dv[p][0] += g2; dv[p][2] += g0;
        } else {
          dv[p][0] += g1; dv[p][1] += g0;
        }
      }
    }

    // right-hand side and RK update of both columns, written as pairs
    unsigned outp[kNC], endp[kNC];
    TC o_lo[kNC], e_lo[kNC];
#pragma unroll
    for (int p = 0; p < 2; p++) {
      const TC lam = p ? mt[9].y : mt[9].x, mu = p ? mt[10].y : mt[10].x, slw = p ? mt[11].y : mt[11].x;
      const TC lam2mu = lam + TC(2) * mu;
      TC h[kNC];
      h[VX] = slw * dv[p][0];
      h[VY] = slw * dv[p][1];
      h[VZ] = slw * dv[p][2];
      h[TXX] = lam2mu * Lg[p][0][0] + lam * (Lg[p][1][1] + Lg[p][2][2]);
      h[TYY] = lam2mu * Lg[p][1][1] + lam * (Lg[p][0][0] + Lg[p][2][2]);
      h[TZZ] = lam2mu * Lg[p][2][2] + lam * (Lg[p][0][0] + Lg[p][1][1]);
      h[TXY] = mu * (Lg[p][0][1] + Lg[p][1][0]);
      h[TXZ] = mu * (Lg[p][0][2] + Lg[p][2][0]);
      h[TYZ] = mu * (Lg[p][1][2] + Lg[p][2][1]);
#pragma unroll
      for (int c = 0; c < kNC; c++) {
        float b0, b1, e0, e1;
        Pair<TS>::unpack(mode == MODE_S0 ? q[c][zc0] : pre_v[c], b0, b1);
        Pair<TS>::unpack(end_v[c], e0, e1);
        const TC base = (p ? b1 : b0) * sc(c);
        const TC base_end = mode == MODE_S0 ? base : (p ? e1 : e0) * sc(c);
        const TC vout = mode == MODE_RHS ? h[c] : base + a.ca * h[c];
        const TC vend = base_end + a.cb * h[c];
        const TC is = (TC)a.isc[c];
        if (p == 0) {
          o_lo[c] = vout * is;
          e_lo[c] = vend * is;
        } else {
          outp[c] = Pair<TS>::pack(o_lo[c], vout * is);
          endp[c] = Pair<TS>::pack(e_lo[c], vend * is);
        }
      }
    }
    // a pair with one owned column (tile edge) stores only that 16-bit half
    const bool wout = mode != MODE_S3, wend = mode != MODE_RHS;
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      if (own0 && own1) {
        if (wout) stp(a.w_out, c, o, outp[c]);
        if (wend) stp(a.w_end, c, o, endp[c]);
      } else if (own0 || own1) {
        const unsigned sh = own0 ? 0u : 16u, ho = own0 ? 0u : 1u;
        auto st1 = [&](TS *base, unsigned v) {
          *reinterpret_cast<unsigned short *>(reinterpret_cast<char *>(base + so + c * cs) +
                                              (o + ho) * (unsigned)sizeof(TS)) = (unsigned short)(v >> sh);
        };
        if (wout) st1(a.w_out, outp[c]);
        if (wend) st1(a.w_end, endp[c]);
      }
    }
  }
}

template <typename TS, int TXP, int TY, int MINB, bool CART>
static hipError_t launch_h2_t(const StageArgs<TS, float> &a, const MarchSet &b, int dx, int dy, int dz, int nshots,
                              hipStream_t st) {
  constexpr int TX = 2 * TXP;
  MarchGrid g{};
  int acc = 0;
  for (int q = 0; q < b.nbox; q++) {
    const int x0 = b.box[q][0] & ~1;
    g.wg_beg[q] = acc;
    g.ntx[q] = (b.box[q][1] - x0 + TX) / TX;
    g.nty[q] = (b.box[q][3] - b.box[q][2] + TY) / TY;
    g.nzc[q] = (b.box[q][5] - b.box[q][4] + b.zchunk) / b.zchunk;
    acc += g.ntx[q] * g.nty[q] * g.nzc[q];
  }
  g.wg_beg[b.nbox] = acc;
  const long nwg = 8L * ((acc + 7) / 8) * nshots;  // see the remap at the top of the kernel
  if (nwg == 0) return hipSuccess;
  dim3 grid((unsigned)nwg), block(TXP * TY);
  const int sel = dx * 4 + dy * 2 + dz;
#define CGFD_L(X, Y, Z) hipLaunchKernelGGL((k_march_h2<X, Y, Z, TS, TXP, TY, MINB, CART>), grid, block, 0, st, a, b, g)
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

// shape 7: 64x8 points (32x8 threads), shape 8: 64x16 points (32x16 threads)
template <typename TS>
hipError_t launch_interior_h2(const StageArgs<TS, float> &a, const MarchSet &b, int dx, int dy, int dz, int nshots,
                              hipStream_t st) {
  if (a.met_const)
    return b.shape == 8 ? launch_h2_t<TS, 32, 16, 1, true>(a, b, dx, dy, dz, nshots, st)
                        : launch_h2_t<TS, 32, 8, 2, true>(a, b, dx, dy, dz, nshots, st);
  return b.shape == 8 ? launch_h2_t<TS, 32, 16, 1, false>(a, b, dx, dy, dz, nshots, st)
                      : launch_h2_t<TS, 32, 8, 2, false>(a, b, dx, dy, dz, nshots, st);
}

template hipError_t launch_interior_h2<__half>(const StageArgs<__half, float> &, const MarchSet &, int, int, int,
                                               int, hipStream_t);
template hipError_t launch_interior_h2<__hip_bfloat16>(const StageArgs<__hip_bfloat16, float> &, const MarchSet &,
                                                       int, int, int, int, hipStream_t);

}  // namespace cgfd
