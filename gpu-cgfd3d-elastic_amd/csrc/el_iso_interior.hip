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
// Per plane, every global load a thread needs (the queue's leading plane,
// its share of the halo, the 12 coefficients and the RK levels it updates)
// is issued before the first barrier, so one memory latency is exposed per
// plane instead of one per dependent load.
//
// On a Cartesian grid every metric component is one constant and only the
// diagonal ones (xi_x, eta_y, zeta_z) are non-zero (CART = true): the metric
// comes in as arguments instead of nine HBM streams, and each derivative
// maps to one physical direction (6 of the 9 products per component drop).
//
// The PML variant (PML = true) sweeps whole x-y planes, PML face slabs
// included: a point inside a face box recomputes the nine derivatives along
// that face's axis from the same LDS tile or z queue (bit-identical to the
// sums above it) and applies the CFS-PML ADE terms and memory-variable
// update.  y and z faces are uniform across a wavefront (a wave spans one
// or two rows of one plane); only the tiles at the two x ends diverge.  The
// interior variant (PML = false) sweeps a box outside every face box.
// Points within 4 rows of a free surface go to the per-point kernel.
#include "el_iso_device.h"
#include "el_iso_launch.h"
#include "el_iso_pml.h"

#include <type_traits>

namespace cgfd {

template <int DX, int DY, int DZ, typename TS, typename TC, int TX, int TY, int MINB, bool CART, bool PML, int MODE>
__global__ __launch_bounds__(TX *TY, MINB) void k_march(StageArgs<TS, TC> a, MarchSet b, MarchGrid g) {
  constexpr int NT = TX * TY;
  constexpr int HXL = (DX == 0) ? 3 : 1;
  constexpr int HYL = (DY == 0) ? 3 : 1;
  constexpr int LW = TX + 4, LH = TY + 4;
  constexpr int zlo = (DZ == 0) ? -3 : -1;
  constexpr int zc0 = -zlo;                       // queue slot of the centre plane
  // Halo cells per component: 4 rows of TX (y halo) and 4 columns of TY (x
  // halo).  A thread owns one y-halo cell and one x-halo cell, the same for
  // every component it loads: NT / YH components per pass of the y halo
  // (wave-uniform component, so each load has a scalar base) and NT / XH per
  // pass of the x halo.
  constexpr int YH = 4 * TX, XH = 4 * TY;
  static_assert(NT % YH == 0 && NT % XH == 0, "tile must hold whole halo passes");
  constexpr int YPC = NT / YH, XPC = NT / XH;           // components per pass
  constexpr int NPY = (kNC + YPC - 1) / YPC;            // y-halo passes
  constexpr int NPX = (kNC + XPC - 1) / XPC;            // x-halo passes
  constexpr int NPER = NPY + NPX;                       // halo loads per thread
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
  const int lx = tid % TX, ly = tid / TX;
  const int i0 = bi0 + tx_t * TX, j0 = bj0 + ty_t * TY;
  const int i = i0 + lx, j = j0 + ly;
  const int kbeg = bk0 + zc * b.zchunk;
  int kend = kbeg + b.zchunk - 1;
  if (kend > bk1) kend = bk1;
  const bool owner = (i <= bi1) && (j <= bj1);
  // face boxes this tile's columns and z chunk meet (uniform bit mask): the
  // PML block runs only on planes that meet one
  unsigned tile_faces = 0;
  if constexpr (PML) {
    const int ti1 = i0 + TX - 1 < bi1 ? i0 + TX - 1 : bi1, tj1 = j0 + TY - 1 < bj1 ? j0 + TY - 1 : bj1;
#pragma unroll
    for (int fi = 0; fi < 6; fi++) {
      const FaceArgs<TS, TC> &F = a.f[fi];
      if (F.on && F.i1 <= ti1 && F.i2 >= i0 && F.j1 <= tj1 && F.j2 >= j0 && F.k1 <= kend && F.k2 >= kbeg)
        tile_faces |= 1u << fi;
    }
  }
  // lanes past the last ghost column/row only feed LDS cells no owner reads;
  // clamping their address keeps every load unconditional and in bounds
  const int ic = i <= L.ni + 2 ? i : L.ni + 2, jc = j <= L.nj + 2 ? j : L.nj + 2;
  const long so = (long)shot * a.shot_stride;
  const unsigned upz = (unsigned)L.pz, upx = (unsigned)L.px;
  const unsigned col = (unsigned)(jc + kNG) * upx + (unsigned)(ic + L.xa);

  // uniform per-component bases with 32-bit byte offsets: scalar-base addressing
  auto wq = [&](int c) -> const TS * { return a.w_in + so + c * cs; };
  auto sc = [&](int c) -> TC { return St<TS>::scaled ? (TC)a.sc[c] : TC(1); };
  auto gld = [&](int c, unsigned off) -> TC {
    return (TC)St<TS>::ld(&at_b(wq(c), off * (unsigned)sizeof(TS))) * sc(c);
  };
  // Prefetched values stay raw (16-bit storage: the loaded bits) until the
  // plane that uses them: a conversion next to its load would make the wave
  // wait for the whole prefetch before computing the current plane.
  using RT = typename St<TS>::raw;
  auto graw = [&](int c, unsigned off) -> RT { return St<TS>::ld_raw(&at_b(wq(c), off * (unsigned)sizeof(TS))); };
  auto cv = [&](int c, RT r) -> TC { return (TC)St<TS>::cvt(r) * sc(c); };

  // this thread's halo cells (positions fixed over the sweep)
  const int ycb = tid / YH, yr = tid % YH;   // y halo: first component, cell
  const int xcb = tid / XH, xr = tid % XH;   // x halo: first component, cell
  unsigned y_off, x_off;
  int y_lds, x_lds;
  {
    const int h = yr / TX, cc = yr % TX;
    const int lj = (h < HYL) ? h : TY + h;
    int hj = j0 - HYL + lj;
    int hi = i0 + cc;
    hi = hi <= L.ni + 2 ? hi : L.ni + 2;
    hj = hj <= L.nj + 2 ? hj : L.nj + 2;
    y_lds = lj * LW + cc + HXL;
    y_off = (unsigned)(hj + kNG) * upx + (unsigned)(hi + L.xa);
  }
  {
    const int h = xr % 4, row = xr / 4;
    const int li = (h < HXL) ? h : TX + h;
    int hi = i0 - HXL + li;
    int hj = j0 + row;
    hi = hi <= L.ni + 2 ? hi : L.ni + 2;
    hj = hj <= L.nj + 2 ? hj : L.nj + 2;
    x_lds = (row + HYL) * LW + li;
    x_off = (unsigned)(hj + kNG) * upx + (unsigned)(hi + L.xa);
  }
  // component of halo load p (y passes first), clamped; valid when < kNC
  auto hcomp = [&](int p) -> int { return p < NPY ? p * YPC + ycb : (p - NPY) * XPC + xcb; };
  TC *ldsf = &lds[0][0][0][0];
  constexpr int LBUF = kNC * LH * LW;  // one LDS plane buffer

  // scale of each halo load's component on this lane, by selects over the
  // kernel arguments (a per-lane index into them would be a memory load)
  TC hsc[NPER];
#pragma unroll
  for (int p = 0; p < NPER; p++) {
    const int c = hcomp(p);
    TC v = sc(0);
#pragma unroll
    for (int q = 1; q < kNC; q++) v = c == q ? sc(q) : v;
    hsc[p] = v;
  }

  // register queues: planes k+zlo .. k+zlo+4
  TC q[kNC][5];
#pragma unroll
  for (int c = 0; c < kNC; c++)
#pragma unroll
    for (int s = 0; s < 4; s++) q[c][s + 1] = gld(c, (unsigned)(kbeg + zlo + s + kNG) * upz + col);

  // Everything plane k needs from HBM/L2, fetched one plane ahead so its
  // latency overlaps the previous plane's arithmetic.
  // The RK mode is a template argument: every load a mode needs is then
  // unconditional, so no 16-bit value is converted inside a branch right
  // after its load (which makes the wave wait there for every load in
  // flight, the next plane's prefetch included).
  constexpr int mode = MODE;
  // W^n at the point: S12 and S3; U1 and U2 at the point: S3 (folded into
  // one partial sum per component as soon as they arrive)
  constexpr bool need_pre = mode == MODE_S12 || mode == MODE_S3, need_u = mode == MODE_S3;
  // The PML form fetches the RK levels for the current plane (not one plane
  // ahead): the 18 registers it saves are what its face update needs.
  constexpr bool LATE_RK = PML;
  // With FP32/FP64 storage the RK levels W^n / END and the stage outputs are
  // each touched once per stage: they bypass L2 (non-temporal) and leave it
  // to the planes the stencil re-reads (sweep 1.93 -> 1.78 ms on the
  // topography workload).  16-bit storage gained nothing (its loads lost).
  auto ld_rk = [&](const TS *p) -> RT {
    if constexpr (St<TS>::scaled) return St<TS>::ld_raw(p);
    else return St<TS>::ld_raw_nt(p);
  };
  auto st_w = [&](TS *p, TC v) {
    if constexpr (St<TS>::scaled) St<TS>::st(p, v);
    else St<TS>::st_nt(p, v);
  };
  RT n_lead[kNC], n_hv[NPER], n_pre[kNC], n_u1[kNC], n_u2[kNC];
  TC n_met[12];
#pragma unroll
  for (int c = 0; c < kNC; c++) { n_pre[c] = RT(0); n_u1[c] = RT(0); n_u2[c] = RT(0); }
  // this plane's RK levels at the point (raw)
  auto load_rk = [&](unsigned qb, RT *pre, RT *u1, RT *u2) {
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      pre[c] = need_pre ? ld_rk(&at_b(a.w_pre + so + c * cs, qb)) : RT(0);
      if constexpr (need_u) {
        u1[c] = ld_rk(&at_b(a.w_end + so + c * cs, qb));
        u2[c] = ld_rk(&at_b(a.w_u2 + so + c * cs, qb));
      }
    }
  };
  auto issue = [&](int kk) {
    const unsigned pk = (unsigned)(kk + kNG) * upz;
    const unsigned o = pk + col;
#pragma unroll
    for (int c = 0; c < kNC; c++) n_lead[c] = graw(c, (unsigned)(kk + zlo + 4 + kNG) * upz + col);
#pragma unroll
    for (int p = 0; p < NPER; p++) {
      const int c = hcomp(p);
      n_hv[p] = graw(c < kNC ? c : kNC - 1, pk + (p < NPY ? y_off : x_off));
    }
    const unsigned ob = o * (unsigned)sizeof(TC);
#pragma unroll
    for (int m = 0; m < 9; m++) n_met[m] = CART ? a.cmet[m] : (at_b(a.met + (XIX + m) * cs, ob));
#pragma unroll
    for (int m = 0; m < 3; m++) n_met[9 + m] = (at_b(a.med + m * cs, ob));
    const unsigned qb = o * (unsigned)sizeof(TS);
    if constexpr (!LATE_RK && need_pre) load_rk(qb, n_pre, n_u1, n_u2);
  };
  // Every priming load done before the loop (vmcnt(0); the other counters
  // untouched): otherwise the wait analysis, merging the loop entry with its
  // back edge, makes each plane wait for most of the next plane's prefetch
  // before its arithmetic.
  __builtin_amdgcn_s_waitcnt(0x0F70);
  issue(kbeg);

  for (int k = kbeg; k <= kend; k++) {
    const unsigned pk = (unsigned)(k + kNG) * upz;
    const unsigned o = pk + col;
    const unsigned qb = o * (unsigned)sizeof(TS);
    TC hv[NPER], mt[12], pre_v[kNC], part_v[kNC];
#pragma unroll
    for (int c = 0; c < kNC; c++) {
#pragma unroll
      for (int s = 0; s < 4; s++) q[c][s] = q[c][s + 1];
      q[c][4] = cv(c, n_lead[c]);
      if constexpr (!LATE_RK) {
        pre_v[c] = need_pre ? cv(c, n_pre[c]) : TC(0);
        part_v[c] = need_u ? rk_partial(a.rk_w, pre_v[c], cv(c, n_u1[c]), cv(c, n_u2[c])) : TC(0);
      }
    }
#pragma unroll
    for (int p = 0; p < NPER; p++) hv[p] = (TC)St<TS>::cvt(n_hv[p]) * hsc[p];
#pragma unroll
    for (int m = 0; m < 12; m++) mt[m] = n_met[m];
    // LATE_RK: this plane's RK levels, issued before the next plane's
    // prefetch (memory counters complete in order), converted where used
    RT pre_r[kNC], u1_r[kNC], u2_r[kNC];
    if constexpr (LATE_RK && need_pre) load_rk(qb, pre_r, u1_r, u2_r);
    if (k < kend) issue(k + 1);
    const TC xix = mt[0], xiy = mt[1], xiz = mt[2], etx = mt[3], ety = mt[4], etz = mt[5];
    const TC ztx = mt[6], zty = mt[7], ztz = mt[8], lam = mt[9], mu = mt[10], slw = mt[11];
    // ---- stage the centre plane and its halo in LDS (double buffered) ----
    const int buf = (k - kbeg) & 1;
#pragma unroll
    for (int c = 0; c < kNC; c++) lds[buf][c][ly + HYL][lx + HXL] = q[c][zc0];
#pragma unroll
    for (int p = 0; p < NPER; p++) {
      const int c = hcomp(p);
      if (c < kNC) ldsf[buf * LBUF + c * (LH * LW) + (p < NPY ? y_lds : x_lds)] = hv[p];
    }
    __syncthreads();

    // each component's derivatives fold straight into the velocity gradient
    // (V) or the stress divergence (T): 12 accumulators instead of 27 values
    TC Lg[3][3];                       // Lg[a][b] = dV_a / dx_b
    TC dv[3] = {TC(0), TC(0), TC(0)};  // dT_ab / dx_b
#pragma unroll
    for (int c = 0; c < kNC; c++) {
      TC dx = TC(0), dy = TC(0), dz = TC(0);
#pragma unroll
      for (int n = 0; n < 5; n++) {
        dx += Drp<TC, DX>::c(n) * lds[buf][c][ly + HYL][lx + n];
        dy += Drp<TC, DY>::c(n) * lds[buf][c][ly + n][lx + HXL];
        dz += Drp<TC, DZ>::c(n) * q[c][n];
      }
      const TC g0 = CART ? xix * dx : xix * dx + etx * dy + ztx * dz;
      const TC g1 = CART ? ety * dy : xiy * dx + ety * dy + zty * dz;
      const TC g2 = CART ? ztz * dz : xiz * dx + etz * dy + ztz * dz;
      if (c < 3) {
        Lg[c][0] = g0; Lg[c][1] = g1; Lg[c][2] = g2;
      } else if (c == TXX) {
        dv[0] += g0;
      } else if (c == TYY) {
        dv[1] += g1;
      } else if (c == TZZ) {
        dv[2] += g2;
      } else if (c == TYZ) {
        dv[1] += g2; dv[2] += g1;
      } else if (c == TXZ) {
        dv[0] += g2; dv[2] += g0;
      } else {
        dv[0] += g1; dv[1] += g0;
      }
    }

    if (owner) {
      const TC lam2mu = lam + TC(2) * mu;
      TC h[kNC];
      h[VX] = slw * dv[0];
      h[VY] = slw * dv[1];
      h[VZ] = slw * dv[2];
      h[TXX] = lam2mu * Lg[0][0] + lam * (Lg[1][1] + Lg[2][2]);
      h[TYY] = lam2mu * Lg[1][1] + lam * (Lg[0][0] + Lg[2][2]);
      h[TZZ] = lam2mu * Lg[2][2] + lam * (Lg[0][0] + Lg[1][1]);
      h[TXY] = mu * (Lg[0][1] + Lg[1][0]);
      h[TXZ] = mu * (Lg[0][2] + Lg[2][0]);
      h[TYZ] = mu * (Lg[1][2] + Lg[2][1]);
      unsigned plane_faces = 0;
      if constexpr (PML) {
#pragma unroll
        for (int fi = 0; fi < 6; fi++)
          if (((tile_faces >> fi) & 1u) && k >= a.f[fi].k1 && k <= a.f[fi].k2) plane_faces |= 1u << fi;
      }
      if (PML && plane_faces) {
        // Every face box holding the point, in the reference's face order
        // (sv_curv_col_el_iso_gpu.cu:613-1147).  The compiler-only memory
        // clobber makes the derivative sums below read LDS again instead of
        // keeping the 90 x/y tap values of the loop above live until here;
        // the face loop stays rolled (one copy of its body).
        __asm__ volatile("" ::: "memory");
        // one body per face axis: straight-line derivative sums (LDS rows,
        // LDS columns or the z queue)
        auto face = [&](auto dimc, int fi) {
          constexpr int DIM = decltype(dimc)::value;
          TC Dd[kNC];
#pragma unroll
          for (int c = 0; c < kNC; c++) {
            TC d = TC(0);
#pragma unroll
            for (int n = 0; n < 5; n++) {
              if constexpr (DIM == 0) d += Drp<TC, DX>::c(n) * lds[buf][c][ly + HYL][lx + n];
              else if constexpr (DIM == 1) d += Drp<TC, DY>::c(n) * lds[buf][c][ly + n][lx + HXL];
              else d += Drp<TC, DZ>::c(n) * q[c][n];
            }
            Dd[c] = d;
          }
          const TC e1 = DIM == 0 ? xix : DIM == 1 ? etx : ztx;
          const TC e2 = DIM == 0 ? xiy : DIM == 1 ? ety : zty;
          const TC e3 = DIM == 0 ? xiz : DIM == 1 ? etz : ztz;
          pml_face_march<DIM, mode, TS, TC>(a, fi, shot, i, j, k, Dd, e1, e2, e3, lam, mu, lam2mu, slw, h);
        };
#pragma unroll 1
        for (unsigned fm = plane_faces; fm; fm &= fm - 1) {
          const int fi = __builtin_ctz(fm);  // uniform: scalar loads of the face's arguments
          if (!pml_in_face(a, fi, i, j, k)) continue;
          if (fi < 2) face(std::integral_constant<int, 0>{}, fi);
          else if (fi < 4) face(std::integral_constant<int, 1>{}, fi);
          else face(std::integral_constant<int, 2>{}, fi);
        }
      }
#pragma unroll
      for (int c = 0; c < kNC; c++) {
        TS *wout = a.w_out + so + c * cs;
        TS *wend = a.w_end + so + c * cs;
        const TC is = St<TS>::scaled ? (TC)a.isc[c] : TC(1);
        TC pre_c = TC(0);
        if constexpr (need_pre) pre_c = LATE_RK ? cv(c, pre_r[c]) : pre_v[c];
        if constexpr (mode == MODE_S12) {
          st_w(&at_b(wout, qb), (pre_c + a.ca * h[c]) * is);
        } else if constexpr (mode == MODE_S0) {
          st_w(&at_b(wout, qb), (q[c][zc0] + a.ca * h[c]) * is);
        } else if constexpr (mode == MODE_S3) {
          TC part;
          if constexpr (LATE_RK) part = rk_partial(a.rk_w, pre_c, cv(c, u1_r[c]), cv(c, u2_r[c]));
          else part = part_v[c];
          st_w(&at_b(wend, qb), rk_final(pre_c, part, a.rk_w[2], q[c][zc0], a.cb * h[c]) * is);
        } else {
          st_w(&at_b(wout, qb), h[c] * is);
        }
      }
    }
    // no second barrier: the next plane writes the other LDS buffer, and the
    // one after that is only written once every wave has passed the next barrier
  }
}

// one RK mode's k_march for the operator directions sel = 4 dx + 2 dy + dz
template <typename TS, typename TC, int TX, int TY, int MINB, bool CART, bool PML, int MODE>
static void launch_march_m(const StageArgs<TS, TC> &a, const MarchSet &b, const MarchGrid &g, dim3 grid, dim3 block,
                           int sel, hipStream_t st) {
  // 16-bit storage, constant metric, PML form, 32x8 tile: S1/S2 compiled for
  // three waves per SIMD (168 VGPRs instead of 174, 40 bytes of spill) ran
  // its sweep 15 % faster on 800x800x400.  The same bound spills 84-336
  // bytes in the other variants (S3, the general metric, 64x4) and was
  // slower or no faster there.
  constexpr int MB = (sizeof(TS) == 2 && CART && PML && MODE == MODE_S12 && TX == 32 && TY == 8) ? 3 : MINB;
#define CGFD_L(X, Y, Z) \
  hipLaunchKernelGGL((k_march<X, Y, Z, TS, TC, TX, TY, MB, CART, PML, MODE>), grid, block, 0, st, a, b, g)
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

template <typename TS, typename TC, int TX, int TY, int MINB, bool CART, bool PML>
static hipError_t launch_march_t(const StageArgs<TS, TC> &a, const MarchSet &b, int dx, int dy, int dz,
                                 int nshots, hipStream_t st) {
  MarchGrid g{};
  int acc = 0;
  for (int q = 0; q < b.nbox; q++) {
    g.wg_beg[q] = acc;
    g.ntx[q] = (b.box[q][1] - b.box[q][0] + TX) / TX;
    g.nty[q] = (b.box[q][3] - b.box[q][2] + TY) / TY;
    g.nzc[q] = (b.box[q][5] - b.box[q][4] + b.zchunk) / b.zchunk;
    acc += g.ntx[q] * g.nty[q] * g.nzc[q];
  }
  g.wg_beg[b.nbox] = acc;
  const long nwg = 8L * ((acc + 7) / 8) * nshots;  // see the remap at the top of the kernel
  if (nwg == 0) return hipSuccess;
  dim3 grid((unsigned)nwg), block(TX * TY);
  const int sel = dx * 4 + dy * 2 + dz;
  switch (a.mode) {
    case MODE_S0: launch_march_m<TS, TC, TX, TY, MINB, CART, PML, MODE_S0>(a, b, g, grid, block, sel, st); break;
    case MODE_S12: launch_march_m<TS, TC, TX, TY, MINB, CART, PML, MODE_S12>(a, b, g, grid, block, sel, st); break;
    case MODE_S3: launch_march_m<TS, TC, TX, TY, MINB, CART, PML, MODE_S3>(a, b, g, grid, block, sel, st); break;
    default: launch_march_m<TS, TC, TX, TY, MINB, CART, PML, MODE_RHS>(a, b, g, grid, block, sel, st); break;
  }
  return hipGetLastError();
}

// tile shapes the autotuner chooses from (TX x TY threads, one per column)
#if !defined(CGFD_TYPE) || CGFD_TYPE == 0
int interior_num_shapes() { return 4; }

void interior_shape(int id, int *tx, int *ty) {
  static const int T[4][2] = {{64, 4}, {32, 8}, {64, 8}, {32, 16}};
  id = id < 0 ? 0 : id > 3 ? 3 : id;
  *tx = T[id][0];
  *ty = T[id][1];
}
#endif

template <typename TS, typename TC, bool PML>
static hipError_t launch_march_shape(const StageArgs<TS, TC> &a, const MarchSet &b, int dx, int dy, int dz,
                                     int nshots, hipStream_t st) {
  if (a.met_const) {
    switch (b.shape) {
      case 0: return launch_march_t<TS, TC, 64, 4, 2, true, PML>(a, b, dx, dy, dz, nshots, st);
      case 1: return launch_march_t<TS, TC, 32, 8, 2, true, PML>(a, b, dx, dy, dz, nshots, st);
      case 3: return launch_march_t<TS, TC, 32, 16, 1, true, PML>(a, b, dx, dy, dz, nshots, st);
      default: return launch_march_t<TS, TC, 64, 8, 1, true, PML>(a, b, dx, dy, dz, nshots, st);
    }
  }
  switch (b.shape) {
    case 0: return launch_march_t<TS, TC, 64, 4, 2, false, PML>(a, b, dx, dy, dz, nshots, st);
    case 1: return launch_march_t<TS, TC, 32, 8, 2, false, PML>(a, b, dx, dy, dz, nshots, st);
    case 3: return launch_march_t<TS, TC, 32, 16, 1, false, PML>(a, b, dx, dy, dz, nshots, st);
    default: return launch_march_t<TS, TC, 64, 8, 1, false, PML>(a, b, dx, dy, dz, nshots, st);
  }
}

template <typename TS, typename TC>
hipError_t launch_interior(const StageArgs<TS, TC> &a, const MarchSet &b, int dx, int dy, int dz, int nshots,
                           hipStream_t st) {
  if (b.nbox < 1) return hipSuccess;
  if (b.pml) return launch_march_shape<TS, TC, true>(a, b, dx, dy, dz, nshots, st);
  return launch_march_shape<TS, TC, false>(a, b, dx, dy, dz, nshots, st);
}

#define CGFD_INST(TS, TC)                                                                                   \
  template hipError_t launch_interior<TS, TC>(const StageArgs<TS, TC> &, const MarchSet &, int, int, int, int, \
                                              hipStream_t);
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

}  // namespace cgfd
