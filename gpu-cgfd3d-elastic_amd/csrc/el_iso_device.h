// el_iso_device.h -- device-side building blocks of the gfx950 el_iso path:
// storage conversion, MacCormack/DRP operator constants, the padded HBM
// layout and the by-value argument blocks of the stage kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <type_traits>

namespace cgfd {

constexpr int kNG = 3;    // ghost points per side (src/forward/fd_t.cu:64-66)
constexpr int kNC = 9;    // wavefield components
enum Cmp { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };
enum Met { JAC = 0, XIX, XIY, XIZ, ETX, ETY, ETZ, ZTX, ZTY, ZTZ, NMET };
// RK modes of a stage kernel.  The classical RK4 step is evaluated from the
// stored stage values U1..U3 (U_{i+1} = W^n + a_i dt K_i):
//   S0:  U1 = W + a0 dt K0                      (input W^n)
//   S12: U2 = W + a1 dt K1, U3 = W + a2 dt K2   (input U1 / U2, W^n read at the point)
//   S3:  W^{n+1} = W + w1 (U1 - W) + w2 (U2 - W) + w3 (U3 - W) + b3 dt K3
//        with w_i = b_{i-1} / a_{i-1} = 1/3, 2/3, 1/3 (input U3; written over U1)
// which equals W + dt (b0 K0 + b1 K1 + b2 K2 + b3 K3).  MODE_RHS writes the
// right-hand side alone (the per-stage operator).
enum Mode { MODE_RHS = 0, MODE_S0 = 1, MODE_S12 = 2, MODE_S3 = 3 };

// ---- storage traits ---------------------------------------------------------
template <typename TS> struct St;
// ld/st: ordinary accesses; ld_nt/st_nt: non-temporal (streamed once, kept
// out of the way of the wavefield planes that the stencil re-reads from L2)
// raw / ld_raw / cvt: the loaded bits kept unconverted (a prefetch buffer
// converts where the value is used, one plane later)
template <> struct St<float> {
  static constexpr bool scaled = false;
  using raw = float;
  __device__ __forceinline__ static raw ld_raw(const float *p) { return *p; }
  __device__ __forceinline__ static raw ld_raw_nt(const float *p) { return __builtin_nontemporal_load(p); }
  __device__ __forceinline__ static float cvt(raw r) { return r; }
  __device__ __forceinline__ static float ld(const float *p) { return *p; }
  __device__ __forceinline__ static void st(float *p, float v) { *p = v; }
  __device__ __forceinline__ static float ld_nt(const float *p) { return __builtin_nontemporal_load(p); }
  __device__ __forceinline__ static void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
};
template <> struct St<double> {
  static constexpr bool scaled = false;
  using raw = double;
  __device__ __forceinline__ static raw ld_raw(const double *p) { return *p; }
  __device__ __forceinline__ static raw ld_raw_nt(const double *p) { return __builtin_nontemporal_load(p); }
  __device__ __forceinline__ static double cvt(raw r) { return r; }
  __device__ __forceinline__ static double ld(const double *p) { return *p; }
  __device__ __forceinline__ static void st(double *p, double v) { *p = v; }
  __device__ __forceinline__ static double ld_nt(const double *p) { return __builtin_nontemporal_load(p); }
  __device__ __forceinline__ static void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
};
template <> struct St<__half> {
  static constexpr bool scaled = true;
  using raw = unsigned short;
  __device__ __forceinline__ static raw ld_raw(const __half *p) { return *reinterpret_cast<const unsigned short *>(p); }
  __device__ __forceinline__ static raw ld_raw_nt(const __half *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
  }
  __device__ __forceinline__ static float cvt(raw r) { return __half2float(__ushort_as_half(r)); }
  __device__ __forceinline__ static float ld(const __half *p) { return __half2float(*p); }
  __device__ __forceinline__ static void st(__half *p, float v) { *p = __float2half(v); }
  __device__ __forceinline__ static float ld_nt(const __half *p) {
    const unsigned short u = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
    return __half2float(__ushort_as_half(u));
  }
  __device__ __forceinline__ static void st_nt(__half *p, float v) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(v)), reinterpret_cast<unsigned short *>(p));
  }
};
template <> struct St<__hip_bfloat16> {
  static constexpr bool scaled = true;
  using raw = unsigned short;
  __device__ __forceinline__ static raw ld_raw(const __hip_bfloat16 *p) {
    return *reinterpret_cast<const unsigned short *>(p);
  }
  __device__ __forceinline__ static raw ld_raw_nt(const __hip_bfloat16 *p) {
    return __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
  }
  // bfloat16 is the upper half of a float
  __device__ __forceinline__ static float cvt(raw r) { return __uint_as_float((unsigned)r << 16); }
  __device__ __forceinline__ static float ld(const __hip_bfloat16 *p) { return __bfloat162float(*p); }
  __device__ __forceinline__ static void st(__hip_bfloat16 *p, float v) { *p = __float2bfloat16(v); }
  __device__ __forceinline__ static float ld_nt(const __hip_bfloat16 *p) {
    // bfloat16 is the upper half of a float
    const unsigned short u = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p));
    return __uint_as_float((unsigned)u << 16);
  }
  __device__ __forceinline__ static void st_nt(__hip_bfloat16 *p, float v) {
    const __hip_bfloat16 b = __float2bfloat16(v);
    unsigned short u;
    __builtin_memcpy(&u, &b, 2);
    __builtin_nontemporal_store(u, reinterpret_cast<unsigned short *>(p));
  }
};

// non-temporal load of a model coefficient
template <typename T> __device__ __forceinline__ T ld_nt(const T &r) { return __builtin_nontemporal_load(&r); }

// Load/store through a uniform base pointer and a 32-bit byte offset: the
// compiler then addresses with a scalar base plus one 32-bit VGPR offset
// (valid while a component stays below 4 GiB).
template <typename T> __device__ __forceinline__ const T &at_b(const T *base, unsigned byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <typename T> __device__ __forceinline__ T &at_b(T *base, unsigned byte_off) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off);
}

// ---- the S3 combination (see Mode), one evaluation order for every kernel ----
// partial = w1 (U1 - W) + w2 (U2 - W); W^{n+1} = W + ((partial + w3 (U3 - W)) + b3 dt K3)
template <typename TC> __device__ __forceinline__ TC rk_partial(const TC *w, TC y0, TC u1, TC u2) {
  return w[0] * (u1 - y0) + w[1] * (u2 - y0);
}
template <typename TC> __device__ __forceinline__ TC rk_final(TC y0, TC partial, TC w3, TC u3, TC cbh) {
  return y0 + ((partial + w3 * (u3 - y0)) + cbh);
}

// ---- operator constants (src/forward/fd_t.cu:76-148) ----------------------
// DIR 0 = backward {-3,-2,-1,0,1}, DIR 1 = forward {-1,0,1,2,3}.
template <typename TC, int DIR> struct Drp;
template <typename TC> struct Drp<TC, 0> {
  static constexpr int lo = -3;
  __device__ __forceinline__ static TC c(int n) {
    return n == 0 ? TC(-0.041686285) : n == 1 ? TC(0.333411808) : n == 2 ? TC(-1.250117713)
         : n == 3 ? TC(0.666745142) : TC(0.291647047);
  }
};
template <typename TC> struct Drp<TC, 1> {
  static constexpr int lo = -1;
  __device__ __forceinline__ static TC c(int n) {
    return n == 0 ? TC(-0.291647047) : n == 1 ? TC(-0.666745142) : n == 2 ? TC(1.250117713)
         : n == 3 ? TC(-0.333411808) : TC(0.041686285);
  }
};

// ---- padded HBM layout --------------------------------------------------------
// Physical point (i,j,k), 0-based, lives at (k+3)*pz + (j+3)*px + (i+xa) within a
// component; xa >= 3 aligns i = 0 to 128 bytes, px is the padded row pitch.
struct Layout {
  int ni, nj, nk;
  int xa;
  long px, pz, cs;
  __host__ __device__ __forceinline__ long off(int i, int j, int k) const {
    return (long)(k + kNG) * pz + (long)(j + kNG) * px + (long)(i + xa);
  }
};

// ---- kernel argument blocks ------------------------------------------------------
template <typename TS, typename TM> struct FaceArgs {
  int on, dim;
  int i1, i2, j1, j2, k1, k2;   // physical inclusive box
  int bi, bj;                   // box extents along i and j
  long fsize;                   // points in the box
  long shot_stride;             // elements between shots of this face's buffer
  const TM *A, *B, *D;          // nlay+1 profile values
  // shot-0 level bases, component stride fsize: stage input, W^n, output,
  // U1 (S3 reads it and writes W^{n+1} over it) and U2 (read by S3)
  TS *a_in, *a_pre, *a_out, *a_end, *a_u2;
};

// boxes one per-point launch may cover (the boundary shell splits into at most 26)
constexpr int kMaxBox = 26;

template <typename TS, typename TM> struct StageArgs {
  Layout L;
  const TM *met;   // NMET components, stride L.cs
  const TM *med;   // lam, mu, 1/rho, stride L.cs
  const TM *mat;   // 18 components [c][j][i] on the physical top plane
  int free_top;
  int met_const;   // 1: one constant per metric component, off-diagonal ones zero (Cartesian grid)
  TM cmet[9];      // xi_x..zeta_z when met_const (off-diagonal entries exactly 0)
  FaceArgs<TS, TM> f[6];
  // shot-0 level bases: stage input, W^n, output, U1 (S3: read, then
  // overwritten by W^{n+1}) and U2 (S3)
  TS *w_in, *w_pre, *w_out, *w_end, *w_u2;
  long shot_stride;
  int mode;
  TM ca, cb;      // a_i dt and b_i dt of the stage
  TM rk_w[3];     // w1, w2, w3 of the S3 combination
  float sc[kNC], isc[kNC];  // storage scale (value = stored*sc) and its inverse
  int nbox;
  int box[kMaxBox][6];      // i1,i2,j1,j2,k1,k2 inclusive, physical
  long box_end[kMaxBox];    // cumulative point counts
  const int *shot_active;   // per-shot 0/1, or null
};

// Source contributions grouped by grid point (CSR), all shots of a session.
template <typename TS, typename TM> struct SrcArgs {
  Layout L;
  int npt;
  const int *pt_shot;       // [npt]
  const long *pt_off;       // [npt] layout offset
  const int *pt_beg;        // [npt+1] into contributions
  const int *ct_src;        // [nct] global source id
  const TM *ct_coef;        // [nct]
  const int *s_itb, *s_ite; // [nsrc_total]
  const long *s_tab;        // [nsrc_total] offset of the source's (it,stage) table
  const int *s_kind;        // [nsrc_total] bit0 force, bit1 moment
  const TM *F;              // [3][ntab]
  const TM *M;              // [6][ntab]
  long ntab;
  const TM *met;            // for jac
  const TM *med;            // for 1/rho
  TS *w_out, *w_end;   // stage output (S0..S2) / W^{n+1} (S3) of each shot
  long shot_stride;
  int mode, istage;
  TM ca, cb;
  const int *clock;         // current global step
  const int *shot_start;    // per-shot start step
  const int *shot_active;
  int only_shot;            // >= 0: inject this shot only, at step fixed_it
  int fixed_it;
  float isc[kNC];
};

template <typename TS, typename TC> struct RecvArgs {
  int nrec;                 // receivers over all shots
  const int *r_shot;
  const long *r_off;        // [nrec][8] layout offsets
  const TC *r_frac;         // [nrec][3]
  const long *r_out;        // [nrec] offset of the receiver's [9][pitch] block in seismo
  const int *r_nt;          // [nrec] samples kept
  int ring;                 // 0: pitch nt, sample it in column it; > 0: pitch ring, column it % ring
  TC *seismo;
  const TS *w_pre;
  long shot_stride, cs;
  const int *clock;
  const int *shot_start;
  const int *shot_active;
  float sc[kNC];
};

}  // namespace cgfd
