// cgfd_abi.cpp -- session management and the C ABI of the gfx950 el_iso path.
//
// The session plays the role of the device half of drv_rk_curv_col_allstep
// (src/forward/drv_rk_curv_col.cu:28-567): it owns the device copies of the
// metric, medium, free-surface matrices, PML profiles and memory variables,
// source tables and receivers (init_*_device, src/forward/alloc.cu:9-419),
// and runs the RK4 time loop with the reference's step/stage semantics.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cgfd_el_iso.h"
#include "cgfd_internal.h"
#include "el_iso_device.h"
#include "el_iso_launch.h"

using namespace cgfd;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return fail(CGFD_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// RK4 coefficients (src/forward/fd_t.cu:36-49) and operator pair flags (fd_t.cu:184-194)
const double kRkA[3] = {0.5, 0.5, 1.0};
const double kRkB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
const int kFlags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                          {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};

// host float -> half / bfloat16, round to nearest even
uint16_t f2h(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t mant = x & 0x7fffffu;
  int exp = (int)((x >> 23) & 0xff);
  if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0));
  int e = exp - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    mant |= 0x800000u;
    int shift = 14 - e;
    uint32_t half = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1), mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1))) half++;
    return (uint16_t)(sign | half);
  }
  uint32_t half = (uint32_t)(e << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (half & 1))) half++;
  return (uint16_t)(sign | half);
}
float h2f(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  int exp = (h >> 10) & 0x1f;
  uint32_t mant = h & 0x3ffu;
  uint32_t x;
  if (exp == 0) {
    if (mant == 0) x = sign;
    else {
      exp = 1;
      while (!(mant & 0x400u)) { mant <<= 1; exp--; }
      mant &= 0x3ffu;
      x = sign | (uint32_t)(exp + 127 - 15) << 23 | (mant << 13);
    }
  } else if (exp == 31) x = sign | 0x7f800000u | (mant << 13);
  else x = sign | (uint32_t)(exp + 127 - 15) << 23 | (mant << 13);
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}
uint16_t f2bf(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  if ((x & 0x7f800000u) == 0x7f800000u && (x & 0x7fffffu)) return (uint16_t)((x >> 16) | 0x40);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;
  return (uint16_t)(x >> 16);
}
float bf2f(uint16_t b) {
  uint32_t x = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

double rd(int dtype, const void *p, size_t i) {
  return dtype == CGFD_DT_F64 ? ((const double *)p)[i] : (double)((const float *)p)[i];
}

struct FaceHost {
  bool on = false;
  int nlay = 0;
  int i1 = 0, i2 = -1, j1 = 0, j2 = -1, k1 = 0, k2 = -1;
  long fsize = 0;
  void *coef = nullptr;   // A,B,D (model type)
  void *aux = nullptr;    // nshots x 5 levels x 9 x fsize (storage type)
};

struct SrcHost {
  int nsrc = 0, force_on = 0, moment_on = 0, max_nt = 0, max_ext = 0;
  std::vector<int> itb, ite, ext_num;
  std::vector<long long> ext_indx;
  std::vector<double> ext_coef, F, M;
};

struct RecvHost {
  int nrecv = 0, nt = 0;
  std::vector<long long> indx8;
  std::vector<double> frac;
};

}  // namespace

struct cgfd_session {
  cgfd_desc d{};
  int dev = 0;
  hipStream_t st = nullptr;
  Layout L{};
  size_t es = 4, ms = 4;   // storage / model element sizes
  long level_elems = 0;    // 9 * cs
  long shot_elems = 0;     // 4 * level_elems
  void *wav = nullptr;     // nshots x 4 levels
  void *rhs = nullptr;     // nshots x 1 level (lazy)
  void *met = nullptr, *med = nullptr, *mat = nullptr;
  bool have_met = false, have_med = false;
  bool met_const = false;  // all metric components constant (Cartesian grid)
  double cmet[9] = {0};
  int free_top = 0;
  FaceHost face[6];
  int slot[4] = {0, 1, 2, 3};  // physical level slot of W^n and of the stage values U1, U2, U3
  float sc[9], isc[9];
  std::vector<SrcHost> src;
  std::vector<RecvHost> rec;
  // device source CSR (grow-only buffers: capacities in bytes)
  int s_npt = 0;
  size_t s_cap[11] = {0};
  int *s_pt_shot = nullptr, *s_pt_beg = nullptr, *s_ct_src = nullptr, *s_itb = nullptr, *s_ite = nullptr,
      *s_kind = nullptr;
  long *s_pt_off = nullptr, *s_tab = nullptr;
  void *s_ct_coef = nullptr, *s_F = nullptr, *s_M = nullptr;
  long s_ntab = 0;
  bool src_dirty = false;
  // device receivers
  int r_n = 0;
  int *r_shot = nullptr, *r_nt = nullptr;
  long *r_off = nullptr, *r_out = nullptr;
  void *r_frac = nullptr, *seismo = nullptr;
  long seismo_elems = 0;
  std::vector<long> seismo_off;  // per shot: shot * seismo_stride
  long seismo_stride = 0;        // elements reserved per shot
  size_t r_cap[5] = {0};
  bool rec_dirty = false;
  // Seismogram ring (ring > 0): the device keeps `ring` samples per trace;
  // runs are cut into chunks of ring/2 steps and after each chunk its
  // samples are copied on st_copy into the pinned host record (per shot
  // [nrecv][9][nt] at host_off), overlapping the next chunk.  A chunk waits
  // for the copy of the chunk two back, whose ring columns it reuses.
  int ring = 0;
  hipStream_t st_copy = nullptr;
  hipEvent_t ev_chunk = nullptr;
  hipEvent_t ev_copied[2] = {nullptr, nullptr};
  long copied_chunk[2] = {-1, -1};
  void *host_seismo = nullptr;  // hipHostMalloc, nshots * host_stride samples
  long host_stride = 0;
  std::vector<int> start_h, active_h;  // host copies of shot_start / shot_active
  int *clock = nullptr, *shot_start = nullptr, *shot_active = nullptr;
  // interior tile choice (-1: not tuned yet)
  int tile_shape = -1, tile_zchunk = 48;
  bool tile_fixed = false;
  int split = -1;  // CGFD_SPLIT_*, chosen by the autotuner (-1: not yet: planes)
  bool split_fixed = false;  // cgfd_set_split pinned it
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev;
  hipEvent_t ev_run0 = nullptr, ev_run1 = nullptr;
  size_t ev_used = 0;
  // 8-step periods captured as graphs, one per orientation of the W/U1 slots;
  // any change of kernel arguments bumps the epoch and retires them
  bool use_graph = true;
  unsigned long epoch = 1;
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    unsigned long epoch = 0;
    double dt = 0;
    int slot0 = -1;
  } graph[2];
  std::vector<char> rec_changed;  // per shot: receivers replaced since the last rebuild
  void *gather = nullptr;         // device buffer of cgfd_fetch_subvolume
  void *peak = nullptr;           // peak ground motion maps, nshots x 18 x ni x nj (model type)
  size_t gather_bytes = 0;
  double last_run_ms = 0, last_int_avg_ms = 0;
  int last_int_n = 0;
  std::vector<void *> allocs;
};

namespace {

int dalloc(cgfd_session *s, void **p, size_t bytes) {
  if (bytes == 0) { *p = nullptr; return CGFD_OK; }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(CGFD_E_NOMEM, "insufficient HBM: hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  s->allocs.push_back(*p);
  return CGFD_OK;
}

void dfree(cgfd_session *s, void *&p) {
  if (!p) return;
  auto it = std::find(s->allocs.begin(), s->allocs.end(), p);
  if (it != s->allocs.end()) s->allocs.erase(it);
  (void)hipFree(p);
  p = nullptr;
}

// Device buffer of at least `need` bytes, reallocated only when it grows, so
// that rebuilding small tables between runs does not free memory (a free
// waits for the whole device, including other sessions' work).
int ensure(cgfd_session *s, void **p, size_t *cap, size_t need) {
  if (*p && need <= *cap) return CGFD_OK;
  if (*p) dfree(s, *p);
  *cap = 0;
  const size_t want = std::max<size_t>(need + need / 2, 256);
  int r = dalloc(s, p, want);
  if (r) return r;
  *cap = want;
  return CGFD_OK;
}

// stream-ordered copies; the session stream never waits on other streams
hipError_t h2d(cgfd_session *s, void *dst, const void *src, size_t n) {
  hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s->st);
  return e != hipSuccess ? e : hipStreamSynchronize(s->st);
}
hipError_t d2h(cgfd_session *s, void *dst, const void *src, size_t n) {
  hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s->st);
  return e != hipSuccess ? e : hipStreamSynchronize(s->st);
}

void drop_graphs(cgfd_session *s) {
  for (auto &g : s->graph) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = cgfd_session::GraphSlot{};
  }
}

// kernel arguments changed: captured periods no longer apply
void invalidate(cgfd_session *s) { s->epoch++; }

// the kernels address one component (or one level of a face's memory
// variables) with 32-bit byte offsets from a uniform base
constexpr unsigned long long kMaxComponentBytes = 1ull << 32;

Layout make_layout(int ni, int nj, int nk, size_t es) {
  Layout L;
  L.ni = ni; L.nj = nj; L.nk = nk;
  const int align = (int)(128 / es);  // elements per 128 bytes
  L.xa = align;                        // i = 0 at a 128-byte boundary, ghosts before it
  long need = (long)L.xa + ni + kNG;
  L.px = (need + align - 1) / align * align;
  L.pz = L.px * (nj + 2 * kNG);
  L.cs = L.pz * (nk + 2 * kNG);
  return L;
}

long ref_to_dev(const Layout &L, long long ref) {
  const long nx = L.ni + 2 * kNG, ny = L.nj + 2 * kNG;
  long ir = (long)(ref % nx);
  long t = (long)(ref / nx);
  long jr = t % ny, kr = t / ny;
  return kr * L.pz + jr * L.px + (ir - kNG + L.xa);
}

void face_box(const cgfd_session *s, int f, int nlay, FaceHost &F) {
  const int ni = s->d.ni, nj = s->d.nj, nk = s->d.nk;
  F.i1 = 0; F.i2 = ni - 1; F.j1 = 0; F.j2 = nj - 1; F.k1 = 0; F.k2 = nk - 1;
  switch (f) {
    case 0: F.i2 = nlay; break;
    case 1: F.i1 = ni - 1 - nlay; break;
    case 2: F.j2 = nlay; break;
    case 3: F.j1 = nj - 1 - nlay; break;
    case 4: F.k2 = nlay; break;
    default: F.k1 = nk - 1 - nlay; break;
  }
  F.fsize = (long)(F.i2 - F.i1 + 1) * (F.j2 - F.j1 + 1) * (F.k2 - F.k1 + 1);
}

// copy host model values (any dtype) into a device buffer of the model type
int put_model(cgfd_session *s, void *dst, const std::vector<double> &v) {
  if (s->ms == 8) {
    HIPCHK(h2d(s, dst, v.data(), v.size() * 8));
  } else {
    std::vector<float> f(v.size());
    for (size_t i = 0; i < v.size(); i++) f[i] = (float)v[i];
    HIPCHK(h2d(s, dst, f.data(), f.size() * 4));
  }
  return CGFD_OK;
}

// reference-layout component -> padded device-layout host buffer (double)
void ref_to_dev_comp(const Layout &L, int dtype, const void *src, size_t comp, std::vector<double> &out,
                     size_t out_base) {
  const long nx = L.ni + 2 * kNG, ny = L.nj + 2 * kNG, nz = L.nk + 2 * kNG;
  const size_t sc = (size_t)nx * ny * nz;
  for (long k = 0; k < nz; k++)
    for (long j = 0; j < ny; j++)
      for (long i = 0; i < nx; i++)
        out[out_base + k * L.pz + j * L.px + (i - kNG + L.xa)] = rd(dtype, src, comp * sc + (k * ny + j) * nx + i);
}

// storage-typed device write of n values given as double (scaled by inverse scale per component)
int put_storage(cgfd_session *s, void *dst, const std::vector<double> &v, long comp_stride) {
  const size_t n = v.size();
  if (s->d.storage == CGFD_STORE_F64) {
    HIPCHK(h2d(s, dst, v.data(), n * 8));
  } else if (s->d.storage == CGFD_STORE_F32) {
    std::vector<float> f(n);
    for (size_t i = 0; i < n; i++) f[i] = (float)v[i];
    HIPCHK(h2d(s, dst, f.data(), n * 4));
  } else {
    std::vector<uint16_t> h(n);
    const bool bf = s->d.storage == CGFD_STORE_BF16;
    for (size_t i = 0; i < n; i++) {
      int c = comp_stride > 0 ? (int)(i / comp_stride) % 9 : 0;
      float x = (float)v[i] * s->isc[c];
      h[i] = bf ? f2bf(x) : f2h(x);
    }
    HIPCHK(h2d(s, dst, h.data(), n * 2));
  }
  return CGFD_OK;
}

int get_storage(cgfd_session *s, const void *src, size_t n, long comp_stride, std::vector<double> &v) {
  v.resize(n);
  if (s->d.storage == CGFD_STORE_F64) {
    HIPCHK(d2h(s, v.data(), src, n * 8));
  } else if (s->d.storage == CGFD_STORE_F32) {
    std::vector<float> f(n);
    HIPCHK(d2h(s, f.data(), src, n * 4));
    for (size_t i = 0; i < n; i++) v[i] = f[i];
  } else {
    std::vector<uint16_t> h(n);
    HIPCHK(d2h(s, h.data(), src, n * 2));
    const bool bf = s->d.storage == CGFD_STORE_BF16;
    for (size_t i = 0; i < n; i++) {
      int c = comp_stride > 0 ? (int)(i / comp_stride) % 9 : 0;
      v[i] = (double)(bf ? bf2f(h[i]) : h2f(h[i])) * s->sc[c];
    }
  }
  return CGFD_OK;
}

void *wav_level(cgfd_session *s, int shot, int slot) {
  return (char *)s->wav + ((size_t)shot * s->shot_elems + (size_t)slot * s->level_elems) * s->es;
}

int fill_zero(cgfd_session *s, void *p, size_t elems, size_t esz) {
  hipError_t e;
  if (esz == 8) e = launch_fill<double>((double *)p, (long)elems, 0.0, s->st);
  else if (esz == 4) e = launch_fill<float>((float *)p, (long)elems, 0.0f, s->st);
  else e = launch_fill<uint16_t>((uint16_t *)p, (long)elems, (uint16_t)0, s->st);
  if (e != hipSuccess) return fail(CGFD_E_HIP, "fill kernel: %s", hipGetErrorString(e));
  return CGFD_OK;
}

// ---- seismogram ring: pinned host record and per-chunk copies ----
// The host record grows like the device region, keeps the samples of shots
// whose receivers did not change and zeroes the others; pending copies into
// it are drained first.
int host_record(cgfd_session *s, long hstride) {
  if (s->st_copy) HIPCHK(hipStreamSynchronize(s->st_copy));
  const int ns = s->d.nshots;
  const size_t ms = s->ms;
  if (hstride > s->host_stride || !s->host_seismo) {
    const long nst = std::max(hstride, s->host_stride);
    void *p = nullptr;
    if (nst > 0) {
      hipError_t e = hipHostMalloc(&p, (size_t)ns * nst * ms, hipHostMallocDefault);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(CGFD_E_NOMEM, "pinned seismogram record of %zu bytes: %s", (size_t)ns * nst * ms,
                    hipGetErrorString(e));
      }
      std::memset(p, 0, (size_t)ns * nst * ms);
      if (s->host_seismo)
        for (int shot = 0; shot < ns; shot++)
          if (!s->rec_changed[shot])
            std::memcpy((char *)p + (size_t)shot * nst * ms, (const char *)s->host_seismo + (size_t)shot * s->host_stride * ms,
                        (size_t)s->host_stride * ms);
    }
    if (s->host_seismo) (void)hipHostFree(s->host_seismo);
    s->host_seismo = p;
    s->host_stride = nst;
  } else {
    for (int shot = 0; shot < ns; shot++)
      if (s->rec_changed[shot])
        std::memset((char *)s->host_seismo + (size_t)shot * s->host_stride * ms, 0, (size_t)s->host_stride * ms);
  }
  return CGFD_OK;
}

// Copy the samples of steps [a, b) of every active shot from the device
// ring to the host record on st_copy: per shot one strided copy of its
// nrecv*9 traces (two where the window wraps round the ring).
int copy_window(cgfd_session *s, int a, int b) {
  const size_t ms = s->ms;
  for (int shot = 0; shot < s->d.nshots; shot++) {
    const RecvHost &h = s->rec[shot];
    if (h.nrecv == 0 || !s->active_h[shot]) continue;
    long t0 = std::max<long>((long)a - s->start_h[shot], 0);
    const long t1 = std::min<long>((long)b - s->start_h[shot], h.nt);
    while (t0 < t1) {
      const long col = t0 % s->ring, len = std::min<long>(t1 - t0, s->ring - col);
      const char *src = (const char *)s->seismo + ((size_t)s->seismo_off[shot] + col) * ms;
      char *dst = (char *)s->host_seismo + ((size_t)shot * s->host_stride + t0) * ms;
      HIPCHK(hipMemcpy2DAsync(dst, (size_t)h.nt * ms, src, (size_t)s->ring * ms, (size_t)len * ms,
                              (size_t)h.nrecv * kNC, hipMemcpyDeviceToHost, s->st_copy));
      t0 += len;
    }
  }
  return CGFD_OK;
}

// ---- rebuild device source CSR over all shots ----
template <typename TM> int build_sources(cgfd_session *s) {
  s->s_npt = 0;
  std::vector<int> itb, ite, kind;
  std::vector<long> tab;
  std::vector<TM> F[3], M[6];
  std::map<std::pair<int, long>, std::vector<std::pair<int, double>>> pts;
  long ntab = 0;
  int gsrc = 0;
  for (int shot = 0; shot < (int)s->src.size(); shot++) {
    const SrcHost &h = s->src[shot];
    for (int is = 0; is < h.nsrc; is++, gsrc++) {
      itb.push_back(h.itb[is]);
      ite.push_back(h.ite[is]);
      kind.push_back((h.force_on ? 1 : 0) | (h.moment_on ? 2 : 0));
      tab.push_back(ntab);
      const size_t per = (size_t)h.max_nt * 4;
      for (int c = 0; c < 3; c++)
        for (size_t q = 0; q < per; q++)
          F[c].push_back(h.force_on ? (TM)h.F[(size_t)c * h.nsrc * per + (size_t)is * per + q] : TM(0));
      for (int c = 0; c < 6; c++)
        for (size_t q = 0; q < per; q++)
          M[c].push_back(h.moment_on ? (TM)h.M[(size_t)c * h.nsrc * per + (size_t)is * per + q] : TM(0));
      ntab += (long)per;
      for (int e = 0; e < h.ext_num[is]; e++) {
        long long ref = h.ext_indx[(size_t)is * h.max_ext + e];
        long off = ref_to_dev(s->L, ref);
        pts[{shot, off}].push_back({gsrc, h.ext_coef[(size_t)is * h.max_ext + e]});
      }
    }
  }
  const int npt = (int)pts.size();
  s->s_ntab = ntab;
  invalidate(s);
  if (npt == 0) { s->src_dirty = false; return CGFD_OK; }
  std::vector<int> pt_shot, pt_beg, ct_src;
  std::vector<long> pt_off;
  std::vector<TM> ct_coef;
  for (auto &kv : pts) {
    pt_shot.push_back(kv.first.first);
    pt_off.push_back(kv.first.second);
    pt_beg.push_back((int)ct_src.size());
    for (auto &c : kv.second) { ct_src.push_back(c.first); ct_coef.push_back((TM)c.second); }
  }
  pt_beg.push_back((int)ct_src.size());
  int slot = 0;
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    int r = ensure(s, dst, &s->s_cap[slot++], bytes);
    if (r) return r;
    if (bytes) HIPCHK(h2d(s, *dst, src, bytes));
    return CGFD_OK;
  };
  int r;
  if ((r = up((void **)&s->s_pt_shot, pt_shot.data(), pt_shot.size() * 4))) return r;
  if ((r = up((void **)&s->s_pt_off, pt_off.data(), pt_off.size() * 8))) return r;
  if ((r = up((void **)&s->s_pt_beg, pt_beg.data(), pt_beg.size() * 4))) return r;
  if ((r = up((void **)&s->s_ct_src, ct_src.data(), ct_src.size() * 4))) return r;
  if ((r = up(&s->s_ct_coef, ct_coef.data(), ct_coef.size() * sizeof(TM)))) return r;
  if ((r = up((void **)&s->s_itb, itb.data(), itb.size() * 4))) return r;
  if ((r = up((void **)&s->s_ite, ite.data(), ite.size() * 4))) return r;
  if ((r = up((void **)&s->s_kind, kind.data(), kind.size() * 4))) return r;
  if ((r = up((void **)&s->s_tab, tab.data(), tab.size() * 8))) return r;
  std::vector<TM> Fall, Mall;
  for (int c = 0; c < 3; c++) Fall.insert(Fall.end(), F[c].begin(), F[c].end());
  for (int c = 0; c < 6; c++) Mall.insert(Mall.end(), M[c].begin(), M[c].end());
  if ((r = up(&s->s_F, Fall.data(), Fall.size() * sizeof(TM)))) return r;
  if ((r = up(&s->s_M, Mall.data(), Mall.size() * sizeof(TM)))) return r;
  s->s_npt = npt;
  s->src_dirty = false;
  return CGFD_OK;
}

// Rebuild the receiver tables of all shots.  Each shot owns a fixed-stride
// seismogram region; shots whose receivers did not change keep what they
// recorded, and while the stride suffices nothing is reallocated, so a shot
// can be given new receivers while the others keep running.
template <typename TC> int build_receivers(cgfd_session *s) {
  std::vector<int> shot_of, nt_of;
  std::vector<long> off, out;
  std::vector<TC> frac;
  long stride = 0, hstride = 0;
  auto pitch = [&](const RecvHost &h) { return (long)(s->ring > 0 ? s->ring : h.nt); };
  for (const RecvHost &h : s->rec) {
    stride = std::max(stride, (long)h.nrecv * kNC * pitch(h));
    hstride = std::max(hstride, (long)h.nrecv * kNC * h.nt);
  }
  int r;
  if (s->ring > 0 && (r = host_record(s, hstride))) return r;
  const bool grow = stride > s->seismo_stride || !s->seismo;
  const long new_stride = grow ? std::max(stride, s->seismo_stride) : s->seismo_stride;
  for (int shot = 0; shot < (int)s->rec.size(); shot++) {
    const RecvHost &h = s->rec[shot];
    long pos = (long)shot * new_stride;
    for (int r = 0; r < h.nrecv; r++) {
      shot_of.push_back(shot);
      nt_of.push_back(h.nt);
      for (int q = 0; q < 8; q++) off.push_back(ref_to_dev(s->L, h.indx8[(size_t)r * 8 + q]));
      for (int q = 0; q < 3; q++) frac.push_back((TC)h.frac[(size_t)r * 3 + q]);
      out.push_back(pos);
      pos += (long)kNC * pitch(h);
    }
  }
  s->r_n = (int)shot_of.size();
  invalidate(s);
  if (new_stride > 0 && grow) {
    void *old = s->seismo;
    const long old_stride = s->seismo_stride;
    s->seismo = nullptr;
    if ((r = dalloc(s, &s->seismo, (size_t)s->d.nshots * new_stride * sizeof(TC)))) return r;
    if ((r = fill_zero(s, s->seismo, (size_t)s->d.nshots * new_stride, sizeof(TC)))) return r;
    if (old) {
      for (int shot = 0; shot < (int)s->rec.size(); shot++) {
        const size_t n = (size_t)s->rec[shot].nrecv * kNC * pitch(s->rec[shot]);
        if (s->rec_changed[shot] || n == 0) continue;
        HIPCHK(hipMemcpyAsync((TC *)s->seismo + (size_t)shot * new_stride, (const TC *)old + (size_t)shot * old_stride,
                              n * sizeof(TC), hipMemcpyDeviceToDevice, s->st));
      }
      HIPCHK(hipStreamSynchronize(s->st));
      dfree(s, old);
    }
    s->seismo_stride = new_stride;
  } else if (new_stride > 0) {
    for (int shot = 0; shot < (int)s->rec.size(); shot++)
      if (s->rec_changed[shot] && (r = fill_zero(s, (TC *)s->seismo + (size_t)shot * new_stride,
                                                  (size_t)new_stride, sizeof(TC))))
        return r;
  }
  s->seismo_off.assign(s->rec.size(), 0);
  for (int shot = 0; shot < (int)s->rec.size(); shot++) s->seismo_off[shot] = (long)shot * s->seismo_stride;
  s->seismo_elems = (long)s->d.nshots * s->seismo_stride;
  std::fill(s->rec_changed.begin(), s->rec_changed.end(), 0);
  s->rec_dirty = false;
  if (s->r_n == 0) return CGFD_OK;
  int slot = 0;
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    int rr = ensure(s, dst, &s->r_cap[slot++], bytes);
    if (rr) return rr;
    HIPCHK(h2d(s, *dst, src, bytes));
    return CGFD_OK;
  };
  if ((r = up((void **)&s->r_shot, shot_of.data(), shot_of.size() * 4))) return r;
  if ((r = up((void **)&s->r_nt, nt_of.data(), nt_of.size() * 4))) return r;
  if ((r = up((void **)&s->r_off, off.data(), off.size() * 8))) return r;
  if ((r = up((void **)&s->r_out, out.data(), out.size() * 8))) return r;
  if ((r = up(&s->r_frac, frac.data(), frac.size() * sizeof(TC)))) return r;
  return CGFD_OK;
}

// ---- stage argument assembly ----
// Level slots of a stage: input, W^n, output, U1 (written with W^{n+1} by
// S3) and U2 (read by S3); rhs_mode sends the output to the RHS level.
template <typename TS, typename TC>
StageArgs<TS, TC> stage_args(cgfd_session *s, int slot_in, int slot_pre, int slot_out, int slot_end, int slot_u2,
                             bool rhs_mode) {
  StageArgs<TS, TC> a{};
  a.L = s->L;
  a.met = (const TC *)s->met;
  a.med = (const TC *)s->med;
  a.mat = (const TC *)s->mat;
  a.free_top = s->free_top;
  a.met_const = s->met_const ? 1 : 0;
  for (int c = 0; c < 9; c++) a.cmet[c] = (TC)s->cmet[c];
  const long aux_lvl_shot = 5;
  for (int f = 0; f < 6; f++) {
    FaceHost &F = s->face[f];
    FaceArgs<TS, TC> &A = a.f[f];
    A.on = F.on ? 1 : 0;
    A.dim = f / 2;
    if (!F.on) continue;
    A.i1 = F.i1; A.i2 = F.i2; A.j1 = F.j1; A.j2 = F.j2; A.k1 = F.k1; A.k2 = F.k2;
    A.bi = F.i2 - F.i1 + 1;
    A.bj = F.j2 - F.j1 + 1;
    A.fsize = F.fsize;
    A.shot_stride = aux_lvl_shot * kNC * F.fsize;
    const TC *cf = (const TC *)F.coef;
    A.A = cf;
    A.B = cf + (F.nlay + 1);
    A.D = cf + 2 * (F.nlay + 1);
    TS *base = (TS *)F.aux;
    auto lv = [&](int sl) { return base + (long)sl * kNC * F.fsize; };
    A.a_in = lv(slot_in);
    A.a_pre = lv(slot_pre);
    A.a_out = rhs_mode ? lv(4) : lv(slot_out);
    A.a_end = lv(slot_end);
    A.a_u2 = lv(slot_u2);
  }
  TS *w = (TS *)s->wav;
  a.w_in = w + (long)slot_in * s->level_elems;
  a.w_pre = w + (long)slot_pre * s->level_elems;
  a.w_out = rhs_mode ? (TS *)s->rhs : w + (long)slot_out * s->level_elems;
  a.w_end = w + (long)slot_end * s->level_elems;
  a.w_u2 = w + (long)slot_u2 * s->level_elems;
  a.shot_stride = rhs_mode ? 0 : s->shot_elems;
  // S3 weights of the stage values: w_i = b_{i-1} / a_{i-1}
  for (int q = 0; q < 3; q++) a.rk_w[q] = (TC)(kRkB[q] / kRkA[q]);
  for (int c = 0; c < kNC; c++) { a.sc[c] = s->sc[c]; a.isc[c] = s->isc[c]; }
  a.shot_active = s->shot_active;
  return a;
}

// Work split of one stage.  CGFD_SPLIT_PLANES: the z-marching kernel sweeps
// whole x-y planes from the bottom up to the last plane below the
// free-surface rows, applying the CFS-PML terms inside the face boxes, and one
// per-point launch takes the 4 free-surface rows.  CGFD_SPLIT_CORE: the same
// planes in two sweeps, the PML-free form over a core box inside every face
// box whose x extent starts and ends on tile boundaries (so its rows load
// whole 128-byte lines), and the PML form over the up to six rim boxes around
// it; the PML-free form needs far fewer registers (FP32 32x8 tile: 108 / 128
// / 164 VGPRs against 160 / 168 / 212 in S0 / S1-S2 / S3), so more waves
// hide its latency.  CGFD_SPLIT_SHELL (or CGFD_FLAG_PML_POINT): the sweep
// keeps to the box inside every face box and the per-point launch takes the
// whole shell (face slabs and surface rows); the per-point launch also
// covers everything for tiny grids or the plain method.  (Splitting that
// shell into per-axis single-face launches measured slower: the x slabs are
// only nlay+1 points wide, and four launches add four tails.)
struct BoxList {
  int n = 0;
  int box[kMaxBox][6];
  long end[kMaxBox];
  void add(int a0, int a1, int b0, int b1, int c0, int c1) {
    if (a1 < a0 || b1 < b0 || c1 < c0) return;
    const long prev = n ? end[n - 1] : 0;
    box[n][0] = a0; box[n][1] = a1; box[n][2] = b0; box[n][3] = b1; box[n][4] = c0; box[n][5] = c1;
    end[n] = prev + (long)(a1 - a0 + 1) * (b1 - b0 + 1) * (c1 - c0 + 1);
    n++;
  }
};

struct Shell {
  MarchSet inner;  // PML-free sweep (core or shell split) or the whole-plane sweep
  MarchSet rim;    // core split: PML-form sweep of the boxes around the core
  BoxList point;
};

// z chunk of the rim sweep: a third of the core's, at least 4 planes.  The
// rim is latency-bound (PML form, 2-3 waves per SIMD), and more, shorter
// workgroups balance its uneven work better than the extra priming planes
// cost: on 800x800x400 (core chunk 17) a step ran 1.1 % (FP32) / 1.4 % (FP16)
// faster than with the core's chunk, and 5 % / 7 % faster than with twice it.
int rim_zchunk(int core_zchunk) { return std::max(4, core_zchunk / 3); }

Shell make_shell(cgfd_session *s) {
  Shell sh{};
  const int ni = s->d.ni, nj = s->d.nj, nk = s->d.nk;
  auto lay = [&](int f) { return s->face[f].on ? s->face[f].nlay + 1 : 0; };
  const int ib0 = lay(0), ib1 = ni - 1 - lay(1);
  const int jb0 = lay(2), jb1 = nj - 1 - lay(3);
  const int kb0 = lay(4);
  const int ktopm = s->free_top ? nk - 5 : nk - 1;  // highest plane the z-march may sweep
  const int kb1 = ktopm - (s->free_top ? 0 : lay(5));
  const bool plain = (s->d.flags & CGFD_FLAG_PLAIN) != 0;
  auto addm = [](MarchSet &m, int a0, int a1, int b0, int b1, int c0, int c1) {
    if (a1 >= a0 && b1 >= b0 && c1 >= c0) {
      int q = m.nbox++;
      m.box[q][0] = a0; m.box[q][1] = a1; m.box[q][2] = b0; m.box[q][3] = b1; m.box[q][4] = c0; m.box[q][5] = c1;
    }
  };
  if (plain || ib1 < ib0 || jb1 < jb0 || kb1 < kb0) {
    sh.point.add(0, ni - 1, 0, nj - 1, 0, nk - 1);
    return sh;
  }
  sh.inner.zchunk = s->tile_zchunk;
  sh.inner.shape = s->tile_shape < 0 ? 0 : s->tile_shape;
  sh.rim.zchunk = rim_zchunk(sh.inner.zchunk);
  sh.rim.shape = sh.inner.shape;
  sh.rim.pml = 1;
  const bool any_face = ib0 > 0 || ib1 < ni - 1 || jb0 > 0 || jb1 < nj - 1 || kb0 > 0 || kb1 < ktopm;
  if (any_face && !(s->d.flags & CGFD_FLAG_PML_POINT) && s->split != CGFD_SPLIT_SHELL) {
    if (s->split == CGFD_SPLIT_CORE) {
      // core: x from the first tile boundary at or past the x1 face box to
      // the last one before the x2 face box, y likewise on the tile height
      int tx = 0, ty = 0;
      interior_shape(sh.inner.shape, &tx, &ty);
      const int ca0 = (ib0 + tx - 1) / tx * tx, ca1 = ca0 + (ib1 + 1 - ca0) / tx * tx - 1;
      const int cb0 = (jb0 + ty - 1) / ty * ty, cb1 = cb0 + (jb1 + 1 - cb0) / ty * ty - 1;
      if (ca1 >= ca0 && cb1 >= cb0) {
        addm(sh.inner, ca0, ca1, cb0, cb1, kb0, kb1);
        addm(sh.rim, 0, ni - 1, 0, cb0 - 1, 0, ktopm);          // y1 side, whole x
        addm(sh.rim, 0, ni - 1, cb1 + 1, nj - 1, 0, ktopm);     // y2 side, whole x
        addm(sh.rim, 0, ca0 - 1, cb0, cb1, 0, ktopm);           // x1 side
        addm(sh.rim, ca1 + 1, ni - 1, cb0, cb1, 0, ktopm);      // x2 side
        addm(sh.rim, ca0, ca1, cb0, cb1, 0, kb0 - 1);           // below the core
        addm(sh.rim, ca0, ca1, cb0, cb1, kb1 + 1, ktopm);       // above it (no free surface)
        sh.point.add(0, ni - 1, 0, nj - 1, ktopm + 1, nk - 1);
        return sh;
      }
    }
    sh.inner.pml = 1;
    addm(sh.inner, 0, ni - 1, 0, nj - 1, 0, ktopm);
    sh.point.add(0, ni - 1, 0, nj - 1, ktopm + 1, nk - 1);
    return sh;
  }
  addm(sh.inner, ib0, ib1, jb0, jb1, kb0, kb1);
  sh.point.add(0, ib0 - 1, 0, nj - 1, 0, nk - 1);
  sh.point.add(ib1 + 1, ni - 1, 0, nj - 1, 0, nk - 1);
  sh.point.add(ib0, ib1, 0, jb0 - 1, 0, nk - 1);
  sh.point.add(ib0, ib1, jb1 + 1, nj - 1, 0, nk - 1);
  sh.point.add(ib0, ib1, jb0, jb1, 0, kb0 - 1);
  sh.point.add(ib0, ib1, jb0, jb1, kb1 + 1, nk - 1);
  return sh;
}

template <typename TS, typename TC> void set_boxes(StageArgs<TS, TC> &a, const BoxList &l) {
  a.nbox = l.n;
  for (int b = 0; b < l.n; b++) {
    for (int q = 0; q < 6; q++) a.box[b][q] = l.box[b][q];
    a.box_end[b] = l.end[b];
  }
}

// one stage over the whole grid: z-marching sweep (timed when requested), then
// the per-point boxes
template <typename TS, typename TC>
int launch_stage(cgfd_session *s, const Shell &sh, StageArgs<TS, TC> &a, int dx, int dy, int dz, int nshots,
                 hipEvent_t ev0, hipEvent_t ev1) {
  if (sh.inner.nbox) {
    if (ev0) HIPCHK(hipEventRecord(ev0, s->st));
    hipError_t e = launch_interior<TS, TC>(a, sh.inner, dx, dy, dz, nshots, s->st);
    if (e != hipSuccess) return fail(CGFD_E_HIP, "interior kernel: %s", hipGetErrorString(e));
    if (sh.rim.nbox) {
      e = launch_interior<TS, TC>(a, sh.rim, dx, dy, dz, nshots, s->st);
      if (e != hipSuccess) return fail(CGFD_E_HIP, "rim kernel: %s", hipGetErrorString(e));
    }
    if (ev1) HIPCHK(hipEventRecord(ev1, s->st));
  }
  if (sh.point.n) {
    set_boxes(a, sh.point);
    hipError_t e = launch_point<TS, TC>(a, dx, dy, dz, nshots, s->st);
    if (e != hipSuccess) return fail(CGFD_E_HIP, "point kernel: %s", hipGetErrorString(e));
  }
  return CGFD_OK;
}

// Tile autotuner: per (shape, z chunk), time the interior sweeps of one full
// step's stage sequence (S0, S12, S12, S3 with the operator directions of
// pair 0) and keep the fastest; the stages differ in traffic (30 to 57 words
// per cell), so tuning on one mode alone can pick a tile that loses on the
// others.  The RK coefficients are zero, and every level the sweeps write
// (U1, U2, U3) is rewritten by the next step before it is read, so
// tuning between steps leaves the state intact.  Each work split (core,
// whole planes, shell) is timed with its best tile and the fastest pair is
// kept.  Results are cached per (grid, storage, shots, flags, metric form,
// PML layout, pinned split) for the life of the process.
struct TileKey {
  int ni, nj, nk, storage, nshots, flags, met_const, free_top;
  std::array<int, 6> nlay;  // PML layers per face (0: none)
  int pinned;               // split pinned by cgfd_set_split, -1: automatic
  bool operator<(const TileKey &o) const {
    return std::tie(ni, nj, nk, storage, nshots, flags, met_const, free_top, nlay, pinned) <
           std::tie(o.ni, o.nj, o.nk, o.storage, o.nshots, o.flags, o.met_const, o.free_top, o.nlay, o.pinned);
  }
};
struct TileChoice {
  int shape, zchunk, split;
};
// shared by every session of the process; sessions may tune on several
// threads at once (the shot server's worker and callers of their own)
std::map<TileKey, TileChoice> g_tile_cache;
std::mutex g_tile_mu;

template <typename TS, typename TC> int autotune_t(cgfd_session *s) {
  TileKey key{s->d.ni, s->d.nj, s->d.nk, s->d.storage, s->d.nshots, s->d.flags, s->met_const ? 1 : 0, s->free_top,
              {}, s->split_fixed ? s->split : -1};
  for (int f = 0; f < 6; f++) key.nlay[f] = s->face[f].on ? s->face[f].nlay : 0;
  {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    auto hit = g_tile_cache.find(key);
    if (hit != g_tile_cache.end()) {
      s->tile_shape = hit->second.shape;
      s->tile_zchunk = hit->second.zchunk;
      s->split = hit->second.split;
      return CGFD_OK;
    }
  }
  auto remember = [&](TileChoice c) {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_cache[key] = c;
  };
  const int pinned = s->split_fixed ? s->split : -1;
  s->tile_shape = 0;
  s->split = CGFD_SPLIT_PLANES;
  const Shell planes = make_shell(s);
  if (!planes.inner.nbox) {  // no sweep at all (plain method, tiny grid)
    if (pinned >= 0) s->split = pinned;
    remember({s->tile_shape, s->tile_zchunk, s->split});
    return CGFD_OK;
  }
  const int PRE = s->slot[0], U1 = s->slot[1], U2 = s->slot[2], U3 = s->slot[3];
  const int sin[4] = {PRE, U1, U2, U3}, sout[4] = {U1, U2, U3, U1};
  const int smode[4] = {MODE_S0, MODE_S12, MODE_S12, MODE_S3};
  StageArgs<TS, TC> sa[4];
  for (int q = 0; q < 4; q++) {
    sa[q] = stage_args<TS, TC>(s, sin[q], PRE, sout[q], U1, U2, false);
    sa[q].mode = smode[q];
    sa[q].ca = TC(0);
    sa[q].cb = TC(0);
    // time every shot slot, active or not: a batched session is tuned for
    // its full batch even when its first run has one shot active (a server
    // group warmed by one request).  The stages write only the stage-value
    // levels of each slot; W^n (and its PML memory) is read, never written.
    sa[q].shot_active = nullptr;
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // minimum over reps 1..2 of one full step's four stages of `launch`
  auto timed = [&](auto launch, float &tmin) -> int {
    tmin = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      HIPCHK(hipEventRecord(e0, s->st));
      for (int q = 0; q < 4; q++) {
        const int d = (kFlags[0][0] + q) % 2;
        hipError_t e = launch(q, d);
        if (e != hipSuccess) return fail(CGFD_E_HIP, "autotune launch: %s", hipGetErrorString(e));
      }
      HIPCHK(hipEventRecord(e1, s->st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) tmin = std::min(tmin, ms);
    }
    return CGFD_OK;
  };
  // Work splits: whole-plane sweep (PML terms inside) and, when the grid has
  // PML faces, the PML-free core with the PML-form rim, and the sweep inside
  // the face boxes with the per-point kernel on the slabs.  Each is timed
  // with its best tile plus its per-point launch (a pinned split alone).
  const bool faces = planes.inner.pml && !(s->d.flags & CGFD_FLAG_PML_POINT);
  std::vector<int> splits;
  if (pinned >= 0) splits.push_back(faces ? pinned : CGFD_SPLIT_PLANES);
  else if (faces) splits = {CGFD_SPLIT_CORE, CGFD_SPLIT_PLANES, CGFD_SPLIT_SHELL};
  else splits = {CGFD_SPLIT_PLANES};
  float best_total = 1e30f;
  int bs = 0, bz = 24, bsplit = splits[0];
  for (int split : splits) {
    s->split = split;
    s->tile_shape = 0;
    const Shell sh = make_shell(s);
    const int nz = sh.inner.box[0][5] - sh.inner.box[0][4] + 1;
    // z chunks near 16..48 planes that split the swept planes evenly: every
    // chunk re-reads 4 planes below it, so a short last chunk costs almost a
    // full one (196 planes: 17, 20, 25, 28, 33, 49 instead of 16 .. 48,
    // whose last chunk has 4 planes)
    constexpr int kNZ = 8;
    const int ztarget[kNZ] = {16, 20, 24, 28, 32, 48, 64, 96};
    int zcands[kNZ];
    for (int zi = 0; zi < kNZ; zi++) {
      const int n = std::max(1, (nz + ztarget[zi] / 2) / ztarget[zi]);
      zcands[zi] = (nz + n - 1) / n;
    }
    float best = 1e30f;
    int ts = 0, tz = 24;
    for (int shape = 0; shape < interior_num_shapes(); shape++) {
      s->tile_shape = shape;
      const Shell shs = make_shell(s);  // the core's bounds follow the tile shape
      for (int zi = 0; zi < kNZ; zi++) {
        const int zc = zcands[zi];
        if (zi > 0 && zc == zcands[zi - 1]) continue;
        MarchSet ib = shs.inner, rb = shs.rim;
        ib.zchunk = zc;
        rb.zchunk = rim_zchunk(zc);
        float t;
        int r = timed(
            [&](int q, int d) {
              hipError_t e = launch_interior<TS, TC>(sa[q], ib, d, d, d, s->d.nshots, s->st);
              if (e == hipSuccess && rb.nbox) e = launch_interior<TS, TC>(sa[q], rb, d, d, d, s->d.nshots, s->st);
              return e;
            },
            t);
        if (r) return r;
        if (t < best) { best = t; ts = shape; tz = zc; }
      }
    }
    float tp = 0;
    if (sh.point.n) {
      int r = timed(
          [&](int q, int d) {
            StageArgs<TS, TC> pa = sa[q];
            set_boxes(pa, sh.point);
            return launch_point<TS, TC>(pa, d, d, d, s->d.nshots, s->st);
          },
          tp);
      if (r) return r;
    }
    if (best + tp < best_total) { best_total = best + tp; bs = ts; bz = tz; bsplit = split; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->tile_shape = bs;
  s->tile_zchunk = bz;
  s->split = bsplit;
  remember({bs, bz, bsplit});
  invalidate(s);
  return CGFD_OK;
}

// Enqueue steps it0 .. it0+nsteps-1 of the RK4 loop on the session stream:
// receivers, four stages (interior sweep, PML slabs / free-surface rows,
// source) and the clock advance per step, then the W <-> U1 swap.
template <typename TS, typename TC>
int enqueue_steps(cgfd_session *s, const Shell &sh, int it0, int nsteps, TC dt, bool timed) {
  for (int n = 0; n < nsteps; n++) {
    const int it = it0 + n;
    // receivers read W^n before the step (drv_rk_curv_col.cu:195-216)
    if (s->r_n > 0) {
      RecvArgs<TS, TC> ra{};
      ra.nrec = s->r_n; ra.r_shot = s->r_shot; ra.r_off = s->r_off; ra.r_frac = (const TC *)s->r_frac;
      ra.r_out = s->r_out; ra.r_nt = s->r_nt; ra.seismo = (TC *)s->seismo; ra.ring = s->ring;
      ra.w_pre = (const TS *)s->wav + (long)s->slot[0] * s->level_elems;
      ra.shot_stride = s->shot_elems; ra.cs = s->L.cs;
      ra.clock = s->clock; ra.shot_start = s->shot_start; ra.shot_active = s->shot_active;
      for (int c = 0; c < kNC; c++) ra.sc[c] = s->sc[c];
      hipError_t e = launch_recv<TS, TC>(ra, s->st);
      if (e != hipSuccess) return fail(CGFD_E_HIP, "receiver kernel: %s", hipGetErrorString(e));
    }
    const int ipair = it % 8;
    // stage values U1..U3; S3 writes W^{n+1} over U1, which then becomes W
    const int PRE = s->slot[0], U1 = s->slot[1], U2 = s->slot[2], U3 = s->slot[3];
    for (int istage = 0; istage < 4; istage++) {
      const int dx = (kFlags[ipair][0] + istage) % 2;
      const int dy = (kFlags[ipair][1] + istage) % 2;
      const int dz = (kFlags[ipair][2] + istage) % 2;
      int in, out, mode;
      if (istage == 0) { in = PRE; out = U1; mode = MODE_S0; }
      else if (istage == 1) { in = U1; out = U2; mode = MODE_S12; }
      else if (istage == 2) { in = U2; out = U3; mode = MODE_S12; }
      else { in = U3; out = U1; mode = MODE_S3; }
      StageArgs<TS, TC> a = stage_args<TS, TC>(s, in, PRE, out, U1, U2, false);
      a.mode = mode;
      // coef = rk_a[istage]*dt, both float in the reference (drv_rk_curv_col.cu:341-342)
      a.ca = (istage < 3) ? (TC)kRkA[istage] * dt : TC(0);
      a.cb = (TC)kRkB[istage] * dt;
      const bool tim = timed && sh.inner.nbox;
      if (tim && s->ev.size() < s->ev_used + 2) {
        s->ev.resize(s->ev_used + 2);
        for (size_t q = s->ev_used; q < s->ev.size(); q++) HIPCHK(hipEventCreate(&s->ev[q]));
      }
      int r = launch_stage<TS, TC>(s, sh, a, dx, dy, dz, s->d.nshots, tim ? s->ev[s->ev_used] : nullptr,
                                   tim ? s->ev[s->ev_used + 1] : nullptr);
      if (r) return r;
      if (tim) s->ev_used += 2;
      if (s->s_npt > 0) {
        SrcArgs<TS, TC> sa{};
        sa.L = s->L; sa.npt = s->s_npt; sa.pt_shot = s->s_pt_shot; sa.pt_off = s->s_pt_off; sa.pt_beg = s->s_pt_beg;
        sa.ct_src = s->s_ct_src; sa.ct_coef = (const TC *)s->s_ct_coef; sa.s_itb = s->s_itb; sa.s_ite = s->s_ite;
        sa.s_tab = s->s_tab; sa.s_kind = s->s_kind; sa.F = (const TC *)s->s_F; sa.M = (const TC *)s->s_M;
        sa.ntab = s->s_ntab; sa.met = (const TC *)s->met; sa.med = (const TC *)s->med;
        sa.w_out = a.w_out; sa.w_end = a.w_end; sa.shot_stride = s->shot_elems;
        sa.mode = mode; sa.istage = istage; sa.ca = a.ca; sa.cb = a.cb;
        sa.clock = s->clock; sa.shot_start = s->shot_start; sa.shot_active = s->shot_active;
        sa.only_shot = -1; sa.fixed_it = 0;
        for (int c = 0; c < kNC; c++) sa.isc[c] = s->isc[c];
        hipError_t e = launch_source<TS, TC>(sa, s->st);
        if (e != hipSuccess) return fail(CGFD_E_HIP, "source kernel: %s", hipGetErrorString(e));
      }
    }
    if (s->peak && s->free_top) {
      // W^{n+1} sits in U1's slot before the swap (drv_rk_curv_col.cu:519-527)
      const TS *wn = (const TS *)wav_level(s, 0, U1), *wo = (const TS *)wav_level(s, 0, PRE);
      hipError_t e = launch_peak<TS, TC>(wn, wo, s->shot_elems, s->L, (TC *)s->peak, dt, s->sc, s->shot_active,
                                         s->d.nshots, s->st);
      if (e != hipSuccess) return fail(CGFD_E_HIP, "peak motion kernel: %s", hipGetErrorString(e));
    }
    hipError_t e = launch_clock_advance(s->clock, s->st);
    if (e != hipSuccess) return fail(CGFD_E_HIP, "clock kernel: %s", hipGetErrorString(e));
    std::swap(s->slot[0], s->slot[1]);  // W^{n+1} becomes W (drv_rk_curv_col.cu:529-541)
  }
  return CGFD_OK;
}

// The graph of one 8-step period (it0 % 8 == 0) for the current W/U1
// orientation; captured on first use.  Eight swaps bring the slots back, and
// the step number lives in device memory, so one capture serves every period.
template <typename TS, typename TC>
int period_graph(cgfd_session *s, const Shell &sh, TC dt, double dt_key, hipGraphExec_t *out) {
  cgfd_session::GraphSlot &g = s->graph[s->slot[0] == 0 ? 0 : 1];
  if (g.exec && g.epoch == s->epoch && g.dt == dt_key && g.slot0 == s->slot[0]) {
    *out = g.exec;
    return CGFD_OK;
  }
  if (g.exec) { (void)hipGraphExecDestroy(g.exec); g.exec = nullptr; }
  const int slot0 = s->slot[0];
  HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
  int r = enqueue_steps<TS, TC>(s, sh, 0, 8, dt, false);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s->st, &graph);
  if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
  if (e != hipSuccess) return fail(CGFD_E_HIP, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) { g.exec = nullptr; return fail(CGFD_E_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
  g.epoch = s->epoch; g.dt = dt_key; g.slot0 = slot0;
  *out = g.exec;
  return CGFD_OK;
}

// Enqueue a run without waiting for it.  Whole 8-step periods go through the
// captured graphs unless per-launch timing is on.
template <typename TS, typename TC> int enqueue_t(cgfd_session *s, int it0, int nsteps, double dt_in) {
  if (!s->tile_fixed && s->tile_shape < 0) {
    int r = autotune_t<TS, TC>(s);
    if (r) return r;
  }
  if (s->src_dirty) { int r = build_sources<TC>(s); if (r) return r; }
  if (s->rec_dirty) { int r = build_receivers<TC>(s); if (r) return r; }
  const Shell sh = make_shell(s);
  const TC dt = (TC)dt_in;
  hipError_t e = launch_fill<int>(s->clock, 1, it0, s->st);
  if (e != hipSuccess) return fail(CGFD_E_HIP, "clock set: %s", hipGetErrorString(e));
  s->ev_used = 0;
  const bool graphs = s->use_graph && !s->timing;
  auto steps = [&](int it, int left) -> int {
    while (left > 0) {
      if (graphs && it % 8 == 0 && left >= 8) {
        hipGraphExec_t g = nullptr;
        int r = period_graph<TS, TC>(s, sh, dt, dt_in, &g);
        if (r) return r;
        HIPCHK(hipGraphLaunch(g, s->st));
        it += 8; left -= 8;
      } else {
        const int n = graphs ? std::min(left, 8 - it % 8) : left;
        int r = enqueue_steps<TS, TC>(s, sh, it, n, dt, s->timing);
        if (r) return r;
        it += n; left -= n;
      }
    }
    return CGFD_OK;
  };
  if (s->ring <= 0 || s->r_n == 0) return steps(it0, nsteps);
  // ring: chunks of ring/2 steps at multiples of ring/2 (whole 8-step
  // periods, as ring is a multiple of 16); chunk q reuses the ring columns
  // of chunk q-2 and so waits for that chunk's copy
  const int C = s->ring / 2;
  int it = it0, left = nsteps;
  while (left > 0) {
    const long q = it / C;
    const int n = (int)std::min<long>(left, (q + 1) * C - it);
    const int par = (int)(q & 1);
    if (s->copied_chunk[par] >= 0 && s->copied_chunk[par] != q)
      HIPCHK(hipStreamWaitEvent(s->st, s->ev_copied[par], 0));
    int r = steps(it, n);
    if (r) return r;
    HIPCHK(hipEventRecord(s->ev_chunk, s->st));
    HIPCHK(hipStreamWaitEvent(s->st_copy, s->ev_chunk, 0));
    if ((r = copy_window(s, it, it + n))) return r;
    HIPCHK(hipEventRecord(s->ev_copied[par], s->st_copy));
    s->copied_chunk[par] = q;
    it += n; left -= n;
  }
  return CGFD_OK;
}

template <typename TS, typename TC> int run_t(cgfd_session *s, int it0, int nsteps, double dt_in) {
  if (!s->ev_run0) { HIPCHK(hipEventCreate(&s->ev_run0)); HIPCHK(hipEventCreate(&s->ev_run1)); }
  HIPCHK(hipEventRecord(s->ev_run0, s->st));
  int r = enqueue_t<TS, TC>(s, it0, nsteps, dt_in);
  if (r) { (void)hipStreamSynchronize(s->st); return r; }
  HIPCHK(hipEventRecord(s->ev_run1, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  if (s->st_copy) HIPCHK(hipStreamSynchronize(s->st_copy));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, s->ev_run0, s->ev_run1));
  s->last_run_ms = ms;
  double tot = 0;
  for (size_t q = 0; q + 1 < s->ev_used; q += 2) {
    float m = 0;
    HIPCHK(hipEventElapsedTime(&m, s->ev[q], s->ev[q + 1]));
    tot += m;
  }
  s->last_int_n = (int)(s->ev_used / 2);
  s->last_int_avg_ms = s->last_int_n ? tot / s->last_int_n : 0.0;
  return CGFD_OK;
}

template <typename TS, typename TC> int onestage_t(cgfd_session *s, int shot, int in_level, int ipair, int istage, int it) {
  if (s->src_dirty) { int r = build_sources<TC>(s); if (r) return r; }
  if (!s->rhs) {
    int r = dalloc(s, &s->rhs, (size_t)s->level_elems * s->es);
    if (r) return r;
    r = fill_zero(s, s->rhs, s->level_elems, s->es);
    if (r) return r;
  }
  const int dx = (kFlags[ipair][0] + istage) % 2;
  const int dy = (kFlags[ipair][1] + istage) % 2;
  const int dz = (kFlags[ipair][2] + istage) % 2;
  const int slot_in = s->slot[in_level];
  StageArgs<TS, TC> a = stage_args<TS, TC>(s, slot_in, s->slot[0], slot_in, s->slot[1], s->slot[2], true);
  // select the shot by offsetting the bases (RHS level holds a single shot)
  const long so = (long)shot * s->shot_elems;
  a.w_in += so; a.w_pre += so; a.w_end += so;
  for (int f = 0; f < 6; f++) {
    if (!a.f[f].on) continue;
    const long fo = (long)shot * a.f[f].shot_stride;
    a.f[f].a_in += fo; a.f[f].a_pre += fo; a.f[f].a_out += fo; a.f[f].a_end += fo;
    a.f[f].shot_stride = 0;
  }
  a.shot_stride = 0;
  a.shot_active = nullptr;
  a.mode = MODE_RHS;
  a.ca = TC(0); a.cb = TC(0);
  const Shell sh = make_shell(s);
  int r = launch_stage<TS, TC>(s, sh, a, dx, dy, dz, 1, nullptr, nullptr);
  if (r) return r;
  if (s->s_npt > 0) {
    // the source kernel scans all shots' points; inject this shot's at step it
    SrcArgs<TS, TC> sa{};
    sa.L = s->L; sa.npt = s->s_npt; sa.pt_shot = s->s_pt_shot; sa.pt_off = s->s_pt_off; sa.pt_beg = s->s_pt_beg;
    sa.ct_src = s->s_ct_src; sa.ct_coef = (const TC *)s->s_ct_coef; sa.s_itb = s->s_itb; sa.s_ite = s->s_ite;
    sa.s_tab = s->s_tab; sa.s_kind = s->s_kind; sa.F = (const TC *)s->s_F; sa.M = (const TC *)s->s_M;
    sa.ntab = s->s_ntab; sa.met = (const TC *)s->met; sa.med = (const TC *)s->med;
    sa.w_out = (TS *)s->rhs; sa.w_end = (TS *)s->rhs;
    sa.shot_stride = 0;
    sa.mode = MODE_RHS; sa.istage = istage; sa.ca = TC(0); sa.cb = TC(0);
    sa.clock = s->clock; sa.shot_start = s->shot_start; sa.shot_active = s->shot_active;
    sa.only_shot = shot; sa.fixed_it = it;
    for (int c = 0; c < kNC; c++) sa.isc[c] = s->isc[c];
    hipError_t e = launch_source<TS, TC>(sa, s->st);
    if (e != hipSuccess) return fail(CGFD_E_HIP, "source kernel: %s", hipGetErrorString(e));
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return CGFD_OK;
}

}  // namespace

void cgfd_internal_set_error(const char *msg) { g_err = msg ? msg : ""; }

// =====================================================================================
extern "C" {

const char *cgfd_last_error(void) { return g_err.c_str(); }

long long cgfd_footprint_bytes(const cgfd_desc *d, int nrecv, int nt_total) {
  if (!d || d->ni <= 0 || d->nj <= 0 || d->nk <= 0) return -1;
  size_t es = d->storage == CGFD_STORE_F64 ? 8 : d->storage == CGFD_STORE_F32 ? 4 : 2;
  size_t ms = d->storage == CGFD_STORE_F64 ? 8 : 4;
  // model arrays share the wavefield's layout (padded for the storage type)
  Layout L = make_layout(d->ni, d->nj, d->nk, es);
  long long ns = d->nshots > 0 ? d->nshots : 1;
  long long wav = ns * 4LL * kNC * L.cs * (long long)es;
  long long model = (long long)(NMET + 3) * L.cs * (long long)ms + 18LL * d->ni * d->nj * (long long)ms;
  // peak-motion maps (when enabled), the sub-volume gather buffer (at most
  // one component) and the source tables (7^3 points, 9 tables of 4 stage
  // values per step, per shot)
  long long extra = ns * 18LL * d->ni * d->nj * (long long)ms + L.cs * (long long)ms +
                    ns * (343LL * 24 + 9LL * 4 * (nt_total > 0 ? nt_total : 0) * (long long)ms);
  // PML memory: 5 faces of 11 layers at most is bounded by 5 levels over the shell
  long long shell = 2LL * 11 * d->nj * d->nk + 2LL * 11 * d->ni * d->nk + 2LL * 11 * d->ni * d->nj;
  long long aux = ns * 5LL * kNC * shell * (long long)es;
  long long rec = ns * (long long)nrecv * kNC * (long long)(nt_total > 0 ? nt_total : 0) * (long long)ms;
  return wav + model + aux + rec + extra;
}

int cgfd_create(const cgfd_desc *d, cgfd_session **out) {
  if (!d || !out) return fail(CGFD_E_ARG, "null descriptor or output pointer");
  *out = nullptr;
  if (d->ni < 1 || d->nj < 1 || d->nk < 1) return fail(CGFD_E_ARG, "grid %d x %d x %d invalid", d->ni, d->nj, d->nk);
  if (d->storage < 0 || d->storage > 3) return fail(CGFD_E_ARG, "storage %d invalid", d->storage);
  if (d->nshots < 1) return fail(CGFD_E_ARG, "nshots %d invalid", d->nshots);
  {
    // kernels address inside one component with 32-bit byte offsets
    const size_t es = d->storage == CGFD_STORE_F64 ? 8 : d->storage == CGFD_STORE_F32 ? 4 : 2;
    const size_t ms = d->storage == CGFD_STORE_F64 ? 8 : 4;
    const Layout L = make_layout(d->ni, d->nj, d->nk, es);
    if ((unsigned long long)L.cs * std::max(es, ms) >= kMaxComponentBytes)
      return fail(CGFD_E_ARG, "grid %d x %d x %d: one component needs %llu bytes, over the 4 GiB kernels address",
                  d->ni, d->nj, d->nk, (unsigned long long)L.cs * std::max(es, ms));
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return fail(CGFD_E_NODEV, "no HIP device available");
  }
  cgfd_session *s = new cgfd_session();
  s->d = *d;
  HIPCHK(hipGetDevice(&s->dev));
  s->es = d->storage == CGFD_STORE_F64 ? 8 : d->storage == CGFD_STORE_F32 ? 4 : 2;
  s->ms = d->storage == CGFD_STORE_F64 ? 8 : 4;
  // wavefield and model share one geometry; aligning i = 0 to 128 bytes in
  // storage elements also aligns the (equal or wider) model elements
  s->L = make_layout(d->ni, d->nj, d->nk, s->es);
  s->level_elems = (long)kNC * s->L.cs;
  s->shot_elems = 4 * s->level_elems;
  for (int c = 0; c < 9; c++) { s->sc[c] = 1.0f; s->isc[c] = 1.0f; }
  auto cleanup = [&](int r) {
    cgfd_destroy(s);
    return r;
  };
  if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(CGFD_E_HIP, "hipStreamCreate failed"));
  int r;
  if ((r = dalloc(s, &s->wav, (size_t)d->nshots * s->shot_elems * s->es))) return cleanup(r);
  if ((r = dalloc(s, &s->met, (size_t)NMET * s->L.cs * s->ms))) return cleanup(r);
  if ((r = dalloc(s, &s->med, (size_t)3 * s->L.cs * s->ms))) return cleanup(r);
  if ((r = dalloc(s, &s->mat, (size_t)18 * d->ni * d->nj * s->ms))) return cleanup(r);
  if ((r = dalloc(s, (void **)&s->clock, sizeof(int)))) return cleanup(r);
  if ((r = dalloc(s, (void **)&s->shot_start, sizeof(int) * d->nshots))) return cleanup(r);
  if ((r = dalloc(s, (void **)&s->shot_active, sizeof(int) * d->nshots))) return cleanup(r);
  if ((r = fill_zero(s, s->wav, (size_t)d->nshots * s->shot_elems, s->es))) return cleanup(r);
  if ((r = fill_zero(s, s->met, (size_t)NMET * s->L.cs, s->ms))) return cleanup(r);
  if ((r = fill_zero(s, s->med, (size_t)3 * s->L.cs, s->ms))) return cleanup(r);
  if ((r = fill_zero(s, s->mat, (size_t)18 * d->ni * d->nj, s->ms))) return cleanup(r);
  if (launch_fill<int>(s->shot_start, d->nshots, 0, s->st) != hipSuccess ||
      launch_fill<int>(s->shot_active, d->nshots, 1, s->st) != hipSuccess ||
      launch_fill<int>(s->clock, 1, 0, s->st) != hipSuccess)
    return cleanup(fail(CGFD_E_HIP, "initial fills failed"));
  s->src.resize(d->nshots);
  s->rec.resize(d->nshots);
  s->rec_changed.assign(d->nshots, 0);
  s->start_h.assign(d->nshots, 0);
  s->active_h.assign(d->nshots, 1);
  if (hipStreamSynchronize(s->st) != hipSuccess) return cleanup(fail(CGFD_E_HIP, "stream sync failed"));
  *out = s;
  return CGFD_OK;
}

void cgfd_destroy(cgfd_session *s) {
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);
  if (s->st_copy) (void)hipStreamSynchronize(s->st_copy);
  drop_graphs(s);
  for (void *p : s->allocs) (void)hipFree(p);
  s->allocs.clear();
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  if (s->ev_run0) (void)hipEventDestroy(s->ev_run0);
  if (s->ev_run1) (void)hipEventDestroy(s->ev_run1);
  if (s->ev_chunk) (void)hipEventDestroy(s->ev_chunk);
  for (hipEvent_t e : s->ev_copied)
    if (e) (void)hipEventDestroy(e);
  if (s->st_copy) (void)hipStreamDestroy(s->st_copy);
  if (s->host_seismo) (void)hipHostFree(s->host_seismo);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

int cgfd_upload_metric(cgfd_session *s, int dtype, const void *metric) {
  if (!s || !metric) return fail(CGFD_E_ARG, "null argument");
  std::vector<double> v((size_t)NMET * s->L.cs, 0.0);
  for (int c = 0; c < NMET; c++) ref_to_dev_comp(s->L, dtype, metric, c, v, (size_t)c * s->L.cs);
  int r = put_model(s, s->met, v);
  if (r) return r;
  s->have_met = true;
  // A Cartesian grid gives one value per metric component everywhere (ghosts
  // too), up to the rounding of the metric computation: its off-diagonal
  // terms come out as noise of ~1e-14 of the diagonal instead of 0.  The
  // grid counts as Cartesian when every component stays within 1e-10 of the
  // largest metric magnitude of its first value and the off-diagonal ones
  // within that of zero; the sweep then takes the first point's diagonal
  // values as arguments and the off-diagonal ones as 0 (a change below
  // 1e-10 relative in any derivative, far under FP32 rounding).  A uniform
  // but sheared grid keeps the per-point arrays, as does
  // CGFD_FLAG_FULL_METRIC.
  const size_t n = (size_t)(s->d.ni + 6) * (s->d.nj + 6) * (s->d.nk + 6);
  double scale = 0.0;
  for (int c = 0; c < 9; c++) scale = std::max(scale, std::fabs(rd(dtype, metric, (size_t)(c + 1) * n)));
  const double tol = 1e-10 * scale;
  bool cst = !(s->d.flags & CGFD_FLAG_FULL_METRIC) && scale > 0.0;
  for (int c = 0; c < 9 && cst; c++) {
    const double v0 = rd(dtype, metric, (size_t)(c + 1) * n);
    const bool diag = c == 0 || c == 4 || c == 8;  // xi_x, eta_y, zeta_z
    if (!diag && std::fabs(v0) > tol) { cst = false; break; }
    for (size_t q = 1; q < n; q++)
      if (std::fabs(rd(dtype, metric, (size_t)(c + 1) * n + q) - v0) > tol) { cst = false; break; }
    s->cmet[c] = diag ? v0 : 0.0;
  }
  s->met_const = cst;
  invalidate(s);
  return CGFD_OK;
}

int cgfd_upload_medium(cgfd_session *s, int dtype, const void *lam, const void *mu, const void *slw) {
  if (!s || !lam || !mu || !slw) return fail(CGFD_E_ARG, "null argument");
  std::vector<double> v((size_t)3 * s->L.cs, 0.0);
  ref_to_dev_comp(s->L, dtype, lam, 0, v, 0);
  ref_to_dev_comp(s->L, dtype, mu, 0, v, s->L.cs);
  ref_to_dev_comp(s->L, dtype, slw, 0, v, 2 * s->L.cs);
  int r = put_model(s, s->med, v);
  if (r) return r;
  s->have_med = true;
  invalidate(s);
  return CGFD_OK;
}

int cgfd_upload_free_surface(cgfd_session *s, int free_top, int dtype, const void *mvx, const void *mvy) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (free_top && s->face[5].on)
    return fail(CGFD_E_ARG, "the top cannot be both a free surface and a PML face");
  s->free_top = free_top ? 1 : 0;
  invalidate(s);
  if (!free_top) return CGFD_OK;
  if (!mvx || !mvy) return fail(CGFD_E_ARG, "free surface enabled without matrices");
  const int ni = s->d.ni, nj = s->d.nj;
  const long nx = ni + 2 * kNG;
  const size_t pl = (size_t)ni * nj;
  std::vector<double> v(18 * pl);
  for (int j = 0; j < nj; j++)
    for (int i = 0; i < ni; i++) {
      size_t ij = (size_t)((j + kNG) * nx + (i + kNG)) * 9;
      for (int q = 0; q < 9; q++) {
        v[(size_t)q * pl + (size_t)j * ni + i] = rd(dtype, mvx, ij + q);
        v[(size_t)(9 + q) * pl + (size_t)j * ni + i] = rd(dtype, mvy, ij + q);
      }
    }
  return put_model(s, s->mat, v);
}

int cgfd_upload_pml(cgfd_session *s, int face, int nlay, int dtype, const void *A, const void *B, const void *D) {
  if (!s || face < 0 || face > 5) return fail(CGFD_E_ARG, "bad session or face %d", face);
  FaceHost &F = s->face[face];
  invalidate(s);
  if (F.coef) dfree(s, F.coef);
  if (F.aux) dfree(s, F.aux);
  F.on = false;
  F.nlay = 0;
  if (nlay <= 0) return CGFD_OK;
  if (face == 5 && s->free_top) return fail(CGFD_E_ARG, "the top cannot be both a free surface and a PML face");
  const int axis_n = face < 2 ? s->d.ni : face < 4 ? s->d.nj : s->d.nk;
  if (nlay + 1 > axis_n) return fail(CGFD_E_ARG, "face %d: %d layers exceed the grid", face, nlay);
  if (!A || !B || !D) return fail(CGFD_E_ARG, "face %d: null profile", face);
  face_box(s, face, nlay, F);
  if ((unsigned long long)kNC * F.fsize * s->es >= kMaxComponentBytes)
    return fail(CGFD_E_ARG, "face %d: %llu bytes of memory variables per level, over the 4 GiB kernels address", face,
                (unsigned long long)kNC * F.fsize * s->es);
  F.nlay = nlay;
  std::vector<double> v(3 * (size_t)(nlay + 1));
  for (int n = 0; n <= nlay; n++) {
    v[n] = rd(dtype, A, n);
    v[(nlay + 1) + n] = rd(dtype, B, n);
    v[2 * (nlay + 1) + n] = rd(dtype, D, n);
  }
  int r;
  if ((r = dalloc(s, &F.coef, v.size() * s->ms))) return r;
  if ((r = put_model(s, F.coef, v))) return r;
  const size_t aux_elems = (size_t)s->d.nshots * 5 * kNC * F.fsize;
  if ((r = dalloc(s, &F.aux, aux_elems * s->es))) return r;
  if ((r = fill_zero(s, F.aux, aux_elems, s->es))) return r;
  HIPCHK(hipStreamSynchronize(s->st));
  F.on = true;
  return CGFD_OK;
}

int cgfd_set_storage_scale(cgfd_session *s, const int exp2[9]) {
  if (!s || !exp2) return fail(CGFD_E_ARG, "null argument");
  if (s->d.storage != CGFD_STORE_F16 && s->d.storage != CGFD_STORE_BF16)
    return fail(CGFD_E_ARG, "storage scales apply to F16/BF16 sessions only");
  for (int c = 0; c < 9; c++) {
    if (exp2[c] < -126 || exp2[c] > 127) return fail(CGFD_E_ARG, "scale exponent %d out of range", exp2[c]);
    s->sc[c] = std::ldexp(1.0f, exp2[c]);
    s->isc[c] = std::ldexp(1.0f, -exp2[c]);
  }
  invalidate(s);
  return CGFD_OK;
}

int cgfd_upload_source(cgfd_session *s, int shot, const cgfd_source_desc *src) {
  if (!s || shot < 0 || shot >= s->d.nshots) return fail(CGFD_E_ARG, "bad session or shot %d", shot);
  SrcHost h;
  if (src && src->nsrc > 0) {
    if (src->max_ext < 1 || src->max_ext > 343 || src->max_nt < 1)
      return fail(CGFD_E_ARG, "source tables: max_ext %d max_nt %d", src->max_ext, src->max_nt);
    if (!src->it_begin || !src->it_end || !src->ext_num || !src->ext_indx || !src->ext_coef)
      return fail(CGFD_E_ARG, "source tables: null array");
    if ((src->force_on && !src->F) || (src->moment_on && !src->M))
      return fail(CGFD_E_ARG, "source tables: missing F or M");
    h.nsrc = src->nsrc; h.force_on = src->force_on; h.moment_on = src->moment_on;
    h.max_nt = src->max_nt; h.max_ext = src->max_ext;
    h.itb.assign(src->it_begin, src->it_begin + h.nsrc);
    h.ite.assign(src->it_end, src->it_end + h.nsrc);
    h.ext_num.assign(src->ext_num, src->ext_num + h.nsrc);
    const size_t ne = (size_t)h.nsrc * h.max_ext;
    h.ext_indx.assign(src->ext_indx, src->ext_indx + ne);
    const long long ntot = (long long)(s->d.ni + 6) * (s->d.nj + 6) * (s->d.nk + 6);
    for (int is = 0; is < h.nsrc; is++) {
      if (h.ext_num[is] < 0 || h.ext_num[is] > h.max_ext) return fail(CGFD_E_ARG, "source %d: ext_num", is);
      if (h.ite[is] - h.itb[is] + 1 > h.max_nt || h.ite[is] < h.itb[is])
        return fail(CGFD_E_ARG, "source %d: steps %d..%d exceed its %d-sample table", is, h.itb[is], h.ite[is], h.max_nt);
      for (int e = 0; e < h.ext_num[is]; e++) {
        long long v = h.ext_indx[(size_t)is * h.max_ext + e];
        if (v < 0 || v >= ntot) return fail(CGFD_E_ARG, "source %d: index %lld outside the grid", is, v);
      }
    }
    h.ext_coef.resize(ne);
    for (size_t q = 0; q < ne; q++) h.ext_coef[q] = rd(src->dtype, src->ext_coef, q);
    const size_t nt4 = (size_t)h.nsrc * h.max_nt * 4;
    if (h.force_on) { h.F.resize(3 * nt4); for (size_t q = 0; q < 3 * nt4; q++) h.F[q] = rd(src->dtype, src->F, q); }
    if (h.moment_on) { h.M.resize(6 * nt4); for (size_t q = 0; q < 6 * nt4; q++) h.M[q] = rd(src->dtype, src->M, q); }
  }
  s->src[shot] = std::move(h);
  s->src_dirty = true;
  return CGFD_OK;
}

int cgfd_upload_receivers(cgfd_session *s, int shot, int nrecv, const long long *indx8, int dtype, const void *frac3,
                          int nt_total) {
  if (!s || shot < 0 || shot >= s->d.nshots) return fail(CGFD_E_ARG, "bad session or shot %d", shot);
  if (nrecv < 0 || (nrecv > 0 && (!indx8 || !frac3 || nt_total < 1))) return fail(CGFD_E_ARG, "receiver arguments");
  RecvHost h;
  h.nrecv = nrecv;
  h.nt = nt_total;
  h.indx8.assign(indx8, indx8 + (size_t)nrecv * 8);
  const long long ntot = (long long)(s->d.ni + 6) * (s->d.nj + 6) * (s->d.nk + 6);
  for (long long v : h.indx8)
    if (v < 0 || v >= ntot) return fail(CGFD_E_ARG, "receiver index %lld outside the grid", v);
  h.frac.resize((size_t)nrecv * 3);
  for (size_t q = 0; q < h.frac.size(); q++) h.frac[q] = rd(dtype, frac3, q);
  s->rec[shot] = std::move(h);
  s->rec_changed[shot] = 1;
  s->rec_dirty = true;
  return CGFD_OK;
}

int cgfd_set_level(cgfd_session *s, int shot, int level, int dtype, const void *w9) {
  if (!s || shot < 0 || shot >= s->d.nshots || level < 0 || level > 3 || !w9)
    return fail(CGFD_E_ARG, "bad set_level arguments");
  std::vector<double> v((size_t)s->level_elems, 0.0);
  for (int c = 0; c < kNC; c++) ref_to_dev_comp(s->L, dtype, w9, c, v, (size_t)c * s->L.cs);
  return put_storage(s, wav_level(s, shot, s->slot[level]), v, s->L.cs);
}

int cgfd_reset_shot(cgfd_session *s, int shot) {
  if (!s || shot < 0 || shot >= s->d.nshots) return fail(CGFD_E_ARG, "bad shot");
  int r = fill_zero(s, wav_level(s, shot, 0), (size_t)s->shot_elems, s->es);
  if (r) return r;
  if (s->peak) {
    const size_t n = (size_t)18 * s->d.ni * s->d.nj;
    r = fill_zero(s, (char *)s->peak + (size_t)shot * n * s->ms, n, s->ms);
    if (r) return r;
  }
  for (int f = 0; f < 6; f++) {
    FaceHost &F = s->face[f];
    if (!F.on) continue;
    r = fill_zero(s, (char *)F.aux + (size_t)shot * 5 * kNC * F.fsize * s->es, (size_t)5 * kNC * F.fsize, s->es);
    if (r) return r;
  }
  HIPCHK(hipStreamSynchronize(s->st));
  return CGFD_OK;
}

int cgfd_run(cgfd_session *s, int it0, int nsteps, double dt) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (!s->have_met || !s->have_med) return fail(CGFD_E_STATE, "metric and medium must be uploaded before running");
  if (nsteps < 0 || it0 < 0) return fail(CGFD_E_ARG, "bad step range");
  if (!(dt > 0.0)) return fail(CGFD_E_ARG, "dt must be positive");
  switch (s->d.storage) {
    case CGFD_STORE_F32: return run_t<float, float>(s, it0, nsteps, dt);
    case CGFD_STORE_F64: return run_t<double, double>(s, it0, nsteps, dt);
    case CGFD_STORE_F16: return run_t<__half, float>(s, it0, nsteps, dt);
    default: return run_t<__hip_bfloat16, float>(s, it0, nsteps, dt);
  }
}

int cgfd_el_iso_onestage(cgfd_session *s, int shot, int in_level, int ipair, int istage, int it) {
  if (!s || shot < 0 || shot >= s->d.nshots) return fail(CGFD_E_ARG, "bad session or shot");
  if (in_level < 0 || in_level > 3 || ipair < 0 || ipair > 7 || istage < 0 || istage > 3)
    return fail(CGFD_E_ARG, "bad level/pair/stage");
  if (!s->have_met || !s->have_med) return fail(CGFD_E_STATE, "metric and medium must be uploaded first");
  switch (s->d.storage) {
    case CGFD_STORE_F32: return onestage_t<float, float>(s, shot, in_level, ipair, istage, it);
    case CGFD_STORE_F64: return onestage_t<double, double>(s, shot, in_level, ipair, istage, it);
    case CGFD_STORE_F16: return onestage_t<__half, float>(s, shot, in_level, ipair, istage, it);
    default: return onestage_t<__hip_bfloat16, float>(s, shot, in_level, ipair, istage, it);
  }
}

int cgfd_read_seismo(cgfd_session *s, int shot, int dtype, void *out) {
  if (!s || shot < 0 || shot >= s->d.nshots || !out) return fail(CGFD_E_ARG, "bad read_seismo arguments");
  if (s->rec_dirty) return fail(CGFD_E_STATE, "receivers uploaded but not yet run");
  const RecvHost &h = s->rec[shot];
  const size_t n = (size_t)h.nrecv * kNC * h.nt;
  if (n == 0) return CGFD_OK;
  if (s->ring > 0) {
    // the host record is complete once the run and its copies are done
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipStreamSynchronize(s->st_copy));
    const char *src = (const char *)s->host_seismo + (size_t)shot * s->host_stride * s->ms;
    for (size_t q = 0; q < n; q++) {
      const double x = s->ms == 8 ? ((const double *)src)[q] : (double)((const float *)src)[q];
      if (dtype == CGFD_DT_F64) ((double *)out)[q] = x;
      else ((float *)out)[q] = (float)x;
    }
    return CGFD_OK;
  }
  const size_t off = (size_t)s->seismo_off[shot];
  if (s->ms == 8) {
    std::vector<double> v(n);
    HIPCHK(d2h(s, v.data(), (double *)s->seismo + off, n * 8));
    for (size_t q = 0; q < n; q++) {
      if (dtype == CGFD_DT_F64) ((double *)out)[q] = v[q];
      else ((float *)out)[q] = (float)v[q];
    }
  } else {
    std::vector<float> v(n);
    HIPCHK(d2h(s, v.data(), (float *)s->seismo + off, n * 4));
    for (size_t q = 0; q < n; q++) {
      if (dtype == CGFD_DT_F64) ((double *)out)[q] = v[q];
      else ((float *)out)[q] = v[q];
    }
  }
  return CGFD_OK;
}

int cgfd_fetch_level(cgfd_session *s, int shot, int level, int dtype, void *w9) {
  if (!s || shot < 0 || shot >= s->d.nshots || level < 0 || level > 4 || !w9)
    return fail(CGFD_E_ARG, "bad fetch_level arguments");
  const void *src;
  if (level == CGFD_LEVEL_RHS) {
    if (!s->rhs) return fail(CGFD_E_STATE, "no RHS level: call cgfd_el_iso_onestage first");
    src = s->rhs;
  } else {
    src = wav_level(s, shot, s->slot[level]);
  }
  HIPCHK(hipStreamSynchronize(s->st));
  std::vector<double> v;
  int r = get_storage(s, src, (size_t)s->level_elems, s->L.cs, v);
  if (r) return r;
  const long nx = s->d.ni + 6, ny = s->d.nj + 6, nz = s->d.nk + 6;
  const size_t sc = (size_t)nx * ny * nz;
  for (int c = 0; c < kNC; c++)
    for (long k = 0; k < nz; k++)
      for (long j = 0; j < ny; j++)
        for (long i = 0; i < nx; i++) {
          double x = v[(size_t)c * s->L.cs + k * s->L.pz + j * s->L.px + (i - kNG + s->L.xa)];
          size_t q = (size_t)c * sc + (k * ny + j) * nx + i;
          if (dtype == CGFD_DT_F64) ((double *)w9)[q] = x;
          else ((float *)w9)[q] = (float)x;
        }
  return CGFD_OK;
}

}  // extern "C"

namespace {
template <typename TS, typename TC>
int subvolume_t(cgfd_session *s, int shot, int level, unsigned mask, int i1, int j1, int k1, int ni, int nj, int nk,
                int di, int dj, int dk, int dtype, void *out) {
  const size_t n = (size_t)ni * nj * nk;
  if (s->gather_bytes < n * sizeof(TC)) {
    if (s->gather) dfree(s, s->gather);
    s->gather_bytes = 0;
    int r = dalloc(s, &s->gather, n * sizeof(TC));
    if (r) return r;
    s->gather_bytes = n * sizeof(TC);
  }
  std::vector<TC> h(n);
  size_t q = 0;
  for (int c = 0; c < kNC; c++) {
    if (!(mask >> c & 1u)) continue;
    const TS *w = (const TS *)wav_level(s, shot, s->slot[level]) + (size_t)c * s->L.cs;
    const TC sc = St<TS>::scaled ? (TC)s->sc[c] : TC(1);
    hipError_t e = launch_gather<TS, TC>(w, s->L, i1, j1, k1, ni, nj, nk, di, dj, dk, sc, (TC *)s->gather, s->st);
    if (e != hipSuccess) return fail(CGFD_E_HIP, "gather kernel: %s", hipGetErrorString(e));
    HIPCHK(d2h(s, h.data(), s->gather, n * sizeof(TC)));
    for (size_t p = 0; p < n; p++, q++) {
      if (dtype == CGFD_DT_F64) ((double *)out)[q] = (double)h[p];
      else ((float *)out)[q] = (float)h[p];
    }
  }
  return CGFD_OK;
}
// ---- the sub-volume fetch that allocates nothing -----------------------------
// scratch: device staging of cgfd_fetch_subvolume_noalloc, a file-scope
// array every GPU the library runs on holds until the process ends.  Host
// code never uses its name as an address: the address on each device comes
// from hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kScratchBytes = 8u << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_address(int dev, void **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(CGFD_E_ARG, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail(CGFD_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CGFD_OK;
}

// The selected components one after another through the scratch staging:
// gather on the session stream, copy straight into the caller's buffer.
template <typename TS, typename TC>
int subvolume_noalloc_t(cgfd_session *s, void *scr, int shot, int level, unsigned mask, int i1, int j1, int k1,
                        int ni, int nj, int nk, int di, int dj, int dk, void *out) {
  const size_t n = (size_t)ni * nj * nk;
  size_t q = 0;
  for (int c = 0; c < kNC; c++) {
    if (!(mask >> c & 1u)) continue;
    const TS *w = (const TS *)wav_level(s, shot, s->slot[level]) + (size_t)c * s->L.cs;
    const TC sc = St<TS>::scaled ? (TC)s->sc[c] : TC(1);
    hipError_t e = launch_gather<TS, TC>(w, s->L, i1, j1, k1, ni, nj, nk, di, dj, dk, sc, (TC *)scr, s->st);
    if (e != hipSuccess) return fail(CGFD_E_HIP, "gather kernel: %s", hipGetErrorString(e));
    HIPCHK(hipMemcpyAsync((TC *)out + q, scr, n * sizeof(TC), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    q += n;
  }
  return CGFD_OK;
}

}  // namespace

extern "C" {

int cgfd_fetch_subvolume_noalloc(cgfd_session *s, int shot, int level, unsigned cmp_mask, int i1, int j1, int k1,
                                 int ni, int nj, int nk, int di, int dj, int dk, int dtype, void *out) {
  if (!s || shot < 0 || shot >= s->d.nshots || level < 0 || level > 3 || !out)
    return fail(CGFD_E_ARG, "bad fetch_subvolume_noalloc arguments");
  if (ni < 1 || nj < 1 || nk < 1 || di < 1 || dj < 1 || dk < 1 || i1 < 0 || j1 < 0 || k1 < 0 ||
      i1 + (long)(ni - 1) * di > s->d.ni - 1 || j1 + (long)(nj - 1) * dj > s->d.nj - 1 ||
      k1 + (long)(nk - 1) * dk > s->d.nk - 1)
    return fail(CGFD_E_ARG, "sub-volume outside the grid");
  if ((cmp_mask & 0x1ffu) == 0 || (cmp_mask & ~0x1ffu)) return fail(CGFD_E_ARG, "component mask %#x", cmp_mask);
  const int want = s->ms == 8 ? CGFD_DT_F64 : CGFD_DT_F32;
  if (dtype != want)
    return fail(CGFD_E_ARG, "fetch_subvolume_noalloc writes the session's model type (%s)",
                want == CGFD_DT_F64 ? "CGFD_DT_F64" : "CGFD_DT_F32");
  const unsigned long long bytes = (unsigned long long)ni * nj * nk * s->ms;
  if (bytes > kScratchBytes)
    return fail(CGFD_E_ARG, "sub-volume of %llu bytes per component exceeds the %zu-byte scratch of the "
                            "non-allocating fetch", bytes, kScratchBytes);
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != s->dev) HIPCHK(hipSetDevice(s->dev));
  void *scr = nullptr;
  int r = scratch_address(s->dev, &scr);
  if (!r) {
    hipError_t e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) r = fail(CGFD_E_HIP, "stream sync: %s", hipGetErrorString(e));
  }
  if (!r) {
    switch (s->d.storage) {
      case CGFD_STORE_F32: r = subvolume_noalloc_t<float, float>(s, scr, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, out); break;
      case CGFD_STORE_F64: r = subvolume_noalloc_t<double, double>(s, scr, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, out); break;
      case CGFD_STORE_F16: r = subvolume_noalloc_t<__half, float>(s, scr, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, out); break;
      default: r = subvolume_noalloc_t<__hip_bfloat16, float>(s, scr, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, out); break;
    }
  }
  if (cur != s->dev) (void)hipSetDevice(cur);
  return r;
}

int cgfd_fetch_subvolume(cgfd_session *s, int shot, int level, unsigned cmp_mask, int i1, int j1, int k1, int ni,
                         int nj, int nk, int di, int dj, int dk, int dtype, void *out) {
  if (!s || shot < 0 || shot >= s->d.nshots || level < 0 || level > 3 || !out)
    return fail(CGFD_E_ARG, "bad fetch_subvolume arguments");
  if (ni < 1 || nj < 1 || nk < 1 || di < 1 || dj < 1 || dk < 1 || i1 < 0 || j1 < 0 || k1 < 0 ||
      i1 + (long)(ni - 1) * di > s->d.ni - 1 || j1 + (long)(nj - 1) * dj > s->d.nj - 1 ||
      k1 + (long)(nk - 1) * dk > s->d.nk - 1)
    return fail(CGFD_E_ARG, "sub-volume outside the grid");
  if ((cmp_mask & 0x1ffu) == 0 || (cmp_mask & ~0x1ffu)) return fail(CGFD_E_ARG, "component mask %#x", cmp_mask);
  HIPCHK(hipStreamSynchronize(s->st));
  switch (s->d.storage) {
    case CGFD_STORE_F32: return subvolume_t<float, float>(s, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, dtype, out);
    case CGFD_STORE_F64: return subvolume_t<double, double>(s, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, dtype, out);
    case CGFD_STORE_F16: return subvolume_t<__half, float>(s, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, dtype, out);
    default: return subvolume_t<__hip_bfloat16, float>(s, shot, level, cmp_mask, i1, j1, k1, ni, nj, nk, di, dj, dk, dtype, out);
  }
}

int cgfd_fetch_pml_aux(cgfd_session *s, int shot, int face, int level, int dtype, void *out) {
  if (!s || shot < 0 || shot >= s->d.nshots || face < 0 || face > 5 || level < 0 || level > 4 || !out)
    return fail(CGFD_E_ARG, "bad fetch_pml_aux arguments");
  FaceHost &F = s->face[face];
  if (!F.on) return fail(CGFD_E_STATE, "face %d has no PML", face);
  const int slot = level == CGFD_LEVEL_RHS ? 4 : s->slot[level];
  const char *src = (const char *)F.aux + ((size_t)shot * 5 + slot) * kNC * F.fsize * s->es;
  HIPCHK(hipStreamSynchronize(s->st));
  std::vector<double> v;
  int r = get_storage(s, src, (size_t)kNC * F.fsize, F.fsize, v);
  if (r) return r;
  for (size_t q = 0; q < v.size(); q++) {
    if (dtype == CGFD_DT_F64) ((double *)out)[q] = v[q];
    else ((float *)out)[q] = (float)v[q];
  }
  return CGFD_OK;
}

int cgfd_set_tile(cgfd_session *s, int shape, int zchunk) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  invalidate(s);
  if (shape < 0) {  // back to automatic selection
    s->tile_fixed = false;
    s->tile_shape = -1;
    return CGFD_OK;
  }
  if (shape >= interior_num_shapes() || zchunk < 1) return fail(CGFD_E_ARG, "tile %d/%d invalid", shape, zchunk);
  s->tile_fixed = true;
  s->tile_shape = shape;
  s->tile_zchunk = zchunk;
  if (!s->split_fixed) s->split = CGFD_SPLIT_PLANES;
  return CGFD_OK;
}

int cgfd_set_split(cgfd_session *s, int split) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (split > CGFD_SPLIT_CORE) return fail(CGFD_E_ARG, "work split %d invalid", split);
  invalidate(s);
  s->split_fixed = split >= 0;
  s->split = split >= 0 ? split : CGFD_SPLIT_PLANES;
  if (!s->tile_fixed) s->tile_shape = -1;  // tune the tile again (under the pinned split)
  return CGFD_OK;
}

int cgfd_get_const_metric(cgfd_session *s, int *const_metric) {
  if (!s || !const_metric) return fail(CGFD_E_ARG, "null argument");
  *const_metric = s->met_const ? 1 : 0;
  return CGFD_OK;
}

int cgfd_get_split(cgfd_session *s, int *split) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (split)
    *split = (s->d.flags & CGFD_FLAG_PML_POINT) ? CGFD_SPLIT_SHELL : s->split < 0 ? CGFD_SPLIT_PLANES : s->split;
  return CGFD_OK;
}
int cgfd_get_tile(cgfd_session *s, int *shape, int *tx, int *ty, int *zchunk) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  int sh = s->tile_shape < 0 ? 0 : s->tile_shape, a = 0, b = 0;
  interior_shape(sh, &a, &b);
  if (shape) *shape = s->tile_shape;
  if (tx) *tx = a;
  if (ty) *ty = b;
  if (zchunk) *zchunk = s->tile_zchunk;
  return CGFD_OK;
}

int cgfd_set_timing(cgfd_session *s, int on) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  s->timing = on != 0;
  return CGFD_OK;
}

int cgfd_set_peak_motion(cgfd_session *s, int on) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  invalidate(s);
  if (!on) {
    if (s->peak) dfree(s, s->peak);
    return CGFD_OK;
  }
  if (s->peak) return CGFD_OK;
  const size_t n = (size_t)s->d.nshots * 18 * s->d.ni * s->d.nj;
  int r = dalloc(s, &s->peak, n * s->ms);
  if (r) return r;
  r = fill_zero(s, s->peak, n, s->ms);
  if (r) return r;
  HIPCHK(hipStreamSynchronize(s->st));
  return CGFD_OK;
}

int cgfd_read_peak_motion(cgfd_session *s, int shot, int dtype, void *out) {
  if (!s || shot < 0 || shot >= s->d.nshots || !out) return fail(CGFD_E_ARG, "bad read_peak_motion arguments");
  if (!s->peak) return fail(CGFD_E_STATE, "peak motion is not enabled");
  const int ni = s->d.ni, nj = s->d.nj, nx = ni + 2 * kNG, ny = nj + 2 * kNG;
  const size_t n = (size_t)ni * nj;
  std::vector<double> v(15 * n);
  if (s->ms == 8) {
    HIPCHK(d2h(s, v.data(), (double *)s->peak + (size_t)shot * 18 * n, 15 * n * 8));
  } else {
    std::vector<float> f(15 * n);
    HIPCHK(d2h(s, f.data(), (float *)s->peak + (size_t)shot * 18 * n, 15 * n * 4));
    for (size_t q = 0; q < f.size(); q++) v[q] = f[q];
  }
  for (int m = 0; m < 15; m++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const bool in = i >= kNG && i < kNG + ni && j >= kNG && j < kNG + nj;
        const double x = in ? v[(size_t)m * n + (size_t)(j - kNG) * ni + (i - kNG)] : 0.0;
        const size_t q = ((size_t)m * ny + j) * nx + i;
        if (dtype == CGFD_DT_F64) ((double *)out)[q] = x;
        else ((float *)out)[q] = (float)x;
      }
  return CGFD_OK;
}

int cgfd_reserve_receivers(cgfd_session *s, int max_nrecv, int max_nt) {
  if (!s || max_nrecv < 0 || max_nt < 0) return fail(CGFD_E_ARG, "bad reserve_receivers arguments");
  const long want = (long)max_nrecv * kNC * (s->ring > 0 ? s->ring : max_nt);
  if (s->ring > 0) {
    int r = host_record(s, (long)max_nrecv * kNC * max_nt);
    if (r) return r;
  }
  if (want <= s->seismo_stride) return CGFD_OK;
  void *old = s->seismo;
  const long old_stride = s->seismo_stride;
  void *p = nullptr;
  int r = dalloc(s, &p, (size_t)s->d.nshots * want * s->ms);
  if (r) return r;
  if ((r = fill_zero(s, p, (size_t)s->d.nshots * want, s->ms))) return r;
  if (old) {
    for (int shot = 0; shot < s->d.nshots; shot++)
      HIPCHK(hipMemcpyAsync((char *)p + (size_t)shot * want * s->ms, (const char *)old + (size_t)shot * old_stride * s->ms,
                            (size_t)old_stride * s->ms, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    dfree(s, old);
  }
  s->seismo = p;
  s->seismo_stride = want;
  // receiver tables point into the seismogram buffer by offset: rebuild them
  s->rec_dirty = true;
  return s->ms == 8 ? build_receivers<double>(s) : build_receivers<float>(s);
}

int cgfd_set_graph(cgfd_session *s, int on) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  s->use_graph = on != 0;
  return CGFD_OK;
}

int cgfd_set_shot(cgfd_session *s, int shot, int active, int it_start) {
  if (!s || shot < 0 || shot >= s->d.nshots) return fail(CGFD_E_ARG, "bad session or shot %d", shot);
  if (active && (it_start < 0 || it_start % 8 != 0))
    return fail(CGFD_E_ARG, "shot start %d must be a non-negative multiple of 8", it_start);
  hipError_t e = launch_fill<int>(s->shot_active + shot, 1, active ? 1 : 0, s->st);
  if (e == hipSuccess && active) e = launch_fill<int>(s->shot_start + shot, 1, it_start, s->st);
  if (e != hipSuccess) return fail(CGFD_E_HIP, "shot state: %s", hipGetErrorString(e));
  s->active_h[shot] = active ? 1 : 0;
  if (active) s->start_h[shot] = it_start;
  return CGFD_OK;
}

int cgfd_run_async(cgfd_session *s, int it0, int nsteps, double dt) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (!s->have_met || !s->have_med) return fail(CGFD_E_STATE, "metric and medium must be uploaded before running");
  if (nsteps < 0 || it0 < 0) return fail(CGFD_E_ARG, "bad step range");
  if (!(dt > 0.0)) return fail(CGFD_E_ARG, "dt must be positive");
  if (s->timing) return fail(CGFD_E_STATE, "per-launch timing needs the synchronous cgfd_run");
  int r;
  switch (s->d.storage) {
    case CGFD_STORE_F32: r = enqueue_t<float, float>(s, it0, nsteps, dt); break;
    case CGFD_STORE_F64: r = enqueue_t<double, double>(s, it0, nsteps, dt); break;
    case CGFD_STORE_F16: r = enqueue_t<__half, float>(s, it0, nsteps, dt); break;
    default: r = enqueue_t<__hip_bfloat16, float>(s, it0, nsteps, dt); break;
  }
  return r;
}

int cgfd_poll(cgfd_session *s) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  hipError_t e = hipStreamQuery(s->st);
  if (e == hipSuccess && s->st_copy) e = hipStreamQuery(s->st_copy);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  return fail(CGFD_E_HIP, "stream query: %s", hipGetErrorString(e));
}

int cgfd_wait(cgfd_session *s) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  HIPCHK(hipStreamSynchronize(s->st));
  if (s->st_copy) HIPCHK(hipStreamSynchronize(s->st_copy));
  return CGFD_OK;
}

int cgfd_set_seismo_ring(cgfd_session *s, int ring_steps) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (ring_steps < 0 || (ring_steps > 0 && ring_steps % 16 != 0))
    return fail(CGFD_E_ARG, "seismogram ring of %d steps: 0 (off) or a positive multiple of 16", ring_steps);
  if (ring_steps == s->ring) return CGFD_OK;
  HIPCHK(hipStreamSynchronize(s->st));
  if (ring_steps > 0 && !s->st_copy) {
    HIPCHK(hipStreamCreateWithFlags(&s->st_copy, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s->ev_chunk, hipEventDisableTiming));
    for (hipEvent_t &e : s->ev_copied) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (s->st_copy) HIPCHK(hipStreamSynchronize(s->st_copy));
  // samples recorded so far are laid out for the old pitch: start over
  dfree(s, s->seismo);
  s->seismo_stride = 0;
  if (s->host_seismo) (void)hipHostFree(s->host_seismo);
  s->host_seismo = nullptr;
  s->host_stride = 0;
  s->ring = ring_steps;
  s->copied_chunk[0] = s->copied_chunk[1] = -1;
  std::fill(s->rec_changed.begin(), s->rec_changed.end(), 1);
  s->rec_dirty = true;
  return CGFD_OK;
}

int cgfd_last_run_timing(cgfd_session *s, double *run_ms, double *interior_avg_ms, int *interior_launches) {
  if (!s) return fail(CGFD_E_ARG, "null session");
  if (run_ms) *run_ms = s->last_run_ms;
  if (interior_avg_ms) *interior_avg_ms = s->last_int_avg_ms;
  if (interior_launches) *interior_launches = s->last_int_n;
  return CGFD_OK;
}

}  // extern "C"
