// cgfd_io.cpp -- SAC and NetCDF-3 classic writers (see cgfd_io.h).  The
// NetCDF file is written directly in the classic format (magic "CDF\1",
// big-endian, record variables interleaved per record), so no netCDF library
// is needed.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cgfd_el_iso.h"
#include "../../include/cgfd_io.h"
#include "cgfd_internal.h"

namespace {

int iofail(int code, const std::string &msg) {
  cgfd_internal_set_error(msg.c_str());
  return code;
}

// ---- SAC ------------------------------------------------------------------------------
// Header: 70 floats, 40 ints, 24 eight-character strings (kevnm takes two);
// word positions from the published SAC v6 header layout.
constexpr int kSacF = 70, kSacI = 40, kSacK = 24;
enum SacF { F_DELTA = 0, F_B = 5, F_E = 6, F_STLA = 31, F_STLO = 32, F_STEL = 33, F_STDP = 34, F_EVLA = 35,
            F_EVLO = 36, F_EVEL = 37, F_EVDP = 38, F_CMPAZ = 57, F_CMPINC = 58 };
enum SacI { I_NVHDR = 6, I_NPTS = 9, I_IFTYPE = 15, I_LEVEN = 35 };
enum SacK { K_KSTNM = 0, K_KCMPNM = 20 };

void put8(char *dst, const char *s, size_t width) {
  // the reference copies with strncpy: the name, then NUL padding
  std::memset(dst, 0, width);
  std::strncpy(dst, s, width);
}

// ---- NetCDF classic --------------------------------------------------------------------
struct BE {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void f32(float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    u32(v);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    for (char c : s) b.push_back((unsigned char)c);
    while (b.size() % 4) b.push_back(0);
  }
};

constexpr uint32_t NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12;
constexpr uint32_t NC_INT = 4, NC_FLOAT = 5;

}  // namespace

struct cgfd_snap {
  FILE *fp = nullptr;
  cgfd_snap_desc d{};
  std::vector<std::string> vars;  // record variables after "time"
  long long data_begin = 0;       // offset of record 0
  long long slab = 0;             // bytes of one variable per record
  long long recsize = 0;
  uint32_t nrec = 0;
  std::vector<unsigned char> buf;
};

struct cgfd_slice {
  FILE *fp = nullptr;
  long long data_begin = 0;  // offset of record 0
  long long slab = 0;        // bytes of one component per record
  long long recsize = 0;
  uint32_t nrec = 0;
  long long nvals = 0;       // values per component per record
};

namespace {
// component names of the wavefield, in storage order (wav_t.cu:100-140)
const char *kCmpName[9] = {"Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"};
}  // namespace

extern "C" {

int cgfd_slice_open(const char *path, int axis, int index_with_ghosts, int n_slow, int n_fast, const int *topo,
                    cgfd_slice **out) {
  if (!path || !out || axis < 0 || axis > 2 || n_slow < 1 || n_fast < 1 || !topo)
    return iofail(CGFD_E_ARG, "cgfd_slice_open: bad arguments");
  *out = nullptr;
  static const char *dims[3][2] = {{"k", "j"}, {"k", "i"}, {"j", "i"}};
  static const char *att[3] = {"i_index_with_ghosts_in_this_thread", "j_index_with_ghosts_in_this_thread",
                               "k_index_with_ghosts_in_this_thread"};
  auto *sl = new cgfd_slice();
  sl->nvals = (long long)n_slow * n_fast;
  sl->slab = 4 * sl->nvals;
  auto header = [&](uint32_t begin0) {
    BE h;
    h.b.insert(h.b.end(), {'C', 'D', 'F', 1});
    h.u32(sl->nrec);
    h.u32(NC_DIMENSION);
    h.u32(3);
    h.name("time"); h.u32(0);
    h.name(dims[axis][0]); h.u32((uint32_t)n_slow);
    h.name(dims[axis][1]); h.u32((uint32_t)n_fast);
    h.u32(NC_ATTRIBUTE);
    h.u32(2);
    h.name(att[axis]); h.u32(NC_INT); h.u32(1); h.u32((uint32_t)index_with_ghosts);
    h.name("coords_of_mpi_topo"); h.u32(NC_INT); h.u32(3);
    for (int q = 0; q < 3; q++) h.u32((uint32_t)topo[q]);
    h.u32(NC_VARIABLE);
    h.u32(10);
    uint32_t off = begin0;
    h.name("time");
    h.u32(1); h.u32(0);
    h.u32(0); h.u32(0);
    h.u32(NC_FLOAT);
    h.u32(4);
    h.u32(off);
    off += 4;
    for (const char *v : kCmpName) {
      h.name(v);
      h.u32(3); h.u32(0); h.u32(1); h.u32(2);
      h.u32(0); h.u32(0);
      h.u32(NC_FLOAT);
      h.u32((uint32_t)sl->slab);
      h.u32(off);
      off += (uint32_t)sl->slab;
    }
    return h.b;
  };
  const std::vector<unsigned char> h0 = header(0);
  sl->data_begin = (long long)h0.size();
  sl->recsize = 4 + 9 * sl->slab;
  if (sl->data_begin + sl->recsize > 0x7fffffffLL) {
    delete sl;
    return iofail(CGFD_E_ARG, "slice record too large for the classic format offsets");
  }
  const std::vector<unsigned char> h1 = header((uint32_t)sl->data_begin);
  sl->fp = std::fopen(path, "wb");
  if (!sl->fp) {
    delete sl;
    return iofail(CGFD_E_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  }
  if (std::fwrite(h1.data(), 1, h1.size(), sl->fp) != h1.size()) {
    std::fclose(sl->fp);
    delete sl;
    return iofail(CGFD_E_ARG, std::string("write failed: ") + path);
  }
  *out = sl;
  return CGFD_OK;
}

int cgfd_slice_append(cgfd_slice *sl, float time, const float *w) {
  if (!sl || !w) return iofail(CGFD_E_ARG, "cgfd_slice_append: bad arguments");
  BE r;
  r.b.reserve((size_t)sl->recsize);
  r.f32(time);
  for (long long q = 0; q < 9 * sl->nvals; q++) r.f32(w[q]);
  const long long off = sl->data_begin + (long long)sl->nrec * sl->recsize;
  if (std::fseek(sl->fp, (long)off, SEEK_SET) != 0 || std::fwrite(r.b.data(), 1, r.b.size(), sl->fp) != r.b.size())
    return iofail(CGFD_E_ARG, "slice write failed");
  sl->nrec++;
  return CGFD_OK;
}

int cgfd_slice_close(cgfd_slice *sl) {
  if (!sl) return iofail(CGFD_E_ARG, "null slice");
  BE n;
  n.u32(sl->nrec);
  bool ok = std::fseek(sl->fp, 4, SEEK_SET) == 0 && std::fwrite(n.b.data(), 1, 4, sl->fp) == 4;
  ok = (std::fclose(sl->fp) == 0) && ok;
  delete sl;
  return ok ? CGFD_OK : iofail(CGFD_E_ARG, "slice close failed");
}

int cgfd_line_sac(const char *dir, const char *line, int seq, const float *seismo, int nt, float dt, float x,
                  float y, float z) {
  if (!dir || !line || !seismo || nt < 1 || seq < 0) return iofail(CGFD_E_ARG, "cgfd_line_sac: bad arguments");
  char name[512];
  std::snprintf(name, sizeof(name), "%s.no%d", line, seq);
  return cgfd_station_sac(dir, name, seismo, nt, dt, x, y, z, 0, 0.0f, 0.0f);
}

int cgfd_sac_write(const char *path, const float *data, int npts, float delta, float b, const float *ev,
                   const float *st, const char *kstnm, const char *kcmpnm) {
  if (!path || (!data && npts > 0) || npts < 0) return iofail(CGFD_E_ARG, "cgfd_sac_write: bad arguments");
  float hf[kSacF];
  int32_t hi[kSacI];
  char hk[kSacK * 8];
  for (float &v : hf) v = -12345.0f;
  for (int32_t &v : hi) v = -12345;
  for (int k = 0; k < kSacK; k++) std::memcpy(hk + 8 * k, "-12345  ", 8);
  // kevnm spans two slots; the reference pads the second with NULs (strncpy of 8 chars into 16)
  std::memset(hk + 8, 0, 16);
  std::memcpy(hk + 8, "-12345  ", 8);
  hf[F_DELTA] = delta;
  hf[F_B] = b;
  hf[F_E] = b + delta * (float)(npts - 1);
  hi[I_NVHDR] = 6;
  hi[I_NPTS] = npts;
  hi[I_IFTYPE] = 1;  // ITIME
  hi[I_LEVEN] = 1;
  const float zero4[4] = {0, 0, 0, 0};
  const float *e = ev ? ev : zero4, *s = st ? st : zero4;
  hf[F_EVLA] = e[0]; hf[F_EVLO] = e[1]; hf[F_EVEL] = e[2]; hf[F_EVDP] = e[3];
  hf[F_STLA] = s[0]; hf[F_STLO] = s[1]; hf[F_STEL] = s[2]; hf[F_STDP] = s[3];
  hf[F_CMPAZ] = 0.0f;
  hf[F_CMPINC] = 0.0f;
  put8(hk + 8 * K_KSTNM, kstnm ? kstnm : "-12345", 8);
  put8(hk + 8 * K_KCMPNM, kcmpnm ? kcmpnm : "-12345", 8);
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return iofail(CGFD_E_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  bool ok = std::fwrite(hf, 4, kSacF, fp) == (size_t)kSacF && std::fwrite(hi, 4, kSacI, fp) == (size_t)kSacI &&
            std::fwrite(hk, 1, sizeof(hk), fp) == sizeof(hk) &&
            (npts == 0 || std::fwrite(data, 4, (size_t)npts, fp) == (size_t)npts);
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? CGFD_OK : iofail(CGFD_E_ARG, std::string("write failed: ") + path);
}

int cgfd_station_sac(const char *dir, const char *name, const float *seismo, int nt, float dt, float x, float y,
                     float z, int with_strain, float lam, float mu) {
  if (!dir || !name || !seismo || nt < 1) return iofail(CGFD_E_ARG, "cgfd_station_sac: bad arguments");
  static const char *cmp[9] = {"Vx", "Vy", "Vz", "Txx", "Tyy", "Tzz", "Tyz", "Txz", "Txy"};
  const float ev[4] = {0, 0, 0, 0};
  const float st[4] = {x, y, z, dt};  // the station depth slot receives dt, as in the reference call
  for (int c = 0; c < 9; c++) {
    std::string p = std::string(dir) + "/" + name + "." + cmp[c] + ".sac";
    int r = cgfd_sac_write(p.c_str(), seismo + (size_t)c * nt, nt, dt, 0.0f, ev, st, "-12345", "-12345");
    if (r) return r;
  }
  if (!with_strain) return CGFD_OK;
  if (!(mu > 0)) return iofail(CGFD_E_ARG, "strain output needs mu > 0");
  // isotropic compliance (io_funcs.cu:1845-1860)
  const float E1 = (lam + mu) / (mu * (3.0f * lam + 2.0f * mu));
  const float E2 = -lam / (2.0f * mu * (3.0f * lam + 2.0f * mu));
  const float E3 = 1.0f / mu;
  std::vector<float> E(6 * (size_t)nt);
  const float *Txx = seismo + 3 * (size_t)nt, *Tyy = seismo + 4 * (size_t)nt, *Tzz = seismo + 5 * (size_t)nt;
  const float *Tyz = seismo + 6 * (size_t)nt, *Txz = seismo + 7 * (size_t)nt, *Txy = seismo + 8 * (size_t)nt;
  for (int it = 0; it < nt; it++) {
    const float E0 = E2 * (Txx[it] + Tyy[it] + Tzz[it]);
    E[0 * (size_t)nt + it] = E0 - (E2 - E1) * Txx[it];
    E[1 * (size_t)nt + it] = E0 - (E2 - E1) * Tyy[it];
    E[2 * (size_t)nt + it] = E0 - (E2 - E1) * Tzz[it];
    E[3 * (size_t)nt + it] = 0.5f * E3 * Tyz[it];
    E[4 * (size_t)nt + it] = 0.5f * E3 * Txz[it];
    E[5 * (size_t)nt + it] = 0.5f * E3 * Txy[it];
  }
  static const char *ecmp[6] = {"Exx", "Eyy", "Ezz", "Eyz", "Exz", "Exy"};
  for (int c = 0; c < 6; c++) {
    std::string p = std::string(dir) + "/" + name + "." + ecmp[c] + ".sac";
    int r = cgfd_sac_write(p.c_str(), E.data() + (size_t)c * nt, nt, dt, 0.0f, ev, st, "-12345", "-12345");
    if (r) return r;
  }
  return CGFD_OK;
}

int cgfd_snap_open(const char *path, const cgfd_snap_desc *d, cgfd_snap **out) {
  if (!path || !d || !out || d->ni < 1 || d->nj < 1 || d->nk < 1)
    return iofail(CGFD_E_ARG, "cgfd_snap_open: bad arguments");
  *out = nullptr;
  auto *sn = new cgfd_snap();
  sn->d = *d;
  if (d->out_vel) sn->vars.insert(sn->vars.end(), {"Vx", "Vy", "Vz"});
  if (d->out_stress) sn->vars.insert(sn->vars.end(), {"Txx", "Tyy", "Tzz", "Txz", "Tyz", "Txy"});
  if (d->out_strain) sn->vars.insert(sn->vars.end(), {"Exx", "Eyy", "Ezz", "Exz", "Eyz", "Exy"});
  sn->slab = 4LL * d->ni * d->nj * d->nk;
  if (sn->slab > 0xffffffffLL) {
    delete sn;
    return iofail(CGFD_E_ARG, "snapshot slab exceeds the 4 GiB limit of the classic format");
  }
  // header with placeholder offsets, then again with the real ones (same size)
  auto header = [&](uint32_t begin0) {
    BE h;
    h.b.insert(h.b.end(), {'C', 'D', 'F', 1});
    h.u32(sn->nrec);
    h.u32(NC_DIMENSION);
    h.u32(4);
    h.name("time"); h.u32(0);
    h.name("k"); h.u32((uint32_t)d->nk);
    h.name("j"); h.u32((uint32_t)d->nj);
    h.name("i"); h.u32((uint32_t)d->ni);
    h.u32(NC_ATTRIBUTE);
    h.u32(4);
    auto att3 = [&](const char *nm, const int *v) {
      h.name(nm);
      h.u32(NC_INT);
      h.u32(3);
      for (int q = 0; q < 3; q++) h.u32((uint32_t)v[q]);
    };
    att3("first_index_to_snapshot_output", d->first_global);
    att3("first_index_in_this_thread_with_ghosts", d->first_local);
    att3("index_stride_in_this_thread", d->stride);
    att3("coords_of_mpi_topo", d->topo);
    h.u32(NC_VARIABLE);
    h.u32((uint32_t)(1 + sn->vars.size()));
    uint32_t off = begin0;
    h.name("time");
    h.u32(1); h.u32(0);
    h.u32(0); h.u32(0);  // no attributes
    h.u32(NC_FLOAT);
    h.u32(4);
    h.u32(off);
    off += 4;
    for (const std::string &v : sn->vars) {
      h.name(v);
      h.u32(4); h.u32(0); h.u32(1); h.u32(2); h.u32(3);
      h.u32(0); h.u32(0);
      h.u32(NC_FLOAT);
      h.u32((uint32_t)sn->slab);
      h.u32(off);
      off += (uint32_t)sn->slab;
    }
    return h.b;
  };
  std::vector<unsigned char> h0 = header(0);
  sn->data_begin = (long long)h0.size();
  sn->recsize = 4 + (long long)sn->vars.size() * sn->slab;
  if (sn->data_begin + sn->recsize > 0x7fffffffLL) {
    delete sn;
    return iofail(CGFD_E_ARG, "snapshot record too large for the classic format offsets");
  }
  std::vector<unsigned char> h1 = header((uint32_t)sn->data_begin);
  sn->fp = std::fopen(path, "wb");
  if (!sn->fp) {
    delete sn;
    return iofail(CGFD_E_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  }
  if (std::fwrite(h1.data(), 1, h1.size(), sn->fp) != h1.size()) {
    std::fclose(sn->fp);
    delete sn;
    return iofail(CGFD_E_ARG, std::string("write failed: ") + path);
  }
  *out = sn;
  return CGFD_OK;
}

int cgfd_snap_append(cgfd_snap *sn, float time, const float *w, const float *lam, const float *mu) {
  if (!sn || !w) return iofail(CGFD_E_ARG, "cgfd_snap_append: bad arguments");
  if (sn->d.out_strain && (!lam || !mu)) return iofail(CGFD_E_ARG, "strain snapshots need lam and mu");
  const size_t n = (size_t)sn->d.ni * sn->d.nj * sn->d.nk;
  BE r;
  r.b.reserve((size_t)sn->recsize);
  r.f32(time);
  auto put = [&](const float *a) {
    for (size_t q = 0; q < n; q++) r.f32(a[q]);
  };
  const float *V[3] = {w, w + n, w + 2 * n};
  const float *Txx = w + 3 * n, *Tyy = w + 4 * n, *Tzz = w + 5 * n, *Tyz = w + 6 * n, *Txz = w + 7 * n,
              *Txy = w + 8 * n;
  if (sn->d.out_vel)
    for (int c = 0; c < 3; c++) put(V[c]);
  if (sn->d.out_stress) {
    put(Txx); put(Tyy); put(Tzz); put(Txz); put(Tyz); put(Txy);
  }
  if (sn->d.out_strain) {
    // variables Exx Eyy Ezz Exz Eyz Exy, each the physical strain of its name
    std::vector<float> E(6 * n);
    for (size_t q = 0; q < n; q++) {
      const float l = lam[q], m = mu[q];
      const float E1 = (l + m) / (m * (3.0f * l + 2.0f * m));
      const float E2 = -l / (2.0f * m * (3.0f * l + 2.0f * m));
      const float E3 = 1.0f / m;
      const float E0 = E2 * (Txx[q] + Tyy[q] + Tzz[q]);
      E[0 * n + q] = E0 - (E2 - E1) * Txx[q];
      E[1 * n + q] = E0 - (E2 - E1) * Tyy[q];
      E[2 * n + q] = E0 - (E2 - E1) * Tzz[q];
      E[3 * n + q] = 0.5f * E3 * Txz[q];
      E[4 * n + q] = 0.5f * E3 * Tyz[q];
      E[5 * n + q] = 0.5f * E3 * Txy[q];
    }
    for (int c = 0; c < 6; c++) put(E.data() + (size_t)c * n);
  }
  const long long off = sn->data_begin + (long long)sn->nrec * sn->recsize;
  if (std::fseek(sn->fp, (long)off, SEEK_SET) != 0 || std::fwrite(r.b.data(), 1, r.b.size(), sn->fp) != r.b.size())
    return iofail(CGFD_E_ARG, "snapshot write failed");
  sn->nrec++;
  return CGFD_OK;
}

int cgfd_snap_close(cgfd_snap *sn) {
  if (!sn) return iofail(CGFD_E_ARG, "null snapshot");
  bool ok = true;
  BE n;
  n.u32(sn->nrec);
  ok = std::fseek(sn->fp, 4, SEEK_SET) == 0 && std::fwrite(n.b.data(), 1, 4, sn->fp) == 4;
  ok = (std::fclose(sn->fp) == 0) && ok;
  delete sn;
  return ok ? CGFD_OK : iofail(CGFD_E_ARG, "snapshot close failed");
}

int cgfd_peak_motion_write(const char *path, int nx, int ny, const float *pg, const int *first_physical,
                           const int *count_physical, const int *topo) {
  if (!path || !pg || nx < 1 || ny < 1 || !first_physical || !count_physical || !topo)
    return iofail(CGFD_E_ARG, "cgfd_peak_motion_write: bad arguments");
  static const char *names[15] = {"PGV", "PGVh", "PGVx", "PGVy", "PGVz", "PGA", "PGAh", "PGAx",
                                  "PGAy", "PGAz", "PGD", "PGDh", "PGDx", "PGDy", "PGDz"};
  const long long vsize = 4LL * nx * ny;
  auto header = [&](uint32_t begin0) {
    BE h;
    h.b.insert(h.b.end(), {'C', 'D', 'F', 1});
    h.u32(0);
    h.u32(NC_DIMENSION);
    h.u32(2);
    h.name("j"); h.u32((uint32_t)ny);
    h.name("i"); h.u32((uint32_t)nx);
    h.u32(NC_ATTRIBUTE);
    h.u32(3);
    auto att = [&](const char *nm, const int *v, int n) {
      h.name(nm);
      h.u32(NC_INT);
      h.u32((uint32_t)n);
      for (int q = 0; q < n; q++) h.u32((uint32_t)v[q]);
    };
    att("global_index_of_first_physical_points", first_physical, 2);
    att("count_index_of_physical_points", count_physical, 2);
    att("coords_of_mpi_topo", topo, 3);
    h.u32(NC_VARIABLE);
    h.u32(15);
    uint32_t off = begin0;
    for (int v = 0; v < 15; v++) {
      h.name(names[v]);
      h.u32(2); h.u32(0); h.u32(1);
      h.u32(0); h.u32(0);
      h.u32(NC_FLOAT);
      h.u32((uint32_t)vsize);
      h.u32(off);
      off += (uint32_t)vsize;
    }
    return h.b;
  };
  const std::vector<unsigned char> h0 = header(0);
  if ((long long)h0.size() + 15 * vsize > 0x7fffffffLL) return iofail(CGFD_E_ARG, "peak motion maps too large");
  const std::vector<unsigned char> h1 = header((uint32_t)h0.size());
  BE d;
  d.b.reserve((size_t)(15 * vsize));
  for (long long q = 0; q < 15LL * nx * ny; q++) d.f32(pg[q]);
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return iofail(CGFD_E_ARG, std::string("cannot open ") + path + ": " + std::strerror(errno));
  bool ok = std::fwrite(h1.data(), 1, h1.size(), fp) == h1.size() &&
            std::fwrite(d.b.data(), 1, d.b.size(), fp) == d.b.size();
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? CGFD_OK : iofail(CGFD_E_ARG, std::string("write failed: ") + path);
}

}  // extern "C"
