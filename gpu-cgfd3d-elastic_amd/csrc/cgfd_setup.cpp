// cgfd_setup.cpp -- FP64 host preparation of the el_iso inputs (see cgfd_setup.h).
#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cgfd_el_iso.h"
#include "../../include/cgfd_setup.h"
#include "cgfd_internal.h"

namespace {

constexpr int NG = 3;

struct G3 {
  int ni, nj, nk, nx, ny, nz;
  long sy, sz, n;
  G3(int a, int b, int c) : ni(a), nj(b), nk(c), nx(a + 2 * NG), ny(b + 2 * NG), nz(c + 2 * NG) {
    sy = nx;
    sz = (long)nx * ny;
    n = sz * nz;
  }
  long at(int i, int j, int k) const { return (long)i + (long)j * sy + (long)k * sz; }
};

// run f(k) for k in [0, n) over hardware threads, deterministic per k
template <typename F> void par_for(int n, F f) {
  unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 32u));
  if (n < 4 || nt == 1) {
    for (int k = 0; k < n; k++) f(k);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      for (int k = (int)t; k < n; k += (int)nt) f(k);
    });
  for (auto &x : th) x.join();
}

// centred 7-point operator of the metric evaluation (src/forward/fd_t.cu:164-170)
const double kFdc[7] = {-0.02084, 0.1667, -0.7709, 0.0, 0.7709, -0.1667, 0.02084};

uint64_t mix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

}  // namespace

extern "C" {

int cgfd_grid_cartesian(int ni, int nj, int nk, double dx, double x0, double dy, double y0, double dz, double z0,
                        double *x, double *y, double *z) {
  if (ni < 1 || nj < 1 || nk < 1 || !x || !y || !z) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        long p = g.at(i, j, k);
        x[p] = x0 + (i - NG) * dx;
        y[p] = y0 + (j - NG) * dy;
        z[p] = z0 + (k - NG) * dz;
      }
  return CGFD_OK;
}

int cgfd_grid_topography(const cgfd_topo_desc *d, double *x, double *y, double *z) {
  if (!d || !x || !y || !z || d->ni < 2 || d->nj < 2 || d->nk < 3 || d->dh <= 0 || d->stretch <= 0)
    return CGFD_E_ARG;
  G3 g(d->ni, d->nj, d->nk);
  const double h = d->dh;
  const double Lx = (d->ni - 1) * h, Ly = (d->nj - 1) * h;
  std::mt19937_64 rng(d->seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> hx, hy, ha, hs;
  for (int q = 0; q < d->nhill; q++) {
    hx.push_back(Lx * (0.15 + 0.7 * U(rng)));
    hy.push_back(Ly * (0.15 + 0.7 * U(rng)));
    ha.push_back(d->amp_min + (d->amp_max - d->amp_min) * U(rng));
    hs.push_back(d->sig_min + (d->sig_max - d->sig_min) * U(rng));
  }
  // Fourier perturbation: modes with wavelengths between min_wavelength and the domain size
  struct Mode { double kx, ky, ph, a; };
  std::vector<Mode> modes;
  const int nmode = 48;
  for (int q = 0; q < nmode; q++) {
    double lam = d->min_wavelength * std::pow(std::max(Lx, Ly) / d->min_wavelength, U(rng));
    double th = 2 * M_PI * U(rng);
    double kk = 2 * M_PI / lam;
    modes.push_back({kk * std::cos(th), kk * std::sin(th), 2 * M_PI * U(rng), 1.0});
  }
  // rms of a sum of unit-amplitude cosines with random phases is sqrt(n/2)
  const double mscale = d->rough_rms / std::sqrt(nmode / 2.0);
  auto surf = [&](double xx, double yy) {
    double s = 0;
    for (size_t q = 0; q < hx.size(); q++) {
      double r2 = (xx - hx[q]) * (xx - hx[q]) + (yy - hy[q]) * (yy - hy[q]);
      s += ha[q] * std::exp(-r2 / (2 * hs[q] * hs[q]));
    }
    for (const Mode &m : modes) s += mscale * std::cos(m.kx * xx + m.ky * yy + m.ph);
    return s;
  };
  // depth below the surface in units of h for unit index m (m = 0 at the surface):
  // spacing grows linearly from 1 (m = 0) to stretch (m = nk-1), extended analytically
  const int nk = d->nk;
  auto spacing = [&](double m) { return 1.0 + (d->stretch - 1.0) * m / (nk - 1); };
  auto depth = [&](double m) {  // integral of spacing from 0 to m (exact for linear spacing)
    return m + (d->stretch - 1.0) * m * m / (2.0 * (nk - 1));
  };
  (void)spacing;
  const double Dref = depth(nk - 1) * h;   // flat-surface thickness
  const double zbot = -Dref;
  std::vector<double> top((size_t)g.nx * g.ny);
  for (int j = 0; j < g.ny; j++)
    for (int i = 0; i < g.nx; i++) top[(size_t)j * g.nx + i] = surf((i - NG) * h, (j - NG) * h);
  par_for(g.nz, [&](int k) {
    const double m = (double)(nk - 1) - (k - NG);   // unit depth index below the surface
    const double frac = depth(m) * h / Dref;        // 0 at the surface, 1 at the bottom
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        long p = g.at(i, j, k);
        double zt = top[(size_t)j * g.nx + i];
        x[p] = (i - NG) * h;
        y[p] = (j - NG) * h;
        z[p] = zt - (zt - zbot) * frac;
      }
  });
  return CGFD_OK;
}

int cgfd_metric_calc(int ni, int nj, int nk, const double *x, const double *y, const double *z, double *metric) {
  if (ni < 1 || nj < 1 || nk < 1 || !x || !y || !z || !metric) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  const long n = g.n;
  std::fill(metric, metric + 10 * n, 0.0);
  par_for(nk, [&](int kk) {
    const int k = kk + NG;
    for (int j = NG; j < NG + nj; j++)
      for (int i = NG; i < NG + ni; i++) {
        const long p = g.at(i, j, k);
        double v[3][3];  // v[a][c]: derivative of coordinate c along axis a
        const long st[3] = {1, g.sy, g.sz};
        const double *cs[3] = {x, y, z};
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int q = 0; q < 7; q++) s += kFdc[q] * cs[c][p + (q - 3) * st[a]];
            v[a][c] = s;
          }
        auto cross = [](const double *A, const double *B, double *C) {
          C[0] = A[1] * B[2] - A[2] * B[1];
          C[1] = A[2] * B[0] - A[0] * B[2];
          C[2] = A[0] * B[1] - A[1] * B[0];
        };
        double c12[3], c23[3], c31[3];
        cross(v[0], v[1], c12);
        cross(v[1], v[2], c23);
        cross(v[2], v[0], c31);
        const double jac = c12[0] * v[2][0] + c12[1] * v[2][1] + c12[2] * v[2][2];
        metric[0 * n + p] = jac;
        for (int c = 0; c < 3; c++) {
          metric[(1 + c) * n + p] = c23[c] / jac;
          metric[(4 + c) * n + p] = c31[c] / jac;
          metric[(7 + c) * n + p] = c12[c] / jac;
        }
      }
  });
  // ghost extrapolation 2 f(edge) - f(mirror), in the order x1,x2,y1,y2,z1,z2
  const int ni1 = NG, ni2 = NG + ni - 1, nj1 = NG, nj2 = NG + nj - 1, nk1 = NG, nk2 = NG + nk - 1;
  for (int c = 0; c < 10; c++) {
    double *a = metric + (long)c * n;
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++) {
        for (int i = 0; i < ni1; i++) a[g.at(i, j, k)] = 2 * a[g.at(ni1, j, k)] - a[g.at(2 * ni1 - i, j, k)];
        for (int i = ni2 + 1; i < g.nx; i++) a[g.at(i, j, k)] = 2 * a[g.at(ni2, j, k)] - a[g.at(2 * ni2 - i, j, k)];
      }
    for (int k = 0; k < g.nz; k++) {
      for (int j = 0; j < nj1; j++)
        for (int i = 0; i < g.nx; i++) a[g.at(i, j, k)] = 2 * a[g.at(i, nj1, k)] - a[g.at(i, 2 * nj1 - j, k)];
      for (int j = nj2 + 1; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) a[g.at(i, j, k)] = 2 * a[g.at(i, nj2, k)] - a[g.at(i, 2 * nj2 - j, k)];
    }
    for (int k = 0; k < nk1; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) a[g.at(i, j, k)] = 2 * a[g.at(i, j, nk1)] - a[g.at(i, j, 2 * nk1 - k)];
    for (int k = nk2 + 1; k < g.nz; k++)
      for (int j = 0; j < g.ny; j++)
        for (int i = 0; i < g.nx; i++) a[g.at(i, j, k)] = 2 * a[g.at(i, j, nk2)] - a[g.at(i, j, 2 * nk2 - k)];
  }
  return CGFD_OK;
}

double cgfd_dt_estimate(int ni, int nj, int nk, const double *x, const double *y, const double *z, const double *lam,
                        const double *mu, const double *rho, double cfl) {
  if (ni < 2 || nj < 2 || nk < 2 || !x || !y || !z || !lam || !mu || !rho) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  std::vector<double> best(g.nz, 1e30);
  par_for(nk - 1, [&](int kk) {
    const int k = kk + NG;
    double b = 1e30;
    for (int j = NG; j < NG + nj - 1; j++)
      for (int i = NG; i < NG + ni - 1; i++) {
        const long p = g.at(i, j, k);
        const double vp = std::sqrt((lam[p] + 2 * mu[p]) / rho[p]);
        const double p0[3] = {x[p], y[p], z[p]};
        double Le = 1e30;
        for (int kc = -1; kc <= 1; kc += 2)
          for (int jc = -1; jc <= 1; jc += 2)
            for (int ic = -1; ic <= 1; ic += 2) {
              const long a = p - ic, bq = p - jc * g.sy, c = p - kc * g.sz;
              const double p1[3] = {x[a], y[a], z[a]}, p2[3] = {x[bq], y[bq], z[bq]}, p3[3] = {x[c], y[c], z[c]};
              const double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
              const double w[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
              const double nrm[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
              const double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
              const double dist = std::fabs(nrm[0] * (p0[0] - p1[0]) + nrm[1] * (p0[1] - p1[1]) +
                                            nrm[2] * (p0[2] - p1[2])) / nn;
              Le = std::min(Le, dist);
            }
        b = std::min(b, cfl * Le / vp);
      }
    best[k] = b;
  });
  return *std::min_element(best.begin(), best.end());
}

// ---- media from a layer (md3lay) file ----------------------------------------------
// Restatement of the reference's layered-model path for elastic_isotropic in
// float, as there: the interface file reader (read_interface_file,
// src/media/media_read_file.cpp:22-260) and the "loc" discretization
// (media_layer2model_el_iso, parametrization_layer_el_iso_loc,
// AssignLayerMediaPara2Point, CalPointValue_layer, media_layer2model.cpp:
// 121-171, 401-533, 1313-1362; BilinearInterpolation, findLast/First-
// GreaterEqualIndex, isEqual, media_utility.cpp:25-180).
namespace {

struct Md3Lay {
  int ni = 0, nx = 0, ny = 0;
  float minx = 0, miny = 0, dx = 0, dy = 0;
  // [layer][ny][nx]: elevation and rho, vp, vs with their gradient and power
  std::vector<float> elev, rho, rho_g, rho_p, vp, vp_g, vp_p, vs, vs_g, vs_p;
};

// The whitespace-separated tokens of a media file; lines starting with '#'
// and empty lines are comments.
int read_tokens(const char *path, std::vector<std::string> &tok, std::string &why) {
  FILE *f = std::fopen(path, "r");
  if (!f) {
    why = std::string("cannot open ") + path;
    return CGFD_E_ARG;
  }
  std::string text;
  char line[8192];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    text += line;
    text += ' ';
  }
  std::fclose(f);
  size_t i = 0;
  while (i < text.size()) {
    while (i < text.size() && std::isspace((unsigned char)text[i])) i++;
    size_t j = i;
    while (j < text.size() && !std::isspace((unsigned char)text[j])) j++;
    if (j > i) tok.emplace_back(text.substr(i, j - i));
    i = j;
  }
  return CGFD_OK;
}

int md3lay_read(const char *path, Md3Lay &L, std::string &why) {
  std::vector<std::string> tok;
  if (read_tokens(path, tok, why) != CGFD_OK) return CGFD_E_ARG;
  size_t t = 0;
  auto next = [&](std::string &out) {
    if (t >= tok.size()) return false;
    out = tok[t++];
    return true;
  };
  std::string w;
  if (!next(w)) { why = "no media_type"; return CGFD_E_ARG; }
  if (w != "elastic_isotropic") {
    why = "media_type " + w + " is not elastic_isotropic";
    return CGFD_E_ARG;
  }
  auto num_i = [&](int &v) { std::string s; if (!next(s)) return false; v = std::atoi(s.c_str()); return true; };
  auto num_f = [&](float &v) { std::string s; if (!next(s)) return false; v = std::strtof(s.c_str(), nullptr); return true; };
  if (!num_i(L.ni) || L.ni < 1) { why = "bad number of interfaces"; return CGFD_E_ARG; }
  if (!num_i(L.nx) || !num_i(L.ny) || !num_f(L.minx) || !num_f(L.miny) || !num_f(L.dx) || !num_f(L.dy)) {
    why = "bad interface mesh line";
    return CGFD_E_ARG;
  }
  if (L.nx < 2 || L.ny < 2) { why = "interface mesh needs NX, NY >= 2"; return CGFD_E_ARG; }
  if (!(L.dx > 0 && L.dy > 0 && std::isfinite(L.dx) && std::isfinite(L.dy) && std::isfinite(L.minx) &&
        std::isfinite(L.miny))) {
    why = "interface mesh needs finite MINX, MINY and DX, DY > 0";
    return CGFD_E_ARG;
  }
  // the header's counts size nothing before the file is known to hold them
  if (10.0 * L.ni * L.nx * L.ny > (double)(tok.size() - t)) {
    why = "insufficient data in the interface file";
    return CGFD_E_ARG;
  }
  const size_t n = (size_t)L.ni * L.nx * L.ny;
  std::vector<float> *cols[10] = {&L.elev, &L.rho, &L.rho_g, &L.rho_p, &L.vp, &L.vp_g, &L.vp_p, &L.vs, &L.vs_g, &L.vs_p};
  for (auto *c : cols) c->resize(n);
  for (size_t q = 0; q < n; q++)
    for (auto *c : cols)
      if (!num_f((*c)[q])) {
        why = "insufficient data in the interface file";
        return CGFD_E_ARG;
      } else if (!std::isfinite((*c)[q])) {
        why = "the interface file holds a value that is not a finite number";
        return CGFD_E_ARG;
      }
  return CGFD_OK;
}

// Bilinear evaluation on an interface mesh of nx x ny nodes from (minx, miny)
// with spacing (dx, dy) (BilinearInterpolation, media_utility.cpp).
struct LayerEval {
  int nx, ny;
  std::vector<float> xv, yv;
  LayerEval(int nx_, int ny_, float minx, float miny, float dx, float dy) : nx(nx_), ny(ny_), xv(nx_), yv(ny_) {
    for (int i = 0; i < nx; i++) xv[i] = minx + dx * i;
    for (int j = 0; j < ny; j++) yv[j] = miny + dy * j;
  }
  // bilinear value of one interface's field at (xq, yq); FLT_MAX left of the mesh
  float bilinear(const float *v, float xq, float yq) const {
    const int ix = (int)((xq - xv[0]) / (xv[1] - xv[0]));
    const int iy = (int)((yq - yv[0]) / (yv[1] - yv[0]));
    if (ix < 0 || iy < 0) return FLT_MAX;
    const int idx = iy * nx + ix;
    if (ix == nx - 1 || iy == ny - 1) return v[idx];
    const float area = (xv[ix + 1] - xv[ix]) * (yv[iy + 1] - yv[iy]);
    float vq = v[idx] * (xv[ix + 1] - xq) * (yv[iy + 1] - yq) + v[idx + 1] * (xq - xv[ix]) * (yv[iy + 1] - yq) +
               v[idx + nx] * (xv[ix + 1] - xq) * (yq - yv[iy]) + v[idx + 1 + nx] * (xq - xv[ix]) * (yq - yv[iy]);
    return vq / area;
  }
};

// the deepest interface at or above z, -1 above the first (elevations listed
// from the top; ties go to the later interface)
int last_ge(float value, const std::vector<float> &x) {
  if (value > x[0]) return -1;
  float dist = FLT_MAX;
  int idx = -1;
  for (size_t i = 0; i < x.size(); i++) {
    if (x[i] == -FLT_MAX) continue;
    const float nd = x[i] - value;
    if (nd < 0) break;
    if (nd >= 0 && nd <= dist) { dist = nd; idx = (int)i; }
  }
  return idx;
}

int first_ge(float value, const std::vector<float> &x) {
  float dist = FLT_MAX;
  int idx = -1;
  for (size_t i = 0; i < x.size(); i++) {
    if (x[i] == -FLT_MAX) continue;
    const float nd = x[i] - value;
    if (nd >= 0 && nd <= dist) {
      if (nd <= 1e-6) return (int)i;
      dist = nd;
      idx = (int)i;
    }
  }
  return idx;
}

// The reference's isEqual(a, b) is `abs(a - b) < FLT_EPSILON` with no std::
// qualification, which binds C's int abs: it holds whenever |a - b| < 1
// (the difference truncates to 0).  Written as that range, so a difference
// beyond the int range (or not a number) is no integer conversion.
bool ref_is_equal(float a, float b) {
  const float d = a - b;
  return d > -1.0f && d < 1.0f;
}

}  // namespace

namespace {

// rho, vp, vs of the layered model at one point (AssignLayerMediaPara2Point +
// CalPointValue_layer for elastic_isotropic); 1: outside the interface mesh,
// 2: vp <= sqrt(2) vs, 3: interpolated interfaces not finite
int layer_point(const Md3Lay &L, const LayerEval &E, float ax, float ay, float az, std::vector<float> &elev, float &r,
                float &vp, float &vs) {
  const float maxx = L.minx + (L.nx - 1) * L.dx, maxy = L.miny + (L.ny - 1) * L.dy;
  if (ax < L.minx || ax > maxx || ay < L.miny || ay > maxy) return 1;
  const size_t slice = (size_t)L.nx * L.ny;
  for (int l = 0; l < L.ni; l++) elev[l] = E.bilinear(&L.elev[l * slice], ax, ay);
  int mi = last_ge(az, elev);
  auto field = [&](const std::vector<float> &v, int l) { return E.bilinear(&v[l * slice], ax, ay); };
  if (mi == -1) {
    // above the top interface: the values at the top of the first layer
    mi = last_ge(elev[0], elev);
    if (mi < 0) return 3;  // interpolated elevations not finite (mesh values overflowed)
    r = field(L.rho, mi);
    vp = field(L.vp, mi);
    vs = field(L.vs, mi);
    return 0;
  }
  auto at = [&](int l, float dz, float &ro, float &p_, float &s_) {
    // pow in double, as the reference's unqualified pow on floats
    p_ = (float)(field(L.vp, l) + std::pow((double)dz, (double)field(L.vp_p, l)) * field(L.vp_g, l));
    s_ = (float)(field(L.vs, l) + std::pow((double)dz, (double)field(L.vs_p, l)) * field(L.vs_g, l));
    ro = (float)(field(L.rho, l) + std::pow((double)dz, (double)field(L.rho_p, l)) * field(L.rho_g, l));
  };
  const float dz = elev[mi] - az;
  at(mi, dz, r, vp, vs);
  if (vp <= vs * std::sqrt(2.0)) return 2;
  if (ref_is_equal(dz, 0.0f)) {
    // on an interface: the mean of the layers below and above it
    const int up = first_ge(az, elev) - 1;
    if (up >= 0) {
      float r2, p2, s2;
      at(up, elev[up] - az, r2, p2, s2);
      r = (float)((r + r2) / 2.0);
      vp = (float)((vp + p2) / 2.0);
      vs = (float)((vs + s2) / 2.0);
    }
  }
  return 0;
}

struct P3 {
  float x, y, z;
};
P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 operator*(P3 a, float p) { return {a.x * p, a.y * p, a.z * p}; }
P3 operator/(P3 a, float p) { return {a.x / p, a.y / p, a.z / p}; }

}  // namespace

extern "C++" {
namespace {

// The equivalent-medium pass shared by layer and grid media files ("har" /
// "ari"; parametrization_{layer,grid}_el_iso_{har,ari},
// media_layer2model.cpp:1963-2174, media_grid2model.cpp:1819-2079, with
// GenerateHalfGrid, GenerateHalfMesh, MeshSubdivide, NumOfValues,
// media_utility.cpp:404-600): cell centres of the grid (the mean of each
// cell's 8 nodes; the last plane of each axis keeps the node) are marked with
// a material number; a node whose dual cell (corners at the 8 surrounding
// centres) meets more than one material takes the average of the 9^3 points
// that subdivide that cell: harmonic for kappa and mu with arithmetic rho, or
// arithmetic for lam, mu and rho.  Grid media files (grid_rules) leave out
// the points on a layer boundary and divide the harmonic rho sum in float.
template <typename Mark, typename Point>
int equivalent_medium(const G3 &g, const double *x, const double *y, const double *z, bool har, int nmat,
                      bool grid_rules, Mark mark, Point point, int nelev, double *lam,
                      double *mu, double *rho) {
  const long n = g.n;
  std::vector<float> hx(n), hy(n), hz(n);
  std::vector<int> mat(n);
  par_for(g.nz, [&](int k) {
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        const long p = g.at(i, j, k);
        if (i < g.nx - 1 && j < g.ny - 1 && k < g.nz - 1) {
          const long q[8] = {p, p + 1, p + g.sy + 1, p + g.sy, p + g.sz, p + g.sz + 1, p + g.sz + g.sy + 1, p + g.sy + g.sz};
          float sx = 0, sy = 0, sz = 0;
          for (int c = 0; c < 8; c++) {
            sx += (float)x[q[c]];
            sy += (float)y[q[c]];
            sz += (float)z[q[c]];
          }
          hx[p] = (float)(sx / 8.0);
          hy[p] = (float)(sy / 8.0);
          hz[p] = (float)(sz / 8.0);
        } else {
          hx[p] = (float)x[p];
          hy[p] = (float)y[p];
          hz[p] = (float)z[p];
        }
      }
  });
  par_for(g.nz, [&](int k) {
    std::vector<float> elev(nelev);
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        const long p = g.at(i, j, k);
        mat[p] = mark(hx[p], hy[p], hz[p], elev);
      }
  });
  constexpr int kSub = 8;  // subdivisions per direction of a dual cell
  std::vector<int> jerr(g.ny, 0);
  par_for(g.ny, [&](int j) {
    if (j < 1 || j > g.ny - 2) return;
    std::vector<float> elev(nelev);
    std::vector<int> tab(nmat);
    for (int k = 1; k < g.nz - 1 && !jerr[j]; k++)
      for (int i = 1; i < g.nx - 1; i++) {
        const long p = g.at(i, j, k);
        const long q[8] = {p - 1 - g.sy - g.sz, p - g.sy - g.sz, p - g.sz, p - 1 - g.sz,
                           p - 1 - g.sy,        p - g.sy,        p,        p - 1};
        // distinct materials among the 8 corners (a corner above every interface: none)
        std::fill(tab.begin(), tab.end(), 0);
        int distinct = 0;
        bool above = false;
        for (int c = 0; c < 8; c++) {
          const int v = mat[q[c]];
          if (v == -1) { above = true; break; }
          if (tab[v]++ == 0) distinct++;
        }
        if (above || distinct <= 1) continue;
        P3 V[8];
        for (int c = 0; c < 8; c++) V[c] = {hx[q[c]], hy[q[c]], hz[q[c]]};
        float a_rho = 0, a_lam = 0, a_mu = 0, h_kappa = 0, h_mu = 0;
        const float ng = (float)kSub;
        constexpr int nsg = (kSub + 1) * (kSub + 1) * (kSub + 1);
        int used = nsg;
        int bad = 0;
        for (int kk = 0; kk <= kSub && !bad; kk++)
          for (int jj = 0; jj <= kSub && !bad; jj++)
            for (int ii = 0; ii <= kSub; ii++) {
              const P3 lx00 = V[0] + (V[1] - V[0]) / ng * (float)ii;
              const P3 lx10 = V[3] + (V[2] - V[3]) / ng * (float)ii;
              const P3 lx01 = V[4] + (V[5] - V[4]) / ng * (float)ii;
              const P3 lx11 = V[7] + (V[6] - V[7]) / ng * (float)ii;
              const P3 lxy0 = lx00 + (lx10 - lx00) / ng * (float)jj;
              const P3 lxy1 = lx01 + (lx11 - lx01) / ng * (float)jj;
              const P3 s = lxy0 + (lxy1 - lxy0) / ng * (float)kk;
              float r, vp, vs;
              bool on;
              if ((bad = point(s.x, s.y, s.z, elev, r, vp, vs, on))) break;
              if (grid_rules && on) {
                used--;
                continue;
              }
              const float m = vs * vs * r;
              const float lm = (float)((double)(vp * vp * r) - 2.0 * (double)m);
              a_rho += r;
              if (har) {
                h_kappa = (float)((double)h_kappa + 1.0 / ((double)lm + 2.0 / 3.0 * (double)m));
                h_mu = (float)((double)h_mu + 1.0 / (double)m);
              } else {
                a_lam += lm;
                a_mu += m;
              }
            }
        if (bad) { jerr[j] = bad; break; }
        if (har) {
          const float hm = (float)(used * 1.0 / (double)h_mu), hk = (float)(used * 1.0 / (double)h_kappa);
          lam[p] = (float)((double)hk - 2.0 / 3.0 * (double)hm);
          mu[p] = hm;
          rho[p] = grid_rules ? a_rho / (float)used : (float)((double)a_rho / (used * 1.0));
        } else {
          lam[p] = (float)((double)a_lam / (used * 1.0));
          mu[p] = (float)((double)a_mu / (used * 1.0));
          rho[p] = (float)((double)a_rho / (used * 1.0));
        }
      }
  });
  for (int j = 0; j < g.ny; j++)
    if (jerr[j]) {
      cgfd_internal_set_error(jerr[j] == 1   ? "grid point outside the interface mesh"
                              : jerr[j] == 2 ? "vp must be larger than sqrt(2) vs in the media file"
                                             : "the media file gives surfaces that are not finite at a grid point");
      return CGFD_E_ARG;
    }
  return CGFD_OK;
}

}  // namespace
}  // extern "C++"

namespace {

// A grid media file in the reference's md3grd format, elastic_isotropic,
// holding only the window of the mesh that covers the grid
// (read_grid_file, media_read_file.cpp:670-925).
struct Md3Grd {
  int nl = 0, ni = 0;  // layers; surfaces (the sum of the per-layer counts)
  int nx = 0, ny = 0;  // window
  float minx = 0, miny = 0, dx = 0, dy = 0;
  std::vector<int> top;       // surface index of each layer's top
  std::vector<char> starts;   // [ni]: surface that is the top of a layer below the first
  std::vector<float> elev, rho, vp, vs;  // [surface][ny][nx]
};

int md3grd_read(const char *path, float xmin, float xmax, float ymin, float ymax, Md3Grd &G, std::string &why) {
  std::vector<std::string> tok;
  if (read_tokens(path, tok, why) != CGFD_OK) return CGFD_E_ARG;
  size_t t = 0;
  auto next = [&](std::string &out) {
    if (t >= tok.size()) return false;
    out = tok[t++];
    return true;
  };
  std::string w;
  if (!next(w)) { why = "no media_type"; return CGFD_E_ARG; }
  if (w != "elastic_isotropic") {
    why = "media_type " + w + " is not elastic_isotropic";
    return CGFD_E_ARG;
  }
  auto num_i = [&](int &v) { std::string s; if (!next(s)) return false; v = std::atoi(s.c_str()); return true; };
  auto num_f = [&](float &v) { std::string s; if (!next(s)) return false; v = std::strtof(s.c_str(), nullptr); return true; };
  if (!num_i(G.nl) || G.nl < 1 || (size_t)G.nl > tok.size() - t) { why = "bad number of layers"; return CGFD_E_ARG; }
  std::vector<int> ngz(G.nl);
  double nsurf = 0;
  for (int l = 0; l < G.nl; l++) {
    if (!num_i(ngz[l]) || ngz[l] < 1) { why = "bad number of surfaces in a layer"; return CGFD_E_ARG; }
    nsurf += ngz[l];
    if (nsurf > (double)(tok.size() - t)) { why = "insufficient data in the media file"; return CGFD_E_ARG; }
    G.top.push_back(G.ni);
    G.ni += ngz[l];
  }
  G.starts.assign(G.ni, 0);
  for (int l = 1; l < G.nl; l++) G.starts[G.top[l]] = 1;
  int NX, NY;
  float MINX, MINY;
  if (!num_i(NX) || !num_i(NY) || !num_f(MINX) || !num_f(MINY) || !num_f(G.dx) || !num_f(G.dy)) {
    why = "bad media mesh line";
    return CGFD_E_ARG;
  }
  if (NX < 2 || NY < 2) { why = "media mesh needs NX, NY >= 2"; return CGFD_E_ARG; }
  if (!(G.dx > 0 && G.dy > 0 && std::isfinite(G.dx) && std::isfinite(G.dy) && std::isfinite(MINX) &&
        std::isfinite(MINY))) {
    why = "media mesh needs finite MINX, MINY and DX, DY > 0";
    return CGFD_E_ARG;
  }
  if (4.0 * G.ni * NX * NY > (double)(tok.size() - t)) { why = "insufficient data in the media file"; return CGFD_E_ARG; }
  // the window [ix0, ix1] x [iy0, iy1] of mesh nodes that covers the grid
  const float fx0 = (xmin - MINX) / G.dx, fy0 = (ymin - MINY) / G.dy;
  const float fx1 = std::ceil((xmax - MINX) / G.dx), fy1 = std::ceil((ymax - MINY) / G.dy);
  if (!(fx0 >= 0) || !(fx1 <= NX - 1)) { why = "the media mesh does not cover the grid in x"; return CGFD_E_ARG; }
  if (!(fy0 >= 0) || !(fy1 <= NY - 1)) { why = "the media mesh does not cover the grid in y"; return CGFD_E_ARG; }
  const int ix0 = (int)fx0, iy0 = (int)fy0, ix1 = (int)fx1, iy1 = (int)fy1;
  G.nx = ix1 - ix0 + 1;
  G.ny = iy1 - iy0 + 1;
  G.minx = (float)ix0 * G.dx + MINX;
  G.miny = (float)iy0 * G.dy + MINY;
  const size_t slice = (size_t)G.nx * G.ny, n = slice * G.ni;
  std::vector<float> *cols[4] = {&G.elev, &G.rho, &G.vp, &G.vs};
  for (auto *c : cols) c->resize(n);
  for (int s = 0; s < G.ni; s++)
    for (int iy = 0; iy < NY; iy++)
      for (int ix = 0; ix < NX; ix++) {
        const bool keep = ix >= ix0 && ix <= ix1 && iy >= iy0 && iy <= iy1;
        const size_t q = s * slice + (size_t)(iy - iy0) * G.nx + (ix - ix0);
        for (auto *c : cols) {
          float v;
          if (!num_f(v)) { why = "insufficient data in the media file"; return CGFD_E_ARG; }
          if (!std::isfinite(v)) { why = "the media file holds a value that is not a finite number"; return CGFD_E_ARG; }
          if (keep) (*c)[q] = v;
        }
      }
  // the last surface of a layer and the first of the next must coincide
  // (checkGridData, media_read_file.cpp:1243-1266)
  for (int l = 1; l < G.nl; l++)
    for (size_t q = 0; q < slice; q++)
      if (G.elev[(G.top[l] - 1) * slice + q] != G.elev[G.top[l] * slice + q]) {
        why = "the last surface of a layer differs from the first of the next";
        return CGFD_E_ARG;
      }
  return CGFD_OK;
}

// rho, vp, vs of the grid medium at one point, and whether the point lies on
// a layer boundary (AssignGridMediaPara2Point + CalPointValue_grid for
// elastic_isotropic, media_grid2model.cpp:424-560): linear in z between the
// two surfaces around the point; above the first surface the values at the
// top; below the last those of the last; within 1 m under the top surface of
// a layer below the first, the mean of that surface and the one above it.
// 1: outside the media mesh, 2: vp <= sqrt(2) vs, 3: interpolated surfaces
// not finite.
int grid_point(const Md3Grd &G, const LayerEval &E, float ax, float ay, float az, std::vector<float> &elev,
               float &r, float &vp, float &vs, bool &on) {
  const float maxx = G.minx + (G.nx - 1) * G.dx, maxy = G.miny + (G.ny - 1) * G.dy;
  if (ax < G.minx || ax > maxx || ay < G.miny || ay > maxy) return 1;
  const size_t slice = (size_t)G.nx * G.ny;
  for (int s = 0; s < G.ni; s++) elev[s] = E.bilinear(&G.elev[s * slice], ax, ay);
  int mi = last_ge(az, elev);
  on = mi >= 0 && G.starts[mi] && ref_is_equal(az, elev[mi]);
  auto field = [&](const std::vector<float> &v, int s) { return E.bilinear(&v[s * slice], ax, ay); };
  if (mi == -1 || mi == G.ni - 1) {
    if (mi == -1) mi = last_ge(elev[0], elev);
    if (mi < 0) return 3;  // interpolated elevations not finite (mesh values overflowed)
    r = field(G.rho, mi);
    vp = field(G.vp, mi);
    vs = field(G.vs, mi);
    return 0;
  }
  float dis_r = 0.5f;
  if (std::fabs(elev[mi + 1] - elev[mi]) > 1e-6) dis_r = (az - elev[mi]) / (elev[mi + 1] - elev[mi]);
  if (mi > 0 && ref_is_equal(elev[mi], az) && G.starts[mi]) {
    mi--;
    dis_r = 0.5f;
  }
  const float p0 = field(G.vp, mi), s0 = field(G.vs, mi), r0 = field(G.rho, mi);
  const float p1 = field(G.vp, mi + 1), s1 = field(G.vs, mi + 1), r1 = field(G.rho, mi + 1);
  vp = p0 + (p1 - p0) * dis_r;
  vs = s0 + (s1 - s0) * dis_r;
  r = r0 + (r1 - r0) * dis_r;
  if (vp <= std::sqrt(2.0) * vs) return 2;
  return 0;
}

}  // namespace

// The file readers run on input read from disk: an allocation that fails or
// a library exception ends the call with an error code instead of crossing
// the C ABI.
extern "C++" {
namespace {
template <typename F> int guarded(F f) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    cgfd_internal_set_error("out of host memory");
    return CGFD_E_NOMEM;
  } catch (const std::exception &e) {
    cgfd_internal_set_error(e.what());
    return CGFD_E_ARG;
  }
}
}  // namespace
}  // extern "C++"

static int cgfd_media_layer_file_body(int ni, int nj, int nk, const double *x, const double *y, const double *z, const char *path,
                          const char *method, double *lam, double *mu, double *rho) {
  if (ni < 1 || nj < 1 || nk < 1 || !x || !y || !z || !path || !method || !lam || !mu || !rho) return CGFD_E_ARG;
  const bool loc = !std::strcmp(method, "loc"), har = !std::strcmp(method, "har"), ari = !std::strcmp(method, "ari");
  if (!loc && !har && !ari) {
    cgfd_internal_set_error("equivalent-medium method must be \"loc\", \"har\" or \"ari\"");
    return CGFD_E_ARG;
  }
  Md3Lay L;
  std::string why;
  if (md3lay_read(path, L, why) != CGFD_OK) {
    cgfd_internal_set_error(why.c_str());
    return CGFD_E_ARG;
  }
  const G3 g(ni, nj, nk);
  const LayerEval E(L.nx, L.ny, L.minx, L.miny, L.dx, L.dy);
  std::vector<int> err(g.nz, 0);
  auto fail = [&]() {
    for (int k = 0; k < g.nz; k++)
      if (err[k]) {
        cgfd_internal_set_error(err[k] == 1   ? "grid point outside the interface mesh"
                                : err[k] == 2 ? "vp must be larger than sqrt(2) vs in the layer file"
                                              : "the layer file gives interfaces that are not finite at a grid point");
        return true;
      }
    return false;
  };
  // "loc": every point's own layer (parametrization_layer_el_iso_loc)
  par_for(g.nz, [&](int k) {
    std::vector<float> elev(L.ni);
    for (int j = 0; j < g.ny && !err[k]; j++)
      for (int i = 0; i < g.nx; i++) {
        const long p = g.at(i, j, k);
        float r, vp, vs;
        if ((err[k] = layer_point(L, E, (float)x[p], (float)y[p], (float)z[p], elev, r, vp, vs))) break;
        const float m = vs * vs * r;
        mu[p] = m;
        lam[p] = (float)((double)(vp * vp * r) - 2.0 * (double)m);
        rho[p] = r;
      }
  });
  if (fail()) return CGFD_E_ARG;
  if (loc) return CGFD_OK;

  return equivalent_medium(g, x, y, z, har, L.ni, false,
                           [&](float hx, float hy, float hz, std::vector<float> &elev) {
                             const size_t slice = (size_t)L.nx * L.ny;
                             for (int l = 0; l < L.ni; l++) elev[l] = E.bilinear(&L.elev[l * slice], hx, hy);
                             int mi = last_ge(hz, elev);
                             if (mi == -1) mi = last_ge(elev[0], elev);
                             return mi;
                           },
                           [&](float px, float py, float pz, std::vector<float> &elev, float &r, float &vp, float &vs,
                               bool &on) {
                             on = false;
                             return layer_point(L, E, px, py, pz, elev, r, vp, vs);
                           },
                           L.ni, lam, mu, rho);
}

static int cgfd_media_grid_file_body(int ni, int nj, int nk, const double *x, const double *y, const double *z, const char *path,
                         const char *method, double *lam, double *mu, double *rho) {
  if (ni < 1 || nj < 1 || nk < 1 || !x || !y || !z || !path || !method || !lam || !mu || !rho) return CGFD_E_ARG;
  const bool loc = !std::strcmp(method, "loc"), har = !std::strcmp(method, "har"), ari = !std::strcmp(method, "ari");
  if (!loc && !har && !ari) {
    cgfd_internal_set_error("equivalent-medium method must be \"loc\", \"har\" or \"ari\"");
    return CGFD_E_ARG;
  }
  const G3 g(ni, nj, nk);
  // the grid's extent in x and y, ghosts included, in float
  float xmin = (float)x[0], xmax = xmin, ymin = (float)y[0], ymax = ymin;
  for (long p = 0; p < g.n; p++) {
    xmin = std::min(xmin, (float)x[p]);
    xmax = std::max(xmax, (float)x[p]);
    ymin = std::min(ymin, (float)y[p]);
    ymax = std::max(ymax, (float)y[p]);
  }
  Md3Grd G;
  std::string why;
  if (md3grd_read(path, xmin, xmax, ymin, ymax, G, why) != CGFD_OK) {
    cgfd_internal_set_error(why.c_str());
    return CGFD_E_ARG;
  }
  const LayerEval E(G.nx, G.ny, G.minx, G.miny, G.dx, G.dy);
  std::vector<int> err(g.nz, 0);
  // "loc": every point's own value (parametrization_grid_el_iso_loc,
  // media_grid2model.cpp:1067-1110)
  par_for(g.nz, [&](int k) {
    std::vector<float> elev(G.ni);
    for (int j = 0; j < g.ny && !err[k]; j++)
      for (int i = 0; i < g.nx; i++) {
        const long p = g.at(i, j, k);
        float r, vp, vs;
        bool on;
        if ((err[k] = grid_point(G, E, (float)x[p], (float)y[p], (float)z[p], elev, r, vp, vs, on))) break;
        const float m = vs * vs * r;
        mu[p] = m;
        lam[p] = (float)((double)(vp * vp * r) - 2.0 * (double)m);
        rho[p] = r;
      }
  });
  for (int k = 0; k < g.nz; k++)
    if (err[k]) {
      cgfd_internal_set_error(err[k] == 1   ? "grid point outside the media mesh"
                              : err[k] == 2 ? "vp must be larger than sqrt(2) vs in the media file"
                                            : "the media file gives surfaces that are not finite at a grid point");
      return CGFD_E_ARG;
    }
  // one layer: nothing to average (the reference falls back to "loc")
  if (loc || G.nl < 2) return CGFD_OK;

  // layers are numbered by the top surface of each (LayerNumberAtPoint,
  // media_grid2model.cpp:1269-1307)
  const size_t slice = (size_t)G.nx * G.ny;
  return equivalent_medium(g, x, y, z, har, G.nl, true,
                           [&](float hx, float hy, float hz, std::vector<float> &elev) {
                             elev.resize(G.nl);
                             for (int l = 0; l < G.nl; l++) elev[l] = E.bilinear(&G.elev[G.top[l] * slice], hx, hy);
                             int mi = last_ge(hz, elev);
                             if (mi == -1) mi = last_ge(elev[0], elev);
                             return mi;
                           },
                           [&](float px, float py, float pz, std::vector<float> &elev, float &r, float &vp, float &vs,
                               bool &on) { return grid_point(G, E, px, py, pz, elev, r, vp, vs, on); },
                           G.ni, lam, mu, rho);
}

static int cgfd_media_bin_file_body(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                        const int order[3], const int size[3], const double spacing[3], const double origin[3],
                        const char *rho_path, const char *vp_path, const char *vs_path, double *lam, double *mu,
                        double *rho) {
  if (ni < 1 || nj < 1 || nk < 1 || !x || !y || !z || !order || !size || !spacing || !origin || !rho_path ||
      !vp_path || !vs_path || !lam || !mu || !rho)
    return CGFD_E_ARG;
  // which file dimension holds x, y and z (order[d] = 0, 1, 2 for x, y, z)
  int dim[3] = {-1, -1, -1};
  for (int d = 0; d < 3; d++) {
    if (order[d] < 0 || order[d] > 2 || dim[order[d]] != -1) {
      cgfd_internal_set_error("bin order must name x (0), y (1) and z (2) once each");
      return CGFD_E_ARG;
    }
    dim[order[d]] = d;
  }
  for (int d = 0; d < 3; d++)
    if (size[d] < 2 || !(spacing[d] != 0) || !std::isfinite(spacing[d]) || !std::isfinite(origin[d])) {
      cgfd_internal_set_error("bin sizes must be at least 2 with finite nonzero spacing and a finite origin");
      return CGFD_E_ARG;
    }
  const G3 g(ni, nj, nk);
  float xmin = (float)x[0], xmax = xmin, ymin = (float)y[0], ymax = ymin, zmin = FLT_MAX, zmax = -FLT_MAX;
  for (long p = 0; p < g.n; p++) {
    xmin = std::min(xmin, (float)x[p]);
    xmax = std::max(xmax, (float)x[p]);
    ymin = std::min(ymin, (float)y[p]);
    ymax = std::max(ymax, (float)y[p]);
    zmin = std::min(zmin, (float)z[p]);
    zmax = std::max(zmax, (float)z[p]);
  }
  const float o[3] = {(float)origin[dim[0]], (float)origin[dim[1]], (float)origin[dim[2]]};
  const float h[3] = {(float)spacing[dim[0]], (float)spacing[dim[1]], (float)spacing[dim[2]]};
  // the window of file nodes that covers the grid in x and y
  // (media_bin2model_el_iso, media_bin2model.cpp:83-137)
  int lo[3], hi[3];
  const float amin[2] = {xmin, ymin}, amax[2] = {xmax, ymax};
  for (int a = 0; a < 2; a++) {
    const int n = size[dim[a]];
    const float flo = amin[a] < o[a] ? 0.0f : (amin[a] - o[a]) / h[a];
    if (!(flo < (float)n) || amax[a] < o[a]) {  // checked before the integer conversion
      cgfd_internal_set_error(a == 0 ? "the bin model does not cover the grid in x"
                                     : "the bin model does not cover the grid in y");
      return CGFD_E_ARG;
    }
    lo[a] = (int)flo;
    const float fhi = std::ceil((amax[a] - o[a]) / h[a]);
    hi[a] = fhi < (float)n ? (int)fhi : n - 1;
    if (hi[a] <= lo[a]) {
      cgfd_internal_set_error("the bin model window covering the grid needs two nodes per axis");
      return CGFD_E_ARG;
    }
  }
  lo[2] = 0;
  hi[2] = size[dim[2]] - 1;
  const float zn = o[2] + h[2] * (size[dim[2]] - 1);
  if (zmin >= std::max(zn, o[2]) || zmax <= std::min(o[2], zn)) {
    cgfd_internal_set_error("the bin model does not cover the grid in z");
    return CGFD_E_ARG;
  }
  const int bn[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] + 1};
  std::vector<float> vec[3];
  for (int a = 0; a < 3; a++) {
    vec[a].resize(bn[a]);
    for (int q = 0; q < bn[a]; q++) vec[a][q] = o[a] + (float)(size_t)(lo[a] + q) * h[a];
  }
  // read the three files: dimension 1 fastest; keep the window as [z][y][x]
  // (read_bin_file, media_read_file.cpp:1274-1320)
  const size_t line = bn[0], slice = (size_t)bn[0] * bn[1], nfile = (size_t)size[0] * size[1] * size[2];
  std::vector<float> fld[3];
  const char *paths[3] = {rho_path, vp_path, vs_path};
  std::vector<float> buf(nfile);
  int flo[3], fhi[3];
  for (int a = 0; a < 3; a++) {
    flo[dim[a]] = lo[a];
    fhi[dim[a]] = hi[a];
  }
  for (int f = 0; f < 3; f++) {
    FILE *fp = std::fopen(paths[f], "rb");
    if (!fp) {
      cgfd_internal_set_error((std::string("cannot open ") + paths[f]).c_str());
      return CGFD_E_ARG;
    }
    const size_t got = std::fread(buf.data(), sizeof(float), nfile, fp);
    std::fclose(fp);
    if (got < nfile) {
      cgfd_internal_set_error((std::string("insufficient data in ") + paths[f]).c_str());
      return CGFD_E_ARG;
    }
    fld[f].resize(slice * bn[2]);
    int i[3];
    for (i[2] = flo[2]; i[2] <= fhi[2]; i[2]++)
      for (i[1] = flo[1]; i[1] <= fhi[1]; i[1]++)
        for (i[0] = flo[0]; i[0] <= fhi[0]; i[0]++)
          fld[f][(i[dim[0]] - lo[0]) + (size_t)(i[dim[1]] - lo[1]) * line + (size_t)i[dim[2]] * slice] =
              buf[i[0] + (size_t)size[0] * (i[1] + (size_t)size[1] * i[2])];
  }
  // trilinear interpolation with the reference's clamping and operation order
  // (TrilinearInterpolation, media_utility.cpp:189-284)
  const float d[3] = {vec[0][1] - vec[0][0], vec[1][1] - vec[1][0], vec[2][1] - vec[2][0]};
  auto tri = [&](const std::vector<float> &v, const int *i0, const int *i2, const float *q) {
    auto at = [&](int a, int b, int c) { return v[a + b * line + c * slice]; };
    if (i0[0] == i2[0] && i0[1] == i2[1] && i0[2] == i2[2]) return at(i0[0], i0[1], i0[2]);
    const float v000 = at(i0[0], i0[1], i0[2]), v002 = at(i0[0], i0[1], i2[2]);
    const float v200 = at(i2[0], i0[1], i0[2]), v202 = at(i2[0], i0[1], i2[2]);
    const float v020 = at(i0[0], i2[1], i0[2]), v022 = at(i0[0], i2[1], i2[2]);
    const float v220 = at(i2[0], i2[1], i0[2]), v222 = at(i2[0], i2[1], i2[2]);
    float v100 = v000, v102 = v002, v122 = v022, v120 = v020;
    if (i0[0] != i2[0]) {
      const float w = (q[0] - vec[0][i0[0]]) / (vec[0][i2[0]] - vec[0][i0[0]]);
      v100 = v000 + w * (v200 - v000);
      v102 = v002 + w * (v202 - v002);
      v122 = v022 + w * (v222 - v022);
      v120 = v020 + w * (v220 - v020);
    }
    float v110 = v100, v112 = v102;
    if (i0[1] != i2[1]) {
      const float w = (q[1] - vec[1][i0[1]]) / (vec[1][i2[1]] - vec[1][i0[1]]);
      v110 = v100 + w * (v120 - v100);
      v112 = v102 + w * (v122 - v102);
    }
    if (i0[2] == i2[2]) return v110;
    return v110 + (q[2] - vec[2][i0[2]]) / (vec[2][i2[2]] - vec[2][i0[2]]) * (v112 - v110);
  };
  par_for(g.nz, [&](int k) {
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        const long p = g.at(i, j, k);
        const float q[3] = {(float)x[p], (float)y[p], (float)z[p]};
        int i0[3], i2[3];
        for (int a = 0; a < 3; a++) {
          const float t = (q[a] - vec[a][0]) / d[a];
          i0[a] = std::min(std::max((int)std::floor(t), 0), bn[a] - 1);
          i2[a] = std::min(std::max((int)std::ceil(t), 0), bn[a] - 1);
        }
        const float r = tri(fld[0], i0, i2, q), vp = tri(fld[1], i0, i2, q), vs = tri(fld[2], i0, i2, q);
        const float m = vs * vs * r;
        mu[p] = m;
        lam[p] = (float)((double)(vp * vp * r) - 2.0 * (double)m);
        rho[p] = r;
      }
  });
  return CGFD_OK;
}

int cgfd_media_homogeneous(int ni, int nj, int nk, double vp, double vs, double rho, double *lam, double *mu,
                           double *rhoa) {
  if (ni < 1 || nj < 1 || nk < 1 || !lam || !mu || !rhoa || vp <= 0 || vs < 0 || rho <= 0) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  const double m = rho * vs * vs, l = rho * vp * vp - 2 * m;
  for (long p = 0; p < g.n; p++) { lam[p] = l; mu[p] = m; rhoa[p] = rho; }
  return CGFD_OK;
}

int cgfd_media_layered(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                       const cgfd_layer_desc *d, double *lam, double *mu, double *rho) {
  (void)x; (void)y;
  if (!d || d->nlayer < 1 || !z || !lam || !mu || !rho || !d->vp || !d->vpvs || !d->rho) return CGFD_E_ARG;
  if (d->nlayer > 1 && !d->z_interface) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  auto layer_of = [&](double zz) {
    int l = 0;
    while (l < d->nlayer - 1 && zz < d->z_interface[l]) l++;
    return l;
  };
  std::vector<float> noise, tmp;
  const bool pert = d->pert_std > 0 && d->corr_len > 0;
  if (pert) { noise.resize(g.n); tmp.resize(g.n); }
  // mean grid spacings for the correlation length in cells
  double hx = std::fabs(x[g.at(NG + 1, NG, NG)] - x[g.at(NG, NG, NG)]);
  double hy = std::fabs(y[g.at(NG, NG + 1, NG)] - y[g.at(NG, NG, NG)]);
  double hz = std::fabs(z[g.at(NG, NG, NG + nk - 1)] - z[g.at(NG, NG, NG)]) / std::max(1, nk - 1);
  for (int l = 0; l < d->nlayer; l++) {
    if (pert) {
      // white noise, seeded per (layer, plane) so results do not depend on threading
      par_for(g.nz, [&](int k) {
        std::mt19937_64 r(mix(d->seed * 1315423911ull + (uint64_t)l * 2654435761ull + (uint64_t)k));
        std::normal_distribution<float> N(0.f, 1.f);
        for (long q = (long)k * g.sz; q < (long)(k + 1) * g.sz; q++) noise[q] = N(r);
      });
      // separable Gaussian smoothing: correlation length = 1/e width
      const double sig[3] = {d->corr_len / std::max(hx, 1e-9) / std::sqrt(2.0),
                             d->corr_len / std::max(hy, 1e-9) / std::sqrt(2.0),
                             d->corr_len / std::max(hz, 1e-9) / std::sqrt(2.0)};
      const int dims[3] = {g.nx, g.ny, g.nz};
      const long strides[3] = {1, g.sy, g.sz};
      for (int a = 0; a < 3; a++) {
        const int R = std::max(1, (int)std::ceil(3 * sig[a]));
        std::vector<float> w(2 * R + 1);
        double ws = 0;
        for (int q = -R; q <= R; q++) ws += (w[q + R] = (float)std::exp(-0.5 * q * q / (sig[a] * sig[a])));
        for (auto &v : w) v = (float)(v / ws);
        par_for(g.nz, [&](int k) {
          for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) {
              const int pos = a == 0 ? i : a == 1 ? j : k;
              const long p = g.at(i, j, k);
              double s = 0;
              for (int q = -R; q <= R; q++) {
                int t = pos + q;
                t = t < 0 ? -t : t >= dims[a] ? 2 * dims[a] - 2 - t : t;  // mirror at the edges
                t = std::min(std::max(t, 0), dims[a] - 1);
                s += w[q + R] * noise[p + (long)(t - pos) * strides[a]];
              }
              tmp[p] = (float)s;
            }
        });
        noise.swap(tmp);
      }
      double s2 = 0;
      for (long p = 0; p < g.n; p++) s2 += (double)noise[p] * noise[p];
      const float inv = (float)(1.0 / std::sqrt(s2 / g.n + 1e-300));
      for (long p = 0; p < g.n; p++) noise[p] *= inv;
    }
    par_for(g.nz, [&](int k) {
      for (long p = (long)k * g.sz; p < (long)(k + 1) * g.sz; p++) {
        if (layer_of(z[p]) != l) continue;
        double f = 1.0;
        if (pert) f = 1.0 + std::max(-d->pert_clip, std::min(d->pert_clip, d->pert_std * (double)noise[p]));
        const double vp = d->vp[l] * f, vs = vp / d->vpvs[l];
        const double r = d->rho[l] > 0 ? d->rho[l] : 310.0 * std::pow(vp, 0.25);
        rho[p] = r;
        mu[p] = r * vs * vs;
        lam[p] = r * vp * vp - 2 * mu[p];
      }
    });
  }
  return CGFD_OK;
}

int cgfd_rho_to_slow(long long n, const double *rho, double *slw) {
  if (n < 0 || !rho || !slw) return CGFD_E_ARG;
  for (long long p = 0; p < n; p++) slw[p] = rho[p] > 1e-10 ? 1.0 / rho[p] : 0.0;
  return CGFD_OK;
}

int cgfd_pml_len_dh(int ni, int nj, int nk, const double *x, const double *y, const double *z, int face, int nlay,
                    double *L0, double *dh) {
  if (face < 0 || face > 5 || nlay < 1 || !x || !y || !z || !L0 || !dh) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  int i1 = NG, i2 = NG + ni - 1, j1 = NG, j2 = NG + nj - 1, k1 = NG, k2 = NG + nk - 1;
  switch (face) {
    case 0: i2 = i1 + nlay; break;
    case 1: i1 = i2 - nlay; break;
    case 2: j2 = j1 + nlay; break;
    case 3: j1 = j2 - nlay; break;
    case 4: k2 = k1 + nlay; break;
    default: k1 = k2 - nlay; break;
  }
  const int dim = face / 2;
  const long st = dim == 0 ? 1 : dim == 1 ? g.sy : g.sz;
  double L = 0;
  long num = 0;
  for (int k = k1; k <= k2 - (dim == 2); k++)
    for (int j = j1; j <= j2 - (dim == 1); j++)
      for (int i = i1; i <= i2 - (dim == 0); i++) {
        const long p = g.at(i, j, k), q = p + st;
        L += std::sqrt((x[q] - x[p]) * (x[q] - x[p]) + (y[q] - y[p]) * (y[q] - y[p]) + (z[q] - z[p]) * (z[q] - z[p]));
        num++;
      }
  *dh = L / (double)num;
  *L0 = *dh * nlay;
  return CGFD_OK;
}

int cgfd_pml_profile(int nlay, int face, double L0, double dh, double vel, double amax, double bmax, double *A,
                     double *B, double *D) {
  if (nlay < 1 || face < 0 || face > 5 || L0 <= 0 || dh <= 0 || !A || !B || !D) return CGFD_E_ARG;
  const int np = nlay + 1;
  const double R = std::pow(10.0, -((std::log10((double)nlay) - 1.0) / std::log10(2.0) + 4.0));
  const double d0 = -vel / (2.0 * L0) * std::log(R) * 3.0;
  const bool low = (face % 2) == 0;
  for (int n = 0; n < np; n++) {
    const double r = n * dh / L0;
    const int i = low ? np - 1 - n : n;
    const double d = d0 * r * r, a = amax * (1.0 - r), b = 1.0 + (bmax - 1.0) * r * r;
    D[i] = d / b;
    A[i] = a + d / b;
    B[i] = 1.0 / b;
  }
  return CGFD_OK;
}

int cgfd_free_surface_matrices(int ni, int nj, int nk, const double *metric, const double *lam, const double *mu,
                               double *mvx, double *mvy) {
  if (!metric || !lam || !mu || !mvx || !mvy) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  const long n = g.n;
  std::fill(mvx, mvx + (long)g.nx * g.ny * 9, 0.0);
  std::fill(mvy, mvy + (long)g.nx * g.ny * 9, 0.0);
  const int k = NG + nk - 1;
  for (int j = NG; j < NG + nj; j++)
    for (int i = NG; i < NG + ni; i++) {
      const long p = g.at(i, j, k);
      double e[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) e[r][c] = metric[(1 + 3 * r + c) * n + p];
      const double l = lam[p], m = mu[p], l2m = l + 2 * m;
      // zero traction on the zeta = const surface: sigma . grad(zeta) = 0 with
      // dV/dzeta unknown; A dV/dzeta = B dV/dxi + C dV/deta
      const double *z3 = e[2];
      double A[3][3], B[3][3], Cm[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          const double zz = z3[0] * z3[0] + z3[1] * z3[1] + z3[2] * z3[2];
          A[r][c] = l * z3[r] * z3[c] + m * z3[c] * z3[r] + (r == c ? m * zz : 0.0);
          const double *x1 = e[0], *x2 = e[1];
          const double dx1 = z3[0] * x1[0] + z3[1] * x1[1] + z3[2] * x1[2];
          const double dx2 = z3[0] * x2[0] + z3[1] * x2[1] + z3[2] * x2[2];
          B[r][c] = -(l * z3[r] * x1[c] + m * z3[c] * x1[r] + (r == c ? m * dx1 : 0.0));
          Cm[r][c] = -(l * z3[r] * x2[c] + m * z3[c] * x2[r] + (r == c ? m * dx2 : 0.0));
        }
      (void)l2m;
      // inverse of A (cofactors, FP64)
      double inv[3][3];
      inv[0][0] = A[1][1] * A[2][2] - A[2][1] * A[1][2];
      inv[0][1] = A[2][1] * A[0][2] - A[0][1] * A[2][2];
      inv[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
      inv[1][0] = A[1][2] * A[2][0] - A[1][0] * A[2][2];
      inv[1][1] = A[0][0] * A[2][2] - A[2][0] * A[0][2];
      inv[1][2] = A[1][0] * A[0][2] - A[0][0] * A[1][2];
      inv[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
      inv[2][1] = A[2][0] * A[0][1] - A[0][0] * A[2][1];
      inv[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
      const double det = inv[0][0] * A[0][0] + inv[0][1] * A[1][0] + inv[0][2] * A[2][0];
      if (!(std::fabs(det) > 0)) return CGFD_E_ARG;
      const long ij = ((long)j * g.nx + i) * 9;
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          double sb = 0, sc = 0;
          for (int q = 0; q < 3; q++) { sb += inv[r][q] * B[q][c]; sc += inv[r][q] * Cm[q][c]; }
          mvx[ij + r * 3 + c] = sb / det;
          mvy[ij + r * 3 + c] = sc / det;
        }
    }
  return CGFD_OK;
}

double cgfd_wavelet(int kind, double t, double c0, double c1) {
  const double tau = t - c1;
  switch (kind) {
    case CGFD_WAVELET_RICKER: {
      const double u = 2 * M_PI * c0 * tau;
      return (1 - u * u / 2) * std::exp(-u * u / 4);
    }
    case CGFD_WAVELET_GAUSSIAN:
      return std::exp(-tau * tau / (c0 * c0)) / (std::sqrt(M_PI) * c0);
    case CGFD_WAVELET_RICKER_DERIV: {
      const double u = 2 * M_PI * c0 * tau;
      return u * (-3 + 0.5 * u * u) * std::exp(-u * u / 4) * M_PI * c0;
    }
    case CGFD_WAVELET_GAUSSIAN_DERIV:
      return std::exp(-tau * tau / (c0 * c0)) / (std::sqrt(M_PI) * c0) * (-2 * tau / (c0 * c0));
    case CGFD_WAVELET_KLAUDER: {
      // linear sweep f1..f2 over T starting t0 = 0.2 s (the reference's fixed
      // parameters; c0, c1 unused); singular at t = t0 as there
      const double ts = 0.2, f1 = 1.5, f2 = 41.5, T = 20.0;
      const double K = (f2 - f1) / T, fm = (f2 + f1) / 2.0;
      return std::sin(M_PI * K * t * (T - (t - ts))) * std::cos(2 * M_PI * fm * (t - ts)) / (M_PI * K * (t - ts));
    }
    default:
      return 0.0;
  }
}

int cgfd_angle2moment(double strike, double dip, double rake, double *m6) {
  if (!m6) return CGFD_E_ARG;
  const double s = strike * M_PI / 180, d = dip * M_PI / 180, r = rake * M_PI / 180;
  // Aki & Richards (x north, y east, z down)
  const double M11 = -(std::sin(d) * std::cos(r) * std::sin(2 * s) + std::sin(2 * d) * std::sin(r) * std::sin(s) * std::sin(s));
  const double M22 = std::sin(d) * std::cos(r) * std::sin(2 * s) - std::sin(2 * d) * std::sin(r) * std::cos(s) * std::cos(s);
  const double M33 = -(M11 + M22);
  const double M12 = std::sin(d) * std::cos(r) * std::cos(2 * s) + 0.5 * std::sin(2 * d) * std::sin(r) * std::sin(2 * s);
  const double M13 = -(std::cos(d) * std::cos(r) * std::cos(s) + std::cos(2 * d) * std::sin(r) * std::sin(s));
  const double M23 = -(std::cos(d) * std::cos(r) * std::sin(s) - std::cos(2 * d) * std::sin(r) * std::cos(s));
  // to x east, y north, z up
  m6[0] = M22;
  m6[1] = M11;
  m6[2] = M33;
  m6[3] = -M13;
  m6[4] = -M23;
  m6[5] = M12;
  return CGFD_OK;
}

namespace {

// Gaussian spreading of a point source over the 7^3 points around node
// (si, sj, sk), the point sitting at (ox, oy, oz) from it: width 1.5 points,
// weights normalised over all 343 before those outside the physical grid
// are dropped (src_cal_norm_delt3d and its use, src_t.cu:614-654).
// Returns the number of points kept.
int spread_gauss(const G3 &g, int si, int sj, int sk, double ox, double oy, double oz, long long *ext_indx,
                 double *ext_coef) {
  const int half = 3;
  const double wid = half / 2.0;
  double w[343], sum = 0;
  int q = 0;
  for (int k = -half; k <= half; k++)
    for (int j = -half; j <= half; j++)
      for (int i = -half; i <= half; i++) {
        auto gs = [&](double t) { return std::exp(-t * t / (wid * wid)) / (std::sqrt(M_PI) * wid); };
        w[q] = gs(i - ox) * gs(j - oy) * gs(k - oz);
        sum += w[q++];
      }
  int ne = 0;
  for (int k = -half; k <= half; k++)
    for (int j = -half; j <= half; j++)
      for (int i = -half; i <= half; i++) {
        const int gi = si + i, gj = sj + j, gk = sk + k;
        if (gi < 0 || gi > g.ni - 1 || gj < 0 || gj > g.nj - 1 || gk < 0 || gk > g.nk - 1) continue;
        const int c = (i + half) + 7 * (j + half) + 49 * (k + half);
        ext_indx[ne] = g.at(gi + NG, gj + NG, gk + NG);
        ext_coef[ne] = w[c] / sum;
        ne++;
      }
  return ne;
}

}  // namespace

int cgfd_source_point(int ni, int nj, int nk, double sx, double sy, double sz, int depth_flag, double t0, double dt,
                      double wavelet_tstart, int wavelet, double c0, double c1, double stf_length, int *it_begin,
                      int *it_end, int *ext_num, long long *ext_indx, double *ext_coef, int *max_nt_out,
                      double *stf) {
  if (ni < 1 || nj < 1 || nk < 1 || dt <= 0 || !it_begin || !it_end || !ext_num || !ext_indx || !ext_coef ||
      !max_nt_out)
    return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  if (depth_flag == 1) sz = (double)(nk - 1) - sz;
  const int si = (int)(sx + 0.5), sj = (int)(sy + 0.5), sk = (int)(sz + 0.5);
  if (si < 0 || si > ni - 1 || sj < 0 || sj > nj - 1 || sk < 0 || sk > nk - 1) return CGFD_E_ARG;
  *ext_num = spread_gauss(g, si, sj, sk, sx - si, sy - sj, sz - sk, ext_indx, ext_coef);
  const int max_nt = (int)(stf_length / dt + 0.5);
  *max_nt_out = max_nt;
  const int ib = (int)((wavelet_tstart - t0) / dt);
  *it_begin = ib;
  *it_end = ib + max_nt - 1;
  if (stf) {
    const double frac[4] = {0.0, 0.5, 0.5, 1.0};
    const double t_shift = wavelet_tstart - (ib * dt + t0);
    for (int it = 0; it < max_nt; it++)
      for (int s = 0; s < 4; s++) stf[it * 4 + s] = cgfd_wavelet(wavelet, it * dt + frac[s] * dt - t_shift, c0, c1);
  }
  return CGFD_OK;
}

int cgfd_receiver_index(int ni, int nj, int nk, double rx, double ry, double rz, int depth_flag, long long *indx8,
                        double *frac3) {
  if (!indx8 || !frac3) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  rx -= 1;
  ry -= 1;
  if (depth_flag == 1) rz = (double)(nk - 1) - rz;
  else rz -= 1;
  // far outside (or not a number): no integer cast of an out-of-range value
  if (!(std::fabs(rx) < 1e9 && std::fabs(ry) < 1e9 && std::fabs(rz) < 1e9)) return CGFD_E_ARG;
  const int ix = (int)std::floor(rx), iy = (int)std::floor(ry), iz = (int)std::floor(rz);
  if (ix < 0 || ix > ni - 1 || iy < 0 || iy > nj - 1 || iz < 0 || iz > nk - 1) return CGFD_E_ARG;
  const long p = g.at(ix + NG, iy + NG, iz + NG);
  indx8[0] = p;
  indx8[1] = p + 1;
  indx8[2] = p + g.sy;
  indx8[3] = p + 1 + g.sy;
  indx8[4] = p + g.sz;
  indx8[5] = p + 1 + g.sz;
  indx8[6] = p + g.sy + g.sz;
  indx8[7] = p + 1 + g.sy + g.sz;
  frac3[0] = rx - ix;
  frac3[1] = ry - iy;
  frac3[2] = rz - iz;
  return CGFD_OK;
}

// ---- receiver lines (io_line_locate, io_funcs.cu:302-417) ----------------------------
// Points start + ipt*incre (0-based global physical indices), ipt < count,
// those inside the grid kept in line order.  Each becomes an on-node
// receiver: fractions 0, so the trilinear weights give the node value
// exactly, the value io_line_keep copies (io_funcs.cu:1669-1700).
int cgfd_line_points(int ni, int nj, int nk, const double *x, const double *y, const double *z, const int *start,
                     const int *incre, int count, int *seq, long long *indx8, double *frac3, double *xyz) {
  if (!start || !incre || count < 0 || !seq || !indx8 || !frac3) return CGFD_E_ARG;
  G3 g(ni, nj, nk);
  int n = 0;
  for (int ipt = 0; ipt < count; ipt++) {
    const long gi = start[0] + (long)ipt * incre[0], gj = start[1] + (long)ipt * incre[1],
               gk = start[2] + (long)ipt * incre[2];
    if (gi < 0 || gi > ni - 1 || gj < 0 || gj > nj - 1 || gk < 0 || gk > nk - 1) continue;
    const long p = g.at((int)gi + NG, (int)gj + NG, (int)gk + NG);
    long long *ix = indx8 + 8L * n;
    ix[0] = p; ix[1] = p + 1; ix[2] = p + g.sy; ix[3] = p + 1 + g.sy;
    ix[4] = p + g.sz; ix[5] = p + 1 + g.sz; ix[6] = p + g.sy + g.sz; ix[7] = p + 1 + g.sy + g.sz;
    frac3[3L * n] = frac3[3L * n + 1] = frac3[3L * n + 2] = 0.0;
    seq[n] = ipt;
    if (xyz && x && y && z) {
      xyz[3L * n] = x[p];
      xyz[3L * n + 1] = y[p];
      xyz[3L * n + 2] = z[p];
    }
    n++;
  }
  return n;
}

// ---- points given by coordinates ---------------------------------------------------
// Restatement of the reference's coordinate location, in float as there
// (its grid arrays are float): depth below the free surface -> z on the
// surface quad holding (x, y) (gd_curv_depth_to_axis, gd_t.cu:1395-1486, with
// fdlib_math_isPoint2InQuad / fdlib_math_rdinterp_2d, fdlib_math.cu:168-250);
// nearest grid node by brute force, the first of the 8 cells around it that
// holds the point (isPointInHexahedron_c, gd_t.cu:1767-1800) and a 101^3
// trilinear sampling of that cell (gd_curv_coord_to_local_indx and
// gd_curv_coord2index_sample, gd_t.cu:1241-1392, 1574-1650); negative
// fractions move to the lower cell (io_recv_read_locate, io_funcs.cu:156-170).
namespace {

struct FGrid {
  G3 g;
  std::vector<float> x, y, z;
  float xmin, xmax, ymin, ymax, zmin, zmax;
  // 8x8x8-node blocks with coordinate bounds, for an exact pruned nearest-node search
  int bx, by, bz;
  std::vector<float> bb;  // per block: xmin, xmax, ymin, ymax, zmin, zmax
  FGrid(int ni, int nj, int nk, const double *X, const double *Y, const double *Z) : g(ni, nj, nk) {
    x.assign(X, X + g.n);
    y.assign(Y, Y + g.n);
    z.assign(Z, Z + g.n);
    xmin = *std::min_element(x.begin(), x.end());
    xmax = *std::max_element(x.begin(), x.end());
    ymin = *std::min_element(y.begin(), y.end());
    ymax = *std::max_element(y.begin(), y.end());
    zmin = *std::min_element(z.begin(), z.end());
    zmax = *std::max_element(z.begin(), z.end());
    bx = (g.nx + 7) / 8;
    by = (g.ny + 7) / 8;
    bz = (g.nz + 7) / 8;
    bb.assign((size_t)bx * by * bz * 6, 0.0f);
    for (int kb = 0; kb < bz; kb++)
      for (int jb = 0; jb < by; jb++)
        for (int ib = 0; ib < bx; ib++) {
          float m[6] = {1e30f, -1e30f, 1e30f, -1e30f, 1e30f, -1e30f};
          for (int k = kb * 8; k < std::min(g.nz, kb * 8 + 8); k++)
            for (int j = jb * 8; j < std::min(g.ny, jb * 8 + 8); j++)
              for (int i = ib * 8; i < std::min(g.nx, ib * 8 + 8); i++) {
                const long p = g.at(i, j, k);
                m[0] = std::min(m[0], x[p]); m[1] = std::max(m[1], x[p]);
                m[2] = std::min(m[2], y[p]); m[3] = std::max(m[3], y[p]);
                m[4] = std::min(m[4], z[p]); m[5] = std::max(m[5], z[p]);
              }
          std::copy(m, m + 6, &bb[(((size_t)kb * by + jb) * bx + ib) * 6]);
        }
  }
  float dist(float sx, float sy, float sz, long p) const {
    return sqrtf((sx - x[p]) * (sx - x[p]) + (sy - y[p]) * (sy - y[p]) + (sz - z[p]) * (sz - z[p]));
  }
  // the first node (k, j, i order) at the least float distance: blocks in
  // order of their distance bound, stopping once a bound exceeds the best
  long nearest(float sx, float sy, float sz) const {
    const size_t nb = (size_t)bx * by * bz;
    std::vector<std::pair<double, size_t>> order(nb);
    for (size_t b = 0; b < nb; b++) {
      const float *m = &bb[b * 6];
      auto gap = [](double v, double lo, double hi) { return v < lo ? lo - v : v > hi ? v - hi : 0.0; };
      const double dx = gap(sx, m[0], m[1]), dy = gap(sy, m[2], m[3]), dz = gap(sz, m[4], m[5]);
      order[b] = {std::sqrt(dx * dx + dy * dy + dz * dz), b};
    }
    std::sort(order.begin(), order.end());
    float best = dist(sx, sy, sz, 0);
    long bestp = 0;
    for (const auto &ob : order) {
      // float rounding of the node distances: compare with margin
      if (ob.first > (double)best * (1.0 + 1e-5) + 1e-6) break;
      const size_t b = ob.second;
      const int ib = (int)(b % bx), jb = (int)((b / bx) % by), kb = (int)(b / ((size_t)bx * by));
      for (int k = kb * 8; k < std::min(g.nz, kb * 8 + 8); k++)
        for (int j = jb * 8; j < std::min(g.ny, jb * 8 + 8); j++)
          for (int i = ib * 8; i < std::min(g.nx, ib * 8 + 8); i++) {
            const long p = g.at(i, j, k);
            const float d = dist(sx, sy, sz, p);
            if (d < best || (d == best && p < bestp)) {
              best = d;
              bestp = p;
            }
          }
    }
    return bestp;
  }
};

float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// the hexahedron test of the reference: the point lies on the inner side of
// the six faces, each taken as the plane through three of its corners
bool in_hexahedron(float px, float py, float pz, const float *vx, const float *vy, const float *vz) {
  static const int F[6][3] = {{0, 4, 6}, {7, 5, 1}, {5, 4, 0}, {2, 6, 7}, {4, 5, 7}, {3, 1, 0}};
  for (int f = 0; f < 6; f++) {
    const int a = F[f][0], b = F[f][1], c = F[f][2];
    const float p2f[3] = {vx[a] - px, vy[a] - py, vz[a] - pz};
    const float A[3] = {vx[b] - vx[a], vy[b] - vy[a], vz[b] - vz[a]};
    const float B[3] = {vx[c] - vx[a], vy[c] - vy[a], vz[c] - vz[a]};
    const float nrm[3] = {A[1] * B[2] - A[2] * B[1], A[2] * B[0] - A[0] * B[2], A[0] * B[1] - A[1] * B[0]};
    const float len = sqrtf(dot3(nrm, nrm));
    const float nu[3] = {nrm[0] / len, nrm[1] / len, nrm[2] / len};
    float sign = dot3(p2f, nu);
    sign /= sqrtf(dot3(p2f, p2f));
    if (sign < 0.0f) return false;
  }
  return true;
}

// corner c of the cell whose first node is (i, j, k): c = n1 + 2 n2 + 4 n3
void cell_corners(const FGrid &G, int i, int j, int k, float *vx, float *vy, float *vz) {
  for (int c = 0; c < 8; c++) {
    const long p = G.g.at(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
    vx[c] = G.x[p];
    vy[c] = G.y[p];
    vz[c] = G.z[p];
  }
}

// the 101^3 sample of the cell nearest to the point: (n1, n2, n3) / 100
void cell_sample(float sx, float sy, float sz, const float *vx, const float *vy, const float *vz, int *s1, int *s2,
                 int *s3) {
  constexpr int N = 100;
  float best = sqrtf((sx - vx[0]) * (sx - vx[0]) + (sy - vy[0]) * (sy - vy[0]) + (sz - vz[0]) * (sz - vz[0]));
  int b1 = 0, b2 = 0, b3 = 0;
  for (int n3 = 0; n3 <= N; n3++) {
    float Lz[2];
    Lz[1] = (float)n3 / (float)N;
    Lz[0] = (float)(1.0 - Lz[1]);
    for (int n2 = 0; n2 <= N; n2++) {
      float Ly[2];
      Ly[1] = (float)n2 / (float)N;
      Ly[0] = (float)(1.0 - Ly[1]);
      for (int n1 = 0; n1 <= N; n1++) {
        float Lx[2];
        Lx[1] = (float)n1 / (float)N;
        Lx[0] = (float)(1.0 - Lx[1]);
        float xp = 0, yp = 0, zp = 0;
        for (int c = 0; c < 8; c++) {
          const float w = Lx[c & 1] * Ly[(c >> 1) & 1] * Lz[c >> 2];
          xp += w * vx[c];
          yp += w * vy[c];
          zp += w * vz[c];
        }
        const float d = sqrtf((sx - xp) * (sx - xp) + (sy - yp) * (sy - yp) + (sz - zp) * (sz - zp));
        if (best > d) {
          best = d;
          b1 = n1;
          b2 = n2;
          b3 = n3;
        }
      }
    }
  }
  *s1 = b1;
  *s2 = b2;
  *s3 = b3;
}

bool in_quad(float px, float py, const float *qx, const float *qy) {
  // corners 0,1,3,2 walk the quad anticlockwise
  const float vx[4] = {qx[0], qx[1], qx[3], qx[2]}, vy[4] = {qy[0], qy[1], qy[3], qy[2]};
  for (int i = 0; i < 4; i++) {
    const int n = (i + 1) % 4;
    const float ex = vx[n] - vx[i], ey = vy[n] - vy[i];
    float sign = ex * (py - vy[i]) - ey * (px - vx[i]);
    sign /= sqrtf(ex * ex + ey * ey);
    if (sign < -1e-6) return false;
  }
  return true;
}

// inverse-distance weights over the 4 quad corners; a point on a corner takes
// that corner's value (the reference reads an unset weight when it is corner 0)
float idw4(float px, float py, const float *qx, const float *qy, const float *qv) {
  float w[4], tot = 0.0f;
  int at = -1;
  for (int c = 0; c < 4; c++) {
    const float d = sqrtf((px - qx[c]) * (px - qx[c]) + (py - qy[c]) * (py - qy[c]));
    if (d < 1e-9) {
      at = c;
      w[c] = 0.0f;
    } else {
      w[c] = (float)(1.0 / d);
      tot += w[c];
    }
  }
  if (at >= 0) {
    for (int c = 0; c < 4; c++) w[c] = 0.0f;
    w[at] = 1.0f;
    tot = 1.0f;
  }
  float v = 0.0f;
  for (int c = 0; c < 4; c++) {
    w[c] = (float)((double)w[c] * (1.0 / (double)tot));
    v += w[c] * qv[c];
  }
  return v;
}

// z of the free surface under (sx, sy): the first top-plane quad holding it,
// visited in the reference's tile order (4 x 4 tiles of the physical columns)
bool surface_z(const FGrid &G, float sx, float sy, float *zs) {
  const G3 &g = G.g;
  const int k = NG + g.nk - 1;
  auto tiles = [](int n, int t) {
    std::vector<std::pair<int, int>> r(4);
    int start = NG;
    for (int q = 0; q < 4; q++) {
      int len = n / 4 + (q < n % 4 ? 1 : 0);
      r[q] = {start, start + len - 1};
      start += len;
    }
    (void)t;
    return r;
  };
  const auto ti = tiles(g.ni, 0), tj = tiles(g.nj, 1);
  for (int jt = 0; jt < 4; jt++)
    for (int it = 0; it < 4; it++)
      for (int j = tj[jt].first; j <= tj[jt].second; j++)
        for (int i = ti[it].first; i <= ti[it].second; i++) {
          // bounds of the 3-D cell (its upper plane is the ghost plane above)
          float xlo = 1e30f, xhi = -1e30f, ylo = 1e30f, yhi = -1e30f;
          for (int c = 0; c < 8; c++) {
            const long p = g.at(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2));
            xlo = std::min(xlo, G.x[p]); xhi = std::max(xhi, G.x[p]);
            ylo = std::min(ylo, G.y[p]); yhi = std::max(yhi, G.y[p]);
          }
          if (sx < xlo || sx > xhi || sy < ylo || sy > yhi) continue;
          float qx[4], qy[4], qz[4];
          for (int c = 0; c < 4; c++) {
            const long p = g.at(i + (c & 1), j + (c >> 1), k);
            qx[c] = G.x[p];
            qy[c] = G.y[p];
            qz[c] = G.z[p];
          }
          if (in_quad(sx, sy, qx, qy)) {
            *zs = idw4(sx, sy, qx, qy, qz);
            return true;
          }
        }
  return false;
}

}  // namespace

// One point: depth -> z where asked (the z used is returned in *zout), then
// the nearest physical node m (0-based) and the signed shift of the point
// from it (gd_curv_coord_to_glob_indx_gpu, gd_t.cu:1190-1392); false when
// the point is outside the grid.
static bool locate_raw(const FGrid &G, float sx, float sy, float sz, bool depth, int *m, float *inc, float *zout) {
  const int ni = G.g.ni, nj = G.g.nj, nk = G.g.nk;
  // depth -> z; as in the reference, a point under no surface quad keeps
  // its depth value as z (and is then normally outside the grid)
  float zs;
  if (depth && sx >= G.xmin && sx <= G.xmax && sy >= G.ymin && sy <= G.ymax && surface_z(G, sx, sy, &zs))
    sz = zs - sz;
  if (zout) *zout = sz;
  for (int d = 0; d < 3; d++) { m[d] = -1000; inc[d] = 0.0f; }
  if (sx < G.xmin || sx > G.xmax || sy < G.ymin || sy > G.ymax || sz < G.zmin || sz > G.zmax) return false;
  const long p = G.nearest(sx, sy, sz);
  const int mi = (int)(p % G.g.sy), mj = (int)((p / G.g.sy) % G.g.ny), mk = (int)(p / G.g.sz);
  if (mi < NG || mi > NG + ni - 1 || mj < NG || mj > NG + nj - 1 || mk < NG || mk > NG + nk - 1) return false;
  for (int c = 0; c < 8; c++) {
    const int ci = mi - 1 + (c & 1), cj = mj - 1 + ((c >> 1) & 1), ck = mk - 1 + (c >> 2);
    float vx[8], vy[8], vz[8];
    cell_corners(G, ci, cj, ck, vx, vy, vz);
    if (!in_hexahedron(sx, sy, sz, vx, vy, vz)) continue;
    int s1, s2, s3;
    cell_sample(sx, sy, sz, vx, vy, vz, &s1, &s2, &s3);
    inc[0] = ((float)ci + (float)s1 / 100.0f) - (float)mi;
    inc[1] = ((float)cj + (float)s2 / 100.0f) - (float)mj;
    inc[2] = ((float)ck + (float)s3 / 100.0f) - (float)mk;
    break;
  }
  m[0] = mi - NG; m[1] = mj - NG; m[2] = mk - NG;
  return true;
}

int cgfd_locate_points(int ni, int nj, int nk, const double *x, const double *y, const double *z, int n,
                       const double *pts, const int *depth_flag, int *idx, double *frac, int *found) {
  if (ni < 2 || nj < 2 || nk < 2 || !x || !y || !z || n < 0 || (n > 0 && (!pts || !idx || !frac || !found)))
    return CGFD_E_ARG;
  const FGrid G(ni, nj, nk, x, y, z);
  par_for(n, [&](int q) {
    int m[3];
    float inc[3];
    found[q] = locate_raw(G, (float)pts[3 * q], (float)pts[3 * q + 1], (float)pts[3 * q + 2],
                          depth_flag && depth_flag[q], m, inc, nullptr) ? 1 : 0;
    for (int d = 0; d < 3; d++) {
      idx[3 * q + d] = -1000;
      frac[3 * q + d] = 0.0;
    }
    if (!found[q]) return;
    // negative shifts move to the lower cell (io_recv_read_locate, io_funcs.cu:156-170)
    for (int d = 0; d < 3; d++) {
      int id = m[d];
      float f = inc[d];
      if (f < 0.0f) {
        f = (float)(1.0 + f);
        id -= 1;
      }
      idx[3 * q + d] = id;
      frac[3 * q + d] = f;
    }
  });
  return CGFD_OK;
}

// ---- station files (io_recv_read_locate, io_funcs.cu:24-299) ----------------------
namespace {

// next line that is neither empty nor a comment, newline removed
// (io_get_nextline, io_funcs.cu:2131-2154)
bool next_line(FILE *fp, char *buf, int len) {
  do {
    if (!std::fgets(buf, len, fp)) return false;
  } while (buf[0] == '#' || buf[0] == '\n');
  const size_t n = std::strlen(buf);
  if (n > 0 && buf[n - 1] == '\n') buf[n - 1] = '\0';
  return true;
}

void corners8(const G3 &g, int i, int j, int k, long long *ix) {
  const long p = g.at(i + NG, j + NG, k + NG);
  ix[0] = p; ix[1] = p + 1; ix[2] = p + g.sy; ix[3] = p + 1 + g.sy;
  ix[4] = p + g.sz; ix[5] = p + 1 + g.sz; ix[6] = p + g.sy + g.sz; ix[7] = p + 1 + g.sy + g.sz;
}

}  // namespace

static int cgfd_station_file_body(const char *path, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                      int max_n, cgfd_station *out, int *n_file) {
  if (!path || max_n < 0 || (max_n > 0 && !out) || ni < 2 || nj < 2 || nk < 2) return CGFD_E_ARG;
  FILE *fp = std::fopen(path, "rt");
  if (!fp) {
    cgfd_internal_set_error((std::string("cannot open station file ") + path).c_str());
    return CGFD_E_ARG;
  }
  char line[500];
  int total = -1;
  if (!next_line(fp, line, sizeof(line)) || std::sscanf(line, "%d", &total) != 1 || total < 0) {
    std::fclose(fp);
    cgfd_internal_set_error("station file: the first line must hold the number of stations");
    return CGFD_E_ARG;
  }
  if (n_file) *n_file = total;
  const int n = std::min(total, max_n);
  std::vector<cgfd_station> st((size_t)n);
  for (int r = 0; r < total; r++) {
    char name[500];
    int fc = 0, fd = 0;
    float v[3] = {0, 0, 0};
    if (!next_line(fp, line, sizeof(line)) ||
        std::sscanf(line, "%499s %d %d %g %g %g", name, &fc, &fd, &v[0], &v[1], &v[2]) != 6) {
      std::fclose(fp);
      cgfd_internal_set_error(("station file: station line " + std::to_string(r + 1) + " unreadable").c_str());
      return CGFD_E_ARG;
    }
    if (r >= n) continue;
    cgfd_station &s = st[(size_t)r];
    std::memset(&s, 0, sizeof(s));
    std::snprintf(s.name, sizeof(s.name), "%s", name);
    s.by_coord = fc;
    s.depth = fd;
    for (int d = 0; d < 3; d++) s.in[d] = v[d];
  }
  std::fclose(fp);
  const G3 g(ni, nj, nk);
  // coordinate stations: located together
  std::vector<int> cq;
  for (int r = 0; r < n; r++)
    if (st[(size_t)r].by_coord == 1) cq.push_back(r);
  if (!cq.empty()) {
    if (!x || !y || !z) {
      cgfd_internal_set_error("station file: stations by coordinates need the grid coordinates");
      return CGFD_E_ARG;
    }
    const int m = (int)cq.size();
    std::vector<double> pts(3 * (size_t)m);
    std::vector<int> dfl((size_t)m), idx(3 * (size_t)m), found((size_t)m);
    std::vector<double> fr(3 * (size_t)m);
    for (int q = 0; q < m; q++) {
      for (int d = 0; d < 3; d++) pts[3 * (size_t)q + d] = st[(size_t)cq[q]].in[d];
      dfl[(size_t)q] = st[(size_t)cq[q]].depth == 1 ? 1 : 0;
    }
    int r = cgfd_locate_points(ni, nj, nk, x, y, z, m, pts.data(), dfl.data(), idx.data(), fr.data(), found.data());
    if (r) return r;
    const FGrid G(ni, nj, nk, x, y, z);
    for (int q = 0; q < m; q++) {
      cgfd_station &s = st[(size_t)cq[q]];
      // the reference keeps the point with its depth turned into z
      float sx = (float)s.in[0], sy = (float)s.in[1], sz = (float)s.in[2], zs;
      if (s.depth == 1 && sx >= G.xmin && sx <= G.xmax && sy >= G.ymin && sy <= G.ymax && surface_z(G, sx, sy, &zs))
        sz = zs - sz;
      s.xyz[0] = sx; s.xyz[1] = sy; s.xyz[2] = sz;
      if (!found[(size_t)q]) continue;
      s.inside = 1;
      corners8(g, idx[3 * (size_t)q], idx[3 * (size_t)q + 1], idx[3 * (size_t)q + 2], s.indx8);
      for (int d = 0; d < 3; d++) s.frac[d] = fr[3 * (size_t)q + d];
    }
  }
  // grid-index stations
  for (int r = 0; r < n; r++) {
    cgfd_station &s = st[(size_t)r];
    if (s.by_coord == 1) continue;
    if (cgfd_receiver_index(ni, nj, nk, s.in[0], s.in[1], s.in[2], s.depth, s.indx8, s.frac) != CGFD_OK) {
      std::memset(s.indx8, 0, sizeof(s.indx8));
      s.frac[0] = s.frac[1] = s.frac[2] = 0.0;
      continue;
    }
    s.inside = 1;
    if (x && y && z) {
      s.xyz[0] = (float)x[s.indx8[0]];
      s.xyz[1] = (float)y[s.indx8[0]];
      s.xyz[2] = (float)z[s.indx8[0]];
    }
  }
  for (int r = 0; r < n; r++) out[r] = st[(size_t)r];
  return n;
}

// ---- source files (src_read_locate_file, src_t.cu:143-795) -------------------------
namespace {

// piecewise Lagrange interpolation of order `order` on the uniform samples
// t[i] = t_start + i dt (LagInterp_Piecewise_1d, interp.cu:220-298): the
// order+1 samples around the nearest one, shifted inside at the ends
double lag_piecewise(const std::vector<double> &t, const std::vector<double> &v, int order, double t_start,
                     double dt, double ti) {
  const int n = (int)t.size();
  int lower, upper;
  if (order % 2 == 0) { lower = order / 2; upper = order / 2; }
  else { lower = order / 2; upper = order / 2 + 1; }
  // nearest sample, clamped to [0, n-1] before the integer cast (an index
  // below 0 selects the same first order+1 samples as 0 does)
  double fi = (ti - t_start) / dt + 0.5;
  if (!(fi >= 0.0)) fi = 0.0;
  if (fi > n - 1) fi = n - 1;
  int indx = (int)fi;
  int a, b;
  if (indx - lower >= 0 && indx + upper <= n - 1) { a = indx - lower; b = indx + upper; }
  else if (indx - lower < 0) { a = 0; b = indx + upper + (lower - indx); }
  else { a = indx - lower - (indx + upper - (n - 1)); b = n - 1; }
  a = std::max(a, 0);
  b = std::min(b, n - 1);
  double li = 0.0;
  for (int i = a; i <= b; i++) {
    double l = 1.0;
    for (int j = a; j <= b; j++)
      if (j != i) l *= (ti - t[j]) / (t[i] - t[j]);
    li += l * v[i];
  }
  return li;
}

int wavelet_kind(const char *name) {
  if (!std::strcmp(name, "ricker")) return CGFD_WAVELET_RICKER;
  if (!std::strcmp(name, "gaussian")) return CGFD_WAVELET_GAUSSIAN;
  if (!std::strcmp(name, "ricker_deriv")) return CGFD_WAVELET_RICKER_DERIV;
  if (!std::strcmp(name, "gaussian_deriv")) return CGFD_WAVELET_GAUSSIAN_DERIV;
  if (!std::strcmp(name, "klauder")) return CGFD_WAVELET_KLAUDER;
  return -1;
}

// strike, dip, rake, mu, D, A -> tensor mu D A * unit double couple (src_muDA_to_moment, src_t.cu:995-1012)
void muda_to_moment(double strike, double dip, double rake, double mu, double D, double A, double *m6) {
  double u[6];
  cgfd_angle2moment(strike, dip, rake, u);
  const double m0 = mu * D * A;
  for (int c = 0; c < 6; c++) m6[c] = m0 * u[c];
}

}  // namespace

static int cgfd_source_file_body(const char *path, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                     const double *mu, double t0, double dt, cgfd_source_file_info *info, int *it_begin,
                     int *it_end, int *ext_num, long long *ext_indx, double *ext_coef, double *F, double *M,
                     int *node, double *shift) {
  if (!path || !info || ni < 2 || nj < 2 || nk < 2 || !(dt > 0)) return CGFD_E_ARG;
  auto bad = [&](const std::string &m) {
    cgfd_internal_set_error(("source file " + std::string(path) + ": " + m).c_str());
    return CGFD_E_ARG;
  };
  FILE *fp = std::fopen(path, "rt");
  if (!fp) return bad("cannot open");
  struct Closer { FILE *f; ~Closer() { std::fclose(f); } } closer{fp};
  char line[500];
  cgfd_source_file_info I;
  std::memset(&I, 0, sizeof(I));
  if (!next_line(fp, line, sizeof(line))) return bad("no event name");
  std::snprintf(I.event, sizeof(I.event), "%s", line);
  if (!next_line(fp, line, sizeof(line)) || std::sscanf(line, "%d", &I.nsrc_file) != 1 || I.nsrc_file <= 0)
    return bad("the source count must be positive");
  float fa = 0, fb = 0;
  int na = 1;
  if (!next_line(fp, line, sizeof(line)) || std::sscanf(line, "%d", &I.stf_given) != 1) return bad("no time-function line");
  if (I.stf_given == 0) {
    std::sscanf(line, "%d %f", &I.stf_given, &fa);
    I.stf_length = fa;
  } else if (I.stf_given == 1) {
    std::sscanf(line, "%d %f %d", &I.stf_given, &fb, &na);
    I.stf_dt = fb;
    I.stf_nt = na;
    if (na < 1 || !(fb > 0)) return bad("sampled time function needs dt > 0 and nt >= 1");
  } else {
    return bad("time function given by name (0) or values (1)");
  }
  if (!next_line(fp, line, sizeof(line)) || std::sscanf(line, "%d %d", &I.cmp_type, &I.mechanism) != 2)
    return bad("no component line");
  if (I.cmp_type < 1 || I.cmp_type > 3) return bad("component type 1, 2 or 3");
  if (!next_line(fp, line, sizeof(line)) || std::sscanf(line, "%d %d", &I.by_coord, &I.depth) != 2)
    return bad("no location line");
  const int ns = I.nsrc_file;
  std::vector<float> loc;  // grown line by line: the count in the file sizes nothing
  for (int q = 0; q < ns; q++) {
    float v[3];
    if (!next_line(fp, line, sizeof(line)) || std::sscanf(line, "%f %f %f", &v[0], &v[1], &v[2]) != 3)
      return bad("location line " + std::to_string(q + 1) + " unreadable");
    loc.insert(loc.end(), v, v + 3);
  }
  // locate: nearest node and signed shift
  const G3 g(ni, nj, nk);
  std::vector<int> nd(3 * (size_t)ns, -1000);
  std::vector<double> sh(3 * (size_t)ns, 0.0);
  std::vector<char> here((size_t)ns, 0);
  if (I.by_coord == 1) {
    if (!x || !y || !z) return bad("locations by coordinates need the grid coordinates");
    const FGrid G(ni, nj, nk, x, y, z);
    for (int q = 0; q < ns; q++) {
      int m[3];
      float inc[3];
      if (locate_raw(G, loc[3 * (size_t)q], loc[3 * (size_t)q + 1], loc[3 * (size_t)q + 2], I.depth == 1, m, inc,
                     nullptr))
        for (int d = 0; d < 3; d++) { nd[3 * (size_t)q + d] = m[d]; sh[3 * (size_t)q + d] = inc[d]; }
    }
  } else {
    const int tot[3] = {ni, nj, nk};
    for (int q = 0; q < ns; q++) {
      float v[3] = {loc[3 * (size_t)q], loc[3 * (size_t)q + 1], loc[3 * (size_t)q + 2]};
      if (I.depth == 1) v[2] = (float)(nk - 1) - v[2];
      for (int d = 0; d < 3; d++) {
        if (!(std::fabs(v[d]) < 1e9f)) {  // far outside or not a number
          nd[3 * (size_t)q + d] = -1000;
          continue;
        }
        const int id = (int)(v[d] + 0.5f);
        sh[3 * (size_t)q + d] = v[d] - (float)id;
        nd[3 * (size_t)q + d] = (id < 0 || id > tot[d] - 1) ? -1000 : id;
      }
    }
  }
  // kept: the node inside the grid (the 7^3 spreading then meets it)
  int kept = 0;
  for (int q = 0; q < ns; q++) {
    const int *m = &nd[3 * (size_t)q];
    here[(size_t)q] = m[0] >= 0 && m[1] >= 0 && m[2] >= 0;
    kept += here[(size_t)q];
  }
  I.nsrc = kept;
  if (kept > 0) {
    if (I.cmp_type == 1 || I.cmp_type == 3) I.force_on = 1;
    else I.moment_on = 1;  // cmp_type 2 (3 never reaches here, as in the reference)
  }
  const double mnt_d = I.stf_given == 0 ? I.stf_length / dt + 0.5 : (I.stf_nt - 1) * I.stf_dt / dt + 0.5;
  if (!(mnt_d >= 1.0 - (I.stf_given == 1) && mnt_d < 1073741824.0))
    return bad("the time function must last at least one step and fewer than 2^30");
  I.max_nt = (int)mnt_d + (I.stf_given == 1);
  *info = I;
  if (!it_begin && !it_end && !ext_num && !ext_indx && !ext_coef && !F && !M) return CGFD_OK;
  if (!it_begin || !it_end || !ext_num || !ext_indx || !ext_coef || (I.force_on && !F) || (I.moment_on && !M))
    return bad("table arrays missing");
  const int mnt = I.max_nt;
  const double stage[4] = {0.0, 0.5, 0.5, 1.0};
  // values that persist from one source to the next, as in the reference
  double coefs[10] = {0}, wstart = 0;
  int wkind = -1;
  double fc[3] = {0, 0, 0}, mc[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> samp[9];  // grown as sample lines are read (zero before)
  int ks = 0;
  for (int q = 0; q < ns; q++) {
    double tstart = 0;
    int kind = -1;
    const int *m = &nd[3 * (size_t)q];
    const long long pnode = here[(size_t)q] ? g.at(m[0] + NG, m[1] + NG, m[2] + NG) : 0;
    auto mu_at = [&](double v) -> double {
      if (v >= 0.0) return v;
      if (!mu) return std::nan("");
      return mu[pnode];
    };
    auto read_cmp = [&](const char *ln, double *f3, double *m6) {
      float v[9];
      int got;
      if (I.cmp_type == 1) {
        got = std::sscanf(ln, "%f %f %f", &v[0], &v[1], &v[2]);
        for (int c = 0; c < got; c++) f3[c] = v[c];
      } else if (I.cmp_type == 2) {
        got = std::sscanf(ln, "%f %f %f %f %f %f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]);
        for (int c = 0; c < got; c++) m6[c] = v[c];
      } else {
        got = std::sscanf(ln, "%f %f %f %f %f %f %f %f %f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7],
                          &v[8]);
        for (int c = 0; c < got; c++) (c < 3 ? f3[c] : m6[c - 3]) = v[c];
      }
    };
    if (I.stf_given == 0) {
      if (!next_line(fp, line, sizeof(line))) return bad("time-function line of source " + std::to_string(q + 1));
      // parsed only for a kept source: the start time and coefficients
      // persist from the last kept source, as in the reference
      if (here[(size_t)q]) {
        char name[500] = {0};
        float ts = 0, c[10];
        const int got = std::sscanf(line, "%f %499s %f %f %f %f %f %f %f %f %f %f", &ts, name, &c[0], &c[1], &c[2],
                                    &c[3], &c[4], &c[5], &c[6], &c[7], &c[8], &c[9]);
        if (got < 2) return bad("time-function line of source " + std::to_string(q + 1));
        for (int i = 0; i + 2 < got; i++) coefs[i] = c[i];
        wstart = ts;
        wkind = wavelet_kind(name);
        if (wkind < 0) return bad(std::string("wavelet ") + name + " not implemented");
      }
      tstart = wstart;
      kind = wkind;
      if (!next_line(fp, line, sizeof(line))) return bad("component line of source " + std::to_string(q + 1));
      if (here[(size_t)q]) {
        read_cmp(line, fc, mc);
        if (I.moment_on && I.mechanism == 1) muda_to_moment(mc[0], mc[1], mc[2], mu_at(mc[3]), mc[4], mc[5], mc);
      }
    } else {
      if (!next_line(fp, line, sizeof(line))) return bad("start time of source " + std::to_string(q + 1));
      float ts = 0;
      std::sscanf(line, "%f", &ts);
      tstart = ts;
      for (int it = 0; it < I.stf_nt; it++) {
        if (!next_line(fp, line, sizeof(line))) return bad("time-function samples of source " + std::to_string(q + 1));
        if (!here[(size_t)q]) continue;
        if (samp[0].size() <= (size_t)it)
          for (auto &v : samp) v.resize((size_t)it + 1, 0.0);
        double f3[3], m6[6];
        for (int c = 0; c < 3; c++) f3[c] = samp[c][(size_t)it];
        for (int c = 0; c < 6; c++) m6[c] = samp[3 + c][(size_t)it];
        read_cmp(line, f3, m6);
        if (I.moment_on && I.mechanism == 1)
          muda_to_moment(m6[0], m6[1], m6[2], mu_at(m6[3]), m6[4], m6[5], m6);
        for (int c = 0; c < 3; c++) samp[c][(size_t)it] = f3[c];
        for (int c = 0; c < 6; c++) samp[3 + c][(size_t)it] = m6[c];
      }
    }
    if (!here[(size_t)q]) continue;
    // spreading, start step, per-(step, stage) values
    ext_num[ks] = spread_gauss(g, m[0], m[1], m[2], sh[3 * (size_t)q], sh[3 * (size_t)q + 1], sh[3 * (size_t)q + 2],
                               ext_indx + 343 * (size_t)ks, ext_coef + 343 * (size_t)ks);
    const double ib_d = (tstart - t0) / dt;
    if (!(std::fabs(ib_d) < 1073741824.0)) return bad("start time of source " + std::to_string(q + 1) + " out of range");
    const int ib = (int)ib_d;
    it_begin[ks] = ib;
    it_end[ks] = ib + mnt - 1;
    const double t_shift = tstart - (ib * dt + t0);
    std::vector<double> tin;
    if (I.stf_given == 1) {
      tin.resize((size_t)I.stf_nt);
      for (int it = 0; it < I.stf_nt; it++) tin[(size_t)it] = tstart + it * I.stf_dt;
    }
    for (int it = ib; it <= ib + mnt - 1; it++) {
      const int r = it - ib;
      for (int st = 0; st < 4; st++) {
        const size_t o = ((size_t)ks * mnt + r) * 4 + st;
        double fv[3], mv[6];
        if (I.stf_given == 0) {
          const double t = r * dt + stage[st] * dt - t_shift;
          const double v = cgfd_wavelet(kind, t, coefs[0], coefs[1]);
          for (int c = 0; c < 3; c++) fv[c] = v * fc[c];
          for (int c = 0; c < 6; c++) mv[c] = v * mc[c];
        } else {
          // the reference evaluates the samples at the absolute step it here
          const double t = it * dt + stage[st] * dt - t_shift;
          for (int c = 0; c < 3; c++) fv[c] = lag_piecewise(tin, samp[c], 3, tstart, I.stf_dt, t);
          for (int c = 0; c < 6; c++) mv[c] = lag_piecewise(tin, samp[3 + c], 3, tstart, I.stf_dt, t);
        }
        if (I.force_on)
          for (int c = 0; c < 3; c++) F[(size_t)c * kept * mnt * 4 + o] = fv[c];
        if (I.moment_on)
          for (int c = 0; c < 6; c++) M[(size_t)c * kept * mnt * 4 + o] = mv[c];
      }
    }
    if (node)
      for (int d = 0; d < 3; d++) node[3 * ks + d] = m[d];
    if (shift)
      for (int d = 0; d < 3; d++) shift[3 * ks + d] = sh[3 * (size_t)q + d];
    ks++;
  }
  return CGFD_OK;
}

int cgfd_storage_scales(int storage, int ni, int nj, int nk, const double *metric, const double *lam,
                        const double *mu, const double *slw, const cgfd_source_desc *src, double dt, int *exp2) {
  if (!exp2 || ni < 1 || nj < 1 || nk < 1 || !(dt > 0)) return CGFD_E_ARG;
  for (int c = 0; c < 9; c++) exp2[c] = 0;
  if (storage != CGFD_STORE_F16) return CGFD_OK;
  if (!metric || !lam || !mu || !slw) return CGFD_E_ARG;
  const G3 g(ni, nj, nk);
  // impedances over the physical grid
  double zs_min = HUGE_VAL, zp_max = 0.0;
  for (int k = NG; k < NG + nk; k++)
    for (int j = NG; j < NG + nj; j++)
      for (int i = NG; i < NG + ni; i++) {
        const long q = g.at(i, j, k);
        if (!(slw[q] > 0)) continue;
        const double rho = 1.0 / slw[q];
        if (mu[q] > 0) zs_min = std::min(zs_min, std::sqrt(mu[q] * rho));
        zp_max = std::max(zp_max, std::sqrt((lam[q] + 2.0 * mu[q]) * rho));
      }
  double s_inj = 0.0, v_inj = 0.0;  // injected stress [Pa] and velocity [m/s]
  if (src && src->nsrc > 0) {
    auto rdv = [&](const void *p, size_t i) {
      return src->dtype == CGFD_DT_F64 ? ((const double *)p)[i] : (double)((const float *)p)[i];
    };
    const size_t per = (size_t)src->max_nt * 4;
    for (int is = 0; is < src->nsrc; is++) {
      double mint = 0.0, fint = 0.0;  // time integrals of max |M_c| and max |F_c| (stage 0 samples)
      const int nst = std::min(src->max_nt, src->it_end[is] - src->it_begin[is] + 1);
      for (int it = 0; it < nst; it++) {
        const size_t q = (size_t)is * per + (size_t)it * 4;
        double mm = 0.0, ff = 0.0;
        if (src->moment_on && src->M)
          for (int c = 0; c < 6; c++) mm = std::max(mm, std::fabs(rdv(src->M, (size_t)c * src->nsrc * per + q)));
        if (src->force_on && src->F)
          for (int c = 0; c < 3; c++) ff = std::max(ff, std::fabs(rdv(src->F, (size_t)c * src->nsrc * per + q)));
        mint += mm * dt;
        fint += ff * dt;
      }
      for (int e = 0; e < src->ext_num[is]; e++) {
        const long long q = src->ext_indx[(size_t)is * src->max_ext + e];
        if (q < 0 || q >= g.n) return CGFD_E_ARG;
        const double coef = std::fabs(rdv(src->ext_coef, (size_t)is * src->max_ext + e));
        const double jac = std::fabs(metric[q]);
        if (!(jac > 0)) return CGFD_E_ARG;
        s_inj += mint * coef / jac;
        v_inj += fint * coef * std::fabs(slw[q]) / jac;
      }
    }
  }
  const double s_bound = s_inj + v_inj * zp_max;
  const double v_bound = (zs_min < HUGE_VAL ? s_inj / zs_min : 0.0) + v_inj;
  auto ex = [](double b) -> int {
    if (!(b > 0) || !std::isfinite(b)) return 0;
    int e = (int)std::ceil(std::log2(b / 4096.0));
    return std::max(-126, std::min(127, e));
  };
  const int ev = ex(v_bound), es = ex(s_bound);
  for (int c = 0; c < 3; c++) exp2[c] = ev;
  for (int c = 3; c < 9; c++) exp2[c] = es;
  return CGFD_OK;
}

}  // extern "C"

// ---- guarded C entry points of the file readers ----------------------------------
int cgfd_media_layer_file(int ni, int nj, int nk, const double *x, const double *y, const double *z, const char *path,
                          const char *method, double *lam, double *mu, double *rho) {
  return guarded([&] { return cgfd_media_layer_file_body(ni, nj, nk, x, y, z, path, method, lam, mu, rho); });
}

int cgfd_media_grid_file(int ni, int nj, int nk, const double *x, const double *y, const double *z, const char *path,
                         const char *method, double *lam, double *mu, double *rho) {
  return guarded([&] { return cgfd_media_grid_file_body(ni, nj, nk, x, y, z, path, method, lam, mu, rho); });
}

int cgfd_media_bin_file(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                        const int order[3], const int size[3], const double spacing[3], const double origin[3],
                        const char *rho_path, const char *vp_path, const char *vs_path, double *lam, double *mu,
                        double *rho) {
  return guarded([&] {
    return cgfd_media_bin_file_body(ni, nj, nk, x, y, z, order, size, spacing, origin, rho_path, vp_path, vs_path, lam,
                                    mu, rho);
  });
}

int cgfd_station_file(const char *path, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                      int max_n, cgfd_station *out, int *n_file) {
  return guarded([&] { return cgfd_station_file_body(path, ni, nj, nk, x, y, z, max_n, out, n_file); });
}

int cgfd_source_file(const char *path, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                     const double *mu, double t0, double dt, cgfd_source_file_info *info, int *it_begin,
                     int *it_end, int *ext_num, long long *ext_indx, double *ext_coef, double *F, double *M,
                     int *node, double *shift) {
  return guarded([&] {
    return cgfd_source_file_body(path, ni, nj, nk, x, y, z, mu, t0, dt, info, it_begin, it_end, ext_num, ext_indx,
                                 ext_coef, F, M, node, shift);
  });
}
