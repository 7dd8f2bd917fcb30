// el_iso_launch.h -- host-callable launchers of the el_iso kernels.
#pragma once
#include "el_iso_device.h"

namespace cgfd {

template <typename TS, typename TC>
hipError_t launch_point(const StageArgs<TS, TC> &a, int dx, int dy, int dz, int nshots, hipStream_t st);

template <typename TS, typename TC> hipError_t launch_source(const SrcArgs<TS, TC> &a, hipStream_t st);

template <typename TS, typename TC> hipError_t launch_recv(const RecvArgs<TS, TC> &a, hipStream_t st);

hipError_t launch_clock_advance(int *clock, hipStream_t st);

// peak ground motion maps of every shot after one step (free surface only)
template <typename TS, typename TC>
hipError_t launch_peak(const TS *w_new, const TS *w_old, long shot_stride, const Layout &L, TC *pg, TC dt,
                       const float *sc, const int *shot_active, int nshots, hipStream_t st);

// one component's points (i1 + n*di, j1 + m*dj, k1 + l*dk), physical indices,
// into out [nk][nj][ni], multiplied by scale
template <typename TS, typename TC>
hipError_t launch_gather(const TS *w, const Layout &L, int i1, int j1, int k1, int ni, int nj, int nk, int di,
                         int dj, int dk, TC scale, TC *out, hipStream_t st);

template <typename T> hipError_t launch_fill(T *p, long n, T v, hipStream_t st);

// Boxes swept by the z-marching kernel ([i0,i1]x[j0,j1]x[k0,k1], physical,
// inclusive).  Every point must lie at least 4 rows below a free surface.
// With pml set the kernel applies the CFS-PML terms to points inside a face
// box; without it no point of the set may lie in one.
struct MarchSet {
  int nbox;
  int box[6][6];
  int zchunk;  // planes per workgroup along z
  int shape;   // tile shape id, see interior_shape()
  int pml;     // the boxes reach into PML face boxes
};

// per-box tile counts, filled by the launcher for the chosen tile shape
struct MarchGrid {
  int ntx[6], nty[6], nzc[6];
  int wg_beg[7];
};

int interior_num_shapes();
void interior_shape(int id, int *tx, int *ty);

template <typename TS, typename TC>
hipError_t launch_interior(const StageArgs<TS, TC> &a, const MarchSet &b, int dx, int dy, int dz, int nshots,
                           hipStream_t st);

}  // namespace cgfd
