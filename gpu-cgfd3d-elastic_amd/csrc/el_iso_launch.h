// el_iso_launch.h -- host-callable launchers of the el_iso kernels.
#pragma once
#include "el_iso_device.h"

namespace cgfd {

template <typename TS, typename TC>
hipError_t launch_point(const StageArgs<TS, TC> &a, int dx, int dy, int dz, int nshots, hipStream_t st);

template <typename TS, typename TC> hipError_t launch_source(const SrcArgs<TS, TC> &a, hipStream_t st);

template <typename TS, typename TC> hipError_t launch_recv(const RecvArgs<TS, TC> &a, hipStream_t st);

hipError_t launch_clock_advance(int *clock, hipStream_t st);

template <typename T> hipError_t launch_fill(T *p, long n, T v, hipStream_t st);

// z-marching interior kernel over the box [i0,i1]x[j0,j1]x[k0,k1] (physical,
// inclusive); the box must keep every point at least 4 rows below the top and
// outside every PML face box.
struct InteriorBox {
  int i0, i1, j0, j1, k0, k1;
  int zchunk;  // planes per workgroup along z
  int shape;   // tile shape id, see interior_shape()
};

int interior_num_shapes();
void interior_shape(int id, int *tx, int *ty);

template <typename TS, typename TC>
hipError_t launch_interior(const StageArgs<TS, TC> &a, const InteriorBox &b, int dx, int dy, int dz, int nshots,
                           hipStream_t st);

}  // namespace cgfd
