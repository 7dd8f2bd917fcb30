// hbm_stream.hip -- the achievable HBM bandwidth of this device, measured
// with the simplest streaming kernels that reach it: a read-only reduction
// and a copy, both 16 bytes per lane and access (dwordx4), non-temporal, a
// grid of many workgroups striding over the whole buffer so every XCD and
// every HBM stack stays busy.  The sweep kernels are compared with these
// ceilings (no reference counterpart: a measurement utility).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "../../include/cgfd_el_iso.h"
#include "cgfd_internal.h"

namespace {

constexpr int kBlock = 256;
typedef float v4f __attribute__((ext_vector_type(4)));   // one 16-byte access
constexpr int kUnroll = 4;   // independent 16-byte loads in flight per lane

__global__ __launch_bounds__(kBlock) void k_stream_read(const v4f *__restrict__ src, long n4, float *partial) {
  const long stride = (long)gridDim.x * kBlock;
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float acc = 0.0f;
  for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
    v4f v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kUnroll; u++) acc += (v[u].x + v[u].y) + (v[u].z + v[u].w);
  }
  for (; i < n4; i += stride) {
    const v4f v = __builtin_nontemporal_load(src + i);
    acc += (v.x + v.y) + (v.z + v.w);
  }
  // one value per thread keeps the loads live; written through vector stores
  partial[(long)blockIdx.x * kBlock + threadIdx.x] = acc;
}

__global__ __launch_bounds__(kBlock) void k_stream_copy(const v4f *__restrict__ src, v4f *__restrict__ dst,
                                                        long n4) {
  const long stride = (long)gridDim.x * kBlock;
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n4; i += kUnroll * stride) {
    v4f v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < kUnroll; u++) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n4; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// 4-byte loads per lane (the granularity of the sweep kernels' loads), plain
// cached loads, 8 independent loads in flight per lane
constexpr int kUnrollDw = 8;
__global__ __launch_bounds__(kBlock) void k_stream_read_dw(const float *__restrict__ src, long n, float *partial) {
  const long stride = (long)gridDim.x * kBlock;
  long i = (long)blockIdx.x * kBlock + threadIdx.x;
  float acc = 0.0f;
  for (; i + (kUnrollDw - 1) * stride < n; i += kUnrollDw * stride) {
    float v[kUnrollDw];
#pragma unroll
    for (int u = 0; u < kUnrollDw; u++) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < kUnrollDw; u++) acc += v[u];
  }
  for (; i < n; i += stride) acc += src[i];
  partial[(long)blockIdx.x * kBlock + threadIdx.x] = acc;
}

__global__ void k_stream_fill(v4f *p, long n4) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    p[i] = v4f{1.0f, 2.0f, 3.0f, 4.0f};
}

}  // namespace

extern "C" int cgfd_hbm_stream(long long bytes, int reps, double *read_gbps, double *copy_gbps) {
  if (bytes < (1LL << 20) || reps < 1 || !read_gbps || !copy_gbps) {
    cgfd_internal_set_error("cgfd_hbm_stream: bytes >= 1 MiB, reps >= 1 and both outputs required");
    return CGFD_E_ARG;
  }
  const long n4 = (long)(bytes / 16);
  const int nblk = 8 * 256 * 8;  // 8 workgroups per CU slot pass over 256 CUs, a multiple of the 8 XCDs
  v4f *a = nullptr, *b = nullptr;
  float *part = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CGFD_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == CGFD_OK) {
      char buf[256];
      snprintf(buf, sizeof(buf), "cgfd_hbm_stream: %s: %s", what, hipGetErrorString(e));
      cgfd_internal_set_error(buf);
      rc = e == hipErrorOutOfMemory ? CGFD_E_NOMEM : CGFD_E_HIP;
    }
    return rc == CGFD_OK;
  };
  if (chk(hipMalloc(&a, (size_t)n4 * 16), "hipMalloc") && chk(hipMalloc(&b, (size_t)n4 * 16), "hipMalloc") &&
      chk(hipMalloc(&part, (size_t)nblk * kBlock * sizeof(float)), "hipMalloc") &&
      chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream") && chk(hipEventCreate(&e0), "event") &&
      chk(hipEventCreate(&e1), "event")) {
    hipLaunchKernelGGL(k_stream_fill, dim3(4096), dim3(256), 0, st, a, n4);
    hipLaunchKernelGGL(k_stream_fill, dim3(4096), dim3(256), 0, st, b, n4);
    // warm once, then time reps launches of each
    hipLaunchKernelGGL(k_stream_read, dim3(nblk), dim3(kBlock), 0, st, a, n4, part);
    hipLaunchKernelGGL(k_stream_copy, dim3(nblk), dim3(kBlock), 0, st, a, b, n4);
    float ms = 0;
    if (chk(hipEventRecord(e0, st), "record")) {
      for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream_read, dim3(nblk), dim3(kBlock), 0, st, a, n4, part);
      if (chk(hipEventRecord(e1, st), "record") && chk(hipEventSynchronize(e1), "sync") &&
          chk(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
        *read_gbps = (double)n4 * 16 * reps / (ms * 1e-3) / 1e9;
    }
    if (rc == CGFD_OK && chk(hipEventRecord(e0, st), "record")) {
      for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream_copy, dim3(nblk), dim3(kBlock), 0, st, a, b, n4);
      if (chk(hipEventRecord(e1, st), "record") && chk(hipEventSynchronize(e1), "sync") &&
          chk(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
        *copy_gbps = 2.0 * (double)n4 * 16 * reps / (ms * 1e-3) / 1e9;   // read + write bytes
    }
    chk(hipGetLastError(), "launch");
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (part) (void)hipFree(part);
  return rc;
}

extern "C" int cgfd_hbm_stream_dword(long long bytes, int reps, double *read_gbps) {
  if (bytes < (1LL << 20) || reps < 1 || !read_gbps) {
    cgfd_internal_set_error("cgfd_hbm_stream_dword: bytes >= 1 MiB, reps >= 1 and the output required");
    return CGFD_E_ARG;
  }
  const long n = (long)(bytes / 4);
  const int nblk = 8 * 256 * 8;
  float *a = nullptr, *part = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CGFD_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == CGFD_OK) {
      char buf[256];
      snprintf(buf, sizeof(buf), "cgfd_hbm_stream_dword: %s: %s", what, hipGetErrorString(e));
      cgfd_internal_set_error(buf);
      rc = e == hipErrorOutOfMemory ? CGFD_E_NOMEM : CGFD_E_HIP;
    }
    return rc == CGFD_OK;
  };
  if (chk(hipMalloc(&a, (size_t)n * 4), "hipMalloc") &&
      chk(hipMalloc(&part, (size_t)nblk * kBlock * sizeof(float)), "hipMalloc") &&
      chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream") && chk(hipEventCreate(&e0), "event") &&
      chk(hipEventCreate(&e1), "event")) {
    hipLaunchKernelGGL(k_stream_fill, dim3(4096), dim3(256), 0, st, (v4f *)a, n / 4);
    hipLaunchKernelGGL(k_stream_read_dw, dim3(nblk), dim3(kBlock), 0, st, a, n, part);
    float ms = 0;
    if (chk(hipEventRecord(e0, st), "record")) {
      for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream_read_dw, dim3(nblk), dim3(kBlock), 0, st, a, n, part);
      if (chk(hipEventRecord(e1, st), "record") && chk(hipEventSynchronize(e1), "sync") &&
          chk(hipEventElapsedTime(&ms, e0, e1), "elapsed"))
        *read_gbps = (double)n * 4 * reps / (ms * 1e-3) / 1e9;
    }
    chk(hipGetLastError(), "launch");
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (part) (void)hipFree(part);
  return rc;
}
