// pfdtd_probe.hip -- HBM streaming probes: the read, write and read+write ceilings of the part
// that the update kernels are measured against. Plain streaming kernels, 16 bytes per lane per
// access, four accesses in flight per lane, one pass over the arrays per launch.
#include "../../include/pfdtd.h"
#include "pfdtd_internal.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;
constexpr int64_t kPerBlock = pfdtd::kProbeBlockElems;  // float4 elements per block
static_assert(kPerBlock == (int64_t)kBlock * kUnroll, "probe block shape");

__global__ __launch_bounds__(kBlock) void probe_read(const float4 *__restrict__ a,
                                                     float *__restrict__ sink, float never) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock + threadIdx.x;
  float4 v[kUnroll];
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) v[k] = a[base + k * kBlock];
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) acc += (v[k].x + v[k].y) + (v[k].z + v[k].w);
  if (acc == never) sink[threadIdx.x] = acc;  // keeps the loads; `never` is not a sum here
}

__global__ __launch_bounds__(kBlock) void probe_write(float4 *__restrict__ a, float s) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) a[base + k * kBlock] = make_float4(s, s, s, s);
}

__global__ __launch_bounds__(kBlock) void probe_copy(const float4 *__restrict__ a,
                                                     float4 *__restrict__ b) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock + threadIdx.x;
  float4 v[kUnroll];
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) v[k] = a[base + k * kBlock];
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) b[base + k * kBlock] = v[k];
}

// a = b + s * c
__global__ __launch_bounds__(kBlock) void probe_triad(float4 *__restrict__ a,
                                                      const float4 *__restrict__ b,
                                                      const float4 *__restrict__ c, float s) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock + threadIdx.x;
  float4 u[kUnroll], w[kUnroll];
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) {
    u[k] = b[base + k * kBlock];
    w[k] = c[base + k * kBlock];
  }
#pragma unroll
  for (int k = 0; k < kUnroll; ++k)
    a[base + k * kBlock] = make_float4(u[k].x + s * w[k].x, u[k].y + s * w[k].y,
                                       u[k].z + s * w[k].z, u[k].w + s * w[k].w);
}

// a = s * a + b, in place: the access pattern of the leapfrog update (P_past overwritten)
__global__ __launch_bounds__(kBlock) void probe_update(float4 *__restrict__ a,
                                                       const float4 *__restrict__ b, float s) {
  const int64_t base = (int64_t)blockIdx.x * kPerBlock + threadIdx.x;
  float4 u[kUnroll], w[kUnroll];
#pragma unroll
  for (int k = 0; k < kUnroll; ++k) {
    u[k] = a[base + k * kBlock];
    w[k] = b[base + k * kBlock];
  }
#pragma unroll
  for (int k = 0; k < kUnroll; ++k)
    a[base + k * kBlock] = make_float4(s * u[k].x + w[k].x, s * u[k].y + w[k].y,
                                       s * u[k].z + w[k].z, s * u[k].w + w[k].w);
}

}  // namespace

namespace pfdtd {

hipError_t launch_hbm_probe(int kind, float4 *a, float4 *b, float4 *c, float *sink,
                            int64_t blocks, hipStream_t st) {
  const dim3 grid((unsigned)blocks), block(kBlock);
  switch (kind) {
  case PFDTD_PROBE_READ:
    hipLaunchKernelGGL(probe_read, grid, block, 0, st, a, sink, -1.f);
    break;
  case PFDTD_PROBE_WRITE:
    hipLaunchKernelGGL(probe_write, grid, block, 0, st, a, 0.5f);
    break;
  case PFDTD_PROBE_COPY:
    hipLaunchKernelGGL(probe_copy, grid, block, 0, st, a, b);
    break;
  case PFDTD_PROBE_TRIAD:
    hipLaunchKernelGGL(probe_triad, grid, block, 0, st, a, b, c, 0.5f);
    break;
  default:
    hipLaunchKernelGGL(probe_update, grid, block, 0, st, a, b, 0.5f);
    break;
  }
  return hipGetLastError();
}

}  // namespace pfdtd
