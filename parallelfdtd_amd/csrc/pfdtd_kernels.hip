// pfdtd_kernels.hip -- gfx950 kernels of the SRL_FORWARD time-stepping core.
//
// The pressure update is a bandwidth-bound 7-point stencil: per node it reads the position
// and material bytes, P at the node and its six face neighbours, and P_past, and writes
// P_next over P_past (src/kernels/kernels3d.cu:485-529). No MFMA is involved.
//
// Floating-point contraction is disabled in every kernel so each product and sum is a
// separately rounded IEEE operation in the reference's operand order; with the per-(K, m)
// coefficients built on the host in the same order, results are bit-identical to a CPU
// evaluation of the reference expression.
#include "pfdtd_device.h"

#include <algorithm>

namespace pfdtd {

// ---------------------------------------------------------------------------------------
// SRL_FORWARD update: one tile per block (srl_tile, pfdtd_device.h)
// ---------------------------------------------------------------------------------------
template <typename S, int RY, int SCHEME = 0>
__global__ __launch_bounds__(kTX) void srl_update_kernel(
    const uint8_t *__restrict__ pos, const uint8_t *__restrict__ mat, const S *__restrict__ P,
    S *__restrict__ Pp, const typename CoefOf<typename Fmt<S>::C>::T *__restrict__ coef,
    unsigned ncol, typename CoefOf<typename Fmt<S>::C>::T air, int air_uniform,
    typename Fmt<S>::C lam2, int X, int Y, int Z, int zc) {
  typedef typename CoefOf<typename Fmt<S>::C>::T CE;
  extern __shared__ __align__(16) unsigned char s_dyn[];
  CE *s_coef = reinterpret_cast<CE *>(s_dyn);
  stage_coef(s_coef, coef, ncol);
  srl_tile<S, RY, kTX, SCHEME>(pos, mat, P, Pp, s_coef, ncol, air, air_uniform, lam2, X, Y, Z,
                               zc, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}

// ---------------------------------------------------------------------------------------
// SRL_FORWARD update of a batch of B source configurations on one domain
// ---------------------------------------------------------------------------------------
// Fields hold B values per voxel, slot innermost ([voxel][slot]): a block of 256 threads
// covers 256/B consecutive voxels of one row, lane b of each B-lane group owns slot b. All
// B slots of a voxel share its position/material bytes and coefficients, so those bytes are
// read once per voxel (2 + 12*B bytes per voxel-step in FP32). Each thread marches RY rows of
// its (x, slot) column through `zc` slices with the same z-neighbour queue as the unbatched
// kernel. lgB = log2(B), B in 1..256.
template <typename S, int RY>
__global__ __launch_bounds__(kTX) void srl_update_batch_kernel(
    const uint8_t *__restrict__ pos, const uint8_t *__restrict__ mat, const S *__restrict__ P,
    S *__restrict__ Pp, const typename CoefOf<typename Fmt<S>::C>::T *__restrict__ coef,
    unsigned ncol, typename CoefOf<typename Fmt<S>::C>::T air, int air_uniform,
    typename Fmt<S>::C lam2, int X, int Y, int Z, int zc, int lgB) {
#pragma clang fp contract(off)
  typedef typename Fmt<S>::C C;
  typedef typename CoefOf<C>::T CE;
  extern __shared__ __align__(16) unsigned char s_dyn[];
  CE *s_coef = reinterpret_cast<CE *>(s_dyn);
  stage_coef(s_coef, coef, ncol);
  const int B = 1 << lgB;
  const int b = threadIdx.x & (B - 1);
  const int x = blockIdx.x * (kTX >> lgB) + (int)(threadIdx.x >> lgB);
  const int ybase = blockIdx.y * RY;
  if (x >= X) return;
  const int z0 = 1 + (int)blockIdx.z * zc;
  const int z1 = min(z0 + zc, Z - 1);
  const int64_t sX = X;
  const int64_t XY = (int64_t)X * (int64_t)Y;
  const int64_t eX = (int64_t)B;           // element stride of x
  const int64_t eZ = XY * B;               // of z
  const bool xin = (x > 0) && (x < X - 1);

  bool live[RY], upd[RY];
  int64_t row[RY];  // voxel index of (x, y_r) within a slice
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int y = ybase + r;
    live[r] = y < Y;
    upd[r] = live[r] && xin && y > 0 && y < Y - 1;
    row[r] = (int64_t)min(y, Y - 1) * sX + x;
  }
  const int64_t halo_lo = (int64_t)max(ybase - 1, 0) * sX + x;
  const int64_t halo_hi = (int64_t)min(ybase + RY, Y - 1) * sX + x;

  C pm[RY], pc[RY];
  int64_t base = (int64_t)z0 * XY;  // voxel index of the slice
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    pm[r] = Fmt<S>::ld(P, (base - XY + row[r]) * B + b);
    pc[r] = Fmt<S>::ld(P, (base + row[r]) * B + b);
  }
  for (int z = z0; z < z1; ++z, base += XY) {
    C pn[RY], past[RY], xp[RY], xm[RY];
    unsigned pb[RY], mb[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int64_t v = base + row[r];
      const int64_t e = v * B + b;
      pn[r] = Fmt<S>::ld(P, e + eZ);
      past[r] = Fmt<S>::ld(Pp, e);
      pb[r] = pos[v];
      mb[r] = mat[v];
      xp[r] = Fmt<S>::ld(P, e + eX);
      xm[r] = Fmt<S>::ld(P, e - eX);
    }
    const C ylo = Fmt<S>::ld(P, (base + halo_lo) * B + b);
    const C yhi = Fmt<S>::ld(P, (base + halo_hi) * B + b);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const C yp = (r + 1 < RY) ? pc[r + 1] : yhi;
      const C ym = (r > 0) ? pc[r - 1] : ylo;
      C s = pn[r] + pm[r];
      s = s + yp;
      s = s + ym;
      s = s + xp[r];
      s = s + xm[r];
      const C ret =
          srl_node<C>(pb[r], mb[r], pc[r], s, past[r], s_coef, ncol, air, air_uniform, lam2);
      if (live[r]) Fmt<S>::st(Pp, (base + row[r]) * B + b, upd[r] ? ret : (C)0);
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pm[r] = pc[r];
      pc[r] = pn[r];
    }
  }
}

UpdateGeom plan_update_batch(int X, int Y, int Z, int ry, int zc, int lgB) {
  UpdateGeom g = plan_update(X, Y, Z, ry, zc);
  const int nx = kTX >> lgB;
  g.grid.x = (X + nx - 1) / nx;
  g.lgB = lgB;
  return g;
}

UpdateGeom plan_update(int X, int Y, int Z, int ry, int zc) {
  UpdateGeom g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  g.ry = ry;
  const int zi = Z - 2;
  g.zc = zc < 1 ? 1 : (zc > zi ? zi : zc);
  const int bx = (X + kTX - 1) / kTX;
  const int by = (Y + ry - 1) / ry;
  const int nz = (zi + g.zc - 1) / g.zc;
  g.grid = dim3(bx, by, nz);
  g.block = dim3(kTX, 1, 1);
  g.lgB = 0;
  return g;
}

template <typename S, typename CT, typename C>
static hipError_t launch_update(const UpdateGeom &g, const UpdateArgs<S, CT, C> &a,
                                hipStream_t stream) {
  if ((a.scheme != 0 && a.scheme != 2) || (a.scheme == 2 && g.lgB > 0) || a.ncol < 1 ||
      a.ncol > 256)
    return hipErrorInvalidValue;
  const unsigned ncol = (unsigned)a.ncol;
  const size_t lds = sizeof(CT) * kStageRows * (size_t)ncol;
  if (g.lgB > 0) {
    switch (g.ry) {
#define PFDTD_RYB_CASE(R)                                                                    \
  case R:                                                                                    \
    hipLaunchKernelGGL((srl_update_batch_kernel<S, R>), g.grid, g.block, lds, stream, a.pos, \
                       a.mat, a.P, a.Pp, a.coef, ncol, a.air, a.air_uniform, a.lam2, g.X, g.Y, \
                       g.Z, g.zc, g.lgB);                                                    \
    break;
      PFDTD_RYB_CASE(1)
      PFDTD_RYB_CASE(2)
      PFDTD_RYB_CASE(4)
      PFDTD_RYB_CASE(8)
#undef PFDTD_RYB_CASE
    default:
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (a.scheme == 2) {
    switch (g.ry) {
#define PFDTD_RYC_CASE(R)                                                                    \
  case R:                                                                                    \
    hipLaunchKernelGGL((srl_update_kernel<S, R, 2>), g.grid, g.block, lds, stream, a.pos,    \
                       a.mat, a.P, a.Pp, a.coef, ncol, a.air, a.air_uniform, a.lam2, g.X, g.Y, \
                       g.Z, g.zc);                                                           \
    break;
      PFDTD_RYC_CASE(1)
      PFDTD_RYC_CASE(2)
      PFDTD_RYC_CASE(4)
      PFDTD_RYC_CASE(8)
#undef PFDTD_RYC_CASE
    default:
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  switch (g.ry) {
#define PFDTD_RY_CASE(R)                                                                     \
  case R:                                                                                    \
    hipLaunchKernelGGL((srl_update_kernel<S, R>), g.grid, g.block, lds, stream, a.pos, a.mat,\
                       a.P, a.Pp, a.coef, ncol, a.air, a.air_uniform, a.lam2, g.X, g.Y, g.Z,  \
                       g.zc);                                                                \
    break;
    PFDTD_RY_CASE(1)
    PFDTD_RY_CASE(2)
    PFDTD_RY_CASE(4)
    PFDTD_RY_CASE(8)
#undef PFDTD_RY_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_update_f32(const UpdateGeom &g, const UpdateArgs<float, CoefF, float> &a,
                             hipStream_t stream) {
  return launch_update(g, a, stream);
}
hipError_t launch_update_f64(const UpdateGeom &g, const UpdateArgs<double, CoefD, double> &a,
                             hipStream_t stream) {
  return launch_update(g, a, stream);
}
hipError_t launch_update_f16(const UpdateGeom &g, const UpdateArgs<h16, CoefF, float> &a,
                             hipStream_t stream) {
  return launch_update(g, a, stream);
}
hipError_t launch_update_bf16(const UpdateGeom &g, const UpdateArgs<b16, CoefF, float> &a,
                              hipStream_t stream) {
  return launch_update(g, a, stream);
}

// ---------------------------------------------------------------------------------------
// Receiver gather and source injection (one block)
// ---------------------------------------------------------------------------------------
// Gather reads the field just produced (response[step] = P^{step+1}, kernels3d.cu:164-173);
// injection then applies the next step's samples to the same field, in source order, which
// is where the reference's source loop of the next step writes them (kernels3d.cu:93-104).
template <typename S>
__global__ __launch_bounds__(256) void inject_gather_kernel(S *field, InjectArgs a) {
  inject_gather_block<S>(field, a);
}

hipError_t launch_inject_gather(int precision, void *field, const InjectArgs &a,
                                hipStream_t stream) {
  const dim3 grid(1), block(256);
  switch (precision) {
  case 0:
    hipLaunchKernelGGL(inject_gather_kernel<float>, grid, block, 0, stream, (float *)field, a);
    break;
  case 1:
    hipLaunchKernelGGL(inject_gather_kernel<double>, grid, block, 0, stream, (double *)field,
                       a);
    break;
  case 2:
    hipLaunchKernelGGL(inject_gather_kernel<h16>, grid, block, 0, stream, (h16 *)field, a);
    break;
  case 3:
    hipLaunchKernelGGL(inject_gather_kernel<b16>, grid, block, 0, stream, (b16 *)field, a);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Transparent-source signals
// ---------------------------------------------------------------------------------------
// out[s][n] = in[s][n] - sum_{i<n} h[n-i] in[s][i] (getTransparentSourceSample(+Double),
// SimulationParameters.cpp:300-310, :350-360): one thread per sample, the sum accumulated left
// to right in T exactly as the reference's host loop; h is the FP32 grid impulse response
// (0 past its end).
template <typename T>
__global__ __launch_bounds__(256) void transparent_kernel(const T *__restrict__ in,
                                                          const float *__restrict__ h, int n_ir,
                                                          int n_samples, T *__restrict__ out) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * 256 + threadIdx.x;
  const int s = blockIdx.y;
  if (n >= n_samples) return;
  const T *x = in + (int64_t)s * n_samples;
  T acc = (T)0;
  for (int i = 0; i < n; ++i) {
    const int k = n - i;
    const T hk = k < n_ir ? (T)h[k] : (T)0;
    acc += hk * x[i];
  }
  out[(int64_t)s * n_samples + n] = (T)0 + (x[n] - acc);
}

hipError_t launch_transparent(bool f64, const void *in, const float *h, int n_ir, int n_src,
                              int n_samples, void *out, hipStream_t stream) {
  const dim3 grid((n_samples + 255) / 256, n_src), block(256);
  if (f64)
    hipLaunchKernelGGL(transparent_kernel<double>, grid, block, 0, stream, (const double *)in, h,
                       n_ir, n_samples, (double *)out);
  else
    hipLaunchKernelGGL(transparent_kernel<float>, grid, block, 0, stream, (const float *)in, h,
                       n_ir, n_samples, (float *)out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Domain construction and scan
// ---------------------------------------------------------------------------------------
// Label grid -> position/material bytes. Block = 256 consecutive x of one (y, z) row.
__global__ __launch_bounds__(256) void classify_labels_kernel(const uint8_t *__restrict__ label,
                                                              uint8_t *__restrict__ pos,
                                                              uint8_t *__restrict__ mat, int X,
                                                              int Y, int Z) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  const int z = blockIdx.z;
  if (x >= X) return;
  const int64_t XY = (int64_t)X * Y;
  const int64_t i = (int64_t)z * XY + (int64_t)y * X + x;
  if (label[i] != 0) {
    pos[i] = 0;
    mat[i] = 0;
    return;
  }
  const bool inb[6] = {x > 0, x < X - 1, y > 0, y < Y - 1, z > 0, z < Z - 1};
  const int64_t off[6] = {-1, 1, -(int64_t)X, (int64_t)X, -XY, XY};
  int K = 0;
  bool found = false;
  unsigned m = 0;
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    const unsigned lab = inb[d] ? label[i + off[d]] : 1u;
    if (lab == 0) {
      ++K;
    } else if (!found) {
      found = true;
      m = inb[d] ? lab - 1u : 0u;
    }
  }
  pos[i] = (uint8_t)(0x80 | K);
  mat[i] = (uint8_t)(K < 6 ? m : 0);
}

hipError_t launch_classify_labels(const uint8_t *labels, uint8_t *pos, uint8_t *mat, int X,
                                  int Y, int Z, hipStream_t stream) {
  const dim3 grid((X + 255) / 256, Y, Z), block(256);
  hipLaunchKernelGGL(classify_labels_kernel, grid, block, 0, stream, labels, pos, mat, X, Y, Z);
  return hipGetLastError();
}

// Centred-scheme (update type SRL) position byte from the set of air face neighbours
// (bit 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z): an axis with exactly one air neighbour gets its
// DIR bit (0x01 / 0x02 / 0x04) and a SIGN bit (0x10 / 0x20 / 0x40) naming the present neighbour
// the update mirrors (+x, +y, and -z for SIGN_Z), cudaMesh.h:35-44, kernels3d.cu:644-658.
// Returns 0 when some axis has no air neighbour (no centred stencil exists).
__device__ __forceinline__ unsigned centred_byte(unsigned air) {
  unsigned pb = 0x80u;
  const unsigned ax = air & 3u, ay = (air >> 2) & 3u, az = (air >> 4) & 3u;
  if (!ax || !ay || !az) return 0u;
  if (ax == 1u) pb |= 0x01u;
  if (ax == 2u) pb |= 0x11u;
  if (ay == 1u) pb |= 0x02u;
  if (ay == 2u) pb |= 0x22u;
  if (az == 1u) pb |= 0x44u;
  if (az == 2u) pb |= 0x04u;
  return pb;
}

// Label grid -> centred-scheme bytes (same material rule as classify_labels_kernel); voxels
// without a centred stencil are counted in `degenerate` (one atomic per wavefront).
__global__ __launch_bounds__(256) void classify_labels_centred_kernel(
    const uint8_t *__restrict__ label, uint8_t *__restrict__ pos, uint8_t *__restrict__ mat,
    int X, int Y, int Z, unsigned int *degenerate) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  const int z = blockIdx.z;
  bool bad = false;
  if (x < X) {
    const int64_t XY = (int64_t)X * Y;
    const int64_t i = (int64_t)z * XY + (int64_t)y * X + x;
    if (label[i] != 0) {
      pos[i] = 0;
      mat[i] = 0;
    } else {
      const bool inb[6] = {x > 0, x < X - 1, y > 0, y < Y - 1, z > 0, z < Z - 1};
      const int64_t off[6] = {-1, 1, -(int64_t)X, (int64_t)X, -XY, XY};
      unsigned air = 0, m = 0;
      bool found = false;
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        const unsigned lab = inb[d] ? label[i + off[d]] : 1u;
        if (lab == 0)
          air |= 1u << d;
        else if (!found) {
          found = true;
          m = inb[d] ? lab - 1u : 0u;
        }
      }
      unsigned pb = centred_byte(air);
      if (!pb) {
        bad = true;
        // keep the DIR/SIGN bits of the axes that have a stencil
        pb = 0x80u;
        const unsigned ax = air & 3u, ay = (air >> 2) & 3u, az = (air >> 4) & 3u;
        if (ax == 1u) pb |= 0x01u;
        if (ax == 2u) pb |= 0x11u;
        if (ay == 1u) pb |= 0x02u;
        if (ay == 2u) pb |= 0x22u;
        if (az == 1u) pb |= 0x44u;
        if (az == 2u) pb |= 0x04u;
      }
      pos[i] = (uint8_t)pb;
      mat[i] = (uint8_t)(air == 0x3Fu ? 0u : m);
    }
  }
  const uint64_t b = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(degenerate, (unsigned)__popcll(b));
}

hipError_t launch_classify_labels_centred(const uint8_t *labels, uint8_t *pos, uint8_t *mat,
                                          int X, int Y, int Z, unsigned int *degenerate,
                                          hipStream_t stream) {
  const dim3 grid((X + 255) / 256, Y, Z), block(256);
  hipLaunchKernelGGL(classify_labels_centred_kernel, grid, block, 0, stream, labels, pos, mat, X,
                     Y, Z, degenerate);
  return hipGetLastError();
}

// toKowalczyk (src/kernels/cudaMesh.cu:363-480) applied in place to voxelizer codes: each code
// names its air directions; the byte follows from them as in centred_byte.
__global__ __launch_bounds__(256) void codes_to_centred_kernel(uint8_t *pos, uint8_t *mat,
                                                               int64_t n) {
  // air-direction sets of codes 1..27 (bit 0 -x "Left", 1 +x "Right", 2 -y "In", 3 +y "Out",
  // 4 -z "Down", 5 +z "Up")
  constexpr uint8_t kAir[28] = {
      0x00, 0x15, 0x16, 0x19, 0x1A, 0x25, 0x26, 0x29, 0x2A, 0x17, 0x1B, 0x1D, 0x1E, 0x27,
      0x2B, 0x2D, 0x2E, 0x35, 0x36, 0x39, 0x3A, 0x1F, 0x3B, 0x37, 0x3E, 0x3D, 0x2F, 0x3F};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const unsigned k = pos[i];
    if (k == 0) {
      pos[i] = 0;
      mat[i] = 0;
    } else if (k <= 27) {
      pos[i] = (uint8_t)centred_byte(kAir[k]);
    }
  }
}

hipError_t launch_codes_to_centred(uint8_t *pos, uint8_t *mat, int64_t n, hipStream_t stream) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(codes_to_centred_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, pos,
                     mat, n);
  return hipGetLastError();
}

// toBilbao (src/kernels/cudaMesh.cu:328-361) applied in place to voxelizer codes.
__global__ __launch_bounds__(256) void codes_to_bilbao_kernel(uint8_t *pos, uint8_t *mat,
                                                              int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const unsigned k = pos[i];
    if (k == 0) {
      pos[i] = 0;
      mat[i] = 0;
    } else if (k <= 8) {
      pos[i] = 0x83;
    } else if (k <= 20) {
      pos[i] = 0x84;
    } else if (k <= 26) {
      pos[i] = 0x85;
    } else if (k == 27) {
      pos[i] = 0x86;
    }
  }
}

hipError_t launch_codes_to_bilbao(uint8_t *pos, uint8_t *mat, int64_t n, hipStream_t stream) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(codes_to_bilbao_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, pos,
                     mat, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void pack_pm_kernel(const uint8_t *__restrict__ pos,
                                                      const uint8_t *__restrict__ mat,
                                                      uint16_t *__restrict__ pm, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    pm[i] = (uint16_t)(pos[i] | ((unsigned)mat[i] << 8));
}

hipError_t launch_pack_pm(const uint8_t *pos, const uint8_t *mat, uint16_t *pm, int64_t n,
                          hipStream_t stream) {
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
  hipLaunchKernelGGL(pack_pm_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, pos, mat, pm,
                     n);
  return hipGetLastError();
}

// Counts air/boundary voxels (calcBoundaries, cudaMesh.cu:500-514; air value 0x86 for the
// forward scheme, 0x80 for the centred one, cudaMesh.cu:176-186) and checks the
// preconditions the update relies on. Blocks stride over (y, z) rows, threads over x; each
// thread keeps private counters and a 64-bit set of the material bytes below 64 it saw, the
// wavefront reduces them with cross-lane shuffles, and only one atomic per wavefront and
// counter reaches LDS (one per block reaches global memory).
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ unsigned long long wave_or(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

__global__ __launch_bounds__(256) void scan_domain_kernel(const uint8_t *__restrict__ pos,
                                                          const uint8_t *__restrict__ mat, int X,
                                                          int Y, int Z, unsigned air_value,
                                                          int check_k, DomainScan *out) {
  __shared__ unsigned long long s_cnt[4];
  __shared__ unsigned int s_mat[8];
  if (threadIdx.x < 8) s_mat[threadIdx.x] = 0;
  if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long air = 0, bnd = 0, shell = 0, badk = 0, mlo = 0;
  const int64_t rows = (int64_t)Y * Z;
  for (int64_t row = blockIdx.x; row < rows; row += gridDim.x) {
    const int y = (int)(row % Y);
    const int z = (int)(row / Y);
    const bool edge_row = (y == 0 || y == Y - 1 || z == 0 || z == Z - 1);
    const uint8_t *prow = pos + row * X;
    const uint8_t *mrow = mat + row * X;
    for (int x = threadIdx.x; x < X; x += 256) {
      const unsigned p = prow[x];
      const unsigned m = mrow[x];
      air += (p == air_value);
      bnd += (p != 0u && p != air_value);
      if (p >> 7) {
        badk += check_k && ((p & 0x7Fu) > 6u);
        shell += (edge_row || x == 0 || x == X - 1);
      }
      if (m < 64u)
        mlo |= 1ull << m;
      else
        atomicOr(&s_mat[m >> 5], 1u << (m & 31u));
    }
  }
  air = wave_sum(air);
  bnd = wave_sum(bnd);
  shell = wave_sum(shell);
  badk = wave_sum(badk);
  mlo = wave_or(mlo);
  if ((threadIdx.x & 63) == 0) {
    if (air) atomicAdd(&s_cnt[0], air);
    if (bnd) atomicAdd(&s_cnt[1], bnd);
    if (shell) atomicAdd(&s_cnt[2], shell);
    if (badk) atomicAdd(&s_cnt[3], badk);
    if (mlo & 0xFFFFFFFFull) atomicOr(&s_mat[0], (unsigned)(mlo & 0xFFFFFFFFull));
    if (mlo >> 32) atomicOr(&s_mat[1], (unsigned)(mlo >> 32));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_cnt[0]) atomicAdd(&out->air, s_cnt[0]);
    if (s_cnt[1]) atomicAdd(&out->boundary, s_cnt[1]);
    if (s_cnt[2]) atomicAdd(&out->shell_inside, (unsigned)min(s_cnt[2], 0xFFFFFFFFull));
    if (s_cnt[3]) atomicAdd(&out->bad_k, (unsigned)min(s_cnt[3], 0xFFFFFFFFull));
  }
  if (threadIdx.x < 8 && s_mat[threadIdx.x]) atomicOr(&out->mat_present[threadIdx.x], s_mat[threadIdx.x]);
}

hipError_t launch_scan_domain(const uint8_t *pos, const uint8_t *mat, int X, int Y, int Z,
                              int scheme, DomainScan *d_out, hipStream_t stream) {
  const int64_t rows = (int64_t)Y * Z;
  const int blocks = (int)std::min<int64_t>(rows, 8192);
  const unsigned air_value = scheme == 2 ? 0x80u : 0x86u;
  hipLaunchKernelGGL(scan_domain_kernel, dim3(blocks), dim3(256), 0, stream, pos, mat, X, Y, Z,
                     air_value, scheme == 2 ? 0 : 1, d_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Capture
// ---------------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void capture_slice_kernel(const S *__restrict__ field,
                                                            const uint8_t *__restrict__ pos,
                                                            int X, int Y, int Z, int64_t slice,
                                                            int orientation, float inv_scale,
                                                            float *out_p, uint8_t *out_pos) {
  // Plane coordinates (u fastest, v): xy -> (x, y), xz -> (x, z), yz -> (y, z).
  const int U = orientation == 2 ? Y : X;
  const int V = orientation == 0 ? Y : Z;
  const int64_t n = (int64_t)U * V;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * 256) {
    const int u = (int)(t % U);
    const int v = (int)(t / U);
    int64_t x, y, z;
    if (orientation == 0) {
      x = u;
      y = v;
      z = slice;
    } else if (orientation == 1) {
      x = u;
      y = slice;
      z = v;
    } else {
      x = slice;
      y = u;
      z = v;
    }
    const int64_t i = (z * Y + y) * X + x;
    out_p[t] = (float)Fmt<S>::ld(field, i) * inv_scale;
    out_pos[t] = pos[i];
  }
}

template <typename S>
__global__ __launch_bounds__(256) void field_to_f32_kernel(const S *__restrict__ field,
                                                           int64_t n, float inv_scale,
                                                           float *out) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * 256)
    out[t] = (float)Fmt<S>::ld(field, t) * inv_scale;
}

static unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

hipError_t launch_capture_slice(int precision, const void *field, const uint8_t *pos, int X,
                                int Y, int Z, int64_t slice, int orientation, int scale_exp,
                                float *out_p, uint8_t *out_pos, hipStream_t stream) {
  const int U = orientation == 2 ? Y : X;
  const int V = orientation == 0 ? Y : Z;
  const dim3 grid(grid_for((int64_t)U * V)), block(256);
  const float inv_scale = ldexpf(1.f, -scale_exp);
  switch (precision) {
  case 0:
    hipLaunchKernelGGL(capture_slice_kernel<float>, grid, block, 0, stream, (const float *)field,
                       pos, X, Y, Z, slice, orientation, inv_scale, out_p, out_pos);
    break;
  case 1:
    hipLaunchKernelGGL(capture_slice_kernel<double>, grid, block, 0, stream,
                       (const double *)field, pos, X, Y, Z, slice, orientation, inv_scale, out_p,
                       out_pos);
    break;
  case 2:
    hipLaunchKernelGGL(capture_slice_kernel<h16>, grid, block, 0, stream, (const h16 *)field,
                       pos, X, Y, Z, slice, orientation, inv_scale, out_p, out_pos);
    break;
  case 3:
    hipLaunchKernelGGL(capture_slice_kernel<b16>, grid, block, 0, stream, (const b16 *)field,
                       pos, X, Y, Z, slice, orientation, inv_scale, out_p, out_pos);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_field_to_f32(int precision, const void *field, int64_t n, int scale_exp,
                               float *out, hipStream_t stream) {
  const dim3 grid(grid_for(n)), block(256);
  const float inv_scale = ldexpf(1.f, -scale_exp);
  switch (precision) {
  case 0:
    hipLaunchKernelGGL(field_to_f32_kernel<float>, grid, block, 0, stream, (const float *)field,
                       n, inv_scale, out);
    break;
  case 1:
    hipLaunchKernelGGL(field_to_f32_kernel<double>, grid, block, 0, stream,
                       (const double *)field, n, inv_scale, out);
    break;
  case 2:
    hipLaunchKernelGGL(field_to_f32_kernel<h16>, grid, block, 0, stream, (const h16 *)field, n,
                       inv_scale, out);
    break;
  case 3:
    hipLaunchKernelGGL(field_to_f32_kernel<b16>, grid, block, 0, stream, (const b16 *)field, n,
                       inv_scale, out);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pfdtd
