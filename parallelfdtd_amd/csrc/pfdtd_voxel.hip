// pfdtd_voxel.hip -- solid voxelization of a closed triangle mesh into a label grid
// (0 = air inside the mesh, m+1 = solid voxel of material m), on the GPU.
//
// Replaces the reference's call into the external hakarlss/Voxelizer library
// (voxelizeGeometry, src/kernels/voxelizationUtils.cu:58-131, called from App::initializeMesh,
// src/App.cpp:178-187), which is not part of the reference's sources. README.md:100 gives
// its method: intersection tests decide inside/outside, so meshes must be watertight. The
// definition implemented here:
//  * frame: vertices minus their per-axis minimum (float subtraction, as
//    GeometryHandler::initialize, src/base/GeometryHandler.cpp:64-73), then fixed point
//    q = llrint(v * 4096 / dx) in double arithmetic;
//  * node (i,j,k) sits at ((i-1)dx, (j-1)dx, (k-1)dx) -- the +1 padding offset of
//    Position::getElementIdx (src/base/SrcRec.cpp:26-34); floor(extent/dx) + 3 nodes per
//    axis, rounded up to a multiple of pad[a] at the high end (padWithZeros,
//    src/kernels/cudaMesh.cu:253-326, with App.cpp:193's (32,4,1));
//  * a node is air when the x-ray through its centre crosses the mesh an odd number of
//    times before it: triangle t crosses column (y,z) when the column point lies in its yz
//    projection (exact integer edge functions; a point on an edge belongs to the triangle
//    whose CCW edge points down, or left-to-right when horizontal, so shared edges count
//    once), at x* = xA + (wB (xB-xA) + wC (xC-xA)) / W in double without contraction; the
//    crossing flips every node with x >= x*;
//  * the same rasterisation along y and z records, for every link between neighbouring
//    nodes, the lowest index of a triangle crossing it; a solid node's label is 1 + the
//    material of the crossing on the link to its first air neighbour (-x,+x,-y,+y,-z,+z),
//    1 when it has no air neighbour or the link recorded no crossing.
//
// Kernels (one launch each, 64-bit indices):
//  vox_raster_kernel  -- one work item per (triangle, candidate column) pair, found by binary
//                        search over the per-triangle running column count; parity events
//                        are XORed into a bit grid (x rows of 32-bit words), link crossings
//                        min-reduced with 32-bit atomics;
//  vox_prefix_kernel  -- one thread per x row: prefix XOR of the events -> air bits;
//  vox_label_kernel   -- one thread per node (3D grid over x-row blocks, y, z): air bit, first
//                        air neighbour, link material.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pfdtd_internal.h"

namespace pfdtd {

namespace {

constexpr int64_t kVoxSub = 4096;
constexpr int64_t kVoxMaxDim = 32768;
constexpr uint32_t kVoxNone = 0xFFFFFFFFu;

struct VoxRaster {
  const int32_t *q;     // fixed-point vertices [n_vert][3]
  const uint32_t *tri;  // [n_tri][3]
  const int64_t *offs;  // [n_tri + 1] running count of candidate columns
  const int32_t *box;   // [n_tri][3]: first u node, first v node, u node count
  int64_t n_tri;
  int64_t total;
  int axis;  // ray axis; the projection plane is (u, v) = ((axis+1)%3, (axis+2)%3)
  int64_t dim[3];
  int64_t words;   // 32-bit words per x row of the bit grid
  uint32_t *bits;  // parity events (ray axis x only), else null
  uint32_t *link;  // lowest crossing triangle per link along this axis
};

__device__ __forceinline__ int64_t vox_edge(int64_t ux, int64_t uy, int64_t vx, int64_t vy,
                                            int64_t px, int64_t py) {
  return (vx - ux) * (py - uy) - (vy - uy) * (px - ux);
}

__device__ __forceinline__ bool vox_incl(int64_t w, int64_t ex, int64_t ey) {
  return w > 0 || (w == 0 && (ey < 0 || (ey == 0 && ex > 0)));
}

__global__ __launch_bounds__(256) void vox_raster_kernel(VoxRaster r) {
#pragma clang fp contract(off)
  const int au = (r.axis + 1) % 3, av = (r.axis + 2) % 3;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < r.total; g += stride) {
    int64_t lo = 0, hi = r.n_tri;  // offs[lo] <= g < offs[hi]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (r.offs[mid] <= g) lo = mid;
      else hi = mid;
    }
    const int64_t t = lo;
    const int64_t local = g - r.offs[t];
    const int64_t nu_cnt = r.box[3 * t + 2];
    const int64_t nu = r.box[3 * t] + local % nu_cnt;
    const int64_t nv = r.box[3 * t + 1] + local / nu_cnt;
    const int32_t *A = r.q + 3 * (int64_t)r.tri[3 * t];
    const int32_t *B = r.q + 3 * (int64_t)r.tri[3 * t + 1];
    const int32_t *C = r.q + 3 * (int64_t)r.tri[3 * t + 2];
    const int64_t ax = A[au], ay = A[av], bx = B[au], by = B[av], cx = C[au], cy = C[av];
    const int64_t area = vox_edge(ax, ay, bx, by, cx, cy);
    if (area == 0) continue;
    const int64_t s = area > 0 ? 1 : -1;
    const int64_t px = (nu - 1) * kVoxSub, py = (nv - 1) * kVoxSub;
    const int64_t wa = s * vox_edge(bx, by, cx, cy, px, py);
    const int64_t wb = s * vox_edge(cx, cy, ax, ay, px, py);
    const int64_t wc = s * vox_edge(ax, ay, bx, by, px, py);
    if (!vox_incl(wa, s * (cx - bx), s * (cy - by)) || !vox_incl(wb, s * (ax - cx), s * (ay - cy)) ||
        !vox_incl(wc, s * (bx - ax), s * (by - ay)))
      continue;
    const double W = (double)(s * area);
    const double num = (double)wb * (double)(B[r.axis] - A[r.axis]) +
                       (double)wc * (double)(C[r.axis] - A[r.axis]);
    const double h = (double)A[r.axis] + num / W;
    int64_t n0 = (int64_t)ceil(h / (double)kVoxSub) + 1;
    if (n0 < 1) n0 = 1;
    if (n0 >= r.dim[r.axis]) continue;
    int64_t c[3];
    c[r.axis] = n0;
    c[au] = nu;
    c[av] = nv;
    const int64_t idx = (c[2] * r.dim[1] + c[1]) * r.dim[0] + c[0];
    if (r.bits)
      atomicXor(r.bits + (c[2] * r.dim[1] + c[1]) * r.words + (c[0] >> 5), 1u << (c[0] & 31));
    atomicMin(r.link + idx, (uint32_t)t);
  }
}

__global__ __launch_bounds__(256) void vox_prefix_kernel(uint32_t *bits, int64_t rows,
                                                         int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += stride) {
    uint32_t carry = 0;
    uint32_t *row = bits + r * words;
    for (int64_t w = 0; w < words; ++w) {
      uint32_t x = row[w];
      x ^= x << 1;
      x ^= x << 2;
      x ^= x << 4;
      x ^= x << 8;
      x ^= x << 16;
      x ^= carry;
      row[w] = x;
      carry = (x & 0x80000000u) ? 0xFFFFFFFFu : 0u;
    }
  }
}

// grid (ceil(X/256), Y, Z): one thread per node of an x row, no index divisions
__global__ __launch_bounds__(256) void vox_label_kernel(const uint32_t *bits, const uint32_t *link,
                                                        const uint8_t *tmat, int64_t X, int64_t Y,
                                                        int64_t Z, int64_t words, uint8_t *labels) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= X) return;
  const int64_t y = blockIdx.y, z = blockIdx.z;
  const int64_t N = X * Y * Z;
  const int64_t idx = (z * Y + y) * X + x;
  auto air = [&](int64_t xx, int64_t yy, int64_t zz) -> bool {
    return (bits[(zz * Y + yy) * words + (xx >> 5)] >> (xx & 31)) & 1u;
  };
  if (air(x, y, z)) {
    labels[idx] = 0;
    return;
  }
  uint8_t lab = 1;
  const bool ok[6] = {x > 0, x + 1 < X, y > 0, y + 1 < Y, z > 0, z + 1 < Z};
  const int64_t nx[6] = {x - 1, x + 1, x, x, x, x};
  const int64_t ny[6] = {y, y, y - 1, y + 1, y, y};
  const int64_t nz[6] = {z, z, z, z, z - 1, z + 1};
  const int64_t lk[6] = {idx, idx + 1, N + idx, N + idx + X, 2 * N + idx, 2 * N + idx + X * Y};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (!ok[k] || !air(nx[k], ny[k], nz[k])) continue;
    const uint32_t t = link[lk[k]];
    if (t != kVoxNone) lab = (uint8_t)(1 + (tmat ? tmat[t] : 0));
    break;
  }
  labels[idx] = lab;
}

inline unsigned grid_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 65536));
}

template <typename T>
hipError_t upload(T **dst, const T *src, size_t n, hipStream_t s) {
  hipError_t e = hipMalloc((void **)dst, std::max<size_t>(n, 1) * sizeof(T));
  if (e == hipSuccess && n) e = hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, s);
  return e;
}

}  // namespace

int voxel_prepare(const float *v, int64_t nv, float dx, const int32_t *pad, int64_t dims[3],
                  std::vector<int32_t> *q) {
  if (!v || nv <= 0 || !(dx > 0.f) || !std::isfinite(dx)) return -1;
  float mn[3] = {v[0], v[1], v[2]};
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float x = v[3 * i + a];
      if (!std::isfinite(x)) return -1;
      mn[a] = std::min(mn[a], x);
    }
  if (q) q->resize((size_t)(3 * nv));
  int64_t ext[3] = {0, 0, 0};
  const double lim = (double)(kVoxMaxDim - 3) * kVoxSub;
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float r = v[3 * i + a] - mn[a];
      const double s = ((double)r * (double)kVoxSub) / (double)dx;
      if (!(s <= lim)) return -1;
      const long long k = std::llrint(s);
      if (q) (*q)[(size_t)(3 * i + a)] = (int32_t)k;
      ext[a] = std::max<int64_t>(ext[a], k);
    }
  for (int a = 0; a < 3; ++a) {
    const int64_t p = (pad && pad[a] > 0) ? pad[a] : 1;
    int64_t n = ext[a] / kVoxSub + 3;
    n = (n + p - 1) / p * p;
    if (n > kVoxMaxDim) return -1;
    dims[a] = n;
  }
  return 0;
}

hipError_t voxelize_labels(const std::vector<int32_t> &q, const uint32_t *tri, int64_t n_tri,
                           const uint8_t *tmat, const int64_t dims[3], uint8_t *d_labels,
                           hipStream_t s) {
  const int64_t X = dims[0], Y = dims[1], Z = dims[2], N = X * Y * Z;
  const int64_t words = (X + 31) / 32;
  int32_t *d_q = nullptr, *d_box = nullptr;
  uint32_t *d_tri = nullptr, *d_bits = nullptr, *d_link = nullptr;
  int64_t *d_offs = nullptr;
  uint8_t *d_tmat = nullptr;
  // candidate columns per triangle and ray axis: the node columns inside its projected bbox
  std::vector<int64_t> offs[3];
  std::vector<int32_t> box[3];
  for (int axis = 0; axis < 3; ++axis) {
    const int au = (axis + 1) % 3, av = (axis + 2) % 3;
    offs[axis].assign((size_t)n_tri + 1, 0);
    box[axis].assign((size_t)(3 * n_tri), 0);
    for (int64_t t = 0; t < n_tri; ++t) {
      int64_t umin = INT64_MAX, umax = 0, vmin = INT64_MAX, vmax = 0;
      for (int k = 0; k < 3; ++k) {
        const int32_t *P = q.data() + 3 * (int64_t)tri[3 * t + k];
        umin = std::min<int64_t>(umin, P[au]);
        umax = std::max<int64_t>(umax, P[au]);
        vmin = std::min<int64_t>(vmin, P[av]);
        vmax = std::max<int64_t>(vmax, P[av]);
      }
      const int64_t u0 = (umin + kVoxSub - 1) / kVoxSub + 1;
      const int64_t u1 = std::min<int64_t>(umax / kVoxSub + 1, dims[au] - 1);
      const int64_t v0 = (vmin + kVoxSub - 1) / kVoxSub + 1;
      const int64_t v1 = std::min<int64_t>(vmax / kVoxSub + 1, dims[av] - 1);
      const int64_t cu = std::max<int64_t>(0, u1 - u0 + 1), cv = std::max<int64_t>(0, v1 - v0 + 1);
      box[axis][(size_t)(3 * t)] = (int32_t)u0;
      box[axis][(size_t)(3 * t + 1)] = (int32_t)v0;
      box[axis][(size_t)(3 * t + 2)] = (int32_t)std::max<int64_t>(cu, 1);
      offs[axis][(size_t)t + 1] = offs[axis][(size_t)t] + cu * cv;
    }
  }
  hipError_t e = upload(&d_q, q.data(), q.size(), s);
  if (e == hipSuccess) e = upload(&d_tri, tri, (size_t)(3 * n_tri), s);
  if (e == hipSuccess && tmat) e = upload(&d_tmat, tmat, (size_t)n_tri, s);
  if (e == hipSuccess) e = hipMalloc((void **)&d_offs, sizeof(int64_t) * 3 * ((size_t)n_tri + 1));
  if (e == hipSuccess) e = hipMalloc((void **)&d_box, sizeof(int32_t) * 9 * std::max<size_t>((size_t)n_tri, 1));
  if (e == hipSuccess) e = hipMalloc((void **)&d_bits, sizeof(uint32_t) * (size_t)(Y * Z * words));
  if (e == hipSuccess) e = hipMalloc((void **)&d_link, sizeof(uint32_t) * 3 * (size_t)N);
  for (int axis = 0; axis < 3 && e == hipSuccess; ++axis) {
    e = hipMemcpyAsync(d_offs + axis * (n_tri + 1), offs[axis].data(), sizeof(int64_t) * (n_tri + 1),
                       hipMemcpyHostToDevice, s);
    if (e == hipSuccess && n_tri)
      e = hipMemcpyAsync(d_box + axis * 3 * n_tri, box[axis].data(), sizeof(int32_t) * 3 * n_tri,
                         hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = hipMemsetAsync(d_bits, 0, sizeof(uint32_t) * (size_t)(Y * Z * words), s);
  if (e == hipSuccess) e = hipMemsetAsync(d_link, 0xFF, sizeof(uint32_t) * 3 * (size_t)N, s);
  for (int axis = 0; axis < 3 && e == hipSuccess; ++axis) {
    VoxRaster r;
    r.q = d_q;
    r.tri = d_tri;
    r.offs = d_offs + axis * (n_tri + 1);
    r.box = d_box + axis * 3 * n_tri;
    r.n_tri = n_tri;
    r.total = offs[axis][(size_t)n_tri];
    r.axis = axis;
    r.dim[0] = X;
    r.dim[1] = Y;
    r.dim[2] = Z;
    r.words = words;
    r.bits = axis == 0 ? d_bits : nullptr;
    r.link = d_link + axis * N;
    if (r.total > 0) {
      vox_raster_kernel<<<grid_for(r.total), 256, 0, s>>>(r);
      e = hipGetLastError();
    }
  }
  if (e == hipSuccess) {
    vox_prefix_kernel<<<grid_for(Y * Z), 256, 0, s>>>(d_bits, Y * Z, words);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const dim3 grid((unsigned)((X + 255) / 256), (unsigned)Y, (unsigned)Z);
    vox_label_kernel<<<grid, 256, 0, s>>>(d_bits, d_link, d_tmat, X, Y, Z, words, d_labels);
    e = hipGetLastError();
  }
  const hipError_t es = hipStreamSynchronize(s);
  if (e == hipSuccess) e = es;
  for (void *p : {(void *)d_q, (void *)d_tri, (void *)d_tmat, (void *)d_offs, (void *)d_box,
                  (void *)d_bits, (void *)d_link})
    if (p) (void)hipFree(p);
  return e;
}

}  // namespace pfdtd
