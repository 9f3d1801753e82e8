// pfdtd_device.h -- device-side helpers shared by the kernel translation units: field storage
// formats (load / store / round to the storage precision) and coefficient types.
#pragma once

#include "pfdtd_internal.h"

namespace pfdtd {

// Fmt<S>: compute type C of storage type S, loads, stores, and round(v) = the value a store
// followed by a load would give (identity for FP32/FP64).
template <typename S>
struct Fmt;

template <>
struct Fmt<float> {
  typedef float C;
  static __device__ __forceinline__ float val(float s) { return s; }
  static __device__ __forceinline__ float raw(float v) { return v; }
  static __device__ __forceinline__ float ld(const float *p, int64_t i) { return p[i]; }
  static __device__ __forceinline__ void st(float *p, int64_t i, float v) { p[i] = v; }
  static __device__ __forceinline__ float round(float v) { return v; }
};

template <>
struct Fmt<double> {
  typedef double C;
  static __device__ __forceinline__ double val(double s) { return s; }
  static __device__ __forceinline__ double raw(double v) { return v; }
  static __device__ __forceinline__ double ld(const double *p, int64_t i) { return p[i]; }
  static __device__ __forceinline__ void st(double *p, int64_t i, double v) { p[i] = v; }
  static __device__ __forceinline__ double round(double v) { return v; }
};

template <>
struct Fmt<h16> {
  typedef float C;
  static __device__ __forceinline__ float ld(const h16 *p, int64_t i) {
    return (float)__builtin_bit_cast(_Float16, p[i].bits);
  }
  static __device__ __forceinline__ uint16_t bits(float v) {
    // The value must be rounded to FP32 first and then to binary16 (round to nearest even):
    // without the register barrier the compiler folds a producing multiply into
    // v_fma_mixlo_f16, which rounds the exact product once, straight to binary16. The barrier
    // is not volatile, so the compiler may still evaluate it speculatively (branch-free selects).
    asm("" : "+v"(v));
    return __builtin_bit_cast(uint16_t, (_Float16)v);
  }
  static __device__ __forceinline__ void st(h16 *p, int64_t i, float v) { p[i].bits = bits(v); }
  static __device__ __forceinline__ float val(h16 s) {
    return (float)__builtin_bit_cast(_Float16, s.bits);
  }
  static __device__ __forceinline__ h16 raw(float v) { return h16{bits(v)}; }
  static __device__ __forceinline__ float round(float v) {
    return (float)__builtin_bit_cast(_Float16, bits(v));
  }
};

template <>
struct Fmt<b16> {
  typedef float C;
  static __device__ __forceinline__ float ld(const b16 *p, int64_t i) {
    return __uint_as_float((uint32_t)p[i].bits << 16);
  }
  static __device__ __forceinline__ uint16_t bits(float v) {
    const uint32_t x = __float_as_uint(v);
    const uint32_t rne = (x + 0x7FFFu + ((x >> 16) & 1u)) >> 16;  // round to nearest even
    const uint32_t qnan = (x >> 16) | 0x40u;                       // NaN stays a (quiet) NaN
    // a select, not a branch: a lane-divergent branch here made every rounding a masked region
    return (uint16_t)((x & 0x7FFFFFFFu) > 0x7F800000u ? qnan : rne);
  }
  static __device__ __forceinline__ void st(b16 *p, int64_t i, float v) { p[i].bits = bits(v); }
  static __device__ __forceinline__ float val(b16 s) { return __uint_as_float((uint32_t)s.bits << 16); }
  static __device__ __forceinline__ b16 raw(float v) { return b16{bits(v)}; }
  static __device__ __forceinline__ float round(float v) {
    return __uint_as_float((uint32_t)bits(v) << 16);
  }
};

// Global-memory access through a wave-uniform base and a 32-bit byte offset: one scalar base
// plus one vector offset per access, no 64-bit vector address arithmetic.
#define PFDTD_GLOBAL __attribute__((address_space(1)))
template <typename T>
struct Scalar {
  typedef T type;
};
template <>
struct Scalar<h16> {
  typedef uint16_t type;
};
template <>
struct Scalar<b16> {
  typedef uint16_t type;
};
template <typename T>
__device__ __forceinline__ T gld(const T *base, uint32_t e) {
  typedef typename Scalar<T>::type U;
  const U u = *(const PFDTD_GLOBAL U *)((const PFDTD_GLOBAL char *)base + e * (uint32_t)sizeof(T));
  return __builtin_bit_cast(T, u);
}
template <typename T>
__device__ __forceinline__ void gst(T *base, uint32_t e, T v) {
  typedef typename Scalar<T>::type U;
  *(PFDTD_GLOBAL U *)((PFDTD_GLOBAL char *)base + e * (uint32_t)sizeof(T)) =
      __builtin_bit_cast(U, v);
}

template <typename C>
struct CoefOf;
template <>
struct CoefOf<float> {
  typedef CoefF T;
};
template <>
struct CoefOf<double> {
  typedef CoefD T;
};

// Copies the coefficient rows a domain can address -- ncol material columns of each of the
// first kStageRows rows of the [row][256] device table -- into the block's LDS (row stride
// ncol), then waits for the whole block. Gathers from LDS neither go to L2 nor share the
// vector-memory counter with the field loads. Every thread of the block must call it.
template <typename CE>
__device__ __forceinline__ void stage_coef(CE *__restrict__ s, const CE *__restrict__ g,
                                           unsigned ncol) {
  for (unsigned i = threadIdx.x; i < kStageRows * ncol; i += blockDim.x)
    s[i] = g[(i / ncol) * 256u + i % ncol];
  __syncthreads();
}

// One SRL_FORWARD node update (kernels3d.cu:508-528) from the neighbour sum s (already summed
// in the order z+1, z-1, y+1, y-1, x+1, x-1), the node's value pc and past value past.
// coef is the block's LDS copy of the coefficient table (stage_coef), cs material columns
// per row.
template <typename C>
__device__ __forceinline__ C srl_node(unsigned pb, unsigned mb, C pc, C s, C past,
                                      const typename CoefOf<C>::T *__restrict__ coef, unsigned cs,
                                      const typename CoefOf<C>::T &air, int air_uniform, C lam2) {
#pragma clang fp contract(off)
  typename CoefOf<C>::T e = air;
  if (pb != 0x86u || !air_uniform) e = coef[min(pb & 0x7Fu, 6u) * cs + mb];
  const C a = (C)(pb >> 7) * e.inv;
  return a * ((e.tk * pc + lam2 * s) - e.omb * past);
}

// ---------------------------------------------------------------------------------------
// SRL_FORWARD update of one tile: row-segment tiles marching short z-chunks
// ---------------------------------------------------------------------------------------
// A tile is 256 threads = four wavefronts side by side along x, so one tile row is a
// contiguous 1 KB (FP32) segment and every wavefront access is a full 256 B run. Each thread
// owns RY consecutive y rows of its x column and marches them through `zc` slices, keeping
// P[z-1] and P[z] of every row in registers (the z-neighbour queue); per slice only P[z+1],
// P_past and the two byte arrays are new, the y-neighbours of inner rows come from registers
// and the halo rows y0-1, y0+RY and the x-neighbours are cache hits of neighbouring threads'
// loads. Short z-chunks keep the set of concurrently streamed slices small so the y-halo rows
// and the slices above and below are still in L2 / Infinity Cache when they are re-read.
// Air voxels (position byte 0x86) take their coefficients from kernel arguments; every other
// voxel gathers its (K, material) entry from the table. Tile (bx, by, bz) covers x in
// [TW bx, TW bx + TW), y in [R by, R by + R) with R = RY * 256 / TW, z in
// [1 + zc bz, 1 + zc bz + zc). TW = 256 (one 1 KB FP32 row segment per tile row) suits wide
// rooms; TW = 64 (four wavefronts stacked in y) wastes fewer lanes on rooms whose width is
// not a multiple of 256.
constexpr int kTX = 256;

// Centred-boundary node update, fdtd3dStdKowalczykMaterials (kernels3d.cu:609-665): the
// position byte holds DIR_X/Y/Z (bits 0-2: the axis has one missing neighbour) and SIGN_X/Y/Z
// (bits 4-6: which present neighbour stands in for it: +x, +y, and -z for SIGN_Z). The
// coefficient entry of (number of DIR bits D, material) holds {1/(1+beta), beta-1, 2-6 lam2}
// with beta = table * lambda * D, evaluated on the host in the reference's order.
template <typename C>
__device__ __forceinline__ C centred_node(unsigned pb, unsigned mb, C pc, C zp, C zm, C yp, C ym,
                                          C xp, C xm, C past,
                                          const typename CoefOf<C>::T *__restrict__ coef,
                                          unsigned cs, const typename CoefOf<C>::T &air, C lam2) {
#pragma clang fp contract(off)
  const unsigned dxb = pb & 1u, dyb = (pb >> 1) & 1u, dzb = (pb >> 2) & 1u;
  typename CoefOf<C>::T e = air;
  if (pb != 0x80u) e = coef[(dxb + dyb + dzb) * cs + mb];
  const C sx = (pb & 0x10u) ? xp : xm;  // p_x[sign_x]
  const C sy = (pb & 0x20u) ? yp : ym;  // p_y[sign_y]
  const C sz = (pb & 0x40u) ? zm : zp;  // p_z[sign_z]: p_z[1] is z-1
  C sb = sx * (C)dxb;
  sb = sb + sy * (C)dyb;
  sb = sb + sz * (C)dzb;
  // (p_x[0]+p_x[1]+p_y[0]+p_y[1]+p_z[0]+p_z[1]+S_boundary)*lambda^2, left to right
  C s = xm + xp;
  s = s + ym;
  s = s + yp;
  s = s + zp;
  s = s + zm;
  s = s + sb;
  s = s * lam2;
  C t = s - e.tk * pc;
  t = t + e.omb * past;
  t = t * (C)(pb >> 7);
  return t * e.inv;
}

// centred_node for an air voxel (position byte 0x80: inside, no DIR or SIGN bit), whatever its
// material byte: the entry is `air`, the boundary sum takes p_x[0], p_y[0], p_z[0] times 0, and
// the final factor (pb >> 7) is 1, a multiplication that changes no value. The same operations
// in the same order otherwise, so the same bits (signed zeros included); only the byte decode
// and the table gather go.
template <typename C>
__device__ __forceinline__ C centred_air(C pc, C zp, C zm, C yp, C ym, C xp, C xm, C past,
                                         const typename CoefOf<C>::T &air, C lam2) {
#pragma clang fp contract(off)
  C sb = xm * (C)0;
  sb = sb + ym * (C)0;
  sb = sb + zp * (C)0;
  C s = xm + xp;
  s = s + ym;
  s = s + yp;
  s = s + zp;
  s = s + zm;
  s = s + sb;
  s = s * lam2;
  C t = s - air.tk * pc;
  t = t + air.omb * past;
  return t * air.inv;
}

// SCHEME 0: forward (Bilbao) boundaries, update types SRL_FORWARD and SHARED; SCHEME 2: the
// centred (Kowalczyk) boundaries of update type SRL.
template <typename S, int RY, int TW = kTX, int SCHEME = 0>
__device__ __forceinline__ void srl_tile(
    const uint8_t *__restrict__ pos, const uint8_t *__restrict__ mat, const S *__restrict__ P,
    S *__restrict__ Pp, const typename CoefOf<typename Fmt<S>::C>::T *__restrict__ coef,
    unsigned cs, typename CoefOf<typename Fmt<S>::C>::T air, int air_uniform,
    typename Fmt<S>::C lam2, int X, int Y, int Z, int zc, int bx, int by, int bz) {
#pragma clang fp contract(off)
  typedef typename Fmt<S>::C C;
  // TW threads along x; 256 / TW row bands of RY rows each stacked in y
  const int x = bx * TW + (int)(threadIdx.x % TW);
  const int ybase = (by * (kTX / TW) + (int)(threadIdx.x / TW)) * RY;
  if (x >= X) return;
  const int z0 = 1 + bz * zc;
  const int z1 = min(z0 + zc, Z - 1);
  const int64_t sX = X;
  const int64_t XY = (int64_t)X * (int64_t)Y;
  const bool xin = (x > 0) && (x < X - 1);

  // Row r is live when ybase + r < Y; it is updated when additionally 0 < y < Y-1 and x is
  // interior. Rows past the grid edge are clamped to Y-1 for loads and never stored.
  bool live[RY], upd[RY];
  int64_t row[RY];
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
  int64_t base = (int64_t)z0 * XY;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    pm[r] = Fmt<S>::ld(P, base - XY + row[r]);
    pc[r] = Fmt<S>::ld(P, base + row[r]);
  }
  for (int z = z0; z < z1; ++z, base += XY) {
    C pn[RY], past[RY], xp[RY], xm[RY];
    unsigned pb[RY], mb[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int64_t i = base + row[r];
      pn[r] = Fmt<S>::ld(P, i + XY);
      past[r] = Fmt<S>::ld(Pp, i);
      pb[r] = pos[i];
      mb[r] = mat[i];
      xp[r] = Fmt<S>::ld(P, i + 1);
      xm[r] = Fmt<S>::ld(P, i - 1);
    }
    const C ylo = Fmt<S>::ld(P, base + halo_lo);
    const C yhi = Fmt<S>::ld(P, base + halo_hi);
    // centred scheme: a wave whose rows are all air (position byte 0x80) skips the decode and
    // the table gather (centred_air is the same arithmetic)
    bool wave_air = false;
    if constexpr (SCHEME == 2) {
      bool lane_air = true;
#pragma unroll
      for (int r = 0; r < RY; ++r) lane_air = lane_air && pb[r] == 0x80u;
      wave_air = __all(lane_air);
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const C yp = (r + 1 < RY) ? pc[r + 1] : yhi;
      const C ym = (r > 0) ? pc[r - 1] : ylo;
      C ret;
      if (SCHEME == 2) {
        if (wave_air)
          ret = centred_air<C>(pc[r], pn[r], pm[r], yp, ym, xp[r], xm[r], past[r], air, lam2);
        else
          ret = centred_node<C>(pb[r], mb[r], pc[r], pn[r], pm[r], yp, ym, xp[r], xm[r], past[r],
                                coef, cs, air, lam2);
      } else {
        // Sum order z+1, z-1, y+1, y-1, x+1, x-1 (kernels3d.cu:516-525).
        C s = pn[r] + pm[r];
        s = s + yp;
        s = s + ym;
        s = s + xp[r];
        s = s + xm[r];
        typename CoefOf<C>::T e = air;
        if (pb[r] != 0x86u || !air_uniform) e = coef[min(pb[r] & 0x7Fu, 6u) * cs + mb[r]];
        const C a = (C)(pb[r] >> 7) * e.inv;
        ret = a * ((e.tk * pc[r] + lam2 * s) - e.omb * past[r]);
      }
      if (live[r]) Fmt<S>::st(Pp, base + row[r], upd[r] ? ret : (C)0);
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pm[r] = pc[r];
      pc[r] = pn[r];
    }
  }
}

// Receiver gather and source injection for one field, executed by one block (all threads
// call it). Gather reads the field just produced (response[step] = P^{step+1},
// kernels3d.cu:164-173); injection then applies the next step's samples to the same field, in
// source order, which is where the reference's source loop of the next step writes them
// (kernels3d.cu:93-104).
template <typename S>
__device__ __forceinline__ void inject_gather_block(S *field, const InjectArgs &a) {
#pragma clang fp contract(off)
  typedef typename Fmt<S>::C C;
  if (a.gather_step >= 0) {
    const double inv_scale = ldexp(1.0, -a.scale_exp);
    for (int r = threadIdx.x; r < a.n_rec; r += blockDim.x) {
      const int64_t idx = a.rec_idx[r];
      if (idx >= 0)
        a.resp[(int64_t)r * a.resp_steps + a.gather_step] =
            (double)Fmt<S>::ld(field, idx) * inv_scale;
    }
  }
  __syncthreads();
  if (a.inject_step < 0) return;
  const C scale = (C)ldexp(1.0, a.scale_exp);
  auto apply = [&](int s) {
    const int64_t idx = a.src_idx[s];
    if (idx < 0) return;
    C smp = (C)0;
    if (a.inject_step < a.n_samples) {
      const int64_t off = (int64_t)s * a.n_samples + a.inject_step;
      if (sizeof(C) == sizeof(double))
        smp = (C)a.sig_d[off];
      else
        smp = (C)a.sig_f[off];
    }
    smp = smp * scale;
    const bool hard = a.src_kind[s] == 0 || a.soft_overwrite;
    if (a.dfield) {
      // the difference form: the sample changes P and keeps the past value P - D
      S *dfield = (S *)a.dfield;
      const C p = Fmt<S>::ld(field, idx);
      const C past = p - Fmt<S>::ld(dfield, idx);
      const C pn = Fmt<S>::round(hard ? smp : p + smp);
      Fmt<S>::st(field, idx, pn);
      Fmt<S>::st(dfield, idx, pn - past);
    } else if (hard) {
      Fmt<S>::st(field, idx, smp);
    } else {
      Fmt<S>::st(field, idx, Fmt<S>::ld(field, idx) + smp);
    }
  };
  if (a.distinct) {
    // distinct elements: the order of application cannot change any value
    for (int s = threadIdx.x; s < a.n_src; s += blockDim.x) apply(s);
  } else if (threadIdx.x == 0) {
    for (int s = 0; s < a.n_src; ++s) apply(s);  // index order: the last hard source wins
  }
}

}  // namespace pfdtd
