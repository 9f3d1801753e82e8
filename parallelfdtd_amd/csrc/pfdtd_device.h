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
    // v_fma_mixlo_f16, which rounds the exact product once, straight to binary16.
    asm volatile("" : "+v"(v));
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
    uint32_t x = __float_as_uint(v);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x40u);  // quiet NaN
    x += 0x7FFFu + ((x >> 16) & 1u);  // round to nearest even
    return (uint16_t)(x >> 16);
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

// One SRL_FORWARD node update (kernels3d.cu:508-528) from the neighbour sum s (already summed
// in the order z+1, z-1, y+1, y-1, x+1, x-1), the node's value pc and past value past.
template <typename C>
__device__ __forceinline__ C srl_node(unsigned pb, unsigned mb, C pc, C s, C past,
                                      const typename CoefOf<C>::T *__restrict__ coef,
                                      const typename CoefOf<C>::T &air, int air_uniform, C lam2) {
#pragma clang fp contract(off)
  typename CoefOf<C>::T e = air;
  if (pb != 0x86u || !air_uniform) e = coef[min(pb & 0x7Fu, 7u) * 256u + mb];
  const C a = (C)(pb >> 7) * e.inv;
  return a * ((e.tk * pc + lam2 * s) - e.omb * past);
}

}  // namespace pfdtd
