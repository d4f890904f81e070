// Device-side storage traits and 16-byte chunk access for the SoA layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gbx.h"

namespace dev {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

template <int P> struct Traits;
template <> struct Traits<GBX_F64> {
  typedef double acc;
  typedef double st;
  typedef d2 vec;
  static constexpr int R = 2, NCH = 12, BYTES = 8;
};
template <> struct Traits<GBX_F32> {
  typedef float acc;
  typedef float st;
  typedef f4 vec;
  static constexpr int R = 4, NCH = 6, BYTES = 4;
};
template <> struct Traits<GBX_F16> {
  typedef float acc;
  typedef _Float16 st;
  typedef h8 vec;
  static constexpr int R = 8, NCH = 3, BYTES = 2;
};
template <> struct Traits<GBX_BF16> {
  typedef float acc;
  typedef __bf16 st;
  typedef b8 vec;
  static constexpr int R = 8, NCH = 3, BYTES = 2;
};

// Chunk t (16 bytes) of a buffer: R reals of storage precision P.
template <int P, typename A>
__device__ __forceinline__ void load_chunk(const void *base, uint64_t t, A *v)
{
  typedef typename Traits<P>::vec V;
  const V x = ((const V *)base)[t];
#pragma unroll
  for (int j = 0; j < Traits<P>::R; j++) v[j] = (A)x[j];
}

template <int P>
__device__ __forceinline__ void load_chunk_f64(const void *base, uint64_t t, double *v)
{
  load_chunk<P, double>(base, t, v);
}

template <int P, typename A>
__device__ __forceinline__ void store_chunk(void *base, uint64_t t, const A *v, bool nt)
{
  typedef typename Traits<P>::vec V;
  typedef typename Traits<P>::st S;
  V x;
#pragma unroll
  for (int j = 0; j < Traits<P>::R; j++) x[j] = (S)v[j];
  V *p = ((V *)base) + t;
  if (nt) __builtin_nontemporal_store(x, p);
  else *p = x;
}

template <int P>
__device__ __forceinline__ void store_real(void *base, uint64_t i, double v)
{
  ((typename Traits<P>::st *)base)[i] = (typename Traits<P>::st)v;
}

template <int P>
__device__ __forceinline__ double load_real(const void *base, uint64_t i)
{
  return (double)((const typename Traits<P>::st *)base)[i];
}

}  // namespace dev
