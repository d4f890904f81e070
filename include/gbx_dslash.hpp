// gbx_dslash.hpp -- C++ host-side mirror of the reference entry point
//
//   template<class Simd>
//   double dslash_kernel(int nrep, Simd *Up, Simd *outp, Simd *inp,
//                        uint64_t *nbr, uint64_t nsite, uint64_t Ls, uint8_t *prm);
//                                                  (WilsonKernelsHand.h:17-29)
//
// Same name, arguments and meaning, so host code written against the
// reference (bench.cc:148-164) compiles unchanged when it includes this
// header instead of WilsonKernelsHand.h.  Simd is any type with a static
// Nsimd() and a scalar_type whose size says double or float complex -- the
// reference's vComplexD / vComplexF qualify (SimdVector.h:86-91), and so
// does gbx::vcomplex<T, N> below for code without the reference's headers.
// The arrays are the reference's host layouts; the work runs on the GPU
// through the C-ABI (gbx.h).  Returns microseconds of applications
// 1..nrep-1 as the reference does; on failure returns a negative value and
// leaves the library's message in gbx_last_error().  dslash_kernel_dag is
// the dagger (the reference has none).
//
// The reference's compile-time layout switches are honoured as it defines
// them: -DRRII selects the RRII lane layout (bench.cc:105-106), and
// DATA_SIMD < Nsimd() means vectors expanded from DATA_SIMD data lanes
// (bench.cc:1-3, 98-138) whose tables and permutes stay those of DATA_SIMD.
#pragma once

#include <cstdint>
#include <cstdio>
#include <type_traits>

#include "gbx.h"

namespace gbx {

// Minimal stand-in for the reference's Simd<complex<T>, V> storage type.
template <class T, int N> struct vcomplex {
  T v[2 * N];
  struct scalar_type {
    T re, im;
  };
  static constexpr int Nsimd() { return N; }
};

template <class Simd>
inline double dslash_kernel_impl(int nrep, int dag, Simd *Up, Simd *outp, Simd *inp, uint64_t *nbr, uint64_t nsite,
                                 uint64_t Ls, uint8_t *prm)
{
  typedef typename Simd::scalar_type S;
  static_assert(sizeof(S) == 16 || sizeof(S) == 8, "scalar_type must be a double or float complex");
  const int nsimd = Simd::Nsimd();
#ifdef DATA_SIMD
  const int data_nsimd = (DATA_SIMD < nsimd) ? (int)(DATA_SIMD) : nsimd;
#else
  const int data_nsimd = nsimd;
#endif
#ifdef RRII
  const int rrii = 1;
#else
  const int rrii = 0;
#endif
  double usec = 0.0;
  int rc;
  if (sizeof(S) == 16)
    rc = gbx_dslash_kernel_expand_f64(nrep, nsimd, data_nsimd, rrii, dag, (const double *)Up, (double *)outp,
                                      (const double *)inp, nbr, nsite, Ls, prm, &usec);
  else
    rc = gbx_dslash_kernel_expand_f32(nrep, nsimd, data_nsimd, rrii, dag, (const float *)Up, (float *)outp,
                                      (const float *)inp, nbr, nsite, Ls, prm, &usec);
  if (rc != GBX_OK) {
    std::fprintf(stderr, "dslash_kernel: %s: %s\n", gbx_strerror(rc), gbx_last_error());
    return -(double)rc;
  }
  return usec;
}

}  // namespace gbx

template <class Simd>
inline double dslash_kernel(int nrep, Simd *Up, Simd *outp, Simd *inp, uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                            uint8_t *prm)
{
  return gbx::dslash_kernel_impl(nrep, 0, Up, outp, inp, nbr, nsite, Ls, prm);
}

template <class Simd>
inline double dslash_kernel_dag(int nrep, Simd *Up, Simd *outp, Simd *inp, uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                                uint8_t *prm)
{
  return gbx::dslash_kernel_impl(nrep, 1, Up, outp, inp, nbr, nsite, Ls, prm);
}
