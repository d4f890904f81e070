// ORACLE helper -- test infrastructure only, never part of the product.
//
// Thin C-ABI driver around the reference's own CPU Hand kernel
// dslash_kernel_cpu (WilsonKernelsHandCpu.h:385-451), compiled from the
// reference sources where they lie (see oracle/Makefile, target `ref`), with
// the generic backend (-DGEN -DGEN_SIMD_WIDTH=W; arch/generic/Simd_generic.h).
// Nsimd is W/16 complex lanes for double and W/8 for float.  The output goes
// to oracle/_ref/ only; nothing here is copied from the reference.
#include <cassert>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <iostream>

#include "Simd.h"
#include "WilsonKernelsHand.h"

extern "C" {

int gbxref_nsimd_d(void) { return vComplexD::Nsimd(); }
int gbxref_nsimd_f(void) { return vComplexF::Nsimd(); }

// Same argument list as dslash_kernel<vComplexD> (WilsonKernelsHand.h:25-29).
double gbxref_dslash_d(int nrep, double *U, double *out, double *in, uint64_t *nbr, uint64_t nsite,
                       uint64_t Ls, uint8_t *prm)
{
  return dslash_kernel<vComplexD>(nrep, (vComplexD *)U, (vComplexD *)out, (vComplexD *)in, nbr, nsite, Ls, prm);
}

double gbxref_dslash_f(int nrep, float *U, float *out, float *in, uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                       uint8_t *prm)
{
  return dslash_kernel<vComplexF>(nrep, (vComplexF *)U, (vComplexF *)out, (vComplexF *)in, nbr, nsite, Ls, prm);
}
}
