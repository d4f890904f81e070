// Device building blocks of the hopping term: spin projection, SU(3) x
// half-spinor, reconstruction.  Operator and leg order as the reference Hand
// kernel (WilsonKernelsHandCpu.h:426-446): leg d reads the neighbour in
// direction d (Xp Yp Zp Tp Xm Ym Zm Tm) and link U[x][d]; non-dagger uses the
// projector *M_PROJ on forward legs and *P_PROJ on backward legs
// (WilsonKernelsHandCpu.h:94-159, 165-303); the dagger swaps them.
#pragma once

#include "layout_dev.h"

namespace dev {

enum { PXP = 0, PYP, PZP, PTP, PXM, PYM, PZM, PTM };

__device__ __forceinline__ float dfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double dfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Projector kind used on leg `dir`.
template <bool DAG, int DIR> struct LegProj {
  static constexpr int value = DAG ? DIR : (DIR ^ 4);
};

// v[2*(sp*3+c) + {0,1}] = re/im of spin sp colour c of the neighbour spinor.
// h[0..2] = upper half-spinor row 0, h[3..5] = row 1 (re in hr, im in hi).
template <int K, typename A>
__device__ __forceinline__ void project(const A *v, A *hr, A *hi)
{
#define PR(sp, c) v[2 * ((sp) * 3 + (c))]
#define PI(sp, c) v[2 * ((sp) * 3 + (c)) + 1]
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (K == PXP) {  // h0 = p0 + i p3, h1 = p1 + i p2
      hr[c] = PR(0, c) - PI(3, c); hi[c] = PI(0, c) + PR(3, c);
      hr[3 + c] = PR(1, c) - PI(2, c); hi[3 + c] = PI(1, c) + PR(2, c);
    } else if (K == PYP) {  // h0 = p0 - p3, h1 = p1 + p2
      hr[c] = PR(0, c) - PR(3, c); hi[c] = PI(0, c) - PI(3, c);
      hr[3 + c] = PR(1, c) + PR(2, c); hi[3 + c] = PI(1, c) + PI(2, c);
    } else if (K == PZP) {  // h0 = p0 + i p2, h1 = p1 - i p3
      hr[c] = PR(0, c) - PI(2, c); hi[c] = PI(0, c) + PR(2, c);
      hr[3 + c] = PR(1, c) + PI(3, c); hi[3 + c] = PI(1, c) - PR(3, c);
    } else if (K == PTP) {  // h0 = p0 + p2, h1 = p1 + p3
      hr[c] = PR(0, c) + PR(2, c); hi[c] = PI(0, c) + PI(2, c);
      hr[3 + c] = PR(1, c) + PR(3, c); hi[3 + c] = PI(1, c) + PI(3, c);
    } else if (K == PXM) {  // h0 = p0 - i p3, h1 = p1 - i p2
      hr[c] = PR(0, c) + PI(3, c); hi[c] = PI(0, c) - PR(3, c);
      hr[3 + c] = PR(1, c) + PI(2, c); hi[3 + c] = PI(1, c) - PR(2, c);
    } else if (K == PYM) {  // h0 = p0 + p3, h1 = p1 - p2
      hr[c] = PR(0, c) + PR(3, c); hi[c] = PI(0, c) + PI(3, c);
      hr[3 + c] = PR(1, c) - PR(2, c); hi[3 + c] = PI(1, c) - PI(2, c);
    } else if (K == PZM) {  // h0 = p0 - i p2, h1 = p1 + i p3
      hr[c] = PR(0, c) + PI(2, c); hi[c] = PI(0, c) - PR(2, c);
      hr[3 + c] = PR(1, c) - PI(3, c); hi[3 + c] = PI(1, c) + PR(3, c);
    } else {  // PTM: h0 = p0 - p2, h1 = p1 - p3
      hr[c] = PR(0, c) - PR(2, c); hi[c] = PI(0, c) - PI(2, c);
      hr[3 + c] = PR(1, c) - PR(3, c); hi[3 + c] = PI(1, c) - PI(3, c);
    }
  }
#undef PR
#undef PI
}

// The reference's arithmetic for U x chi: every complex product
// (ar*br - ai*bi, ar*bi + ai*br) rounded term by term, summed over the
// columns left to right, t = U_i0 chi_0; t += U_i1 chi_1; t += U_i2 chi_2
// (WilsonKernelsHandCpu.h:62-90, arch/generic/Simd_generic.h:153-155).
// No multiply-add is fused, so the result is the reference's to the bit.
template <typename A>
__device__ __forceinline__ void su3_mul_exact(const A *U, const A *hr, const A *hi, A *ur, A *ui)
{
#pragma clang fp contract(off)
#pragma unroll
  for (int sp = 0; sp < 2; sp++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      A re = U[(i * 3) * 2] * hr[sp * 3] - U[(i * 3) * 2 + 1] * hi[sp * 3];
      A im = U[(i * 3) * 2] * hi[sp * 3] + U[(i * 3) * 2 + 1] * hr[sp * 3];
#pragma unroll
      for (int j = 1; j < 3; j++) {
        const A pr = U[(i * 3 + j) * 2] * hr[sp * 3 + j] - U[(i * 3 + j) * 2 + 1] * hi[sp * 3 + j];
        const A pi = U[(i * 3 + j) * 2] * hi[sp * 3 + j] + U[(i * 3 + j) * 2 + 1] * hr[sp * 3 + j];
        re = re + pr;
        im = im + pi;
      }
      ur[sp * 3 + i] = re;
      ui[sp * 3 + i] = im;
    }
}

// u[sp*3+i] = sum_j U[i][j] h[sp*3+j]; U row-major complex, 18 reals.
// MUL_CHAIN (FP64 and FP32 storage): one multiply + five fused multiply-adds
// per real output (72 ops per leg, 840 VALU ops per 5D site instead of 1032
// for the mul/fma/add form the compiler derives from the plain expression).
// MUL_PLAIN (16-bit storage): the plain expression, whose register schedule
// keeps those kernels free of spills.  MUL_EXACT: the reference's rounding.
enum { MUL_PLAIN = 0, MUL_CHAIN = 1, MUL_EXACT = 2 };
template <int MUL, typename A>
__device__ __forceinline__ void su3_mul(const A *U, const A *hr, const A *hi, A *ur, A *ui)
{
  if constexpr (MUL == MUL_EXACT) {
    su3_mul_exact(U, hr, hi, ur, ui);
    return;
  }
  if constexpr (MUL == MUL_PLAIN) {
#pragma unroll
    for (int sp = 0; sp < 2; sp++)
#pragma unroll
      for (int i = 0; i < 3; i++) {
        A re = U[(i * 3) * 2] * hr[sp * 3] - U[(i * 3) * 2 + 1] * hi[sp * 3];
        A im = U[(i * 3) * 2] * hi[sp * 3] + U[(i * 3) * 2 + 1] * hr[sp * 3];
#pragma unroll
        for (int j = 1; j < 3; j++) {
          re += U[(i * 3 + j) * 2] * hr[sp * 3 + j] - U[(i * 3 + j) * 2 + 1] * hi[sp * 3 + j];
          im += U[(i * 3 + j) * 2] * hi[sp * 3 + j] + U[(i * 3 + j) * 2 + 1] * hr[sp * 3 + j];
        }
        ur[sp * 3 + i] = re;
        ui[sp * 3 + i] = im;
      }
    return;
  }
#pragma unroll
  for (int sp = 0; sp < 2; sp++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      A re = U[(i * 3) * 2] * hr[sp * 3];
      A im = U[(i * 3) * 2] * hi[sp * 3];
      re = dfma(-U[(i * 3) * 2 + 1], hi[sp * 3], re);
      im = dfma(U[(i * 3) * 2 + 1], hr[sp * 3], im);
#pragma unroll
      for (int j = 1; j < 3; j++) {
        re = dfma(U[(i * 3 + j) * 2], hr[sp * 3 + j], re);
        im = dfma(U[(i * 3 + j) * 2], hi[sp * 3 + j], im);
        re = dfma(-U[(i * 3 + j) * 2 + 1], hi[sp * 3 + j], re);
        im = dfma(U[(i * 3 + j) * 2 + 1], hr[sp * 3 + j], im);
      }
      ur[sp * 3 + i] = re;
      ui[sp * 3 + i] = im;
    }
}

// Accumulate (FIRST: assign) the reconstructed 4-spinor of kind K.
// u0 = (ur[0..2], ui[0..2]), u1 = (ur[3..5], ui[3..5]); r index sp*3+c.
template <int K, bool FIRST, typename A>
__device__ __forceinline__ void recon(const A *ur, const A *ui, A *rr, A *ri)
{
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const A u0r = ur[c], u0i = ui[c], u1r = ur[3 + c], u1i = ui[3 + c];
    if (FIRST) {
      rr[c] = u0r; ri[c] = u0i;
      rr[3 + c] = u1r; ri[3 + c] = u1i;
      if (K == PXP) {  // r2 = -i u1, r3 = -i u0
        rr[6 + c] = u1i; ri[6 + c] = -u1r;
        rr[9 + c] = u0i; ri[9 + c] = -u0r;
      } else {  // PXM: r2 = i u1, r3 = i u0
        rr[6 + c] = -u1i; ri[6 + c] = u1r;
        rr[9 + c] = -u0i; ri[9 + c] = u0r;
      }
      continue;
    }
    rr[c] += u0r; ri[c] += u0i;
    rr[3 + c] += u1r; ri[3 + c] += u1i;
    if (K == PXP) {  // r2 -= i u1, r3 -= i u0
      rr[6 + c] += u1i; ri[6 + c] -= u1r;
      rr[9 + c] += u0i; ri[9 + c] -= u0r;
    } else if (K == PXM) {  // r2 += i u1, r3 += i u0
      rr[6 + c] -= u1i; ri[6 + c] += u1r;
      rr[9 + c] -= u0i; ri[9 + c] += u0r;
    } else if (K == PYP) {  // r2 += u1, r3 -= u0
      rr[6 + c] += u1r; ri[6 + c] += u1i;
      rr[9 + c] -= u0r; ri[9 + c] -= u0i;
    } else if (K == PYM) {  // r2 -= u1, r3 += u0
      rr[6 + c] -= u1r; ri[6 + c] -= u1i;
      rr[9 + c] += u0r; ri[9 + c] += u0i;
    } else if (K == PZP) {  // r2 -= i u0, r3 += i u1
      rr[6 + c] += u0i; ri[6 + c] -= u0r;
      rr[9 + c] -= u1i; ri[9 + c] += u1r;
    } else if (K == PZM) {  // r2 += i u0, r3 -= i u1
      rr[6 + c] -= u0i; ri[6 + c] += u0r;
      rr[9 + c] += u1i; ri[9 + c] -= u1r;
    } else if (K == PTP) {  // r2 += u0, r3 += u1
      rr[6 + c] += u0r; ri[6 + c] += u0i;
      rr[9 + c] += u1r; ri[9 + c] += u1i;
    } else {  // PTM: r2 -= u0, r3 -= u1
      rr[6 + c] -= u0r; ri[6 + c] -= u0i;
      rr[9 + c] -= u1r; ri[9 + c] -= u1i;
    }
  }
}

// Third row of an SU(3) matrix from the first two: row2 = conj(row0 x row1).
template <typename A>
__device__ __forceinline__ void su3_row2(A *U)
{
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int a = (j + 1) % 3, b = (j + 2) % 3;
    const A ar = U[a * 2], ai = U[a * 2 + 1], br = U[6 + b * 2], bi = U[6 + b * 2 + 1];
    const A cr = U[b * 2], ci = U[b * 2 + 1], dr = U[6 + a * 2], di = U[6 + a * 2 + 1];
    const A xr = (ar * br - ai * bi) - (cr * dr - ci * di);
    const A xi = (ar * bi + ai * br) - (cr * di + ci * dr);
    U[12 + j * 2] = xr;
    U[12 + j * 2 + 1] = -xi;
  }
}

// Stage the links (and neighbour slots) of workgroup sites q0 .. q0+nsl-1
// into LDS as accumulation-precision 3x3 matrices [site][dir][18].
template <int P, bool C12>
__device__ __forceinline__ void stage_links(const void *__restrict__ gauge, const uint32_t *__restrict__ nbr, uint32_t q0,
                                            uint32_t nsl, typename Traits<P>::acc *sU, uint32_t *sN)
{
  typedef typename Traits<P>::acc A;
  typedef typename Traits<P>::st S;
  constexpr int RL = C12 ? 12 : 18;
  const S *g = (const S *)gauge + (uint64_t)q0 * 8 * RL;
  const uint32_t n = nsl * 8 * RL;
  // Links are read once per launch: non-temporal loads keep them from
  // displacing neighbour spinors, which are re-read, in the L2.
  for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint32_t link = e / RL, k = e - link * RL;
    sU[link * 18 + k] = (A)__builtin_nontemporal_load(g + e);
  }
  for (uint32_t e = threadIdx.x; e < nsl * 8; e += blockDim.x) sN[e] = nbr[(uint64_t)q0 * 8 + e];
  if (C12) {
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < nsl * 8; l += blockDim.x) su3_row2(sU + l * 18);
  }
  __syncthreads();
}

// Gather the 24 reals of neighbour slot `nb`, minor index m.
template <int P>
__device__ __forceinline__ void load_nb(const void *__restrict__ in, uint32_t nb, uint32_t M, uint32_t m,
                                        typename Traits<P>::acc *v)
{
  constexpr int NCH = Traits<P>::NCH, R = Traits<P>::R;
  const uint64_t base = (uint64_t)nb * NCH * M + m;
#pragma unroll
  for (int k = 0; k < NCH; k++) load_chunk<P, typename Traits<P>::acc>(in, base + (uint64_t)k * M, v + k * R);
}

// project, multiply by U (18 reals in LDS), reconstruct into r
template <int K, bool FIRST, int MUL, typename A>
__device__ __forceinline__ void leg_compute(const A *v, const A *U, A *rr, A *ri)
{
  A hr[6], hi[6], ur[6], ui[6];
  project<K>(v, hr, hi);
  su3_mul<MUL>(U, hr, hi, ur, ui);
  recon<K, FIRST>(ur, ui, rr, ri);
}

// Legs of a domain part whose ghosts hold projected halves.  Neighbour slot
// n >= ghost0 is half-slot n - ghost0 of the ghost region (NCH/2 chunks per
// half-slot, after the ghost0 whole slots): its chunks k and k + NCH/2 are
// both read from half-slot chunk k (the same lines, so the second read hits
// the cache), v[0..11] then holds the two-spinor, and the leg uses it in
// place of its projection.  Other neighbours are gathered and projected as
// in dhop_site; the arithmetic is the same.
template <int P>
__device__ __forceinline__ bool load_nb_gp(const void *__restrict__ in, uint32_t nb, uint32_t M, uint32_t m,
                                           uint32_t ghost0, typename Traits<P>::acc *v)
{
  constexpr int NCH = Traits<P>::NCH, HCH = NCH / 2, R = Traits<P>::R;
  const bool g = nb >= ghost0;
  const uint64_t full = (uint64_t)nb * NCH * M + m;
  const uint64_t half = (uint64_t)ghost0 * NCH * M + (uint64_t)(g ? nb - ghost0 : 0) * HCH * M + m;
#pragma unroll
  for (int k = 0; k < NCH; k++) {
    const uint64_t t = g ? half + (uint64_t)(k % HCH) * M : full + (uint64_t)k * M;
    load_chunk<P, typename Traits<P>::acc>(in, t, v + k * R);
  }
  return g;
}

template <int K, bool FIRST, int MUL, typename A>
__device__ __forceinline__ void leg_compute_gp(const A *v, bool g, const A *U, A *rr, A *ri)
{
  A hr[6], hi[6], ur[6], ui[6];
  project<K>(v, hr, hi);
#pragma unroll
  for (int j = 0; j < 6; j++) {
    hr[j] = g ? v[2 * j] : hr[j];
    hi[j] = g ? v[2 * j + 1] : hi[j];
  }
  su3_mul<MUL>(U, hr, hi, ur, ui);
  recon<K, FIRST>(ur, ui, rr, ri);
}

// GBX_BASIS_TDIAG fields (psi' = S psi, S = [[1, 1], [1, -1]] / sqrt(2) on
// the spin pairs (0, 2), (1, 3)): there gamma_x,y,z are the chiral ones
// negated, so an x, y or z leg uses the projector kind of the opposite sign,
// and gamma_t = diag(1, 1, -1, -1), so the t legs' projectors 1 +/- gamma_t
// keep twice the upper (kind PTP) or lower (PTM) two spins: such a leg
// gathers NCH/2 chunks, multiplies them by the link and adds twice the
// product to that half of the result.
template <bool DAG, int D, bool TD> struct LegKind {
  static constexpr int value = LegProj<DAG, D>::value ^ ((TD && (D & 3) != 3) ? 4 : 0);
};
template <bool DAG, int D, bool TD> struct LegHalf {
  static constexpr bool value = TD && (D & 3) == 3;              // gathers half a spinor
  static constexpr bool upper = LegProj<DAG, D>::value == PTP;  // which half
};

// Chunks of one spin half (v[0..11] = its 2 spins x 3 colours).
template <int P, bool UPPER>
__device__ __forceinline__ void load_nb_half(const void *__restrict__ in, uint32_t nb, uint32_t M, uint32_t m,
                                             typename Traits<P>::acc *v)
{
  constexpr int NCH = Traits<P>::NCH, R = Traits<P>::R, K0 = UPPER ? 0 : NCH / 2;
  const uint64_t base = (uint64_t)nb * NCH * M + m;
#pragma unroll
  for (int k = 0; k < NCH / 2; k++)
    load_chunk<P, typename Traits<P>::acc>(in, base + (uint64_t)(K0 + k) * M, v + k * R);
}

template <bool UPPER, int MUL, typename A>
__device__ __forceinline__ void leg_half(const A *v, const A *U, A *rr, A *ri)
{
  A hr[6], hi[6], ur[6], ui[6];
#pragma unroll
  for (int j = 0; j < 6; j++) {
    hr[j] = v[2 * j];
    hi[j] = v[2 * j + 1];
  }
  su3_mul<MUL>(U, hr, hi, ur, ui);
  constexpr int o = UPPER ? 0 : 6;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    rr[o + j] = dfma((A)2, ur[j], rr[o + j]);
    ri[o + j] = dfma((A)2, ui[j], ri[o + j]);
  }
}

template <int P, bool DAG, int D, bool TD>
__device__ __forceinline__ void leg_load(const void *__restrict__ in, uint32_t nb, uint32_t M, uint32_t m,
                                         typename Traits<P>::acc *v)
{
  if constexpr (LegHalf<DAG, D, TD>::value) load_nb_half<P, LegHalf<DAG, D, TD>::upper>(in, nb, M, m, v);
  else load_nb<P>(in, nb, M, m, v);
}

// The leg arithmetic runs at raised wave priority (s_setprio 1), the gathers
// at the normal one: a wavefront holding a gathered leg finishes it and
// issues its next gathers sooner, which evens out the stream of gathers
// (FP64: about 1 % on TDIAG fields, 1.7 % on chiral fields; the arithmetic
// and its results are unchanged).
template <bool DAG, int D, bool TD, bool FIRST, int MUL, typename A>
__device__ __forceinline__ void leg_apply(const A *v, const A *U, A *rr, A *ri)
{
  static_assert(!(FIRST && LegHalf<DAG, D, TD>::value), "a half leg never comes first");
  __builtin_amdgcn_s_setprio(1);
  if constexpr (LegHalf<DAG, D, TD>::value) leg_half<LegHalf<DAG, D, TD>::upper, MUL>(v, U, rr, ri);
  else leg_compute<LegKind<DAG, D, TD>::value, FIRST, MUL>(v, U, rr, ri);
  __builtin_amdgcn_s_setprio(0);
}

// Full 8-leg hopping term for (site with links U, neighbours nb[8], minor m).
// PF = 0: gather and compute leg by leg.  PF = 1: the gather of leg d+1 is
// issued before leg d is computed (two gather buffers in registers), which
// keeps a second set of loads in flight per wavefront.  PF = 2: three
// buffers, gathers two legs ahead.  EXACT: the reference's rounding in the
// SU(3) products (FP64 / FP32).  TD: GBX_BASIS_TDIAG fields (LegKind, LegHalf).
template <int P, bool DAG, int PF, bool EXACT = false, bool TD = false>
__device__ __forceinline__ void dhop_site(const void *__restrict__ in, const uint32_t *nb, uint32_t M, uint32_t m,
                                          const typename Traits<P>::acc *U, typename Traits<P>::acc *rr,
                                          typename Traits<P>::acc *ri)
{
  typedef typename Traits<P>::acc A;
  constexpr int CH = EXACT ? MUL_EXACT : (P == GBX_F64 || P == GBX_F32) ? MUL_CHAIN : MUL_PLAIN;
  if (PF == 0) {
    A v[24];
#define GBX_LEG(D, FIRST)                                                  \
    leg_load<P, DAG, D, TD>(in, nb[D], M, m, v);                           \
    leg_apply<DAG, D, TD, FIRST, CH>(v, U + D * 18, rr, ri);
    GBX_LEG(0, true)
    GBX_LEG(1, false)
    GBX_LEG(2, false)
    GBX_LEG(3, false)
    GBX_LEG(4, false)
    GBX_LEG(5, false)
    GBX_LEG(6, false)
    GBX_LEG(7, false)
#undef GBX_LEG
  } else if (PF == 2) {
    A v0[24], v1[24], v2[24];
    leg_load<P, DAG, 0, TD>(in, nb[0], M, m, v0);
    leg_load<P, DAG, 1, TD>(in, nb[1], M, m, v1);
    leg_load<P, DAG, 2, TD>(in, nb[2], M, m, v2);
    leg_apply<DAG, 0, TD, true, CH>(v0, U + 0 * 18, rr, ri);
    leg_load<P, DAG, 3, TD>(in, nb[3], M, m, v0);
    leg_apply<DAG, 1, TD, false, CH>(v1, U + 1 * 18, rr, ri);
    leg_load<P, DAG, 4, TD>(in, nb[4], M, m, v1);
    leg_apply<DAG, 2, TD, false, CH>(v2, U + 2 * 18, rr, ri);
    leg_load<P, DAG, 5, TD>(in, nb[5], M, m, v2);
    leg_apply<DAG, 3, TD, false, CH>(v0, U + 3 * 18, rr, ri);
    leg_load<P, DAG, 6, TD>(in, nb[6], M, m, v0);
    leg_apply<DAG, 4, TD, false, CH>(v1, U + 4 * 18, rr, ri);
    leg_load<P, DAG, 7, TD>(in, nb[7], M, m, v1);
    leg_apply<DAG, 5, TD, false, CH>(v2, U + 5 * 18, rr, ri);
    leg_apply<DAG, 6, TD, false, CH>(v0, U + 6 * 18, rr, ri);
    leg_apply<DAG, 7, TD, false, CH>(v1, U + 7 * 18, rr, ri);
  } else {
    A v0[24], v1[24];
    leg_load<P, DAG, 0, TD>(in, nb[0], M, m, v0);
    leg_load<P, DAG, 1, TD>(in, nb[1], M, m, v1);
    leg_apply<DAG, 0, TD, true, CH>(v0, U + 0 * 18, rr, ri);
    leg_load<P, DAG, 2, TD>(in, nb[2], M, m, v0);
    leg_apply<DAG, 1, TD, false, CH>(v1, U + 1 * 18, rr, ri);
    leg_load<P, DAG, 3, TD>(in, nb[3], M, m, v1);
    leg_apply<DAG, 2, TD, false, CH>(v0, U + 2 * 18, rr, ri);
    leg_load<P, DAG, 4, TD>(in, nb[4], M, m, v0);
    leg_apply<DAG, 3, TD, false, CH>(v1, U + 3 * 18, rr, ri);
    leg_load<P, DAG, 5, TD>(in, nb[5], M, m, v1);
    leg_apply<DAG, 4, TD, false, CH>(v0, U + 4 * 18, rr, ri);
    leg_load<P, DAG, 6, TD>(in, nb[6], M, m, v0);
    leg_apply<DAG, 5, TD, false, CH>(v1, U + 5 * 18, rr, ri);
    leg_load<P, DAG, 7, TD>(in, nb[7], M, m, v1);
    leg_apply<DAG, 6, TD, false, CH>(v0, U + 6 * 18, rr, ri);
    leg_apply<DAG, 7, TD, false, CH>(v1, U + 7 * 18, rr, ri);
  }
}

// dhop_site on a domain part whose ghosts hold projected halves (load_nb_gp).
template <int P, bool DAG, int PF>
__device__ __forceinline__ void dhop_site_gp(const void *__restrict__ in, const uint32_t *nb, uint32_t M, uint32_t m,
                                             uint32_t ghost0, const typename Traits<P>::acc *U,
                                             typename Traits<P>::acc *rr, typename Traits<P>::acc *ri)
{
  typedef typename Traits<P>::acc A;
  constexpr int CH = (P == GBX_F64 || P == GBX_F32) ? MUL_CHAIN : MUL_PLAIN;
  if (PF == 0) {
    A v[24];
#define GBX_LEG(D, FIRST)                                                                  \
    {                                                                                      \
      const bool g = load_nb_gp<P>(in, nb[D], M, m, ghost0, v);                            \
      leg_compute_gp<LegProj<DAG, D>::value, FIRST, CH>(v, g, U + D * 18, rr, ri);         \
    }
    GBX_LEG(0, true)
    GBX_LEG(1, false)
    GBX_LEG(2, false)
    GBX_LEG(3, false)
    GBX_LEG(4, false)
    GBX_LEG(5, false)
    GBX_LEG(6, false)
    GBX_LEG(7, false)
#undef GBX_LEG
  } else if (PF == 2) {
    A v0[24], v1[24], v2[24];
    bool g0 = load_nb_gp<P>(in, nb[0], M, m, ghost0, v0);
    bool g1 = load_nb_gp<P>(in, nb[1], M, m, ghost0, v1);
    bool g2 = load_nb_gp<P>(in, nb[2], M, m, ghost0, v2);
    leg_compute_gp<LegProj<DAG, 0>::value, true, CH>(v0, g0, U + 0 * 18, rr, ri);
    g0 = load_nb_gp<P>(in, nb[3], M, m, ghost0, v0);
    leg_compute_gp<LegProj<DAG, 1>::value, false, CH>(v1, g1, U + 1 * 18, rr, ri);
    g1 = load_nb_gp<P>(in, nb[4], M, m, ghost0, v1);
    leg_compute_gp<LegProj<DAG, 2>::value, false, CH>(v2, g2, U + 2 * 18, rr, ri);
    g2 = load_nb_gp<P>(in, nb[5], M, m, ghost0, v2);
    leg_compute_gp<LegProj<DAG, 3>::value, false, CH>(v0, g0, U + 3 * 18, rr, ri);
    g0 = load_nb_gp<P>(in, nb[6], M, m, ghost0, v0);
    leg_compute_gp<LegProj<DAG, 4>::value, false, CH>(v1, g1, U + 4 * 18, rr, ri);
    g1 = load_nb_gp<P>(in, nb[7], M, m, ghost0, v1);
    leg_compute_gp<LegProj<DAG, 5>::value, false, CH>(v2, g2, U + 5 * 18, rr, ri);
    leg_compute_gp<LegProj<DAG, 6>::value, false, CH>(v0, g0, U + 6 * 18, rr, ri);
    leg_compute_gp<LegProj<DAG, 7>::value, false, CH>(v1, g1, U + 7 * 18, rr, ri);
  } else {
    A v0[24], v1[24];
    bool g0 = load_nb_gp<P>(in, nb[0], M, m, ghost0, v0);
    bool g1 = load_nb_gp<P>(in, nb[1], M, m, ghost0, v1);
    leg_compute_gp<LegProj<DAG, 0>::value, true, CH>(v0, g0, U + 0 * 18, rr, ri);
    g0 = load_nb_gp<P>(in, nb[2], M, m, ghost0, v0);
    leg_compute_gp<LegProj<DAG, 1>::value, false, CH>(v1, g1, U + 1 * 18, rr, ri);
    g1 = load_nb_gp<P>(in, nb[3], M, m, ghost0, v1);
    leg_compute_gp<LegProj<DAG, 2>::value, false, CH>(v0, g0, U + 2 * 18, rr, ri);
    g0 = load_nb_gp<P>(in, nb[4], M, m, ghost0, v0);
    leg_compute_gp<LegProj<DAG, 3>::value, false, CH>(v1, g1, U + 3 * 18, rr, ri);
    g1 = load_nb_gp<P>(in, nb[5], M, m, ghost0, v1);
    leg_compute_gp<LegProj<DAG, 4>::value, false, CH>(v0, g0, U + 4 * 18, rr, ri);
    g0 = load_nb_gp<P>(in, nb[6], M, m, ghost0, v0);
    leg_compute_gp<LegProj<DAG, 5>::value, false, CH>(v1, g1, U + 5 * 18, rr, ri);
    g1 = load_nb_gp<P>(in, nb[7], M, m, ghost0, v1);
    leg_compute_gp<LegProj<DAG, 6>::value, false, CH>(v0, g0, U + 6 * 18, rr, ri);
    leg_compute_gp<LegProj<DAG, 7>::value, false, CH>(v1, g1, U + 7 * 18, rr, ri);
  }
}

// Write the 4-spinor result (optionally out = alpha*center + beta*result).
template <int P, bool AXPBY>
__device__ __forceinline__ void store_site(void *__restrict__ out, const void *__restrict__ center, uint32_t q, uint32_t M,
                                           uint32_t m, const typename Traits<P>::acc *rr, const typename Traits<P>::acc *ri,
                                           typename Traits<P>::acc alpha, typename Traits<P>::acc beta, bool nt)
{
  typedef typename Traits<P>::acc A;
  constexpr int NCH = Traits<P>::NCH, R = Traits<P>::R;
  const uint64_t base = (uint64_t)q * NCH * M + m;
#pragma unroll
  for (int k = 0; k < NCH; k++) {
    A v[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int r = k * R + j;
      v[j] = (r & 1) ? ri[r >> 1] : rr[r >> 1];
    }
    if (AXPBY) {
      A c[R];
      load_chunk<P, A>(center, base + (uint64_t)k * M, c);
#pragma unroll
      for (int j = 0; j < R; j++) v[j] = alpha * c[j] + beta * v[j];
    }
    store_chunk<P, A>(out, base + (uint64_t)k * M, v, nt);
  }
}

}  // namespace dev
