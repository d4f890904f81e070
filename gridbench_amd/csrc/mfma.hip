// Matrix-core forms of the multi-right-hand-side hopping term: FP16 / BF16
// storage with FP32 accumulation (k_dhop_mfma), and FP32 storage on the
// projected half spinors (k_dhop_mfma32, further below).
//
// For one 4D site the whole hopping term is a dense real product
//
//   out[24 x M] = sum over legs d of A_d[24 x 24] * in_{nbr(d)}[24 x M],
//   A_d = (1 -/+ gamma_mu) (x) U_d   (spin (x) colour, complex as 2x2 real),
//
// with M = Ls * nrhs minor columns (s-slice, right-hand side).  Stacking the
// legs along K gives one [24 x 192] * [192 x M] product per site, done with
// v_mfma_f32_16x16x32_{f16,bf16}: two 16-row tiles (rows 24..31 zero) x six
// K-steps per 16 columns.  The B operand needs no repacking: a lane of an
// MFMA holds 8 consecutive K values of one column, and one 16-byte chunk of
// the SoA layout holds 8 consecutive reals of one (site, minor) spinor, so
// every B fragment is a single coalesced 16-byte gather (16 lanes read 256
// contiguous bytes).  Lane group g (lanes 16g .. 16g+15) carries legs 2g and
// 2g+1: K-step s reads leg 2g + s/3, chunk s%3.
//
// The A operand of a site is built once per wavefront in registers from the
// links staged in LDS: each entry is +-Re or +-Im of one link element (the
// spin factor is 0, +-1 or +-i), so the 16-bit A holds the stored link
// exactly (12-real links: the rebuilt third row is rounded once).  The
// result tile gives a lane 4 consecutive output reals of one column: two
// 8-byte stores per lane.
//
// The reference applies the same operator leg by leg
// (WilsonKernelsHandCpu.h:426-446 with the projections and reconstructions
// of lines 94-303); the gamma matrices below are those spin projections
// written as matrices.
#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

using dev::f4;

template <int P> struct MTraits;
template <> struct MTraits<GBX_F16> {
  typedef _Float16 S;
  typedef dev::h8 V8;
  typedef _Float16 V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ f4 mma(V8 a, V8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct MTraits<GBX_BF16> {
  typedef __bf16 S;
  typedef dev::b8 V8;
  typedef __bf16 V4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ f4 mma(V8 a, V8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// gamma_mu (mu = x, y, z, t), chiral basis: row a has its single nonzero in
// column gcol(mu, a), value gre + i gim with (gre, gim) from gval.
__device__ __forceinline__ int gcol(int mu, int a) { return mu < 2 ? 3 - a : a ^ 2; }
__device__ __forceinline__ void gval(int mu, int a, int &re, int &im)
{
  const bool outer = (a == 0 || a == 3);
  switch (mu) {
  case 0: re = 0; im = a < 2 ? 1 : -1; break;   // x: h0 = p0 + i p3 for 1 + gamma_x
  case 1: re = outer ? -1 : 1; im = 0; break;   // y: h0 = p0 - p3
  case 2: re = 0; im = outer ? 1 : -1; break;   // z: h0 = p0 + i p2
  default: re = 1; im = 0; break;               // t: h0 = p0 + p2
  }
}

// Work unit of a wavefront: one site of the workgroup x `mw` minor columns
// (mw a multiple of 16).  Workgroup = `spw` sites x one sweep chunk of `mcs`
// columns; its waves take units round-robin, site-major, and rebuild A only
// when the site changes.  With mcs < M the lattice is swept once per chunk
// (chunk-outer order): each sweep's working set is an mcs-column field, so
// more sites' neighbours fit an XCD's L2 (the links are staged once per chunk).
template <int P, bool DAG, bool C12, bool PF>
__global__ void __launch_bounds__(256)
    k_dhop_mfma(const void *__restrict__ gauge, const uint32_t *__restrict__ nbr, const void *__restrict__ in,
                void *__restrict__ out, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mw, uint32_t mcs,
                uint32_t nwg_sites, uint32_t nwg, uint32_t per_xcd, int xcd, int nt)
{
  typedef MTraits<P> T;
  typedef typename T::V8 V8;
  typedef typename T::V4 V4;
  typedef typename T::S S;
  extern __shared__ __align__(16) unsigned char smem[];
  float *sU = (float *)smem;
  uint32_t *sN = (uint32_t *)(sU + (size_t)spw * 8 * 18);

  const uint32_t L = xcd ? (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (L >= nwg) return;  // uniform over the workgroup, before the barrier
  const uint32_t chunk = L / nwg_sites, q0 = (L - chunk * nwg_sites) * spw;
  const uint32_t nsl = min(spw, nslot - q0);
  dev::stage_links<P, C12>(gauge, nbr, q0, nsl, sU, sN);

  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
  const uint32_t wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  // Output row of this lane in tile r: o = 16 r + i = 6 a + 2 c + sigma.
  int ra[2], rc[2], rs[2], rp[2][2], wre[2][2], wim[2][2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int o = 16 * r + i;
    ra[r] = o < 24 ? o / 6 : -1;
    rc[r] = o < 24 ? (o % 6) / 2 : 0;
    rs[r] = o & 1;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int d = 2 * g + h;
      const int K = DAG ? d : (d ^ 4);  // projector kind: 0-3 (1 + gamma), 4-7 (1 - gamma)
      const int mu = K & 3, sg = K < 4 ? 1 : -1;
      int gr = 0, gi = 0;
      if (ra[r] >= 0) gval(mu, ra[r], gr, gi);
      rp[r][h] = ra[r] >= 0 ? gcol(mu, ra[r]) : -1;
      wre[r][h] = sg * gr;
      wim[r][h] = sg * gi;
    }
  }

  const uint32_t cpw = mcs / mw;  // wave units per site
  const uint32_t units = nsl * cpw;
  const V8 *vin = (const V8 *)in;
  V4 *vout = (V4 *)out;
  uint32_t cur = 0xffffffffu;
  V8 Af[2][6];
  uint32_t nb0 = 0, nb1 = 0;
  for (uint32_t w = wave; w < units; w += nwave) {
    const uint32_t sl = w / cpw, m0 = chunk * mcs + (w - sl * cpw) * mw;
    if (sl != cur) {
      cur = sl;
      nb0 = sN[sl * 8 + 2 * g];
      nb1 = sN[sl * 8 + 2 * g + 1];
#pragma unroll
      for (int r = 0; r < 2; r++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const float *u = sU + (sl * 8 + 2 * g + h) * 18 + rc[r] * 6;
          float dg[6], of[6];
#pragma unroll
          for (int e = 0; e < 3; e++) {
            const float ur = u[2 * e], ui = u[2 * e + 1];
            // real form of z acting on (re, im) of the input: row sigma
            // 0 -> (Re z, -Im z), row 1 -> (Im z, Re z)
            dg[2 * e] = rs[r] ? ui : ur;
            dg[2 * e + 1] = rs[r] ? ur : -ui;
            const float zr = (float)wre[r][h] * ur - (float)wim[r][h] * ui;
            const float zi = (float)wre[r][h] * ui + (float)wim[r][h] * ur;
            of[2 * e] = rs[r] ? zi : zr;
            of[2 * e + 1] = rs[r] ? zr : -zi;
          }
#pragma unroll
          for (int c = 0; c < 3; c++) {
            V8 a;
#pragma unroll
            for (int j = 0; j < 8; j++) {
              const int ep = 8 * c + j, b = ep / 6, k = ep % 6;
              const float v = b == ra[r] ? dg[k] : (b == rp[r][h] ? of[k] : 0.f);
              a[j] = (S)v;
            }
            Af[r][3 * h + c] = a;
          }
        }
    }
    const uint64_t ob0 = ((uint64_t)(q0 + sl) * 3 + (uint32_t)(g >> 1)) * M;
    const uint64_t ob1 = ((uint64_t)(q0 + sl) * 3 + 2) * M;
    const uint64_t ib0 = (uint64_t)nb0 * 3 * M, ib1 = (uint64_t)nb1 * 3 * M;
    V8 B[6];
    auto gather = [&](V8 *dst, uint32_t m) {
#pragma unroll
      for (int s = 0; s < 6; s++) dst[s] = vin[(s < 3 ? ib0 : ib1) + (uint64_t)(s % 3) * M + m];
    };
    gather(B, m0 + i);
    for (uint32_t t = 0; t < mw; t += 16) {
      const uint32_t m = m0 + t + i;
      V8 Bn[6];
      if (PF && t + 16 < mw) gather(Bn, m + 16);
      f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 6; s++) {
        acc0 = T::mma(Af[0][s], B[s], acc0);
        acc1 = T::mma(Af[1][s], B[s], acc1);
      }
      V4 o0, o1;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        o0[j] = (S)acc0[j];
        o1[j] = (S)acc1[j];
      }
      V4 *p0 = vout + (ob0 + m) * 2 + (g & 1);
      V4 *p1 = vout + (ob1 + m) * 2 + (g & 1);
      if (nt) __builtin_nontemporal_store(o0, p0);
      else *p0 = o0;
      if (g < 2) {
        if (nt) __builtin_nontemporal_store(o1, p1);
        else *p1 = o1;
      }
      if (t + 16 < mw) {
        if (PF) {
#pragma unroll
          for (int s = 0; s < 6; s++) B[s] = Bn[s];
        } else {
          gather(B, m + 16);
        }
      }
    }
  }
}

// ------------------------------------------------------------ FP32 form
// FP32 storage: the matrix cores take the projected half spinors,
//
//   out[24 x M] = sum over legs d of B_d[24 x 12] * h_d[12 x M],
//   h_d = P_d psi_{nbr(d)} (the two kept spins),  B_d = R_d (U_d (x) 1),
//
// R_d the reconstruction (identity on spins 0, 1; a phase times one kept
// spin on spins 2, 3), with v_mfma_f32_16x16x4f32: K = 8 legs x 12 reals =
// 96, i.e. 24 K-steps of 4, two 16-row tiles (rows 24..31 zero).  K-slot g of
// every step belongs to lane group g (lanes 16g .. 16g+15): the 12 projected
// reals of leg 2g in steps 0..11 and of leg 2g+1 in steps 12..23, for column
// i = lane & 15.  A lane therefore gathers two whole neighbour spinors of its
// own column (a VALU thread's gathers, 16 lanes reading 256 contiguous bytes
// per chunk), projects them itself and feeds them to the MFMAs with no
// exchange between lanes.  Output tile 0's rows 4g .. 4g+3 and tile 1's rows
// 16+4g .. 19+4g (g < 2) are the 16-byte chunks g and 4+g of the output
// spinor: one 16-byte store each.
//
// The projection h = v_a + w v_b (w in {+-1, +-i}) is formed with
// multiply-adds whose multipliers are 0 or +-1, so each projected real is
// rounded once, as in the VALU kernel; B_d holds +-Re / +-Im of the stored
// link (2x for the half-spinor t legs of GBX_BASIS_TDIAG fields), exactly.

// Projector kind of leg d (dev::LegKind, at run time).
__device__ __forceinline__ int leg_kind32(bool dag, bool td, int d)
{
  const int k = dag ? d : (d ^ 4);
  return (td && (d & 3) != 3) ? (k ^ 4) : k;
}

// One leg as a lane sees it: a GBX_BASIS_TDIAG t leg reads one spin half
// (chunks k0 .. k0+2, used as they are); any other leg reads all 6 chunks and
// forms h_a = v_a + (wr + i wi) v_{b_a} for the kept spins a = 0, 1.
struct Leg32 {
  int half, k0, b[2];
  float wr[2], wi[2];
};

__device__ __forceinline__ Leg32 leg32(bool dag, bool td, int d)
{
  Leg32 p;
  p.half = td && (d & 3) == 3;
  p.k0 = (p.half && (dag ? d : (d ^ 4)) != 3) ? 3 : 0;  // kind PTP keeps the upper half
  const int K = leg_kind32(dag, td, d), mu = K & 3, sg = K < 4 ? 1 : -1;
#pragma unroll
  for (int a = 0; a < 2; a++) {
    int gr, gi;
    gval(mu, a, gr, gi);
    p.b[a] = gcol(mu, a);
    p.wr[a] = (float)(sg * gr);
    p.wi[a] = (float)(sg * gi);
  }
  return p;
}

// Coefficient (complex) of h_d[ah][ch] in out[ao][co]: U[co][ch] times the
// reconstruction's factor, or 0.
__device__ __forceinline__ void coef32(bool dag, bool td, int d, int ao, int ah, const float *u, float &zr, float &zi)
{
  zr = 0.f;
  zi = 0.f;
  if (td && (d & 3) == 3) {  // (1 +- gamma_t) = 2 x one spin half
    const int a = (dag ? d : (d ^ 4)) == 3 ? ah : ah + 2;
    if (ao == a) {
      zr = 2.f * u[0];
      zi = 2.f * u[1];
    }
    return;
  }
  const int K = leg_kind32(dag, td, d), mu = K & 3, sg = K < 4 ? 1 : -1;
  if (ao < 2) {
    if (ah == ao) {
      zr = u[0];
      zi = u[1];
    }
  } else if (ah == gcol(mu, ao)) {  // out[ao] += sg gamma_mu[ao][b] (U h)_b
    int gr, gi;
    gval(mu, ao, gr, gi);
    zr = (float)(sg * gr) * u[0] - (float)(sg * gi) * u[1];
    zi = (float)(sg * gr) * u[1] + (float)(sg * gi) * u[0];
  }
}

template <bool DAG, bool C12, bool TD, bool PF>
__global__ void __launch_bounds__(256)
    k_dhop_mfma32(const void *__restrict__ gauge, const uint32_t *__restrict__ nbr, const void *__restrict__ in,
                  void *__restrict__ out, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mw, uint32_t mcs,
                  uint32_t nwg_sites, uint32_t nwg, uint32_t per_xcd, int xcd, int nt)
{
  extern __shared__ __align__(16) unsigned char smem[];
  float *sU = (float *)smem;
  uint32_t *sN = (uint32_t *)(sU + (size_t)spw * 8 * 18);

  const uint32_t L = xcd ? (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (L >= nwg) return;  // uniform over the workgroup, before the barrier
  const uint32_t chunk = L / nwg_sites, q0 = (L - chunk * nwg_sites) * spw;
  const uint32_t nsl = min(spw, nslot - q0);
  dev::stage_links<GBX_F32, C12>(gauge, nbr, q0, nsl, sU, sN);

  const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
  const uint32_t wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const Leg32 P0 = leg32(DAG, TD, 2 * g), P1 = leg32(DAG, TD, 2 * g + 1);
  const int n0 = P0.half ? 3 : 6, n1 = P1.half ? 3 : 6;

  const uint32_t cpw = mcs / mw;  // wave units per site
  const uint32_t units = nsl * cpw;
  const f4 *vin = (const f4 *)in;
  f4 *vout = (f4 *)out;
  uint32_t cur = 0xffffffffu;
  float Af[2][24];
  uint32_t nb0 = 0, nb1 = 0;

  // Raw chunks of the lane's two legs for column m (half legs: 3 chunks).
  auto gather = [&](f4 *c, uint32_t m) {
    const uint64_t b0 = (uint64_t)nb0 * 6 * M + m + (uint64_t)P0.k0 * M;
    const uint64_t b1 = (uint64_t)nb1 * 6 * M + m + (uint64_t)P1.k0 * M;
#pragma unroll
    for (int k = 0; k < 6; k++)
      if (k < n0) c[k] = vin[b0 + (uint64_t)k * M];
#pragma unroll
    for (int k = 0; k < 6; k++)
      if (k < n1) c[6 + k] = vin[b1 + (uint64_t)k * M];
  };
  // 12 projected reals h[2 (3a + c) + re/im] of one leg from its chunks.
  auto project = [&](const Leg32 &p, const f4 *c, float *h) {
    if (p.half) {  // only its 3 chunks were gathered
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 4; j++) h[4 * k + j] = c[k][j];
      return;
    }
    float v[24];
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
      for (int j = 0; j < 4; j++) v[4 * k + j] = c[k][j];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++) {
        const float pr = p.b[a] == 3 ? v[2 * (9 + cc)] : v[2 * (6 + cc)];
        const float pi = p.b[a] == 3 ? v[2 * (9 + cc) + 1] : v[2 * (6 + cc) + 1];
        const int o = 2 * (3 * a + cc);
        h[o] = __builtin_fmaf(-p.wi[a], pi, __builtin_fmaf(p.wr[a], pr, v[o]));
        h[o + 1] = __builtin_fmaf(p.wi[a], pr, __builtin_fmaf(p.wr[a], pi, v[o + 1]));
      }
  };

  for (uint32_t w = wave; w < units; w += nwave) {
    const uint32_t sl = w / cpw, m0 = chunk * mcs + (w - sl * cpw) * mw;
    if (sl != cur) {
      cur = sl;
      nb0 = sN[sl * 8 + 2 * g];
      nb1 = sN[sl * 8 + 2 * g + 1];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        const int R = 16 * t + i, ao = R / 6, co = (R % 6) / 2, ro = R & 1;
#pragma unroll
        for (int s = 0; s < 24; s++) {
          const int d = 2 * g + s / 12, kk = s % 12, ah = kk / 6, ch = (kk % 6) / 2, rh = kk & 1;
          float zr = 0.f, zi = 0.f;
          if (R < 24) coef32(DAG, TD, d, ao, ah, sU + (sl * 8 + d) * 18 + co * 6 + 2 * ch, zr, zi);
          Af[t][s] = ro == 0 ? (rh == 0 ? zr : -zi) : (rh == 0 ? zi : zr);
        }
      }
    }
    const uint64_t ob0 = ((uint64_t)(q0 + sl) * 6 + (uint32_t)g) * M;
    const uint64_t ob1 = ((uint64_t)(q0 + sl) * 6 + 4 + (uint32_t)(g & 1)) * M;
    f4 C[12];
    gather(C, m0 + i);
    for (uint32_t t = 0; t < mw; t += 16) {
      const uint32_t m = m0 + t + i;
      float h[24];
      project(P0, C, h);
      project(P1, C + 6, h + 12);
      if (t + 16 < mw) {
        if (PF) gather(C, m + 16);  // the next columns' gathers fly during the MFMAs
      }
      f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 24; s++) {
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[0][s], h[s], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[1][s], h[s], acc1, 0, 0, 0);
      }
      f4 *p0 = vout + ob0 + m;
      f4 *p1 = vout + ob1 + m;
      if (nt) __builtin_nontemporal_store(acc0, p0);
      else *p0 = acc0;
      if (g < 2) {
        if (nt) __builtin_nontemporal_store(acc1, p1);
        else *p1 = acc1;
      }
      if (!PF && t + 16 < mw) gather(C, m + 16);
    }
  }
}

template <bool DAG, bool C12, bool TD>
void launch32_pf(bool pf, dim3 grid, dim3 block, size_t lds, hipStream_t st, const void *gauge, const uint32_t *nbr,
                 const void *in, void *out, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mw, uint32_t mcs,
                 uint32_t nwg_sites, uint32_t nwg, uint32_t per_xcd, int xcd, int nt)
{
  if (pf)
    hipLaunchKernelGGL((k_dhop_mfma32<DAG, C12, TD, true>), grid, block, lds, st, gauge, nbr, in, out, nslot, M, spw, mw,
                       mcs, nwg_sites, nwg, per_xcd, xcd, nt);
  else
    hipLaunchKernelGGL((k_dhop_mfma32<DAG, C12, TD, false>), grid, block, lds, st, gauge, nbr, in, out, nslot, M, spw,
                       mw, mcs, nwg_sites, nwg, per_xcd, xcd, nt);
}

void launch32(bool dag, bool c12, bool td, bool pf, dim3 grid, dim3 block, size_t lds, hipStream_t st, const void *gauge,
              const uint32_t *nbr, const void *in, void *out, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mw,
              uint32_t mcs, uint32_t nwg_sites, uint32_t nwg, uint32_t per_xcd, int xcd, int nt)
{
#define GBX_MF32(D_, C_, T_) \
  launch32_pf<D_, C_, T_>(pf, grid, block, lds, st, gauge, nbr, in, out, nslot, M, spw, mw, mcs, nwg_sites, nwg, per_xcd, xcd, nt)
  if (td) {
    if (!dag && !c12) GBX_MF32(false, false, true);
    else if (!dag) GBX_MF32(false, true, true);
    else if (!c12) GBX_MF32(true, false, true);
    else GBX_MF32(true, true, true);
  } else {
    if (!dag && !c12) GBX_MF32(false, false, false);
    else if (!dag) GBX_MF32(false, true, false);
    else if (!c12) GBX_MF32(true, false, false);
    else GBX_MF32(true, true, false);
  }
#undef GBX_MF32
}

template <int P, bool DAG, bool C12>
void launch_pf(bool pf, dim3 grid, dim3 block, size_t lds, hipStream_t st, const void *gauge, const uint32_t *nbr,
               const void *in, void *out, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mw, uint32_t mcs,
               uint32_t nwg_sites, uint32_t nwg, uint32_t per_xcd, int xcd, int nt)
{
  if (pf)
    hipLaunchKernelGGL((k_dhop_mfma<P, DAG, C12, true>), grid, block, lds, st, gauge, nbr, in, out, nslot, M, spw, mw, mcs,
                       nwg_sites, nwg, per_xcd, xcd, nt);
  else
    hipLaunchKernelGGL((k_dhop_mfma<P, DAG, C12, false>), grid, block, lds, st, gauge, nbr, in, out, nslot, M, spw, mw,
                       mcs, nwg_sites, nwg, per_xcd, xcd, nt);
}

template <int P>
void launch_prec(bool dag, bool c12, bool pf, dim3 grid, dim3 block, size_t lds, hipStream_t st, const void *gauge,
                 const uint32_t *nbr, const void *in, void *out, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mw,
                 uint32_t mcs, uint32_t nwg_sites, uint32_t nwg, uint32_t per_xcd, int xcd, int nt)
{
#define GBX_MF(D_, C_) \
  launch_pf<P, D_, C_>(pf, grid, block, lds, st, gauge, nbr, in, out, nslot, M, spw, mw, mcs, nwg_sites, nwg, per_xcd, xcd, nt)
  if (!dag && !c12) GBX_MF(false, false);
  else if (!dag) GBX_MF(false, true);
  else if (!c12) GBX_MF(true, false);
  else GBX_MF(true, true);
#undef GBX_MF
}

}  // namespace

int mfma_supported(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  return u && in && (in->prec == GBX_F16 || in->prec == GBX_BF16 || in->prec == GBX_F32) && in->M % 16 == 0;
}

// kind = 1 configuration: spw sites per workgroup, threads = 64 .. 256
// (one wavefront per 64), mc = minor columns per wavefront unit (a multiple
// of 16 dividing M; 0 = all of M), cblk = columns per lattice sweep (a
// multiple of the unit dividing M; 0 = all of M, one sweep), pf = gather the
// next 16 columns before the current MFMAs, xcd / nt / wgpc as for the VALU
// kernel.
int mfma_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &c, hipStream_t st,
                SlotRange range)
{
  GBX_CHECK(u->prec == in->prec && out->prec == in->prec, GBX_EMISMATCH, "dhop (MFMA): precisions differ");
  GBX_CHECK(mfma_supported(u, in), GBX_EUNSUPPORTED,
            "dhop (MFMA): needs FP32, FP16 or BF16 storage and Ls*nrhs a multiple of 16 (precision %d, Ls*nrhs %u)",
            in->prec, in->M);
  GBX_CHECK(c.threads >= 64 && c.threads <= 256 && c.threads % 64 == 0, GBX_EINVAL, "dhop (MFMA): threads=%d",
            c.threads);
  GBX_CHECK(c.spw >= 1 && c.spw <= 1024, GBX_EINVAL, "dhop (MFMA): spw=%d", c.spw);
  const uint32_t mw = c.mc <= 0 ? in->M : (uint32_t)c.mc;
  GBX_CHECK(mw % 16 == 0 && in->M % mw == 0, GBX_EINVAL, "dhop (MFMA): mc=%d must be a multiple of 16 dividing %u", c.mc,
            in->M);
  const uint32_t mcs = c.cblk <= 0 ? in->M : (uint32_t)c.cblk;
  GBX_CHECK(mcs % mw == 0 && in->M % mcs == 0, GBX_EINVAL,
            "dhop (MFMA): cblk=%d must be a multiple of the unit's %u columns dividing %u", c.cblk, mw, in->M);
  GBX_CHECK(c.wgpc >= 0 && c.wgpc <= 64, GBX_EINVAL, "dhop (MFMA): wgpc=%d", c.wgpc);
  size_t lds = (size_t)c.spw * 8 * 18 * 4 + (size_t)c.spw * 8 * 4;
  GBX_CHECK(lds <= 160 * 1024, GBX_EINVAL, "dhop (MFMA): spw=%d needs %zu B of LDS", c.spw, lds);
  if (c.wgpc > 0) {
    const size_t pad = ((size_t)160 * 1024 / (size_t)c.wgpc) & ~(size_t)511;
    if (pad > lds) lds = pad;
  }
  const uint64_t q0 = range.begin, q1 = std::min<uint64_t>(range.end, u->geom->vol);
  GBX_CHECK(q0 <= q1, GBX_EINVAL, "dhop (MFMA): slot range %llu .. %llu", (unsigned long long)q0, (unsigned long long)q1);
  // The output takes the input's spin basis when it is written whole; a part
  // of it only when it is in that basis already (as for the VALU kernel).
  const bool whole = q0 == 0 && q1 == u->geom->vol;
  GBX_CHECK(whole || out->basis == in->basis, GBX_EMISMATCH,
            "dhop (MFMA): a slot range of an output in another spin basis than the input");
  const uint32_t nslot = (uint32_t)(q1 - q0);
  const uint32_t nwg_sites = (nslot + (uint32_t)c.spw - 1) / (uint32_t)c.spw;
  const uint64_t nwg64 = (uint64_t)nwg_sites * (in->M / mcs);
  GBX_CHECK(nwg64 < (1ull << 31), GBX_EINVAL, "dhop (MFMA): %llu workgroups", (unsigned long long)nwg64);
  out->basis = in->basis;
  if (q0 == q1) return GBX_OK;
  const uint32_t nwg = (uint32_t)nwg64;
  const uint32_t per_xcd = (nwg + 7) / 8;
  const dim3 grid(c.xcd ? per_xcd * 8 : nwg), block((unsigned)c.threads);
  const bool c12 = u->comp12 != 0, pf = c.pf != 0;
  // a slot range offsets the per-slot arrays; neighbour slots stay absolute
  const uint64_t pb = (uint64_t)prec_bytes(u->prec);
  const void *gauge = (const char *)u->d + q0 * 8 * (c12 ? 12 : 18) * pb;
  const uint32_t *nbr = u->geom->nbr + q0 * 8;
  void *dout = (char *)out->d + q0 * (uint64_t)in->M * 24 * pb;
  if (in->prec == GBX_F32)
    launch32(dag != 0, c12, in->basis == GBX_BASIS_TDIAG, pf, grid, block, lds, st, gauge, nbr, in->d, dout, nslot, in->M,
             (uint32_t)c.spw, mw, mcs, nwg_sites, nwg, per_xcd, c.xcd, c.nt);
  else if (in->prec == GBX_F16)
    launch_prec<GBX_F16>(dag != 0, c12, pf, grid, block, lds, st, gauge, nbr, in->d, dout, nslot, in->M, (uint32_t)c.spw,
                         mw, mcs, nwg_sites, nwg, per_xcd, c.xcd, c.nt);
  else
    launch_prec<GBX_BF16>(dag != 0, c12, pf, grid, block, lds, st, gauge, nbr, in->d, dout, nslot, in->M, (uint32_t)c.spw,
                          mw, mcs, nwg_sites, nwg, per_xcd, c.xcd, c.nt);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}
