// Slab-marching hopping kernel (kernel kind 3, whole lattices).
//
// A workgroup owns a tile of sites in three dimensions and a minor chunk, and
// marches along the fourth dimension (the march dimension md) over a segment
// of the lattice, one slab of TS sites at a time.  Each slab's spinors are
// loaded from HBM into LDS once; two slabs are resident (the one computed and
// the next), and the slab after them is loaded into registers while the
// current one is computed, so loads stay in flight across the whole march.
//   * legs along md read LDS (the slab itself, the first slice of the next
//     slab, or the last slice of the previous slab kept in a side buffer);
//   * legs whose neighbour lies inside the tile read LDS;
//   * only legs that leave the tile gather from L2.
// The VALU kernel requests every spinor line from L2 eight times per
// application (once per neighbour); here a line is requested once for the
// slab plus once per leg that leaves a tile.  That request stream, not HBM
// bandwidth, bounds these kernels on gfx950.
//
// Threads: tid = ((site * TS + ts) * MC + mm): tile site `site`, slice ts of
// the slab, minor index m0 + mm.  With TS * MC a multiple of 64 a wavefront
// holds slices of one tile site only, so whether a leg stays in the tile is
// uniform across the wavefront (no divergent loads).
//
// Arithmetic and leg order are those of the VALU kernel (dhop_site with
// the same multiply form), so the result is the VALU kernel's, bit for bit.
//
// LDS (16-byte chunk units unless stated):
//   buf[2][NT3][TS][NCH][MC]   two slabs, chunk-major so that the 16 lanes of
//                              consecutive minor indices read 256 contiguous B
//   side[NT3][NCH][MC]         slice before the slab (backward md legs)
//   sU[NT3*TS][8][18]          the slab's links, accumulation precision
//   sQ[NT3][W]                 slots of every site of the segment, W = seg + TS + 1
//                              (the slice before, the segment, one slab after)
//   sNb[NT3][seg][8]           neighbour slots of every leg (outside legs)
//   sL[NT3][8]                 tile-local neighbour of each leg, -1 outside
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

struct MarchArgs {
  const void *gauge;
  const uint32_t *nbr, *slot_of_site;
  const void *in;
  void *out;
  int L[4];      // lattice extents
  int e[4];      // tile extents; e[md] = TS, the slab thickness
  int md;        // march dimension
  int seg;       // segment length along md (multiple of TS, divides L[md])
  int wrap;      // bit d: the tile spans dimension d (periodic inside the tile)
  int blk[3];    // tiles are visited in blocks of blk[0] x blk[1] x blk[2] tiles (the three tile dimensions)
  uint32_t M, mc, ntile, nseg, nwg, per_xcd;
  int xcd, nt;
};

// Chunks of a leg's neighbour from LDS, chunk stride `cs` (16-byte units):
// all NCH, or the spin half a GBX_BASIS_TDIAG t leg reads.
template <int P, bool DAG, int D, bool TD, typename A, typename V>
__device__ __forceinline__ void lds_leg(const V *p, uint32_t cs, A *v)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  constexpr bool half = dev::LegHalf<DAG, D, TD>::value;
  constexpr int K0 = (half && !dev::LegHalf<DAG, D, TD>::upper) ? NCH / 2 : 0, N = half ? NCH / 2 : NCH;
#pragma unroll
  for (int k = 0; k < N; k++) dev::load_chunk<P, A>(p, (uint64_t)(K0 + k) * cs, v + k * R);
}

template <int P> constexpr int march_max_threads() { return P == GBX_F64 ? 256 : 512; }

// An empty asm that reads and redefines the accumulators: every leg's
// arithmetic is complete before the next leg's loads (otherwise the compiler
// sinks all legs' arithmetic below all loads and keeps eight gathered
// spinors live).
template <typename A>
__device__ __forceinline__ void pin_acc(A *rr, A *ri)
{
  asm volatile("" : "+v"(rr[0]), "+v"(rr[1]), "+v"(rr[2]), "+v"(rr[3]), "+v"(rr[4]), "+v"(rr[5]), "+v"(rr[6]),
               "+v"(rr[7]), "+v"(rr[8]), "+v"(rr[9]), "+v"(rr[10]), "+v"(rr[11]));
  asm volatile("" : "+v"(ri[0]), "+v"(ri[1]), "+v"(ri[2]), "+v"(ri[3]), "+v"(ri[4]), "+v"(ri[5]), "+v"(ri[6]),
               "+v"(ri[7]), "+v"(ri[8]), "+v"(ri[9]), "+v"(ri[10]), "+v"(ri[11]));
}

template <int P, bool DAG, bool TD>
__global__ void __launch_bounds__(march_max_threads<P>()) k_dhop_march(MarchArgs a)
{
  typedef typename dev::Traits<P>::acc A;
  typedef typename dev::Traits<P>::vec V;
  typedef typename dev::Traits<P>::st S;
  constexpr int NCH = dev::Traits<P>::NCH;
  constexpr int CH = (P == GBX_F64 || P == GBX_F32) ? dev::MUL_CHAIN : dev::MUL_PLAIN;
  const int md = a.md, TS = a.e[md];
  // the three tile dimensions, ascending
  int td[3], n3 = 0;
  for (int d = 0; d < 4; d++)
    if (d != md) td[n3++] = d;
  const int NT3 = a.e[td[0]] * a.e[td[1]] * a.e[td[2]];
  const uint32_t MC = a.mc;
  const int seg = a.seg, NS = seg / TS, W = seg + TS + 1;

  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t slab_ch = (uint32_t)NT3 * TS * NCH * MC;  // chunks per slab
  V *sV = (V *)smem;  // the slab buffers and the side buffer, in chunks
  V *buf0 = sV;
  V *buf1 = buf0 + slab_ch;
  const uint32_t side_off = 2 * slab_ch;
  V *side = buf1 + slab_ch;                                          // [NT3][NCH][MC]
  A *sU = (A *)(side + (size_t)NT3 * NCH * MC);                      // [NT3*TS][8][18]
  uint32_t *sQ = (uint32_t *)(sU + (size_t)NT3 * TS * 8 * 18);       // [NT3][W]
  uint32_t *sNb = sQ + (size_t)NT3 * W;                              // [NT3][seg][8]
  int *sL = (int *)(sNb + (size_t)NT3 * seg * 8);                    // [NT3][8]

  const uint32_t Lg = a.xcd ? (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (Lg >= a.nwg) return;  // uniform over the workgroup, before any barrier
  const uint32_t tile = Lg % a.ntile, r1 = Lg / a.ntile, sg = r1 % a.nseg, chunk = r1 / a.nseg;
  const uint32_t m0 = chunk * MC;
  int org[4];  // tile origin: blocks of tiles lexicographic, tiles lexicographic inside a block
  {
    const uint32_t bsz = (uint32_t)(a.blk[0] * a.blk[1] * a.blk[2]);
    uint32_t b = tile / bsz, w = tile - b * bsz;
    for (int j = 0; j < 3; j++) {
      const int d = td[j], nbk = a.L[d] / a.e[d] / a.blk[j];
      const int cb = (int)(b % (uint32_t)nbk), cw = (int)(w % (uint32_t)a.blk[j]);
      org[d] = (cb * a.blk[j] + cw) * a.e[d];
      b /= (uint32_t)nbk;
      w /= (uint32_t)a.blk[j];
    }
    org[md] = (int)sg * seg;
  }
  const int nthr = (int)blockDim.x;

  // ---- prologue: slot and neighbour tables of the segment
  for (int i = threadIdx.x; i < NT3 * W; i += nthr) {
    const int site = i / W, j = i - site * W;
    int c[4];
    int s = site;
    for (int k = 0; k < 3; k++) {
      const int d = td[k];
      c[d] = org[d] + s % a.e[d];
      s /= a.e[d];
    }
    c[md] = ((org[md] - 1 + j) % a.L[md] + a.L[md]) % a.L[md];
    const uint64_t lex = (uint64_t)c[0] + (uint64_t)a.L[0] * (c[1] + (uint64_t)a.L[1] * (c[2] + (uint64_t)a.L[2] * c[3]));
    sQ[i] = a.slot_of_site[lex];
  }
  for (int i = threadIdx.x; i < NT3 * 8; i += nthr) {
    const int site = i >> 3, D = i & 7, mu = D & 3;
    int lnb = -1;
    if (mu != md) {
      int stride = 1, k = 0;
      for (; td[k] != mu; k++) stride *= a.e[td[k]];
      const int cc = (site / stride) % a.e[mu];
      int c2 = cc + (D < 4 ? 1 : -1);
      bool inside = c2 >= 0 && c2 < a.e[mu];
      if (!inside && ((a.wrap >> mu) & 1)) {
        c2 = (c2 + a.e[mu]) % a.e[mu];
        inside = true;
      }
      if (inside) lnb = site + (c2 - cc) * stride;
    }
    sL[i] = lnb;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NT3 * seg * 8; i += nthr) {
    const int site = i / (seg * 8), r = i - site * seg * 8, j = r >> 3, D = r & 7;
    sNb[i] = a.nbr[(uint64_t)sQ[site * W + 1 + j] * 8 + D];
  }

  // this thread's point
  const int tid = threadIdx.x;
  const uint32_t mm = (uint32_t)tid % MC;
  const int ts = (tid / (int)MC) % TS, site = tid / ((int)MC * TS);
  const uint32_t m = m0 + mm;
  const V *gin = (const V *)a.in;
  const S *gU = (const S *)a.gauge;
  const uint32_t own = ((uint32_t)(site * TS + ts) * NCH) * MC + mm;  // own entry in a slab buffer

  // stage slab `sb` (0 .. NS) of the segment: this thread's own chunks
  auto load_slab = [&](int sb, V *regs) {
    const uint32_t q = sQ[site * W + 1 + sb * TS + ts];
#pragma unroll
    for (int k = 0; k < NCH; k++) regs[k] = gin[((uint64_t)q * NCH + k) * a.M + m];
  };
  auto store_slab = [&](V *bufp, const V *regs) {
#pragma unroll
    for (int k = 0; k < NCH; k++) bufp[own + (uint32_t)k * MC] = regs[k];
  };
  {
    V r0[NCH];
    load_slab(0, r0);
    store_slab(buf0, r0);
    if (ts == TS - 1) {  // the slice before the segment into the side buffer
      const uint32_t q = sQ[site * W];
#pragma unroll
      for (int k = 0; k < NCH; k++)
        side[((uint32_t)site * NCH + k) * MC + mm] = gin[((uint64_t)q * NCH + k) * a.M + m];
    }
    load_slab(1, r0);
    store_slab(buf1, r0);
  }
  {
    const S *ul = gU + (uint64_t)sQ[site * W + 1 + ts] * 144;
    A *ud = sU + (uint32_t)(site * TS + ts) * 144;
    for (uint32_t r = mm; r < 144u; r += MC) ud[r] = (A)__builtin_nontemporal_load(ul + r);
  }
  __syncthreads();

  // ---- the march
  for (int s = 0; s < NS; s++) {
    V *cur = (s & 1) ? buf1 : buf0;
    const uint32_t cur_off = (s & 1) ? slab_ch : 0u, nxt_off = (s & 1) ? 0u : slab_ch;
    // (a) prefetch: slab s+2 into registers; the links of slab s+1, each
    // (site, slice) point's 144 reals by the MC lanes of that point (lane mm
    // takes reals mm, mm + MC, ...: at most LPT = 144 / 16 per lane)
    const bool pre = s + 2 <= NS, lnext = s + 1 < NS;
    V regs[NCH];
    if (pre) load_slab(s + 2, regs);
    constexpr int LPT = 9;
    S lk[LPT];
    if (lnext) {
      const S *ul = gU + (uint64_t)sQ[site * W + 1 + (s + 1) * TS + ts] * 144 + mm;
#pragma unroll
      for (int j = 0; j < LPT; j++)
        if (mm + (uint32_t)j * MC < 144u) lk[j] = __builtin_nontemporal_load(ul + j * MC);
    }
    // (b) compute slab s, leg by leg in the reference's order
    const uint32_t q = sQ[site * W + 1 + s * TS + ts];
    const uint32_t *nbq = sNb + ((uint32_t)site * seg + (uint32_t)(s * TS + ts)) * 8;
    const A *Uq = sU + (uint32_t)(site * TS + ts) * 8 * 18;
    A rr[12], ri[12], v[24];
#define GBX_MLEG(D, FIRST)                                                                                             \
    {                                                                                                                  \
      uint32_t off = 0; /* chunk offset from smem of the LDS copy */                                                  \
      bool inl = true;                                                                                                 \
      if (((D) & 3) == md) {                                                                                           \
        if ((D) < 4) off = (ts < TS - 1) ? cur_off + own + (uint32_t)NCH * MC : nxt_off + own - (uint32_t)ts * NCH * MC; \
        else off = (ts > 0) ? cur_off + own - (uint32_t)NCH * MC : side_off + ((uint32_t)site * NCH) * MC + mm;       \
      } else {                                                                                                         \
        const int ln = sL[site * 8 + (D)];                                                                             \
        inl = ln >= 0;                                                                                                 \
        off = cur_off + ((uint32_t)(ln * TS + ts) * NCH) * MC + mm;                                                    \
      }                                                                                                                \
      if (inl) {                                                                                                       \
        lds_leg<P, DAG, D, TD>(sV + off, MC, v);                                                                       \
        dev::leg_apply<DAG, D, TD, FIRST, CH>(v, Uq + (D) * 18, rr, ri);                                               \
      } else {                                                                                                         \
        dev::leg_load<P, DAG, D, TD>(a.in, nbq[D], a.M, m, v);                                                         \
        dev::leg_apply<DAG, D, TD, FIRST, CH>(v, Uq + (D) * 18, rr, ri);                                               \
      }                                                                                                                \
      pin_acc(rr, ri);                                                                                                 \
    }
    GBX_MLEG(0, true)
    GBX_MLEG(1, false)
    GBX_MLEG(2, false)
    GBX_MLEG(3, false)
    GBX_MLEG(4, false)
    GBX_MLEG(5, false)
    GBX_MLEG(6, false)
    GBX_MLEG(7, false)
#undef GBX_MLEG
    dev::store_site<P, false>(a.out, nullptr, q, a.M, m, rr, ri, (A)0, (A)1, a.nt != 0);
    __syncthreads();  // (d) slab s, the side buffer and the links are read
    if (s + 1 < NS) {
      // (e) the last slice of slab s becomes the side slice of slab s+1
      if (ts == TS - 1) {
#pragma unroll
        for (int k = 0; k < NCH; k++) side[((uint32_t)site * NCH + k) * MC + mm] = cur[own + (uint32_t)k * MC];
      }
      // (f) slab s+2 replaces slab s; the links of slab s+1
      if (pre) store_slab(cur, regs);
      {
        A *ud = sU + (uint32_t)(site * TS + ts) * 144 + mm;
#pragma unroll
        for (int j = 0; j < LPT; j++)
          if (mm + (uint32_t)j * MC < 144u) ud[j * MC] = (A)lk[j];
      }
      __syncthreads();
    }
  }
}

template <int P, bool TD>
void launch_p(bool dag, dim3 grid, dim3 block, size_t lds, hipStream_t st, const MarchArgs &a)
{
  if (dag) hipLaunchKernelGGL((k_dhop_march<P, true, TD>), grid, block, lds, st, a);
  else hipLaunchKernelGGL((k_dhop_march<P, false, TD>), grid, block, lds, st, a);
}

struct MarchShape {
  int e[4], md;
};

MarchShape unpack_march(int packed)
{
  MarchShape s;
  for (int d = 0; d < 4; d++) s.e[d] = (packed >> (4 * d)) & 15;
  s.md = (packed >> 16) & 3;
  return s;
}

size_t march_lds(int prec, const MarchShape &s, int mc, int seg)
{
  const size_t acc = prec == GBX_F64 ? 8 : 4;
  const size_t nch = prec == GBX_F64 ? 12 : prec == GBX_F32 ? 6 : 3;
  const int TS = s.e[s.md];
  int nt3 = 1;
  for (int d = 0; d < 4; d++)
    if (d != s.md) nt3 *= s.e[d];
  const size_t W = (size_t)seg + TS + 1;
  return 16 * nch * (size_t)mc * (2 * (size_t)nt3 * TS + nt3) + (size_t)nt3 * TS * 8 * 18 * acc + 4 * (size_t)nt3 * W +
         4 * (size_t)nt3 * seg * 8 + 4 * (size_t)nt3 * 8;
}

// Blocks of tiles a workgroup sequence visits together: about 32 tiles (the
// workgroups resident on one XCD at a time: one per CU), the block with the
// fewest neighbour sites outside it per site inside among the divisors of
// the tile grid, so that the legs that leave one tile mostly land in a tile
// whose workgroup is resident on the same XCD, and hit its L2.
void march_blocks(const int *dims, const MarchShape &s, int blk[3])
{
  int td[3], n = 0, nt[3];
  for (int d = 0; d < 4; d++)
    if (d != s.md) td[n++] = d;
  for (int j = 0; j < 3; j++) nt[j] = dims[td[j]] / s.e[td[j]];
  double best = 1e30;
  blk[0] = blk[1] = blk[2] = 1;
  for (int a = 1; a <= nt[0]; a++)
    for (int b = 1; b <= nt[1]; b++)
      for (int c = 1; c <= nt[2]; c++) {
        if (nt[0] % a || nt[1] % b || nt[2] % c) continue;
        const int tiles = a * b * c;
        if (tiles > 48 || tiles < 16) continue;
        const double x = (double)a * s.e[td[0]], y = (double)b * s.e[td[1]], z = (double)c * s.e[td[2]];
        // halo sites per site of the block's cross-section, a small penalty away from 32 tiles
        const double cost = 2.0 * (y * z + x * z + x * y) / (x * y * z) + 0.002 * std::abs(tiles - 32);
        if (cost < best) {
          best = cost;
          blk[0] = a;
          blk[1] = b;
          blk[2] = c;
        }
      }
}

}  // namespace

int pack_march(const int e[4], int md) { return e[0] | (e[1] << 4) | (e[2] << 8) | (e[3] << 12) | (md << 16); }

// Configuration check for kind 3.
int march_cfg_check(const gbx_kcfg &k, const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  const gbx_geom_s *g = u->geom;
  GBX_CHECK(whole_lattice(g), GBX_EUNSUPPORTED, "march kernel: whole lattices only");
  GBX_CHECK(!u->comp12, GBX_EUNSUPPORTED, "march kernel: 18-real links only");
  GBX_CHECK(u->prec == GBX_F64 || u->prec == GBX_F32, GBX_EUNSUPPORTED, "march kernel: FP64 / FP32 fields");
  const MarchShape s = unpack_march(k.tile);
  for (int d = 0; d < 4; d++)
    GBX_CHECK(s.e[d] >= 1 && g->dims[d] % s.e[d] == 0, GBX_EINVAL, "march kernel: extent %d = %d does not divide %d", d,
              s.e[d], g->dims[d]);
  const int TS = s.e[s.md], Lm = g->dims[s.md];
  const int seg = k.cblk > 0 ? k.cblk : Lm;
  GBX_CHECK(seg % TS == 0 && Lm % seg == 0, GBX_EINVAL, "march kernel: segment %d (slab %d, extent %d)", seg, TS, Lm);
  const uint32_t mc = (k.mc <= 0 || (uint32_t)k.mc > in->M) ? in->M : (uint32_t)k.mc;
  GBX_CHECK(in->M % mc == 0, GBX_EINVAL, "march kernel: minor chunk %u does not divide %u", mc, in->M);
  GBX_CHECK((TS * mc) % 64 == 0, GBX_EINVAL, "march kernel: slab %d x minor chunk %u is not whole wavefronts", TS, mc);
  int nt3 = 1;
  for (int d = 0; d < 4; d++)
    if (d != s.md) nt3 *= s.e[d];
  const int tmax = u->prec == GBX_F64 ? march_max_threads<GBX_F64>() : march_max_threads<GBX_F32>();
  GBX_CHECK(k.threads == nt3 * TS * (int)mc && k.threads <= tmax, GBX_EINVAL,
            "march kernel: threads=%d, the tile needs %d (max %d)", k.threads, nt3 * TS * (int)mc, tmax);
  GBX_CHECK(mc >= 16, GBX_EINVAL, "march kernel: minor chunk %u < 16 (links staged at most 9 reals per thread)", mc);
  const size_t lds = march_lds(u->prec, s, (int)mc, seg);
  GBX_CHECK(lds <= 160 * 1024, GBX_EINVAL, "march kernel: needs %zu B of LDS", lds);
  return GBX_OK;
}

int march_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k,
                 hipStream_t st)
{
  GBX_TRY(march_cfg_check(k, u, in));
  out->basis = in->basis;
  const gbx_geom_s *g = u->geom;
  const MarchShape s = unpack_march(k.tile);
  MarchArgs a;
  a.gauge = u->d;
  a.nbr = g->nbr;
  a.slot_of_site = g->slot_of_site;
  a.in = in->d;
  a.out = out->d;
  for (int d = 0; d < 4; d++) {
    a.L[d] = g->dims[d];
    a.e[d] = s.e[d];
  }
  a.md = s.md;
  a.seg = k.cblk > 0 ? k.cblk : g->dims[s.md];
  a.wrap = 0;
  for (int d = 0; d < 4; d++)
    if (d != s.md && s.e[d] == g->dims[d]) a.wrap |= 1 << d;
  a.M = in->M;
  a.mc = (k.mc <= 0 || (uint32_t)k.mc > in->M) ? in->M : (uint32_t)k.mc;
  march_blocks(g->dims, s, a.blk);
  uint64_t ntile = 1;
  for (int d = 0; d < 4; d++)
    if (d != s.md) ntile *= (uint64_t)(g->dims[d] / s.e[d]);
  a.ntile = (uint32_t)ntile;
  a.nseg = (uint32_t)(g->dims[s.md] / a.seg);
  const uint64_t nwg = ntile * a.nseg * (a.M / a.mc);
  GBX_CHECK(nwg < (1ull << 31), GBX_EINVAL, "march kernel: %llu workgroups", (unsigned long long)nwg);
  a.nwg = (uint32_t)nwg;
  a.per_xcd = (a.nwg + 7) / 8;
  a.xcd = k.xcd;
  a.nt = k.nt;
  const size_t lds = march_lds(u->prec, s, (int)a.mc, a.seg);
  const dim3 grid(k.xcd ? a.per_xcd * 8 : a.nwg), block((unsigned)k.threads);
  const bool td = in->basis == GBX_BASIS_TDIAG;
  if (u->prec == GBX_F64) {
    if (td) launch_p<GBX_F64, true>(dag, grid, block, lds, st, a);
    else launch_p<GBX_F64, false>(dag, grid, block, lds, st, a);
  } else {
    if (td) launch_p<GBX_F32, true>(dag, grid, block, lds, st, a);
    else launch_p<GBX_F32, false>(dag, grid, block, lds, st, a);
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}
