// LDS-tiled hopping kernel (kernel kind 2).
//
// The VALU kernel gathers each of a site's 8 neighbours from L2: every
// spinor line is requested from L2 eight times per application, and those
// requests, not HBM bandwidth, set its speed.  Here a workgroup owns a 4D
// box of sites (the tile, extents tx x ty x tz x tt) and a chunk of mc minor
// indices.  It loads the tile's own spinors once into LDS; a leg whose
// neighbour lies inside the box reads LDS, only legs that leave the box
// gather from L2.  For a 2^4 tile half the legs stay inside, so each spinor
// line is requested 1 + 4 times instead of 8.  GBX_BASIS_TDIAG fields: the
// t legs read half of each spinor, from LDS or memory (dev::LegHalf).
//
// Tiles are visited in the geometry's slot order (sorted by their smallest
// slot), so consecutive workgroups -- and, with xcd = 1, each XCD's share --
// cover a compact region whose halos the XCD's L2 holds.  A site's data is
// contiguous in the SoA layout ([slot][chunk][minor]), so the tile's sites
// need not be consecutive slots.
#include <algorithm>
#include <map>
#include <mutex>

#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

std::mutex g_tile_mu;

struct TileShape {
  int e[4];
  int T() const { return e[0] * e[1] * e[2] * e[3]; }
};

TileShape unpack_tile(int packed)
{
  TileShape s;
  for (int d = 0; d < 4; d++) s.e[d] = (packed >> (4 * d)) & 15;
  return s;
}

// Slots of every tile's sites, tiles ordered by smallest slot, sites in
// tile-lexicographic order (x fastest).  Device array [ntiles][T].
int tile_table(gbx_geom_s *g, int packed, const uint32_t **tab, uint32_t *ntiles)
{
  std::lock_guard<std::mutex> lk(g_tile_mu);
  auto it = g->tile_tabs.find(packed);
  const TileShape s = unpack_tile(packed);
  uint64_t nt = g->vol / (uint64_t)s.T();
  if (it != g->tile_tabs.end()) {
    *tab = it->second;
    *ntiles = (uint32_t)nt;
    return GBX_OK;
  }
  const int *D = g->dims;
  int nb[4];
  for (int d = 0; d < 4; d++) nb[d] = D[d] / s.e[d];
  const int T = s.T();
  std::vector<std::pair<uint32_t, uint32_t>> order(nt);  // (min slot, tile)
  std::vector<uint32_t> slots((size_t)nt * T);
  for (uint64_t t = 0; t < nt; t++) {
    uint64_t r = t;
    int o[4];
    for (int d = 0; d < 4; d++) {
      o[d] = (int)(r % (uint64_t)nb[d]) * s.e[d];
      r /= (uint64_t)nb[d];
    }
    uint32_t mn = UINT32_MAX;
    for (int i = 0; i < T; i++) {
      int c[4], q = i;
      for (int d = 0; d < 4; d++) {
        c[d] = o[d] + q % s.e[d];
        q /= s.e[d];
      }
      const uint64_t site = (uint64_t)c[0] + (uint64_t)D[0] * (c[1] + (uint64_t)D[1] * (c[2] + (uint64_t)D[2] * c[3]));
      const uint32_t slot = g->slot_of_site_h[site];
      slots[t * T + i] = slot;
      mn = std::min(mn, slot);
    }
    order[t] = {mn, (uint32_t)t};
  }
  std::sort(order.begin(), order.end());
  std::vector<uint32_t> sorted((size_t)nt * T);
  for (uint64_t t = 0; t < nt; t++)
    std::copy(slots.begin() + (size_t)order[t].second * T, slots.begin() + (size_t)(order[t].second + 1) * T,
              sorted.begin() + t * T);
  uint32_t *d = nullptr;
  GBX_HIP(hipMalloc(&d, sorted.size() * sizeof(uint32_t)));
  if (hipMemcpy(d, sorted.data(), sorted.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return gbx_fail(GBX_EHIP, "tile table upload failed");
  }
  g->tile_tabs[packed] = d;
  *tab = d;
  *ntiles = (uint32_t)nt;
  return GBX_OK;
}

// Launch bound: FP64 holds 24 accumulation doubles per item, and the
// TDIAG form of FP32 needs more than the 128 VGPRs that 1024-thread
// workgroups leave (at 1024 it spills).
constexpr int tile_max_threads(int prec, bool td) { return prec == GBX_F64 ? 512 : td ? 768 : 1024; }

// A leg's chunks from the LDS copy of the box: all NCH, or the spin half a
// GBX_BASIS_TDIAG t leg reads (dev::LegHalf).
template <int P, bool DAG, int D, bool TD, typename A, typename V>
__device__ __forceinline__ void lds_leg_load(const V *p, A *v)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  constexpr bool half = dev::LegHalf<DAG, D, TD>::value;
  constexpr int K0 = (half && !dev::LegHalf<DAG, D, TD>::upper) ? NCH / 2 : 0, N = half ? NCH / 2 : NCH;
#pragma unroll
  for (int k = 0; k < N; k++) dev::load_chunk<P, A>(p, K0 + k, v + k * R);
}

struct TileArgs {
  const void *gauge;
  const uint32_t *nbr, *tslots;
  const void *in;
  void *out;
  uint32_t ntiles, M, mc, nwg, per_xcd;
  int ex, ey, ez, et, wrap;  // tile extents; bit d of wrap: extent d spans the lattice (periodic inside the tile)
  int xcd, nt;
};

template <int P, bool DAG, bool C12, bool TD = false>
__global__ void __launch_bounds__(tile_max_threads(P, TD)) k_dhop_tile(TileArgs a)
{
  typedef typename dev::Traits<P>::acc A;
  typedef typename dev::Traits<P>::vec V;
  constexpr int NCH = dev::Traits<P>::NCH, RL = C12 ? 12 : 18;
  // LDS holds a (site, minor) point's NCH chunks contiguously, SW chunks
  // apart: one base address per leg, the chunks at immediate offsets.  FP64's
  // 12 chunks are padded to 13 so that 8 consecutive lanes' 16-byte reads
  // fall in distinct banks.
  constexpr int SW = NCH == 12 ? 13 : NCH;
  constexpr int CH = (P == GBX_F64 || P == GBX_F32) ? dev::MUL_CHAIN : dev::MUL_PLAIN;
  const int e[4] = {a.ex, a.ey, a.ez, a.et};
  const int T = e[0] * e[1] * e[2] * e[3];
  extern __shared__ __align__(16) unsigned char smem[];
  V *sIn = (V *)smem;                                       // [T][mc][SW]
  A *sU = (A *)(sIn + (size_t)T * SW * a.mc);               // [T][8][18]
  uint32_t *sN = (uint32_t *)(sU + (size_t)T * 8 * 18);     // [T][8]
  uint32_t *sQ = sN + (size_t)T * 8;                        // [T]
  int *sL = (int *)(sQ + T);                                // [T][8] local neighbour or -1

  const uint32_t L = a.xcd ? (blockIdx.x & 7u) * a.per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (L >= a.nwg) return;  // uniform over the workgroup, before any barrier
  const uint32_t tile = L % a.ntiles, chunk = L / a.ntiles;
  const uint32_t m0 = chunk * a.mc, mcc = min(a.mc, a.M - m0);
  for (int i = threadIdx.x; i < T; i += blockDim.x) sQ[i] = a.tslots[(uint64_t)tile * T + i];
  __syncthreads();
  // The tile's spinors (minor chunk), links and neighbour slots into LDS.
  // Each thread issues NB loads before it waits for any of them: a loop of
  // load-then-store keeps one load per wavefront in flight, and the staging
  // phase then runs at L2/HBM latency rather than bandwidth.  Indexes past
  // the end are clamped to the last element (loaded, not stored).
  constexpr int NB = 8;
  const uint32_t nthr = blockDim.x;
  const V *gin = (const V *)a.in;
  const uint32_t nload = (uint32_t)T * NCH * mcc;
  for (uint32_t w0 = threadIdx.x; w0 < nload; w0 += NB * nthr) {
    V tmp[NB];
    uint32_t dst[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const uint32_t w = min(w0 + j * nthr, nload - 1);
      const uint32_t ik = w / mcc, m = w - ik * mcc;
      const uint32_t i = ik / NCH, k = ik - i * NCH;
      tmp[j] = gin[((uint64_t)sQ[i] * NCH + k) * a.M + m0 + m];
      dst[j] = (i * a.mc + m) * SW + k;
    }
#pragma unroll
    for (int j = 0; j < NB; j++)
      if (w0 + j * nthr < nload) sIn[dst[j]] = tmp[j];
  }
  typedef typename dev::Traits<P>::st S;
  const S *gU = (const S *)a.gauge;
  const uint32_t nlink = (uint32_t)T * 8 * RL;
  for (uint32_t w0 = threadIdx.x; w0 < nlink; w0 += NB * nthr) {
    S tmp[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const uint32_t w = min(w0 + j * nthr, nlink - 1);
      const uint32_t i = w / (8 * RL), r = w - i * (8 * RL);
      tmp[j] = gU[(uint64_t)sQ[i] * 8 * RL + r];
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const uint32_t w = w0 + j * nthr;
      if (w < nlink) {
        const uint32_t i = w / (8 * RL), r = w - i * (8 * RL);
        const uint32_t link = r / RL, k = r - link * RL;
        sU[(i * 8 + link) * 18 + k] = (A)tmp[j];
      }
    }
  }
  for (uint32_t w = threadIdx.x; w < (uint32_t)T * 8; w += blockDim.x) {
    sN[w] = a.nbr[(uint64_t)sQ[w >> 3] * 8 + (w & 7)];
    // Leg d's neighbour inside the box, wrapping only along an extent that
    // spans the whole (periodic) lattice; -1: it leaves the box.
    const int i = (int)(w >> 3), d = (int)(w & 7), mu = d & 3;
    int stride = 1;
    for (int k = 0; k < mu; k++) stride *= e[k];
    const int c = (i / stride) % e[mu];
    int c2 = c + (d < 4 ? 1 : -1);
    bool inside = c2 >= 0 && c2 < e[mu];
    if (!inside && ((a.wrap >> mu) & 1)) {
      c2 = (c2 + e[mu]) % e[mu];
      inside = true;
    }
    sL[w] = inside ? i + (c2 - c) * stride : -1;
  }
  if (C12) {
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < (uint32_t)T * 8; l += blockDim.x) dev::su3_row2(sU + l * 18);
  }
  __syncthreads();

  // A leg whose neighbour is inside the box reads LDS, one that leaves it
  // gathers from L2; each path carries its own copy of the leg arithmetic
  // (a shared tail after the two loads makes the compiler keep both sets of
  // loaded values live across the legs).
  const uint32_t items = (uint32_t)T * mcc;
  for (uint32_t w = threadIdx.x; w < items; w += blockDim.x) {
    const int i = (int)(w / mcc);
    const uint32_t ml = w - (uint32_t)i * mcc, m = m0 + ml;
    A rr[12], ri[12];
    A v[24];
#define GBX_TLEG(D, FIRST)                                                                                             \
    {                                                                                                                  \
      const int ln = sL[i * 8 + (D)];                                                                                  \
      if (ln >= 0) {                                                                                                   \
        lds_leg_load<P, DAG, D, TD>(sIn + ((uint32_t)ln * a.mc + ml) * SW, v);                                         \
        dev::leg_apply<DAG, D, TD, FIRST, CH>(v, sU + (i * 8 + D) * 18, rr, ri);                                       \
      } else {                                                                                                         \
        dev::leg_load<P, DAG, D, TD>(a.in, sN[i * 8 + D], a.M, m, v);                                                  \
        dev::leg_apply<DAG, D, TD, FIRST, CH>(v, sU + (i * 8 + D) * 18, rr, ri);                                       \
      }                                                                                                                \
    }
    GBX_TLEG(0, true)
    GBX_TLEG(1, false)
    GBX_TLEG(2, false)
    GBX_TLEG(3, false)
    GBX_TLEG(4, false)
    GBX_TLEG(5, false)
    GBX_TLEG(6, false)
    GBX_TLEG(7, false)
#undef GBX_TLEG
    dev::store_site<P, false>(a.out, nullptr, sQ[i], a.M, m, rr, ri, (A)0, (A)1, a.nt != 0);
  }
}

template <int P, bool TD = false>
void launch_p(bool dag, bool c12, dim3 grid, dim3 block, size_t lds, hipStream_t st, const TileArgs &a)
{
  if (!dag && !c12) hipLaunchKernelGGL((k_dhop_tile<P, false, false, TD>), grid, block, lds, st, a);
  else if (!dag && c12) hipLaunchKernelGGL((k_dhop_tile<P, false, true, TD>), grid, block, lds, st, a);
  else if (dag && !c12) hipLaunchKernelGGL((k_dhop_tile<P, true, false, TD>), grid, block, lds, st, a);
  else hipLaunchKernelGGL((k_dhop_tile<P, true, true, TD>), grid, block, lds, st, a);
}

// ------------------------------------------------------------ async form
// Persistent variant (kind 2 with pf = 1; FP64 / FP32, 18-real links): each
// workgroup walks a strided sequence of (tile, minor chunk) units with two
// LDS buffers.  While it computes unit k from one buffer, the spinors and
// links of unit k+1 travel into the other by direct global -> LDS loads
// (global_load_lds_dwordx4), which hold no registers.  Those loads are L1
// requests like any other, though, and the L1's in-flight capacity, not
// VGPRs, bounds these kernels: with half the LDS per box (two buffers) this
// form issues more L2 requests than the one-shot kernel and measures slower
// on the 16-RHS FP32 problem (7.3 against 5.3 ms at 24^4 x Ls16).  It is a
// tuner option.  Units are chunk-major and each XCD takes a contiguous eighth
// of them, its workgroups taking consecutive units (so the tiles in flight on
// an XCD are neighbours); with 8 minor chunks an XCD works on its own chunk of
// the whole lattice.  LDS per buffer: spinors [T][NCH][mc] (minor fastest:
// a wavefront's 64 lanes fill 64 consecutive 16-byte items, as the
// instruction requires) and links [T][8][18] (the per-slot global layout,
// copied as is); each area is padded to whole wavefronts of items.  Same
// loads, same arithmetic, same leg order as k_dhop: bit-identical results.
struct AsyncArgs {
  const void *gauge;
  const uint32_t *nbr, *tslots;
  const void *in;
  void *out;
  uint32_t ntiles, M, mc, nunits, units_per_xcd, wg_per_xcd;
  int ex, ey, ez, et, wrap, nt;
};

// 16-byte items of one buffer: spinor and link areas, each a multiple of 64.
struct AsyncLayout {
  uint32_t nspin, nlink, spin_pad, link_pad;
};
inline __host__ __device__ AsyncLayout async_layout(int T, int nch, int link_items, uint32_t mc)
{
  AsyncLayout l;
  l.nspin = (uint32_t)T * (uint32_t)nch * mc;
  l.nlink = (uint32_t)T * (uint32_t)link_items;
  l.spin_pad = (l.nspin + 63) / 64 * 64;
  l.link_pad = (l.nlink + 63) / 64 * 64;
  return l;
}

// Launch bound: at 512 threads FP64 spills (24 accumulation doubles per item).
constexpr int async_max_threads(int prec) { return prec == GBX_F64 ? 256 : 512; }

// A leg's chunks from an LDS area whose chunks lie `stride` items apart.
template <int P, bool DAG, int D, bool TD, typename A, typename V>
__device__ __forceinline__ void lds_leg_load_s(const V *p, uint32_t stride, A *v)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  constexpr bool half = dev::LegHalf<DAG, D, TD>::value;
  constexpr int K0 = (half && !dev::LegHalf<DAG, D, TD>::upper) ? NCH / 2 : 0, N = half ? NCH / 2 : NCH;
#pragma unroll
  for (int k = 0; k < N; k++) dev::load_chunk<P, A>(p, (uint64_t)(K0 + k) * stride, v + k * R);
}

// Start unit u's loads into one buffer (spinor area ds, link area du, slot
// and neighbour tables dq, dn): spinors and links asynchronously (lanes past
// an area's end load its last item into the padding), the tables
// synchronously.
// One wavefront-wide direct load: lane l's 16 bytes at src go to LDS at
// dst + 16 l (dst uniform over the wavefront).  Kept out of templates: the
// host pass rejects the builtin in a dependent context.
__device__ __forceinline__ void lds_dma16(const void *src, void *dst)
{
  typedef __attribute__((address_space(3))) void lds_void;
  __builtin_amdgcn_global_load_lds(src, (lds_void *)dst, 16, 0, 0);
}

template <int P, typename V>
__device__ __forceinline__ void async_issue(const AsyncArgs &a, const AsyncLayout &lay, int T, uint32_t u, V *ds, V *du,
                                            uint32_t *dq, uint32_t *dn)
{
  constexpr int NCH = dev::Traits<P>::NCH;
  constexpr int LV = 8 * 18 * (int)sizeof(typename dev::Traits<P>::acc) / 16;
  const uint32_t nthr = blockDim.x, lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nwave = nthr >> 6;
  const uint32_t mc = a.mc, tile = u % a.ntiles, m0 = (u / a.ntiles) * mc;
  const uint32_t *ts = a.tslots + (uint64_t)tile * T;
  for (uint32_t it0 = wave * 64; it0 < lay.spin_pad; it0 += nwave * 64) {
    const uint32_t it = min(it0 + lane, lay.nspin - 1);
    const uint32_t m = it % mc, ik = it / mc, k = ik % NCH, i = ik / NCH;
    const V *src = (const V *)a.in + ((uint64_t)ts[i] * NCH + k) * a.M + m0 + m;
    lds_dma16(src, ds + it0);
  }
  for (uint32_t it0 = wave * 64; it0 < lay.link_pad; it0 += nwave * 64) {
    const uint32_t it = min(it0 + lane, lay.nlink - 1);
    const uint32_t i = it / LV, r = it - i * LV;
    const V *src = (const V *)a.gauge + (uint64_t)ts[i] * LV + r;
    lds_dma16(src, du + it0);
  }
  for (uint32_t w = threadIdx.x; w < (uint32_t)T * 9; w += nthr) {
    if (w < (uint32_t)T) {
      dq[w] = ts[w];
    } else {
      const uint32_t t = w - T;
      dn[t] = a.nbr[(uint64_t)ts[t >> 3] * 8 + (t & 7)];
    }
  }
}

template <int P, bool DAG, bool TD>
__global__ void __launch_bounds__(async_max_threads(P)) k_dhop_tile_async(AsyncArgs a)
{
  typedef typename dev::Traits<P>::acc A;
  typedef typename dev::Traits<P>::vec V;
  constexpr int NCH = dev::Traits<P>::NCH;
  constexpr int LV = 8 * 18 * (int)sizeof(A) / 16;  // 16-byte link items per site
  constexpr int CH = dev::MUL_CHAIN;
  const int e[4] = {a.ex, a.ey, a.ez, a.et};
  const int T = e[0] * e[1] * e[2] * e[3];
  const uint32_t mc = a.mc;
  const AsyncLayout lay = async_layout(T, NCH, LV, mc);
  extern __shared__ __align__(16) unsigned char smem[];
  V *sIn = (V *)smem;                                  // [2][spin_pad]
  V *sUv = sIn + 2 * (size_t)lay.spin_pad;             // [2][link_pad]
  uint32_t *sQ = (uint32_t *)(sUv + 2 * (size_t)lay.link_pad);  // [2][T]
  uint32_t *sN = sQ + 2 * T;                           // [2][T][8]
  int *sL = (int *)(sN + 2 * 8 * T);                   // [T][8]

  const uint32_t nthr = blockDim.x;
  const uint32_t x = blockIdx.x & 7u, j = blockIdx.x >> 3;
  const uint32_t base = x * a.units_per_xcd;
  const uint32_t end = min(base + a.units_per_xcd, a.nunits);
  if (base + j >= end) return;  // uniform over the workgroup, before any barrier

  // Leg d's neighbour inside the box (wrapping only along an extent that
  // spans the lattice), or -1: the same for every tile.
  for (uint32_t w = threadIdx.x; w < (uint32_t)T * 8; w += nthr) {
    const int i = (int)(w >> 3), d = (int)(w & 7), mu = d & 3;
    int stride = 1;
    for (int k = 0; k < mu; k++) stride *= e[k];
    const int c = (i / stride) % e[mu];
    int c2 = c + (d < 4 ? 1 : -1);
    bool inside = c2 >= 0 && c2 < e[mu];
    if (!inside && ((a.wrap >> mu) & 1)) {
      c2 = (c2 + e[mu]) % e[mu];
      inside = true;
    }
    sL[w] = inside ? i + (c2 - c) * stride : -1;
  }

  async_issue<P>(a, lay, T, base + j, sIn, sUv, sQ, sN);
  __builtin_amdgcn_s_waitcnt(0);  // this thread's loads, the direct-to-LDS ones included, have landed
  __syncthreads();
  const uint32_t items = (uint32_t)T * mc;
  for (uint32_t k = 0;; k++) {
    const uint32_t b = k & 1u, u = base + j + k * a.wg_per_xcd, un = u + a.wg_per_xcd;
    const bool next = un < end;
    if (next) async_issue<P>(a, lay, T, un, sIn + (b ^ 1u) * lay.spin_pad, sUv + (b ^ 1u) * lay.link_pad,
                             sQ + (b ^ 1u) * T, sN + (b ^ 1u) * 8 * T);
    const uint32_t m0 = (u / a.ntiles) * mc;
    const V *cin = sIn + b * lay.spin_pad;
    const A *cU = (const A *)(sUv + b * lay.link_pad);
    const uint32_t *cQ = sQ + b * T, *cN = sN + b * 8 * T;
    for (uint32_t w = threadIdx.x; w < items; w += nthr) {
      const int i = (int)(w / mc);
      const uint32_t ml = w - (uint32_t)i * mc, m = m0 + ml;
      A rr[12], ri[12];
      A v[24];
#define GBX_ALEG(D, FIRST)                                                                                             \
      {                                                                                                                \
        const int ln = sL[i * 8 + (D)];                                                                                \
        if (ln >= 0) {                                                                                                 \
          lds_leg_load_s<P, DAG, D, TD>(cin + ((uint32_t)ln * NCH) * mc + ml, mc, v);                                  \
          dev::leg_apply<DAG, D, TD, FIRST, CH>(v, cU + (i * 8 + D) * 18, rr, ri);                                     \
        } else {                                                                                                       \
          dev::leg_load<P, DAG, D, TD>(a.in, cN[i * 8 + D], a.M, m, v);                                                \
          dev::leg_apply<DAG, D, TD, FIRST, CH>(v, cU + (i * 8 + D) * 18, rr, ri);                                     \
        }                                                                                                              \
      }
      GBX_ALEG(0, true)
      GBX_ALEG(1, false)
      GBX_ALEG(2, false)
      GBX_ALEG(3, false)
      GBX_ALEG(4, false)
      GBX_ALEG(5, false)
      GBX_ALEG(6, false)
      GBX_ALEG(7, false)
#undef GBX_ALEG
      dev::store_site<P, false>(a.out, nullptr, cQ[i], a.M, m, rr, ri, (A)0, (A)1, a.nt != 0);
    }
    __builtin_amdgcn_s_waitcnt(0);  // the next unit's loads have landed (this thread's)
    __syncthreads();                // ... and every thread's; this buffer is free again
    if (!next) break;
  }
}

typedef void (*AsyncKernel)(AsyncArgs);

AsyncKernel async_kernel(int prec, bool dag, bool td)
{
  if (prec == GBX_F64) {
    if (td) return dag ? k_dhop_tile_async<GBX_F64, true, true> : k_dhop_tile_async<GBX_F64, false, true>;
    return dag ? k_dhop_tile_async<GBX_F64, true, false> : k_dhop_tile_async<GBX_F64, false, false>;
  }
  if (td) return dag ? k_dhop_tile_async<GBX_F32, true, true> : k_dhop_tile_async<GBX_F32, false, true>;
  return dag ? k_dhop_tile_async<GBX_F32, true, false> : k_dhop_tile_async<GBX_F32, false, false>;
}

size_t tile_async_lds(int prec, int T, uint32_t mc)
{
  const int nch = prec == GBX_F64 ? 12 : 6, lv = prec == GBX_F64 ? 72 : 36;
  const AsyncLayout l = async_layout(T, nch, lv, mc);
  return 2 * (size_t)(l.spin_pad + l.link_pad) * 16 + 2 * (size_t)T * 9 * 4 + (size_t)T * 8 * 4;
}

size_t tile_lds(int prec, int T, int mc)
{
  const size_t acc = prec == GBX_F64 ? 8 : 4;
  const size_t nch = prec == GBX_F64 ? 12 : prec == GBX_F32 ? 6 : 3;
  const size_t sw = nch == 12 ? 13 : nch;
  return (size_t)T * sw * mc * 16 + (size_t)T * 8 * 18 * acc + (size_t)T * 8 * 4 + (size_t)T * 4 + (size_t)T * 8 * 4;
}

}  // namespace

void tile_tables_free(gbx_geom_s *g)
{
  for (auto &kv : g->tile_tabs) (void)hipFree(kv.second);
  g->tile_tabs.clear();
}

int pack_tile(const int e[4]) { return e[0] | (e[1] << 4) | (e[2] << 8) | (e[3] << 12); }

// Configuration check for kind 2: the tile divides the lattice, its LDS fits.
int tile_cfg_check(const gbx_kcfg &k, const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  const gbx_geom_s *g = u->geom;
  GBX_CHECK(whole_lattice(g), GBX_EUNSUPPORTED, "tile kernel: whole lattices only");
  const TileShape s = unpack_tile(k.tile);
  for (int d = 0; d < 4; d++)
    GBX_CHECK(s.e[d] >= 1 && g->dims[d] % s.e[d] == 0, GBX_EINVAL, "tile kernel: tile extent %d = %d does not divide %d", d,
              s.e[d], g->dims[d]);
  const uint32_t mc = (k.mc <= 0 || (uint32_t)k.mc > in->M) ? in->M : (uint32_t)k.mc;
  GBX_CHECK(k.pf == 0 || k.pf == 1, GBX_EINVAL, "tile kernel: pf=%d (0: one-shot, 1: persistent with async staging)",
            k.pf);
  GBX_CHECK(k.wgpc >= 0 && k.wgpc <= 64, GBX_EINVAL, "tile kernel: wgpc=%d", k.wgpc);
  if (k.pf == 1) {
    GBX_CHECK((u->prec == GBX_F64 || u->prec == GBX_F32) && !u->comp12, GBX_EUNSUPPORTED,
              "tile kernel (async staging): FP64 / FP32 fields with 18-real links");
    GBX_CHECK(in->M % mc == 0, GBX_EINVAL, "tile kernel (async staging): mc=%u must divide Ls*nrhs=%u", mc, in->M);
    GBX_CHECK(k.threads >= 64 && k.threads <= async_max_threads(u->prec) && k.threads % 64 == 0, GBX_EINVAL,
              "tile kernel (async staging): threads=%d (max %d)", k.threads, async_max_threads(u->prec));
    const size_t lds = tile_async_lds(u->prec, s.T(), mc);
    GBX_CHECK(lds <= 160 * 1024, GBX_EINVAL, "tile kernel (async staging): tile of %d sites x %u minor needs %zu B of LDS",
              s.T(), mc, lds);
    return GBX_OK;
  }
  const int tmax = tile_max_threads(u->prec, in->basis == GBX_BASIS_TDIAG);  // the kernel's launch bound
  GBX_CHECK(k.threads >= 64 && k.threads <= tmax && k.threads % 64 == 0, GBX_EINVAL, "tile kernel: threads=%d (max %d)",
            k.threads, tmax);
  const size_t lds = tile_lds(u->prec, s.T(), (int)mc);
  GBX_CHECK(lds <= 160 * 1024, GBX_EINVAL, "tile kernel: tile of %d sites x %u minor needs %zu B of LDS", s.T(), mc, lds);
  return GBX_OK;
}

// The persistent form: as many workgroups per CU as can be resident at once
// (registers, LDS; at most wgpc), so that the static split of the units
// leaves no workgroup waiting for a slot; XCD x's workgroups share its eighth
// of the units.
static int tile_async_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k,
                             hipStream_t st)
{
  gbx_geom_s *g = u->geom;
  const TileShape s = unpack_tile(k.tile);
  AsyncArgs a;
  GBX_TRY(tile_table(g, k.tile, &a.tslots, &a.ntiles));
  a.gauge = u->d;
  a.nbr = g->nbr;
  a.in = in->d;
  a.out = out->d;
  a.M = in->M;
  a.mc = (k.mc <= 0 || (uint32_t)k.mc > in->M) ? in->M : (uint32_t)k.mc;
  const uint64_t nunits = (uint64_t)a.ntiles * (a.M / a.mc);
  GBX_CHECK(nunits < (1ull << 31), GBX_EINVAL, "tile kernel: %llu units", (unsigned long long)nunits);
  a.nunits = (uint32_t)nunits;
  a.units_per_xcd = (a.nunits + 7) / 8;
  const bool td = in->basis == GBX_BASIS_TDIAG;
  const AsyncKernel kern = async_kernel(u->prec, dag != 0, td);
  const size_t lds = tile_async_lds(u->prec, s.T(), a.mc);
  int dev = 0, ncu = 0, fit = 0;
  GBX_HIP(hipGetDevice(&dev));
  GBX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  GBX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, kern, k.threads, lds));
  GBX_CHECK(fit >= 1 && ncu >= 8, GBX_EINVAL, "tile kernel (async staging): %d resident workgroups per CU, %d CUs", fit,
            ncu);
  const int per_cu = k.wgpc > 0 ? std::min(k.wgpc, fit) : fit;
  a.wg_per_xcd = (uint32_t)(ncu / 8 * per_cu);
  a.ex = s.e[0];
  a.ey = s.e[1];
  a.ez = s.e[2];
  a.et = s.e[3];
  a.wrap = 0;
  for (int d = 0; d < 4; d++)
    if (s.e[d] == g->dims[d]) a.wrap |= 1 << d;
  a.nt = k.nt;
  const dim3 grid(8 * a.wg_per_xcd), block((unsigned)k.threads);
  hipLaunchKernelGGL(kern, grid, block, lds, st, a);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

int tile_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k, hipStream_t st)
{
  GBX_TRY(tile_cfg_check(k, u, in));
  out->basis = in->basis;
  if (k.pf == 1) return tile_async_launch(u, out, in, dag, k, st);
  gbx_geom_s *g = u->geom;
  const TileShape s = unpack_tile(k.tile);
  TileArgs a;
  GBX_TRY(tile_table(g, k.tile, &a.tslots, &a.ntiles));
  a.gauge = u->d;
  a.nbr = g->nbr;
  a.in = in->d;
  a.out = out->d;
  a.M = in->M;
  a.mc = (k.mc <= 0 || (uint32_t)k.mc > in->M) ? in->M : (uint32_t)k.mc;
  const uint64_t nwg = (uint64_t)a.ntiles * ((a.M + a.mc - 1) / a.mc);
  GBX_CHECK(nwg < (1ull << 31), GBX_EINVAL, "tile kernel: %llu workgroups", (unsigned long long)nwg);
  a.nwg = (uint32_t)nwg;
  a.per_xcd = (a.nwg + 7) / 8;
  a.ex = s.e[0];
  a.ey = s.e[1];
  a.ez = s.e[2];
  a.et = s.e[3];
  a.wrap = 0;
  for (int d = 0; d < 4; d++)
    if (s.e[d] == g->dims[d]) a.wrap |= 1 << d;
  a.xcd = k.xcd;
  a.nt = k.nt;
  size_t lds = tile_lds(u->prec, s.T(), (int)a.mc);
  if (k.wgpc > 0) {
    const size_t pad = ((size_t)160 * 1024 / (size_t)k.wgpc) & ~(size_t)511;
    if (pad > lds) lds = pad;
  }
  const dim3 grid(k.xcd ? a.per_xcd * 8 : a.nwg), block((unsigned)k.threads);
  if (in->basis == GBX_BASIS_TDIAG) {  // FP64 / FP32 only (gbx_spinor_set_basis)
    if (u->prec == GBX_F64) launch_p<GBX_F64, true>(dag, u->comp12, grid, block, lds, st, a);
    else launch_p<GBX_F32, true>(dag, u->comp12, grid, block, lds, st, a);
    GBX_HIP(hipGetLastError());
    return GBX_OK;
  }
  switch (u->prec) {
  case GBX_F64: launch_p<GBX_F64>(dag, u->comp12, grid, block, lds, st, a); break;
  case GBX_F32: launch_p<GBX_F32>(dag, u->comp12, grid, block, lds, st, a); break;
  case GBX_F16: launch_p<GBX_F16>(dag, u->comp12, grid, block, lds, st, a); break;
  default: launch_p<GBX_BF16>(dag, u->comp12, grid, block, lds, st, a); break;
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}
