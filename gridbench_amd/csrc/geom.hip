// Geometry (device site order + neighbour table) and the layout kernels that
// move fields between host layouts and the device SoA layout.
//
// Device site order: the lattice is cut into blocks (default Lx x 4 x 4 x 4,
// each extent the largest divisor not above the target); slots run block by
// block (blocks lexicographic, x fastest) and lexicographically inside a
// block.  Consecutive slots are therefore spatially compact, so the 8 stencil
// neighbours of a workgroup's sites were touched recently and are still in
// L2 / the memory-side cache.  nbr[slot][mu] holds neighbour slots in the
// reference's direction order Xp Yp Zp Tp Xm Ym Zm Tm (WilsonKernelsHand.h:5-12).
#include <algorithm>
#include <mutex>

#include "gbx_internal.h"
#include "layout_dev.h"

namespace {

std::mutex g_staging_mu;
void *g_staging = nullptr;
size_t g_staging_bytes = 0;

int largest_divisor_le(int n, int cap)
{
  for (int b = cap < n ? cap : n; b > 1; b--)
    if (n % b == 0) return b;
  return 1;
}

uint32_t nblocks(uint64_t n, int threads, uint32_t cap = 1u << 20)
{
  uint64_t b = (n + (uint64_t)threads - 1) / (uint64_t)threads;
  if (b > cap) b = cap;
  if (b == 0) b = 1;
  return (uint32_t)b;
}

// ------------------------------------------------------------------ kernels
// Lexicographic FP64 [nr][V][Ls][24] <-> minor indices m0 .. m0+Mr-1 of the
// device field (Mr = nr*Ls: right-hand sides m0/Ls .. m0/Ls+nr-1).
template <int P>
__global__ void k_pack_lex(const double *__restrict__ lex, void *__restrict__ dst, const uint32_t *__restrict__ site_of_slot,
                           uint64_t total, uint32_t V, uint32_t Ls, uint32_t M, uint32_t m0, uint32_t Mr)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t mm = (uint32_t)(t % Mr);
    const uint64_t qk = t / Mr;
    const uint32_t k = (uint32_t)(qk % NCH);
    const uint32_t q = (uint32_t)(qk / NCH);
    const uint32_t g = site_of_slot[q];
    const uint32_t r = mm / Ls, s = mm % Ls;
    const double *src = lex + (((uint64_t)r * V + g) * Ls + s) * 24 + (uint64_t)k * R;
    double v[R];
#pragma unroll
    for (int j = 0; j < R; j++) v[j] = src[j];
    dev::store_chunk<P>(dst, ((uint64_t)q * NCH + k) * M + m0 + mm, v, false);
  }
}

template <int P>
__global__ void k_unpack_lex(const void *__restrict__ srcd, double *__restrict__ lex, const uint32_t *__restrict__ site_of_slot,
                             uint64_t total, uint32_t V, uint32_t Ls, uint32_t M, uint32_t m0, uint32_t Mr)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t mm = (uint32_t)(t % Mr);
    const uint64_t qk = t / Mr;
    const uint32_t k = (uint32_t)(qk % NCH);
    const uint32_t q = (uint32_t)(qk / NCH);
    const uint32_t g = site_of_slot[q];
    const uint32_t r = mm / Ls, s = mm % Ls;
    double v[R];
    dev::load_chunk_f64<P>(srcd, ((uint64_t)q * NCH + k) * M + m0 + mm, v);
    double *o = lex + (((uint64_t)r * V + g) * Ls + s) * 24 + (uint64_t)k * R;
#pragma unroll
    for (int j = 0; j < R; j++) o[j] = v[j];
  }
}

// Reference virtual-node host layout [nsite*Ls][12][width][2] <-> device.
// The lattice fills nsimd lanes (the data SIMD width); a vector holds `width`
// >= nsimd lanes, of which lanes lane0 .. lane0+nsimd-1 are this lattice's
// (bench.cc:98-138 expands DATA_SIMD lanes to EXPAND_SIMD this way, replica
// nn at lanes nn*DATA_SIMD + n).
struct RefGeom {
  int dims[4], simd[4], rd[4];
  int nsimd;
  int width, lane0;
  int rrii;  // 0: RIRI [..][lane][re,im]; 1: RRII [..][re,im][lane] (bench.cc:105-109)
};

// Position of real `ri` of data lane `lane` of complex element `c`.
__device__ __forceinline__ uint64_t ref_real(uint64_t c, uint32_t lane, uint32_t ri, const RefGeom &rg)
{
  const uint64_t w = (uint64_t)rg.width, l = (uint64_t)rg.lane0 + lane;
  return c * 2 * w + (rg.rrii ? (uint64_t)ri * w + l : l * 2 + ri);
}

__device__ inline void ref_coords(const RefGeom &rg, uint32_t g, uint64_t &o, uint32_t &lane)
{
  uint32_t x[4];
  uint32_t r = g;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    x[d] = r % (uint32_t)rg.dims[d];
    r /= (uint32_t)rg.dims[d];
  }
  o = 0;
  lane = 0;
#pragma unroll
  for (int d = 3; d >= 0; d--) {
    o = o * (uint64_t)rg.rd[d] + x[d] % (uint32_t)rg.rd[d];
    lane = lane * (uint32_t)rg.simd[d] + x[d] / (uint32_t)rg.rd[d];
  }
}

template <int P, typename T>
__global__ void k_pack_ref(const T *__restrict__ ref, void *__restrict__ dst, const uint32_t *__restrict__ site_of_slot,
                           uint64_t total, RefGeom rg, uint32_t Ls)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)(t % Ls);
    const uint64_t qk = t / Ls;
    const uint32_t k = (uint32_t)(qk % NCH);
    const uint32_t q = (uint32_t)(qk / NCH);
    uint64_t o;
    uint32_t lane;
    ref_coords(rg, site_of_slot[q], o, lane);
    double v[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
      const uint32_t ridx = k * R + j;
      v[j] = (double)ref[ref_real((o * Ls + s) * 12 + (ridx >> 1), lane, ridx & 1, rg)];
    }
    dev::store_chunk<P>(dst, t, v, false);
  }
}

template <int P, typename T>
__global__ void k_unpack_ref(const void *__restrict__ srcd, T *__restrict__ ref, const uint32_t *__restrict__ site_of_slot,
                             uint64_t total, RefGeom rg, uint32_t Ls)
{
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)(t % Ls);
    const uint64_t qk = t / Ls;
    const uint32_t k = (uint32_t)(qk % NCH);
    const uint32_t q = (uint32_t)(qk / NCH);
    uint64_t o;
    uint32_t lane;
    ref_coords(rg, site_of_slot[q], o, lane);
    double v[R];
    dev::load_chunk_f64<P>(srcd, t, v);
#pragma unroll
    for (int j = 0; j < R; j++) {
      const uint32_t ridx = k * R + j;
      ref[ref_real((o * Ls + s) * 12 + (ridx >> 1), lane, ridx & 1, rg)] = (T)v[j];
    }
  }
}

// Gauge: lexicographic doubled FP64 [V][8][18] or reference layout
// [nsite][8][9][nsimd][2] -> device [slot][8][Rl] (Rl = 18 or 12 reals).
template <int P>
__global__ void k_pack_gauge_lex(const double *__restrict__ lex, void *__restrict__ dst, const uint32_t *__restrict__ site_of_slot,
                                 uint64_t total, int Rl)
{
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = (uint32_t)(t % Rl);
    const uint64_t qd = t / Rl;
    const uint32_t dir = (uint32_t)(qd % 8);
    const uint32_t q = (uint32_t)(qd / 8);
    const uint32_t g = site_of_slot[q];
    dev::store_real<P>(dst, t, lex[((uint64_t)g * 8 + dir) * 18 + e]);
  }
}

template <int P, typename T>
__global__ void k_pack_gauge_ref(const T *__restrict__ ref, void *__restrict__ dst, const uint32_t *__restrict__ site_of_slot,
                                 uint64_t total, int Rl, RefGeom rg)
{
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = (uint32_t)(t % Rl);
    const uint64_t qd = t / Rl;
    const uint32_t dir = (uint32_t)(qd % 8);
    const uint32_t q = (uint32_t)(qd / 8);
    uint64_t o;
    uint32_t lane;
    ref_coords(rg, site_of_slot[q], o, lane);
    dev::store_real<P>(dst, t, (double)ref[ref_real((o * 8 + dir) * 9 + (e >> 1), lane, e & 1, rg)]);
  }
}

// ------------------------------------------------------------- host helpers
int ref_geom(const gbx_geom_s *g, int nsimd, RefGeom &rg, const gbx_ref_layout &lay = gbx_ref_layout())
{
  rg.rrii = lay.rrii;
  rg.width = lay.width > 0 ? lay.width : nsimd;
  rg.lane0 = lay.lane0;
  GBX_CHECK(rg.width >= nsimd && rg.lane0 >= 0 && rg.lane0 + nsimd <= rg.width, GBX_EINVAL,
            "reference layout: lanes %d..%d outside a %d-lane vector", rg.lane0, rg.lane0 + nsimd - 1, rg.width);
  GBX_CHECK(whole_lattice(g), GBX_EUNSUPPORTED,
            "the reference's virtual-node layout describes a whole lattice, not a red-black half or a domain part");
  rg.nsimd = nsimd;
  for (int d = 0; d < 4; d++) rg.simd[d] = 1;
  GBX_CHECK(nsimd == 1 || nsimd == 2 || nsimd == 4 || nsimd == 8 || nsimd == 16, GBX_EINVAL,
            "nsimd=%d not in {1,2,4,8,16}", nsimd);
  int d = 3;
  for (int n = nsimd; n > 1; n >>= 1) {
    rg.simd[d] *= 2;
    d = (d + 3) % 4;
  }
  for (int e = 0; e < 4; e++) {
    rg.dims[e] = g->dims[e];
    GBX_CHECK(g->dims[e] % rg.simd[e] == 0, GBX_EINVAL, "dim %d = %d not divisible by %d lanes", e, g->dims[e], rg.simd[e]);
    rg.rd[e] = g->dims[e] / rg.simd[e];
  }
  return GBX_OK;
}

int spinor_ok(const gbx_spinor_s *x)
{
  GBX_CHECK(x && x->d, GBX_ENULL, "spinor handle is NULL");
  return GBX_OK;
}

template <template <int> class F, typename... Args>
void dispatch_prec(int p, Args... a)
{
  switch (p) {
  case GBX_F64: F<GBX_F64>::run(a...); break;
  case GBX_F32: F<GBX_F32>::run(a...); break;
  case GBX_F16: F<GBX_F16>::run(a...); break;
  default: F<GBX_BF16>::run(a...); break;
  }
}

template <int P> struct PackLex {
  static void run(const double *lex, void *dst, const uint32_t *sos, uint64_t total, uint32_t V, uint32_t Ls, uint32_t M,
                  uint32_t m0, uint32_t Mr, hipStream_t st)
  {
    hipLaunchKernelGGL(k_pack_lex<P>, dim3(nblocks(total, 256)), dim3(256), 0, st, lex, dst, sos, total, V, Ls, M, m0, Mr);
  }
};
template <int P> struct UnpackLex {
  static void run(const void *src, double *lex, const uint32_t *sos, uint64_t total, uint32_t V, uint32_t Ls, uint32_t M,
                  uint32_t m0, uint32_t Mr, hipStream_t st)
  {
    hipLaunchKernelGGL(k_unpack_lex<P>, dim3(nblocks(total, 256)), dim3(256), 0, st, src, lex, sos, total, V, Ls, M, m0, Mr);
  }
};
template <int P> struct PackGaugeLex {
  static void run(const double *lex, void *dst, const uint32_t *sos, uint64_t total, int Rl, hipStream_t st)
  {
    hipLaunchKernelGGL(k_pack_gauge_lex<P>, dim3(nblocks(total, 256)), dim3(256), 0, st, lex, dst, sos, total, Rl);
  }
};

template <typename T> struct PackRefT {
  template <int P> struct F {
    static void run(const T *ref, void *dst, const uint32_t *sos, uint64_t total, RefGeom rg, uint32_t Ls, hipStream_t st)
    {
      hipLaunchKernelGGL((k_pack_ref<P, T>), dim3(nblocks(total, 256)), dim3(256), 0, st, ref, dst, sos, total, rg, Ls);
    }
  };
};
template <typename T> struct UnpackRefT {
  template <int P> struct F {
    static void run(const void *src, T *ref, const uint32_t *sos, uint64_t total, RefGeom rg, uint32_t Ls, hipStream_t st)
    {
      hipLaunchKernelGGL((k_unpack_ref<P, T>), dim3(nblocks(total, 256)), dim3(256), 0, st, src, ref, sos, total, rg, Ls);
    }
  };
};
template <typename T> struct PackGaugeRefT {
  template <int P> struct F {
    static void run(const T *ref, void *dst, const uint32_t *sos, uint64_t total, int Rl, RefGeom rg, hipStream_t st)
    {
      hipLaunchKernelGGL((k_pack_gauge_ref<P, T>), dim3(nblocks(total, 256)), dim3(256), 0, st, ref, dst, sos, total, Rl, rg);
    }
  };
};

}  // namespace

// ------------------------------------------------------------ staging buffer
int staging_buffer(size_t bytes, void **ptr)
{
  // caller holds g_staging_mu via staging_lock()
  if (bytes > g_staging_bytes) {
    if (g_staging) GBX_HIP(hipFree(g_staging));
    g_staging = nullptr;
    g_staging_bytes = 0;
    GBX_HIP(hipMalloc(&g_staging, bytes));
    g_staging_bytes = bytes;
  }
  *ptr = g_staging;
  return GBX_OK;
}

// ------------------------------------------------------------------- geometry
void geom_retain(gbx_geom_s *g) { g->refcount++; }

void geom_release(gbx_geom_s *g)
{
  if (--g->refcount > 0) return;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  (void)hipFree(g->site_of_slot);
  (void)hipFree(g->slot_of_site);
  (void)hipFree(g->nbr);
  (void)hipFree(g->gsite_of_slot);
  (void)hipFree(g->face_slots);
  tile_tables_free(g);
  gbx_geom_s *owner = g->stream_owner;
  if (g->stream && !owner) (void)hipStreamDestroy(g->stream);
  delete g;
  if (owner) geom_release(owner);
}

extern "C" int gbx_geom_create(const int dims[4], const int block[4], gbx_geom *out)
{
  GBX_CHECK(dims && out, GBX_ENULL, "gbx_geom_create: NULL argument");
  // Default: compact 4D blocks of extent <= 6 per dimension (about 1300
  // sites: the neighbour reuse distance of one block fits an XCD's L2).
  int blk[4];
  for (int d = 0; d < 4; d++) blk[d] = block ? block[d] : largest_divisor_le(dims[d] > 0 ? dims[d] : 1, 6);
  return gbx_geom_create_levels(dims, 1, blk, out);
}

namespace {

int check_dims_levels(const int dims[4], int nlevels, const int *blocks, uint64_t *vol_out)
{
  GBX_CHECK(nlevels >= 1 && nlevels <= GBX_MAX_LEVELS, GBX_EINVAL, "gbx_geom_create_levels: nlevels=%d", nlevels);
  uint64_t vol = 1;
  for (int d = 0; d < 4; d++) {
    GBX_CHECK(dims[d] >= 1 && dims[d] <= 1024, GBX_EINVAL, "gbx_geom_create: dims[%d]=%d", d, dims[d]);
    vol *= (uint64_t)dims[d];
  }
  GBX_CHECK(vol <= (1ull << 30), GBX_EINVAL, "gbx_geom_create: volume %llu too large", (unsigned long long)vol);
  for (int l = 0; l < nlevels; l++)
    for (int d = 0; d < 4; d++) {
      const int b = blocks[l * 4 + d], outer = l + 1 < nlevels ? blocks[(l + 1) * 4 + d] : dims[d];
      GBX_CHECK(b >= 1 && outer % b == 0, GBX_EINVAL, "gbx_geom_create: level %d block[%d]=%d does not divide %d", l, d, b,
                outer);
    }
  *vol_out = vol;
  return GBX_OK;
}

// Slot order: lexicographic over the outermost blocks, then, inside each
// block of level l, lexicographic over its level l-1 sub-blocks, down to
// lexicographic order of the sites of a level-0 block.  Returns the
// full-lattice slot of every lexicographic site.
std::vector<uint32_t> blocked_slots(const int dims[4], int nlevels, const int *blocks, uint64_t vol)
{
  std::vector<uint32_t> slot_of_site(vol);
  for (uint64_t site = 0; site < vol; site++) {
    int x[4];
    uint64_t r = site;
    for (int d = 0; d < 4; d++) {
      x[d] = (int)(r % (uint64_t)dims[d]);
      r /= (uint64_t)dims[d];
    }
    // outermost grid of top-level blocks
    const int *top = blocks + (nlevels - 1) * 4;
    uint64_t slot = 0;
    for (int d = 3; d >= 0; d--) slot = slot * (uint64_t)(dims[d] / top[d]) + (uint64_t)(x[d] / top[d]);
    for (int l = nlevels - 1; l >= 1; l--) {
      const int *bo = blocks + l * 4, *bi = blocks + (l - 1) * 4;
      uint64_t w = 0, n = 1;
      for (int d = 3; d >= 0; d--) {
        w = w * (uint64_t)(bo[d] / bi[d]) + (uint64_t)((x[d] % bo[d]) / bi[d]);
        n *= (uint64_t)(bo[d] / bi[d]);
      }
      slot = slot * n + w;
    }
    uint64_t w = 0, n = 1;
    for (int d = 3; d >= 0; d--) {
      w = w * (uint64_t)blocks[d] + (uint64_t)(x[d] % blocks[d]);
      n *= (uint64_t)blocks[d];
    }
    slot_of_site[site] = (uint32_t)(slot * n + w);
  }
  return slot_of_site;
}

// Lexicographic site of the neighbour of `site` in direction mu (periodic).
uint64_t nbr_site(uint64_t site, int mu, const int dims[4])
{
  const uint64_t stride[4] = {1, (uint64_t)dims[0], (uint64_t)dims[0] * dims[1], (uint64_t)dims[0] * dims[1] * dims[2]};
  const int d = mu & 3;
  const int x = (int)((site / stride[d]) % (uint64_t)dims[d]);
  if (mu < 4) return (x + 1 == dims[d]) ? site - (uint64_t)x * stride[d] : site + stride[d];
  return (x == 0) ? site + (uint64_t)(dims[d] - 1) * stride[d] : site - stride[d];
}

// Checkerboarded lexicographic index of a full-lattice site: (x/2) fastest.
uint32_t half_index(uint64_t site, const int dims[4])
{
  const uint64_t x0 = site % (uint64_t)dims[0];
  return (uint32_t)((site / (uint64_t)dims[0]) * (uint64_t)(dims[0] / 2) + x0 / 2);
}

int site_parity(uint64_t site, const int dims[4])
{
  int p = 0;
  for (int d = 0; d < 4; d++) {
    p += (int)(site % (uint64_t)dims[d]);
    site /= (uint64_t)dims[d];
  }
  return p & 1;
}

gbx_geom_s *geom_new(const int dims[4], int nlevels, const int *blocks, uint64_t vol, int device)
{
  gbx_geom_s *g = new gbx_geom_s();
  g->nlevels = nlevels;
  for (int d = 0; d < 4; d++) {
    g->dims[d] = dims[d];
    g->block[d] = blocks[d];
  }
  for (int l = 0; l < nlevels * 4; l++) g->levels[l] = blocks[l];
  g->vol = vol;
  g->gauge_vol = vol;
  g->device = device;
  return g;
}

// Stream (its own, or that of stream_owner) and device copies of the
// tables; releases g on failure.
int geom_finish(gbx_geom_s *g, const std::vector<uint32_t> &nbr, const std::vector<uint32_t> *gsite)
{
  const uint64_t vol = g->vol;
  auto fail = [&](int s) {
    geom_release(g);
    return s;
  };
  if (g->stream_owner) g->stream = g->stream_owner->stream;
  else if (hipStreamCreate(&g->stream) != hipSuccess) return fail(gbx_fail(GBX_EHIP, "hipStreamCreate failed"));
  if (hipMalloc(&g->site_of_slot, vol * 4) != hipSuccess || hipMalloc(&g->slot_of_site, vol * 4) != hipSuccess ||
      hipMalloc(&g->nbr, vol * 32) != hipSuccess || (gsite && hipMalloc(&g->gsite_of_slot, vol * 4) != hipSuccess))
    return fail(gbx_fail(GBX_ENOMEM, "gbx_geom_create: device allocation failed"));
  if (hipMemcpy(g->site_of_slot, g->site_of_slot_h.data(), vol * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->slot_of_site, g->slot_of_site_h.data(), vol * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(g->nbr, nbr.data(), vol * 32, hipMemcpyHostToDevice) != hipSuccess ||
      (gsite && hipMemcpy(g->gsite_of_slot, gsite->data(), vol * 4, hipMemcpyHostToDevice) != hipSuccess))
    return fail(gbx_fail(GBX_EHIP, "gbx_geom_create: upload failed"));
  return GBX_OK;
}

uint64_t g_next_pair_id = 0;
std::mutex g_pair_mu;

}  // namespace

extern "C" int gbx_geom_create_levels(const int dims[4], int nlevels, const int *blocks, gbx_geom *out)
{
  GBX_CHECK(dims && out && blocks, GBX_ENULL, "gbx_geom_create_levels: NULL argument");
  *out = nullptr;
  uint64_t vol;
  GBX_TRY(check_dims_levels(dims, nlevels, blocks, &vol));
  GBX_TRY(gbx_require_device());
  int device;
  GBX_HIP(hipGetDevice(&device));
  gbx_geom_s *g = geom_new(dims, nlevels, blocks, vol, device);
  g->slot_of_site_h = blocked_slots(dims, nlevels, blocks, vol);
  g->site_of_slot_h.resize(vol);
  for (uint64_t site = 0; site < vol; site++) g->site_of_slot_h[g->slot_of_site_h[site]] = (uint32_t)site;
  std::vector<uint32_t> nbr(vol * 8);
  for (uint64_t q = 0; q < vol; q++)
    for (int mu = 0; mu < 8; mu++) nbr[q * 8 + mu] = g->slot_of_site_h[nbr_site(g->site_of_slot_h[q], mu, dims)];
  GBX_TRY(geom_finish(g, nbr, nullptr));
  *out = g;
  return GBX_OK;
}

// A caller-defined site order: slot_of_site[site] for every lexicographic
// site, a permutation of 0 .. V-1.  The kernels visit slots in order (and,
// with the XCD-contiguous launch, give each XCD one eighth of the slots), so
// the order decides which neighbours are still in an L2 when they are read.
extern "C" int gbx_geom_create_order(const int dims[4], const uint32_t *slot_of_site, gbx_geom *out)
{
  GBX_CHECK(dims && slot_of_site && out, GBX_ENULL, "gbx_geom_create_order: NULL argument");
  *out = nullptr;
  const int one[4] = {1, 1, 1, 1};
  uint64_t vol;
  GBX_TRY(check_dims_levels(dims, 1, one, &vol));
  std::vector<uint32_t> sos(slot_of_site, slot_of_site + vol), site_of(vol, UINT32_MAX);
  for (uint64_t site = 0; site < vol; site++) {
    GBX_CHECK(sos[site] < vol && site_of[sos[site]] == UINT32_MAX, GBX_EINVAL,
              "gbx_geom_create_order: slot_of_site is not a permutation (site %llu)", (unsigned long long)site);
    site_of[sos[site]] = (uint32_t)site;
  }
  GBX_TRY(gbx_require_device());
  int device;
  GBX_HIP(hipGetDevice(&device));
  gbx_geom_s *g = geom_new(dims, 0, one, vol, device);
  g->slot_of_site_h = std::move(sos);
  g->site_of_slot_h = std::move(site_of);
  std::vector<uint32_t> nbr(vol * 8);
  for (uint64_t q = 0; q < vol; q++)
    for (int mu = 0; mu < 8; mu++) nbr[q * 8 + mu] = g->slot_of_site_h[nbr_site(g->site_of_slot_h[q], mu, dims)];
  GBX_TRY(geom_finish(g, nbr, nullptr));  // releases g on failure
  *out = g;
  return GBX_OK;
}

// Red-black pair: the even and odd sites of the lattice, each in the blocked
// order of the full lattice restricted to its parity (so a half's
// consecutive slots stay spatially compact), each half's neighbour table
// pointing into the other half.
extern "C" int gbx_geom_create_rb(const int dims[4], const int block[4], gbx_geom *even, gbx_geom *odd)
{
  GBX_CHECK(dims && even && odd, GBX_ENULL, "gbx_geom_create_rb: NULL argument");
  *even = *odd = nullptr;
  for (int d = 0; d < 4; d++)
    GBX_CHECK(dims[d] >= 2 && dims[d] % 2 == 0, GBX_EINVAL, "gbx_geom_create_rb: dims[%d]=%d must be even", d, dims[d]);
  int blk[4];
  for (int d = 0; d < 4; d++) blk[d] = block ? block[d] : largest_divisor_le(dims[d], 6);
  uint64_t vol;
  GBX_TRY(check_dims_levels(dims, 1, blk, &vol));
  GBX_TRY(gbx_require_device());
  int device;
  GBX_HIP(hipGetDevice(&device));
  const std::vector<uint32_t> full_slot = blocked_slots(dims, 1, blk, vol);
  std::vector<uint32_t> full_site(vol);
  for (uint64_t site = 0; site < vol; site++) full_site[full_slot[site]] = (uint32_t)site;
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(g_pair_mu);
    id = ++g_next_pair_id;
  }
  const uint64_t half = vol / 2;
  gbx_geom_s *h[2];
  std::vector<uint32_t> gsite[2];
  for (int p = 0; p < 2; p++) {
    h[p] = geom_new(dims, 1, blk, half, device);
    h[p]->gauge_vol = vol;
    h[p]->parity = p;
    h[p]->pair_id = id;
    h[p]->site_of_slot_h.reserve(half);
    h[p]->slot_of_site_h.assign(half, 0);
    gsite[p].reserve(half);
  }
  for (uint64_t q = 0; q < vol; q++) {
    const uint64_t site = full_site[q];
    const int p = site_parity(site, dims);
    const uint32_t hi = half_index(site, dims);
    h[p]->slot_of_site_h[hi] = (uint32_t)h[p]->site_of_slot_h.size();
    h[p]->site_of_slot_h.push_back(hi);
    gsite[p].push_back((uint32_t)site);
  }
  std::vector<uint32_t> nbr[2];
  for (int p = 0; p < 2; p++) {
    nbr[p].resize(half * 8);
    for (uint64_t q = 0; q < half; q++)
      for (int mu = 0; mu < 8; mu++)
        nbr[p][q * 8 + mu] = h[1 - p]->slot_of_site_h[half_index(nbr_site(gsite[p][q], mu, dims), dims)];
  }
  int rc = geom_finish(h[0], nbr[0], &gsite[0]);  // releases h[0] on failure
  if (rc != GBX_OK) {
    geom_release(h[1]);
    return rc;
  }
  h[1]->stream_owner = h[0];
  geom_retain(h[0]);
  rc = geom_finish(h[1], nbr[1], &gsite[1]);  // releases h[1] (and its hold on h[0]) on failure
  if (rc != GBX_OK) {
    geom_release(h[0]);
    return rc;
  }
  *even = h[0];
  *odd = h[1];
  return GBX_OK;
}

// The block at `coord` of the lattice cut into grid[0] x .. x grid[3] equal
// blocks, with ghost slots for its cut faces (the local share of a domain-
// decomposed lattice; the faces are exchanged with gbx_halo_pack / _unpack).
namespace {

struct PartTables {
  int ld[4], blk[4];
  uint64_t vol = 0, gvol = 0, ninterior = 0, nghost = 0;
  uint64_t face[4] = {0, 0, 0, 0}, ghost_base[4] = {0, 0, 0, 0}, fs_base[4] = {0, 0, 0, 0};
  std::vector<uint32_t> slot_of_site, site_of_slot, gsite, nbr, fslots;
};

int part_tables(const int dims[4], const int grid[4], const int coord[4], const int block[4], PartTables &t)
{
  GBX_CHECK(dims && grid && coord, GBX_ENULL, "domain part: NULL dims / grid / coord");
  int ncut = 0;
  for (int d = 0; d < 4; d++) {
    GBX_CHECK(dims[d] >= 1 && dims[d] <= 1024, GBX_EINVAL, "domain part: dims[%d]=%d", d, dims[d]);
    GBX_CHECK(grid[d] >= 1 && dims[d] % grid[d] == 0, GBX_EINVAL, "domain part: grid[%d]=%d does not divide %d", d,
              grid[d], dims[d]);
    GBX_CHECK(coord[d] >= 0 && coord[d] < grid[d], GBX_EINVAL, "domain part: coord[%d]=%d of %d", d, coord[d], grid[d]);
    ncut += grid[d] > 1;
  }
  GBX_CHECK(ncut >= 1, GBX_EINVAL, "domain part: the grid cuts no dimension (use gbx_geom_create)");
  int *ld = t.ld, *blk = t.blk;
  for (int d = 0; d < 4; d++) ld[d] = dims[d] / grid[d];
  for (int d = 0; d < 4; d++) blk[d] = block ? block[d] : largest_divisor_le(ld[d], 6);
  uint64_t vol;
  GBX_TRY(check_dims_levels(ld, 1, blk, &vol));
  t.vol = vol;
  t.gvol = 1;
  for (int d = 0; d < 4; d++) t.gvol *= (uint64_t)dims[d];
  GBX_CHECK(t.gvol < (1ull << 32), GBX_EINVAL, "domain part: global volume %llu", (unsigned long long)t.gvol);
  uint64_t gb = vol, fb = 0;
  for (int d = 0; d < 4; d++)
    if (grid[d] > 1) {
      t.face[d] = vol / (uint64_t)ld[d];
      t.ghost_base[d] = gb;
      t.fs_base[d] = fb;
      gb += 2 * t.face[d];
      fb += 2 * t.face[d];
    }
  t.nghost = gb - vol;
  GBX_CHECK(gb < (1ull << 32), GBX_EINVAL, "domain part: %llu slots", (unsigned long long)gb);
  auto coords = [&](uint64_t site, int c[4]) {
    for (int d = 0; d < 4; d++) {
      c[d] = (int)(site % (uint64_t)ld[d]);
      site /= (uint64_t)ld[d];
    }
  };
  // lexicographic index over the three dimensions other than d
  auto face_index = [&](const int c[4], int d) {
    uint64_t f = 0, stride = 1;
    for (int e = 0; e < 4; e++) {
      if (e == d) continue;
      f += (uint64_t)c[e] * stride;
      stride *= (uint64_t)ld[e];
    }
    return f;
  };
  auto on_face = [&](const int c[4]) {
    for (int d = 0; d < 4; d++)
      if (grid[d] > 1 && (c[d] == 0 || c[d] == ld[d] - 1)) return true;
    return false;
  };
  // Blocked order, then stably partitioned: the interior sites (no leg
  // crosses a cut) take slots 0 .. ninterior-1, the face sites the rest,
  // so the interior and the faces are each one slot range.
  const std::vector<uint32_t> bslot = blocked_slots(ld, 1, blk, vol);
  std::vector<uint32_t> border(vol);
  for (uint64_t site = 0; site < vol; site++) border[bslot[site]] = (uint32_t)site;
  t.slot_of_site.assign(vol, 0);
  t.site_of_slot.assign(vol, 0);
  uint64_t n = 0;
  for (int pass = 0; pass < 2; pass++)
    for (uint64_t b = 0; b < vol; b++) {
      const uint32_t site = border[b];
      int c[4];
      coords(site, c);
      if (on_face(c) != (pass == 1)) continue;
      t.slot_of_site[site] = (uint32_t)n;
      t.site_of_slot[n] = site;
      n++;
      if (pass == 0) t.ninterior = n;
    }
  // Neighbours: a leg that crosses the cut of dimension d reads ghost slot
  // ghost_base[d] + f (the previous part's top face) or ghost_base[d] +
  // face[d] + f (the next part's bottom face).
  t.gsite.assign(vol, 0);
  t.nbr.assign(vol * 8, 0);
  t.fslots.assign(fb, 0);
  for (uint64_t q = 0; q < vol; q++) {
    const uint64_t site = t.site_of_slot[q];
    int c[4];
    coords(site, c);
    uint64_t gs = 0, stride = 1;
    for (int d = 0; d < 4; d++) {
      gs += (uint64_t)(c[d] + coord[d] * ld[d]) * stride;
      stride *= (uint64_t)dims[d];
    }
    t.gsite[q] = (uint32_t)gs;
    for (int d = 0; d < 4; d++) {
      if (grid[d] == 1) continue;
      const uint64_t f = face_index(c, d);
      if (c[d] == 0) t.fslots[t.fs_base[d] + f] = (uint32_t)q;
      if (c[d] == ld[d] - 1) t.fslots[t.fs_base[d] + t.face[d] + f] = (uint32_t)q;
    }
    for (int mu = 0; mu < 8; mu++) {
      const int d = mu & 3;
      const bool cut = grid[d] > 1;
      if (cut && mu < 4 && c[d] == ld[d] - 1) t.nbr[q * 8 + mu] = (uint32_t)(t.ghost_base[d] + t.face[d] + face_index(c, d));
      else if (cut && mu >= 4 && c[d] == 0) t.nbr[q * 8 + mu] = (uint32_t)(t.ghost_base[d] + face_index(c, d));
      else t.nbr[q * 8 + mu] = t.slot_of_site[nbr_site(site, mu, ld)];
    }
  }
  return GBX_OK;
}

}  // namespace

// The tables of a domain part, on the host (no device needed): site_of_slot
// and gsite [V_local], nbr [V_local][8], face_slots: for each cut dimension d
// ascending, [2][face_d] with face_d = V_local / (dims[d] / grid[d]).
extern "C" int gbx_part_tables(const int dims[4], const int grid[4], const int coord[4], const int block[4],
                               uint32_t *site_of_slot, uint32_t *gsite, uint32_t *nbr, uint32_t *face_slots,
                               uint64_t *ninterior)
{
  GBX_CHECK(site_of_slot && gsite && nbr && face_slots && ninterior, GBX_ENULL, "gbx_part_tables: NULL argument");
  PartTables t;
  GBX_TRY(part_tables(dims, grid, coord, block, t));
  std::copy(t.site_of_slot.begin(), t.site_of_slot.end(), site_of_slot);
  std::copy(t.gsite.begin(), t.gsite.end(), gsite);
  std::copy(t.nbr.begin(), t.nbr.end(), nbr);
  std::copy(t.fslots.begin(), t.fslots.end(), face_slots);
  *ninterior = t.ninterior;
  return GBX_OK;
}

extern "C" int gbx_geom_create_part(const int dims[4], const int grid[4], const int coord[4], const int block[4],
                                    gbx_geom *out)
{
  GBX_CHECK(out, GBX_ENULL, "gbx_geom_create_part: NULL out");
  *out = nullptr;
  PartTables t;
  GBX_TRY(part_tables(dims, grid, coord, block, t));
  GBX_TRY(gbx_require_device());
  int device;
  GBX_HIP(hipGetDevice(&device));
  gbx_geom_s *g = geom_new(t.ld, 1, t.blk, t.vol, device);
  g->gauge_vol = t.gvol;
  for (int d = 0; d < 4; d++) {
    g->grid[d] = grid[d];
    g->coord[d] = coord[d];
    g->gdims[d] = dims[d];
    g->face[d] = t.face[d];
    g->ghost_base[d] = t.ghost_base[d];
    g->fs_base[d] = t.fs_base[d];
  }
  g->nghost = t.nghost;
  g->ninterior = t.ninterior;
  g->slot_of_site_h = std::move(t.slot_of_site);
  g->site_of_slot_h = std::move(t.site_of_slot);
  GBX_TRY(geom_finish(g, t.nbr, &t.gsite));  // releases g on failure
  if (hipMalloc(&g->face_slots, t.fslots.size() * 4) != hipSuccess ||
      hipMemcpy(g->face_slots, t.fslots.data(), t.fslots.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    geom_release(g);
    return gbx_fail(GBX_ENOMEM, "gbx_geom_create_part: face table");
  }
  *out = g;
  return GBX_OK;
}

// Red-black halves of a domain part: the sites of one global parity of the
// block.  Every local extent must be even, so a block's offsets are even and
// local parity is global parity.  A half's faces hold the face sites of its
// parity: face_d / 2 of them, index f / 2 for the face site f of the part
// (the fastest dimension across a face has even extent, so consecutive
// face indices alternate in parity).  Half p's neighbour table points into
// half 1 - p: its local slots or its ghost slots, which hold the neighbouring
// blocks' face sites of parity 1 - p.
namespace {

int part_rb_tables(const int dims[4], const int grid[4], const int coord[4], const int block[4], PartTables h[2])
{
  PartTables t;
  GBX_TRY(part_tables(dims, grid, coord, block, t));
  for (int d = 0; d < 4; d++)
    GBX_CHECK(t.ld[d] % 2 == 0, GBX_EINVAL, "red-black domain part: local extent %d along %d must be even", t.ld[d], d);
  const uint64_t vol = t.vol, half = vol / 2;
  for (int p = 0; p < 2; p++) {
    PartTables &a = h[p];
    for (int d = 0; d < 4; d++) {
      a.ld[d] = t.ld[d];
      a.blk[d] = t.blk[d];
    }
    a.vol = half;
    a.gvol = t.gvol;
    a.slot_of_site.assign(half, 0);
    a.site_of_slot.clear();
    a.gsite.clear();
    uint64_t gb = half, fb = 0;
    for (int d = 0; d < 4; d++)
      if (t.face[d]) {
        a.face[d] = t.face[d] / 2;
        a.ghost_base[d] = gb;
        a.fs_base[d] = fb;
        gb += 2 * a.face[d];
        fb += 2 * a.face[d];
      }
    a.nghost = gb - half;
    a.fslots.assign(fb, 0);
    a.ninterior = 0;
  }
  // the part's order (interior first) restricted to each parity
  for (uint64_t q = 0; q < vol; q++) {
    const uint32_t site = t.site_of_slot[q];
    const int p = site_parity(site, t.ld);
    PartTables &a = h[p];
    a.slot_of_site[half_index(site, t.ld)] = (uint32_t)a.site_of_slot.size();
    a.site_of_slot.push_back(half_index(site, t.ld));
    a.gsite.push_back(t.gsite[q]);
    if (q < t.ninterior) a.ninterior++;
  }
  // face slots of each half: face site f of the part has parity-face index f / 2
  for (int d = 0; d < 4; d++) {
    if (!t.face[d]) continue;
    for (int side = 0; side < 2; side++)
      for (uint64_t f = 0; f < t.face[d]; f++) {
        const uint32_t site = t.site_of_slot[t.fslots[t.fs_base[d] + side * t.face[d] + f]];
        PartTables &a = h[site_parity(site, t.ld)];
        a.fslots[a.fs_base[d] + side * a.face[d] + f / 2] = a.slot_of_site[half_index(site, t.ld)];
      }
  }
  // neighbours: the part's table, mapped into the partner half's slots
  for (int p = 0; p < 2; p++) {
    PartTables &a = h[p], &b = h[1 - p];
    a.nbr.assign(half * 8, 0);
    for (uint64_t q = 0; q < vol; q++) {
      const uint32_t site = t.site_of_slot[q];
      if (site_parity(site, t.ld) != p) continue;
      const uint32_t hq = a.slot_of_site[half_index(site, t.ld)];
      for (int mu = 0; mu < 8; mu++) {
        const uint32_t n = t.nbr[q * 8 + mu];
        uint32_t m;
        if (n < vol) {
          m = b.slot_of_site[half_index(t.site_of_slot[n], t.ld)];
        } else {
          // ghost slot of the part: (dimension, side, face index f) -> the half's (f / 2)
          int d = 0;
          while (!(t.face[d] && n >= t.ghost_base[d] && n < t.ghost_base[d] + 2 * t.face[d])) d++;
          const uint64_t r = n - t.ghost_base[d], side = r / t.face[d], f = r - side * t.face[d];
          m = (uint32_t)(b.ghost_base[d] + side * b.face[d] + f / 2);
        }
        a.nbr[hq * 8 + mu] = m;
      }
    }
  }
  return GBX_OK;
}

}  // namespace

extern "C" int gbx_geom_create_part_rb(const int dims[4], const int grid[4], const int coord[4], const int block[4],
                                       gbx_geom *even, gbx_geom *odd)
{
  GBX_CHECK(even && odd, GBX_ENULL, "gbx_geom_create_part_rb: NULL out");
  *even = *odd = nullptr;
  PartTables h[2];
  GBX_TRY(part_rb_tables(dims, grid, coord, block, h));
  GBX_TRY(gbx_require_device());
  int device;
  GBX_HIP(hipGetDevice(&device));
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(g_pair_mu);
    id = ++g_next_pair_id;
  }
  gbx_geom_s *g[2];
  for (int p = 0; p < 2; p++) {
    g[p] = geom_new(h[p].ld, 1, h[p].blk, h[p].vol, device);
    g[p]->gauge_vol = h[p].gvol;
    g[p]->parity = p;
    g[p]->pair_id = id;
    for (int d = 0; d < 4; d++) {
      g[p]->grid[d] = grid[d];
      g[p]->coord[d] = coord[d];
      g[p]->gdims[d] = dims[d];
      g[p]->face[d] = h[p].face[d];
      g[p]->ghost_base[d] = h[p].ghost_base[d];
      g[p]->fs_base[d] = h[p].fs_base[d];
    }
    g[p]->nghost = h[p].nghost;
    g[p]->ninterior = h[p].ninterior;
    g[p]->slot_of_site_h = std::move(h[p].slot_of_site);
    g[p]->site_of_slot_h = std::move(h[p].site_of_slot);
  }
  auto upload_faces = [&](int p) -> int {
    if (hipMalloc(&g[p]->face_slots, h[p].fslots.size() * 4) != hipSuccess ||
        hipMemcpy(g[p]->face_slots, h[p].fslots.data(), h[p].fslots.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      return gbx_fail(GBX_ENOMEM, "gbx_geom_create_part_rb: face table");
    }
    return GBX_OK;
  };
  int rc = geom_finish(g[0], h[0].nbr, &h[0].gsite);  // releases g[0] on failure
  if (rc != GBX_OK) {
    geom_release(g[1]);
    return rc;
  }
  g[1]->stream_owner = g[0];
  geom_retain(g[0]);
  rc = geom_finish(g[1], h[1].nbr, &h[1].gsite);  // releases g[1] (and its hold on g[0]) on failure
  if (rc != GBX_OK) {
    geom_release(g[0]);
    return rc;
  }
  rc = upload_faces(0);
  if (rc == GBX_OK) rc = upload_faces(1);
  if (rc != GBX_OK) {
    geom_release(g[1]);  // and its hold on g[0]
    geom_release(g[0]);
    return rc;
  }
  *even = g[0];
  *odd = g[1];
  return GBX_OK;
}

extern "C" int gbx_geom_part_info(gbx_geom g, int grid[4], int coord[4], uint64_t face[4], uint64_t *ninterior)
{
  GBX_CHECK(g && grid && coord && face && ninterior, GBX_ENULL, "gbx_geom_part_info: NULL argument");
  for (int d = 0; d < 4; d++) {
    grid[d] = g->grid[d];
    coord[d] = g->coord[d];
    face[d] = g->face[d];
  }
  *ninterior = g->nghost ? g->ninterior : g->vol;
  return GBX_OK;
}

extern "C" int gbx_geom_parity(gbx_geom g, int *parity)
{
  GBX_CHECK(g && parity, GBX_ENULL, "gbx_geom_parity: NULL argument");
  *parity = g->parity;
  return GBX_OK;
}

extern "C" int gbx_geom_destroy(gbx_geom g)
{
  if (!g) return GBX_OK;
  geom_release(g);
  return GBX_OK;
}

extern "C" int gbx_geom_volume(gbx_geom g, uint64_t *vol)
{
  GBX_CHECK(g && vol, GBX_ENULL, "gbx_geom_volume: NULL argument");
  *vol = g->vol;
  return GBX_OK;
}

extern "C" int gbx_geom_stream(gbx_geom g, void **stream)
{
  GBX_CHECK(g && stream, GBX_ENULL, "gbx_geom_stream: NULL argument");
  *stream = (void *)g->stream;
  return GBX_OK;
}

extern "C" int gbx_sync(gbx_geom g)
{
  GBX_CHECK(g, GBX_ENULL, "gbx_sync: NULL geometry");
  GBX_HIP(hipStreamSynchronize(g->stream));
  return GBX_OK;
}

// --------------------------------------------------------------------- gauge
extern "C" int gbx_gauge_create(gbx_geom g, int prec, int comp12, gbx_gauge *out)
{
  GBX_CHECK(g && out, GBX_ENULL, "gbx_gauge_create: NULL argument");
  GBX_CHECK(prec_valid(prec), GBX_EINVAL, "gbx_gauge_create: precision %d", prec);
  GBX_CHECK(comp12 == 0 || comp12 == 1, GBX_EINVAL, "gbx_gauge_create: comp12 must be 0 or 1");
  *out = nullptr;
  gbx_gauge_s *u = new gbx_gauge_s();
  u->geom = g;
  u->prec = prec;
  u->comp12 = comp12;
  u->bytes = g->vol * 8 * (comp12 ? 12 : 18) * (size_t)prec_bytes(prec);
  if (hipMalloc(&u->d, u->bytes) != hipSuccess) {
    delete u;
    return gbx_fail(GBX_ENOMEM, "gbx_gauge_create: cannot allocate %zu bytes", (size_t)g->vol * 8 * 18 * prec_bytes(prec));
  }
  geom_retain(g);
  *out = u;
  return GBX_OK;
}

extern "C" int gbx_gauge_destroy(gbx_gauge u)
{
  if (!u) return GBX_OK;
  (void)hipStreamSynchronize(u->geom->stream);
  (void)hipFree(u->d);
  geom_release(u->geom);
  delete u;
  return GBX_OK;
}

extern "C" int gbx_gauge_upload(gbx_gauge u, const double *doubled)
{
  GBX_CHECK(u && doubled, GBX_ENULL, "gbx_gauge_upload: NULL argument");
  gbx_geom_s *g = u->geom;
  std::lock_guard<std::mutex> lk(g_staging_mu);
  void *stage;
  const size_t in_bytes = g->gauge_vol * 8 * 18 * sizeof(double);  // a half or a part reads the whole lattice
  GBX_TRY(staging_buffer(in_bytes, &stage));
  GBX_HIP(hipMemcpyAsync(stage, doubled, in_bytes, hipMemcpyHostToDevice, g->stream));
  const int Rl = u->comp12 ? 12 : 18;
  const uint64_t total = g->vol * 8 * (uint64_t)Rl;
  dispatch_prec<PackGaugeLex>(u->prec, (const double *)stage, u->d, gauge_sites(g), total, Rl, g->stream);
  GBX_HIP(hipGetLastError());
  GBX_HIP(hipStreamSynchronize(g->stream));
  return GBX_OK;
}

template <typename T> static int gauge_upload_ref(gbx_gauge u, const T *ref, int nsimd)
{
  GBX_CHECK(u && ref, GBX_ENULL, "gbx_gauge_upload_ref: NULL argument");
  gbx_geom_s *g = u->geom;
  RefGeom rg;
  GBX_TRY(ref_geom(g, nsimd, rg));
  std::lock_guard<std::mutex> lk(g_staging_mu);
  void *stage;
  const size_t in_bytes = g->vol * 8 * 18 * sizeof(T);
  GBX_TRY(staging_buffer(in_bytes, &stage));
  GBX_HIP(hipMemcpyAsync(stage, ref, in_bytes, hipMemcpyHostToDevice, g->stream));
  const int Rl = u->comp12 ? 12 : 18;
  const uint64_t total = g->vol * 8 * (uint64_t)Rl;
  dispatch_prec<PackGaugeRefT<T>::template F>(u->prec, (const T *)stage, u->d, (const uint32_t *)g->site_of_slot, total, Rl, rg,
                                              g->stream);
  GBX_HIP(hipGetLastError());
  GBX_HIP(hipStreamSynchronize(g->stream));
  return GBX_OK;
}

extern "C" int gbx_gauge_upload_ref_f64(gbx_gauge u, const double *ref, int nsimd) { return gauge_upload_ref(u, ref, nsimd); }
extern "C" int gbx_gauge_upload_ref_f32(gbx_gauge u, const float *ref, int nsimd) { return gauge_upload_ref(u, ref, nsimd); }

// -------------------------------------------------------------------- spinor
extern "C" int gbx_spinor_create(gbx_geom g, int Ls, int nrhs, int prec, gbx_spinor *out)
{
  GBX_CHECK(g && out, GBX_ENULL, "gbx_spinor_create: NULL argument");
  *out = nullptr;
  GBX_CHECK(prec_valid(prec), GBX_EINVAL, "gbx_spinor_create: precision %d", prec);
  GBX_CHECK(Ls >= 1 && nrhs >= 1 && (uint64_t)Ls * (uint64_t)nrhs <= 4096, GBX_EINVAL,
            "gbx_spinor_create: Ls=%d nrhs=%d (Ls*nrhs must be in 1..4096)", Ls, nrhs);
  gbx_spinor_s *x = new gbx_spinor_s();
  x->geom = g;
  x->Ls = Ls;
  x->nrhs = nrhs;
  x->prec = prec;
  x->M = (uint32_t)(Ls * nrhs);
  x->bytes = (g->vol + g->nghost) * (size_t)x->M * 24 * (size_t)prec_bytes(prec);
  if (hipMalloc(&x->d, x->bytes) != hipSuccess) {
    const size_t b = x->bytes;
    delete x;
    (void)hipGetLastError();
    return gbx_fail(GBX_ENOMEM, "gbx_spinor_create: cannot allocate %zu bytes", b);
  }
  geom_retain(g);
  *out = x;
  return GBX_OK;
}

extern "C" int gbx_spinor_destroy(gbx_spinor x)
{
  if (!x) return GBX_OK;
  (void)hipStreamSynchronize(x->geom->stream);
  (void)hipFree(x->d);
  geom_release(x->geom);
  delete x;
  return GBX_OK;
}

extern "C" int gbx_spinor_bytes(gbx_spinor x, uint64_t *bytes)
{
  GBX_CHECK(x && bytes, GBX_ENULL, "gbx_spinor_bytes: NULL argument");
  *bytes = x->bytes;
  return GBX_OK;
}

// Right-hand sides r0 .. r0+nr-1 from / to lexicographic FP64 [nr][V][Ls][24].
static int spinor_xfer_lex(gbx_spinor x, int r0, int nr, double *lex, bool up)
{
  GBX_TRY(spinor_ok(x));
  GBX_CHECK(lex, GBX_ENULL, "gbx_spinor_%s: NULL host array", up ? "upload" : "download");
  GBX_CHECK(r0 >= 0 && nr >= 1 && r0 + nr <= x->nrhs, GBX_EINVAL, "gbx_spinor_%s: rhs range %d+%d of %d", up ? "upload" : "download",
            r0, nr, x->nrhs);
  gbx_geom_s *g = x->geom;
  std::lock_guard<std::mutex> lk(g_staging_mu);
  void *stage;
  const uint32_t Mr = (uint32_t)(nr * x->Ls), m0 = (uint32_t)(r0 * x->Ls);
  const size_t n = g->vol * (size_t)Mr * 24;
  GBX_TRY(staging_buffer(n * sizeof(double), &stage));
  const uint64_t total = g->vol * (uint64_t)spinor_nchunk(x->prec) * Mr;
  if (up) {
    GBX_HIP(hipMemcpyAsync(stage, lex, n * sizeof(double), hipMemcpyHostToDevice, g->stream));
    dispatch_prec<PackLex>(x->prec, (const double *)stage, x->d, (const uint32_t *)g->site_of_slot, total, (uint32_t)g->vol,
                           (uint32_t)x->Ls, x->M, m0, Mr, g->stream);
    GBX_HIP(hipGetLastError());
  } else {
    dispatch_prec<UnpackLex>(x->prec, (const void *)x->d, (double *)stage, (const uint32_t *)g->site_of_slot, total,
                             (uint32_t)g->vol, (uint32_t)x->Ls, x->M, m0, Mr, g->stream);
    GBX_HIP(hipGetLastError());
    GBX_HIP(hipMemcpyAsync(lex, stage, n * sizeof(double), hipMemcpyDeviceToHost, g->stream));
  }
  GBX_HIP(hipStreamSynchronize(g->stream));
  return GBX_OK;
}

// Host arrays hold chiral-basis values: a whole-field upload leaves the
// field in that basis, partial uploads and downloads need it there.
extern "C" int gbx_spinor_upload(gbx_spinor x, const double *lex)
{
  GBX_TRY(spinor_xfer_lex(x, 0, x ? x->nrhs : 1, (double *)lex, true));
  x->basis = GBX_BASIS_CHIRAL;
  return GBX_OK;
}

extern "C" int gbx_spinor_download(gbx_spinor x, double *lex)
{
  GBX_TRY(require_chiral("gbx_spinor_download", x));
  return spinor_xfer_lex(x, 0, x ? x->nrhs : 1, lex, false);
}

extern "C" int gbx_spinor_upload_rhs(gbx_spinor x, int r, const double *lex)
{
  GBX_TRY(require_chiral("gbx_spinor_upload_rhs", x));
  return spinor_xfer_lex(x, r, 1, (double *)lex, true);
}

extern "C" int gbx_spinor_download_rhs(gbx_spinor x, int r, double *lex)
{
  GBX_TRY(require_chiral("gbx_spinor_download_rhs", x));
  return spinor_xfer_lex(x, r, 1, lex, false);
}

template <typename T> static int spinor_upload_ref(gbx_spinor x, const T *ref, int nsimd)
{
  GBX_TRY(spinor_ok(x));
  GBX_CHECK(ref, GBX_ENULL, "gbx_spinor_upload_ref: NULL host array");
  GBX_CHECK(x->nrhs == 1, GBX_EINVAL, "gbx_spinor_upload_ref: reference layout holds one right-hand side");
  gbx_geom_s *g = x->geom;
  RefGeom rg;
  GBX_TRY(ref_geom(g, nsimd, rg));
  std::lock_guard<std::mutex> lk(g_staging_mu);
  void *stage;
  const size_t n = g->vol * (size_t)x->Ls * 24;
  GBX_TRY(staging_buffer(n * sizeof(T), &stage));
  GBX_HIP(hipMemcpyAsync(stage, ref, n * sizeof(T), hipMemcpyHostToDevice, g->stream));
  const uint64_t total = g->vol * (uint64_t)spinor_nchunk(x->prec) * x->M;
  dispatch_prec<PackRefT<T>::template F>(x->prec, (const T *)stage, x->d, (const uint32_t *)g->site_of_slot, total, rg,
                                         (uint32_t)x->Ls, g->stream);
  GBX_HIP(hipGetLastError());
  GBX_HIP(hipStreamSynchronize(g->stream));
  return GBX_OK;
}

template <typename T> static int spinor_download_ref(gbx_spinor x, T *ref, int nsimd)
{
  GBX_TRY(spinor_ok(x));
  GBX_CHECK(ref, GBX_ENULL, "gbx_spinor_download_ref: NULL host array");
  GBX_CHECK(x->nrhs == 1, GBX_EINVAL, "gbx_spinor_download_ref: reference layout holds one right-hand side");
  gbx_geom_s *g = x->geom;
  RefGeom rg;
  GBX_TRY(ref_geom(g, nsimd, rg));
  std::lock_guard<std::mutex> lk(g_staging_mu);
  void *stage;
  const size_t n = g->vol * (size_t)x->Ls * 24;
  GBX_TRY(staging_buffer(n * sizeof(T), &stage));
  const uint64_t total = g->vol * (uint64_t)spinor_nchunk(x->prec) * x->M;
  dispatch_prec<UnpackRefT<T>::template F>(x->prec, (const void *)x->d, (T *)stage, (const uint32_t *)g->site_of_slot, total,
                                           rg, (uint32_t)x->Ls, g->stream);
  GBX_HIP(hipGetLastError());
  GBX_HIP(hipMemcpyAsync(ref, stage, n * sizeof(T), hipMemcpyDeviceToHost, g->stream));
  GBX_HIP(hipStreamSynchronize(g->stream));
  return GBX_OK;
}

extern "C" int gbx_spinor_upload_ref_f64(gbx_spinor x, const double *ref, int nsimd)
{
  GBX_TRY(spinor_upload_ref(x, ref, nsimd));
  x->basis = GBX_BASIS_CHIRAL;
  return GBX_OK;
}
extern "C" int gbx_spinor_download_ref_f64(gbx_spinor x, double *ref, int nsimd)
{
  GBX_TRY(require_chiral("gbx_spinor_download_ref_f64", x));
  return spinor_download_ref(x, ref, nsimd);
}
extern "C" int gbx_spinor_upload_ref_f32(gbx_spinor x, const float *ref, int nsimd)
{
  GBX_TRY(spinor_upload_ref(x, ref, nsimd));
  x->basis = GBX_BASIS_CHIRAL;
  return GBX_OK;
}
extern "C" int gbx_spinor_download_ref_f32(gbx_spinor x, float *ref, int nsimd)
{
  GBX_TRY(require_chiral("gbx_spinor_download_ref_f32", x));
  return spinor_download_ref(x, ref, nsimd);
}

extern "C" int gbx_spinor_copy(gbx_spinor dst, gbx_spinor src)
{
  GBX_TRY(spinor_ok(dst));
  GBX_TRY(spinor_ok(src));
  GBX_CHECK(dst->geom == src->geom && dst->Ls == src->Ls && dst->nrhs == src->nrhs, GBX_EMISMATCH,
            "gbx_spinor_copy: different geometry or shape");
  gbx_geom_s *g = dst->geom;
  if (dst == src) return GBX_OK;
  // dst takes src's spin basis, which 16-bit fields do not have but chiral
  GBX_CHECK(src->basis == GBX_BASIS_CHIRAL || dst->prec == GBX_F64 || dst->prec == GBX_F32, GBX_EUNSUPPORTED,
            "gbx_spinor_copy: 16-bit fields hold chiral-basis values only");
  if (dst->prec == src->prec) {
    GBX_HIP(hipMemcpyAsync(dst->d, src->d, dst->bytes, hipMemcpyDeviceToDevice, g->stream));
    dst->basis = src->basis;
    return GBX_OK;
  }
  // The chunked layouts of different precisions place reals differently:
  // convert point by point (all 24 reals of one (slot, m) per thread).
  GBX_TRY(spinor_convert_scaled(dst, src, 1.0, g->stream));
  dst->basis = src->basis;
  return GBX_OK;
}

// Reference-layout arrays already in device memory <-> device objects,
// in any lane layout; enqueued on the geometry's stream, no synchronisation.
template <typename T> static int gauge_pack_ref_dev(gbx_gauge u, const T *dref, int nsimd, const gbx_ref_layout &lay)
{
  GBX_CHECK(u && dref, GBX_ENULL, "gauge_pack_ref_dev: NULL argument");
  gbx_geom_s *g = u->geom;
  RefGeom rg;
  GBX_TRY(ref_geom(g, nsimd, rg, lay));
  const int Rl = u->comp12 ? 12 : 18;
  const uint64_t total = g->vol * 8 * (uint64_t)Rl;
  dispatch_prec<PackGaugeRefT<T>::template F>(u->prec, dref, u->d, (const uint32_t *)g->site_of_slot, total, Rl, rg, g->stream);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

template <typename T> static int spinor_pack_ref_dev(gbx_spinor x, const T *dref, int nsimd, const gbx_ref_layout &lay)
{
  GBX_TRY(spinor_ok(x));
  GBX_CHECK(dref, GBX_ENULL, "spinor_pack_ref_dev: NULL array");
  GBX_CHECK(x->nrhs == 1, GBX_EINVAL, "spinor_pack_ref_dev: reference layout holds one right-hand side");
  gbx_geom_s *g = x->geom;
  RefGeom rg;
  GBX_TRY(ref_geom(g, nsimd, rg, lay));
  const uint64_t total = g->vol * (uint64_t)spinor_nchunk(x->prec) * x->M;
  dispatch_prec<PackRefT<T>::template F>(x->prec, dref, x->d, (const uint32_t *)g->site_of_slot, total, rg, (uint32_t)x->Ls,
                                         g->stream);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

template <typename T> static int spinor_unpack_ref_dev(gbx_spinor x, T *dref, int nsimd, const gbx_ref_layout &lay)
{
  GBX_TRY(spinor_ok(x));
  GBX_CHECK(dref, GBX_ENULL, "spinor_unpack_ref_dev: NULL array");
  GBX_CHECK(x->nrhs == 1, GBX_EINVAL, "spinor_unpack_ref_dev: reference layout holds one right-hand side");
  gbx_geom_s *g = x->geom;
  RefGeom rg;
  GBX_TRY(ref_geom(g, nsimd, rg, lay));
  const uint64_t total = g->vol * (uint64_t)spinor_nchunk(x->prec) * x->M;
  dispatch_prec<UnpackRefT<T>::template F>(x->prec, (const void *)x->d, dref, (const uint32_t *)g->site_of_slot, total, rg,
                                           (uint32_t)x->Ls, g->stream);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

int ref_pack_gauge_dev(gbx_gauge u, const double *d, int nsimd, const gbx_ref_layout &l) { return gauge_pack_ref_dev(u, d, nsimd, l); }
int ref_pack_gauge_dev(gbx_gauge u, const float *d, int nsimd, const gbx_ref_layout &l) { return gauge_pack_ref_dev(u, d, nsimd, l); }
int ref_pack_spinor_dev(gbx_spinor x, const double *d, int nsimd, const gbx_ref_layout &l) { return spinor_pack_ref_dev(x, d, nsimd, l); }
int ref_pack_spinor_dev(gbx_spinor x, const float *d, int nsimd, const gbx_ref_layout &l) { return spinor_pack_ref_dev(x, d, nsimd, l); }
int ref_unpack_spinor_dev(gbx_spinor x, double *d, int nsimd, const gbx_ref_layout &l) { return spinor_unpack_ref_dev(x, d, nsimd, l); }
int ref_unpack_spinor_dev(gbx_spinor x, float *d, int nsimd, const gbx_ref_layout &l) { return spinor_unpack_ref_dev(x, d, nsimd, l); }
