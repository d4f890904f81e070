// Halo exchange of a domain-decomposed lattice: the face pack and the ghost
// slots of a domain part (gbx_geom_create_part).
//
// A part holds its vol local slots followed by ghost slots, two faces' worth
// for every cut dimension.  Its neighbour table sends every leg that crosses
// a cut to a ghost slot, so the Dhop kernels run on a part unchanged once the
// ghosts hold the neighbouring parts' faces.  Along a cut dimension d:
//
//   part c's top face    (coordinate ld-1) -> part c+1's minus ghosts (side 0)
//   part c's bottom face (coordinate 0)    -> part c-1's plus ghosts  (side 1)
//
// with c+1 and c-1 the grid neighbours along d, periodic.  Ghost slot f of a
// side is face site f, f = lexicographic index over the three dimensions
// other than d, so a packed face is exactly the receiving side's ghost range
// and a receive can land there directly.
//
// A face element is one site's whole spinor over all (s, RHS) minor indices:
// nch chunks of M 16-byte units, the same bytes as the site's slot.

#include <algorithm>
#include <functional>
#include <vector>

#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

// dst[f][u] = src[face_slots[f]][u], u over the slot's nch*M 16-byte units.
__global__ __launch_bounds__(256) void k_halo_pack(const uint4 *__restrict__ src, const uint32_t *__restrict__ face_slots,
                                                   uint4 *__restrict__ dst, uint64_t face, uint64_t units)
{
  const uint64_t total = face * units;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t f = i / units, u = i - f * units;
    dst[i] = src[(uint64_t)face_slots[f] * units + u];
  }
}

// Spin-projected face: dst[f][k][m] (k < nch/2) = chunk k of the half
// spinor h = (h0, h1) that projector K gives for face site f, minor m,
// computed in the storage precision exactly as the Dhop kernels compute it.
template <int P, int K>
__global__ __launch_bounds__(256) void k_halo_pack_proj(const void *__restrict__ src, const uint32_t *__restrict__ face_slots,
                                                        void *__restrict__ dst, uint64_t face, uint32_t M)
{
  typedef typename dev::Traits<P>::acc A;
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R, HCH = NCH / 2;
  const uint64_t total = face * M;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t f = i / M, m = i - f * M, q = face_slots[f];
    A v[24];
#pragma unroll
    for (int k = 0; k < NCH; k++) dev::load_chunk<P, A>(src, (q * NCH + k) * M + m, v + k * R);
    A hr[6], hi[6], h[12];
    dev::project<K>(v, hr, hi);
#pragma unroll
    for (int j = 0; j < 6; j++) {
      h[2 * j] = hr[j];
      h[2 * j + 1] = hi[j];
    }
#pragma unroll
    for (int k = 0; k < HCH; k++) dev::store_chunk<P, A>(dst, (f * HCH + k) * M + m, h + k * R, false);
  }
}

// Ghosts of one side from a spin-projected face: spins 0, 1 = the half
// spinor, spins 2, 3 = 0.  Every projector forms h0 = p0 + c p_a and
// h1 = p1 + c' p_b with a, b in {2, 3}, so projecting the ghost gives back
// the half spinor exactly: the kernels' legs read it as they would read the
// neighbour itself.
__global__ __launch_bounds__(256) void k_halo_unpack_proj(const uint4 *__restrict__ src, uint4 *__restrict__ ghost,
                                                          uint64_t face, uint32_t nch, uint32_t M)
{
  const uint32_t hch = nch / 2;
  const uint64_t per = (uint64_t)nch * M, total = face * per;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t f = i / per, r = i - f * per, k = r / M, m = r - k * M;
    ghost[i] = k < hch ? src[(f * hch + k) * M + m] : make_uint4(0u, 0u, 0u, 0u);
  }
}

// All faces of a part in one launch.  Job j covers work items
// [start_j, start_j + n_j) of the flattened index; at most 8 jobs (two faces
// of four dimensions), found by a scan.
struct FaceJob {
  const uint32_t *fs;  // face slots (pack) or unused (unpack)
  const void *src;     // unpack: the received face
  void *dst;           // pack: the face buffer; unpack: the ghost range
  uint64_t start, n, face;
  int K;               // projector (projected pack)
};
struct FaceJobs {
  FaceJob j[8];
  int n;
};

__device__ __forceinline__ int find_job(const FaceJobs &J, uint64_t i)
{
  int k = 0;
  while (k + 1 < J.n && i >= J.j[k + 1].start) k++;
  return k;
}

__global__ __launch_bounds__(256) void k_halo_pack_all(const uint4 *__restrict__ src, FaceJobs J, uint64_t units,
                                                       uint64_t total)
{
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const FaceJob &b = J.j[find_job(J, i)];
    const uint64_t r = i - b.start, f = r / units, u = r - f * units;
    ((uint4 *)b.dst)[r] = src[(uint64_t)b.fs[f] * units + u];
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_halo_pack_all_proj(const void *__restrict__ src, FaceJobs J, uint32_t M,
                                                            uint64_t total)
{
  typedef typename dev::Traits<P>::acc A;
  constexpr int NCH = dev::Traits<P>::NCH, R = dev::Traits<P>::R, HCH = NCH / 2;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const FaceJob &b = J.j[find_job(J, i)];
    const uint64_t r = i - b.start, f = r / M, m = r - f * M, q = b.fs[f];
    A v[24];
#pragma unroll
    for (int k = 0; k < NCH; k++) dev::load_chunk<P, A>(src, (q * NCH + k) * M + m, v + k * R);
    A hr[6], hi[6], h[12];
    switch (b.K) {
    case 0: dev::project<0>(v, hr, hi); break;
    case 1: dev::project<1>(v, hr, hi); break;
    case 2: dev::project<2>(v, hr, hi); break;
    case 3: dev::project<3>(v, hr, hi); break;
    case 4: dev::project<4>(v, hr, hi); break;
    case 5: dev::project<5>(v, hr, hi); break;
    case 6: dev::project<6>(v, hr, hi); break;
    default: dev::project<7>(v, hr, hi); break;
    }
#pragma unroll
    for (int j = 0; j < 6; j++) {
      h[2 * j] = hr[j];
      h[2 * j + 1] = hi[j];
    }
#pragma unroll
    for (int k = 0; k < HCH; k++) dev::store_chunk<P, A>(b.dst, (f * HCH + k) * M + m, h + k * R, false);
  }
}

__global__ __launch_bounds__(256) void k_halo_unpack_all_proj(FaceJobs J, uint32_t nch, uint32_t M, uint64_t total)
{
  const uint32_t hch = nch / 2;
  const uint64_t per = (uint64_t)nch * M;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const FaceJob &b = J.j[find_job(J, i)];
    const uint64_t r = i - b.start, f = r / per, rr = r - f * per, k = rr / M, m = rr - k * M;
    ((uint4 *)b.dst)[r] = k < hch ? ((const uint4 *)b.src)[(f * hch + k) * M + m] : make_uint4(0u, 0u, 0u, 0u);
  }
}

template <int P>
void launch_pack_proj(int K, dim3 grid, hipStream_t st, const void *src, const uint32_t *fs, void *dst, uint64_t face,
                      uint32_t M)
{
  switch (K) {
  case 0: hipLaunchKernelGGL((k_halo_pack_proj<P, 0>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  case 1: hipLaunchKernelGGL((k_halo_pack_proj<P, 1>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  case 2: hipLaunchKernelGGL((k_halo_pack_proj<P, 2>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  case 3: hipLaunchKernelGGL((k_halo_pack_proj<P, 3>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  case 4: hipLaunchKernelGGL((k_halo_pack_proj<P, 4>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  case 5: hipLaunchKernelGGL((k_halo_pack_proj<P, 5>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  case 6: hipLaunchKernelGGL((k_halo_pack_proj<P, 6>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  default: hipLaunchKernelGGL((k_halo_pack_proj<P, 7>), grid, dim3(256), 0, st, src, fs, dst, face, M); break;
  }
}

unsigned grid_for(uint64_t total) { return (unsigned)std::min<uint64_t>((total + 255) / 256, 65536); }

// [b, b + bytes) and the field's allocation share no byte.
bool disjoint(const void *buf, uint64_t bytes, const gbx_spinor_s *x)
{
  const char *b = (const char *)buf, *d = (const char *)x->d;
  return b + bytes <= d || b >= d + x->bytes;
}

int part_ok(gbx_spinor_s *x, int dim, const char *who)
{
  GBX_CHECK(x, GBX_ENULL, "%s: NULL field", who);
  GBX_CHECK(x->geom->nghost != 0, GBX_EINVAL, "%s: the field's geometry is not a domain part", who);
  GBX_CHECK(dim >= 0 && dim < 4 && x->geom->grid[dim] > 1, GBX_EINVAL, "%s: dimension %d is not cut", who, dim);
  return GBX_OK;
}

uint64_t slot_units(const gbx_spinor_s *x) { return (uint64_t)x->M * 24 * prec_bytes(x->prec) / 16; }

}  // namespace

extern "C" int gbx_halo_bytes(gbx_spinor x, int dim, int proj, uint64_t *bytes)
{
  GBX_CHECK(bytes, GBX_ENULL, "gbx_halo_bytes: NULL bytes");
  GBX_TRY(part_ok(x, dim, "gbx_halo_bytes"));
  GBX_CHECK(proj == 0 || proj == 1, GBX_EINVAL, "gbx_halo_bytes: proj=%d", proj);
  *bytes = x->geom->face[dim] * slot_units(x) * 16 / (proj ? 2 : 1);
  return GBX_OK;
}

extern "C" int gbx_halo_ghost(gbx_spinor x, int dim, int side, void **dptr)
{
  GBX_CHECK(dptr, GBX_ENULL, "gbx_halo_ghost: NULL dptr");
  GBX_TRY(part_ok(x, dim, "gbx_halo_ghost"));
  GBX_CHECK(side == 0 || side == 1, GBX_EINVAL, "gbx_halo_ghost: side=%d (0 = minus, 1 = plus)", side);
  const gbx_geom_s *g = x->geom;
  *dptr = (char *)x->d + (g->ghost_base[dim] + (uint64_t)side * g->face[dim]) * slot_units(x) * 16;
  return GBX_OK;
}

extern "C" int gbx_halo_ghost_proj(gbx_spinor x, int dim, int side, void **dptr)
{
  GBX_CHECK(dptr, GBX_ENULL, "gbx_halo_ghost_proj: NULL dptr");
  GBX_TRY(part_ok(x, dim, "gbx_halo_ghost_proj"));
  GBX_CHECK(side == 0 || side == 1, GBX_EINVAL, "gbx_halo_ghost_proj: side=%d (0 = minus, 1 = plus)", side);
  GBX_CHECK(x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED, "gbx_halo_ghost_proj: FP64 / FP32 fields");
  GBX_TRY(require_chiral("gbx_halo_ghost_proj", x));
  const gbx_geom_s *g = x->geom;
  const uint64_t slot = slot_units(x) * 16;
  // half-slots of NCH/2 chunks after the vol whole slots, in the order of the whole-spinor ghosts
  *dptr = (char *)x->d + g->vol * slot + (g->ghost_base[dim] - g->vol + (uint64_t)side * g->face[dim]) * (slot / 2);
  return GBX_OK;
}

// Projector of the leg that reads the face on the receiving part: the top
// face feeds the next part's backward leg (direction dim + 4), the bottom
// face the previous part's forward leg (direction dim); the non-dagger
// operator uses the opposite projector type on each leg, the dagger the same.
static int face_projector(int dim, int face, int dag)
{
  const int dir = face ? dim + 4 : dim;
  return dag ? dir : (dir ^ 4);
}

// Face `face` of dimension dim of x into dbuf (whole or projected), enqueued on st.
static int pack_impl(gbx_spinor_s *x, int dim, int face, int dag, bool proj, void *dbuf, hipStream_t st, const char *who)
{
  GBX_CHECK(dbuf, GBX_ENULL, "%s: NULL buffer", who);
  GBX_TRY(part_ok(x, dim, who));
  GBX_CHECK(face == 0 || face == 1, GBX_EINVAL, "%s: face=%d (0 = bottom, 1 = top)", who, face);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "%s: dag=%d", who, dag);
  if (proj) GBX_TRY(require_chiral(who, x));
  GBX_CHECK(!proj || x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED,
            "%s: projected faces of FP64 / FP32 fields only (16-bit storage is projected in FP32 by the kernels: a "
            "projected FP32 face is no smaller than the full 16-bit one, use gbx_halo_pack)", who);
  GBX_CHECK(((uintptr_t)dbuf & 15) == 0, GBX_EINVAL, "%s: buffer not 16-byte aligned", who);
  const gbx_geom_s *g = x->geom;
  const uint64_t units = slot_units(x), F = g->face[dim];
  GBX_CHECK(disjoint(dbuf, F * units * (proj ? 8 : 16), x), GBX_EINVAL, "%s: buffer overlaps the field", who);
  const uint32_t *fs = g->face_slots + g->fs_base[dim] + (uint64_t)face * F;
  if (!proj) {
    hipLaunchKernelGGL(k_halo_pack, dim3(grid_for(F * units)), dim3(256), 0, st, (const uint4 *)x->d, fs, (uint4 *)dbuf,
                       F, units);
  } else {
    const int K = face_projector(dim, face, dag);
    const dim3 grid(grid_for(F * x->M));
    if (x->prec == GBX_F64) launch_pack_proj<GBX_F64>(K, grid, st, x->d, fs, dbuf, F, x->M);
    else launch_pack_proj<GBX_F32>(K, grid, st, x->d, fs, dbuf, F, x->M);
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

// The ghosts of `side` of dimension dim from dbuf (whole or projected face), enqueued on st.
static int unpack_impl(gbx_spinor_s *x, int dim, int side, bool proj, const void *dbuf, hipStream_t st, const char *who)
{
  GBX_CHECK(dbuf, GBX_ENULL, "%s: NULL buffer", who);
  if (proj) GBX_TRY(require_chiral(who, x));
  void *ghost;
  GBX_TRY(gbx_halo_ghost(x, dim, side, &ghost));
  const gbx_geom_s *g = x->geom;
  const uint64_t units = slot_units(x), F = g->face[dim];
  if (!proj) {
    if (dbuf == ghost) return GBX_OK;  // the face was packed or received in place
    GBX_CHECK(disjoint(dbuf, F * units * 16, x), GBX_EINVAL, "%s: buffer overlaps the field", who);
    GBX_HIP(hipMemcpyAsync(ghost, dbuf, F * units * 16, hipMemcpyDeviceToDevice, st));
    return GBX_OK;
  }
  GBX_CHECK(x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED, "%s: FP64 / FP32 fields", who);
  GBX_CHECK(((uintptr_t)dbuf & 15) == 0, GBX_EINVAL, "%s: buffer not 16-byte aligned", who);
  GBX_CHECK(disjoint(dbuf, F * units * 8, x), GBX_EINVAL, "%s: buffer overlaps the field", who);
  const uint32_t nch = (uint32_t)spinor_nchunk(x->prec);
  hipLaunchKernelGGL(k_halo_unpack_proj, dim3(grid_for(F * nch * x->M)), dim3(256), 0, st, (const uint4 *)dbuf,
                     (uint4 *)ghost, F, nch, x->M);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

extern "C" int gbx_halo_pack(gbx_spinor x, int dim, int face, void *dbuf)
{
  GBX_CHECK(x, GBX_ENULL, "gbx_halo_pack: NULL field");
  return pack_impl(x, dim, face, 0, false, dbuf, x->geom->stream, "gbx_halo_pack");
}

extern "C" int gbx_halo_pack_proj(gbx_spinor x, int dim, int face, int dag, void *dbuf)
{
  GBX_CHECK(x, GBX_ENULL, "gbx_halo_pack_proj: NULL field");
  return pack_impl(x, dim, face, dag, true, dbuf, x->geom->stream, "gbx_halo_pack_proj");
}

extern "C" int gbx_halo_unpack(gbx_spinor x, int dim, int side, const void *dbuf)
{
  GBX_CHECK(x, GBX_ENULL, "gbx_halo_unpack: NULL field");
  return unpack_impl(x, dim, side, false, dbuf, x->geom->stream, "gbx_halo_unpack");
}

extern "C" int gbx_halo_unpack_proj(gbx_spinor x, int dim, int side, const void *dbuf)
{
  GBX_CHECK(x, GBX_ENULL, "gbx_halo_unpack_proj: NULL field");
  return unpack_impl(x, dim, side, true, dbuf, x->geom->stream, "gbx_halo_unpack_proj");
}

// Every face of every cut dimension in one launch: bufs[2*d + face] for cut
// dimensions d (entries of uncut dimensions are ignored).
static int pack_all_impl(gbx_spinor_s *x, int dag, bool proj, void *const *bufs, hipStream_t st, const char *who)
{
  GBX_CHECK(x && bufs, GBX_ENULL, "%s: NULL argument", who);
  GBX_CHECK(x->geom->nghost != 0, GBX_EINVAL, "%s: the field's geometry is not a domain part", who);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "%s: dag=%d", who, dag);
  if (proj) GBX_TRY(require_chiral(who, x));
  GBX_CHECK(!proj || x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED,
            "%s: projected faces of FP64 / FP32 fields only", who);
  const gbx_geom_s *g = x->geom;
  const uint64_t units = slot_units(x);
  const uint64_t per = proj ? (uint64_t)x->M : units;  // work items per face site
  FaceJobs J{};
  uint64_t total = 0;
  for (int d = 0; d < 4; d++) {
    if (g->grid[d] == 1) continue;
    const uint64_t F = g->face[d];
    for (int f = 0; f < 2; f++) {
      void *b = bufs[2 * d + f];
      GBX_CHECK(b, GBX_ENULL, "%s: NULL buffer for dimension %d face %d", who, d, f);
      GBX_CHECK(((uintptr_t)b & 15) == 0, GBX_EINVAL, "%s: buffer not 16-byte aligned", who);
      GBX_CHECK(disjoint(b, F * units * (proj ? 8 : 16), x), GBX_EINVAL, "%s: buffer overlaps the field", who);
      FaceJob &jb = J.j[J.n++];
      jb.fs = g->face_slots + g->fs_base[d] + (uint64_t)f * F;
      jb.dst = b;
      jb.src = nullptr;
      jb.start = total;
      jb.n = F * per;
      jb.face = F;
      jb.K = face_projector(d, f, dag);
      total += jb.n;
    }
  }
  if (!proj) {
    hipLaunchKernelGGL(k_halo_pack_all, dim3(grid_for(total)), dim3(256), 0, st, (const uint4 *)x->d, J, units, total);
  } else if (x->prec == GBX_F64) {
    hipLaunchKernelGGL((k_halo_pack_all_proj<GBX_F64>), dim3(grid_for(total)), dim3(256), 0, st, x->d, J, x->M, total);
  } else {
    hipLaunchKernelGGL((k_halo_pack_all_proj<GBX_F32>), dim3(grid_for(total)), dim3(256), 0, st, x->d, J, x->M, total);
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

// The ghosts of every cut dimension from the received faces, bufs[2*d + side]:
// side 0 (minus) takes the previous part's top face, side 1 the next part's
// bottom face.
static int unpack_all_impl(gbx_spinor_s *x, bool proj, const void *const *bufs, hipStream_t st, const char *who)
{
  GBX_CHECK(x && bufs, GBX_ENULL, "%s: NULL argument", who);
  GBX_CHECK(x->geom->nghost != 0, GBX_EINVAL, "%s: the field's geometry is not a domain part", who);
  if (proj) GBX_TRY(require_chiral(who, x));
  const gbx_geom_s *g = x->geom;
  if (!proj) {
    for (int d = 0; d < 4; d++)
      if (g->grid[d] > 1)
        for (int side = 0; side < 2; side++) GBX_TRY(unpack_impl(x, d, side, false, bufs[2 * d + side], st, who));
    return GBX_OK;
  }
  GBX_CHECK(x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED, "%s: FP64 / FP32 fields", who);
  const uint64_t units = slot_units(x);
  const uint32_t nch = (uint32_t)spinor_nchunk(x->prec);
  FaceJobs J{};
  uint64_t total = 0;
  for (int d = 0; d < 4; d++) {
    if (g->grid[d] == 1) continue;
    const uint64_t F = g->face[d];
    for (int side = 0; side < 2; side++) {
      const void *b = bufs[2 * d + side];
      GBX_CHECK(b, GBX_ENULL, "%s: NULL buffer for dimension %d side %d", who, d, side);
      GBX_CHECK(((uintptr_t)b & 15) == 0, GBX_EINVAL, "%s: buffer not 16-byte aligned", who);
      GBX_CHECK(disjoint(b, F * units * 8, x), GBX_EINVAL, "%s: buffer overlaps the field", who);
      FaceJob &jb = J.j[J.n++];
      jb.fs = nullptr;
      jb.src = b;
      jb.dst = (char *)x->d + (g->ghost_base[d] + (uint64_t)side * F) * units * 16;
      jb.start = total;
      jb.n = F * units;
      jb.face = F;
      jb.K = 0;
      total += jb.n;
    }
  }
  hipLaunchKernelGGL(k_halo_unpack_all_proj, dim3(grid_for(total)), dim3(256), 0, st, J, nch, x->M, total);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

extern "C" int gbx_halo_pack_all(gbx_spinor x, int proj, int dag, void *const bufs[8])
{
  GBX_CHECK(proj == 0 || proj == 1, GBX_EINVAL, "gbx_halo_pack_all: proj=%d", proj);
  GBX_CHECK(x, GBX_ENULL, "gbx_halo_pack_all: NULL field");
  return pack_all_impl(x, dag, proj != 0, bufs, x->geom->stream, "gbx_halo_pack_all");
}

extern "C" int gbx_halo_unpack_all(gbx_spinor x, int proj, const void *const bufs[8])
{
  GBX_CHECK(proj == 0 || proj == 1, GBX_EINVAL, "gbx_halo_unpack_all: proj=%d", proj);
  GBX_CHECK(x, GBX_ENULL, "gbx_halo_unpack_all: NULL field");
  return unpack_all_impl(x, proj != 0, bufs, x->geom->stream, "gbx_halo_unpack_all");
}

// One rank's application on one GPU, timed two ways (HIP events on the
// geometry's stream, nrep applications after one warm-up each):
//  serial:  pack the faces of every cut dimension (one launch), move them,
//           Dhop on the interior slots, then on the face slots, all on one
//           stream;
//  overlap: pack, then move on a second stream while the interior slots
//           run, then the face slots once the faces have landed.
// The move stands in for the transfer between GPUs: a device-to-device copy
// of each packed face straight to where the face launch reads it (the ghost
// slots, or for projected faces the half-slots gbx_halo_ghost_proj names),
// so nothing is unpacked.  The received faces are this part's own, so the
// face sites of `out` hold no meaningful values afterwards; the interior
// sites hold Dhop(in).
namespace {
struct OverlapRes {
  std::vector<void *> buf;  // per cut dimension: the packed bottom and top faces
  hipStream_t cs = nullptr, ms = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  ~OverlapRes()
  {
    // both streams may still use the buffers when a step failed
    if (ms) (void)hipStreamSynchronize(ms);
    if (cs) (void)hipStreamSynchronize(cs);
    for (void *b : buf) (void)hipFree(b);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (cs) (void)hipStreamDestroy(cs);
  }
};
}  // namespace

extern "C" int gbx_halo_overlap_time(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, int proj, int nrep,
                                     double *us_overlap, double *us_serial)
{
  static const char *who = "gbx_halo_overlap_time";
  GBX_CHECK(us_overlap && us_serial, GBX_ENULL, "%s: NULL result", who);
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "%s: nrep=%d", who, nrep);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "%s: dag=%d", who, dag);
  GBX_TRY(dhop_check_operands(u, out, in));
  GBX_CHECK(in->geom->nghost != 0, GBX_EINVAL, "%s: the field's geometry is not a domain part", who);
  GBX_CHECK(!proj || in->prec == GBX_F64 || in->prec == GBX_F32, GBX_EUNSUPPORTED,
            "%s: projected faces of FP64 / FP32 fields only", who);
  if (proj) GBX_TRY(require_chiral(who, in));
  const gbx_geom_s *g = in->geom;
  std::vector<int> cut;
  for (int d = 0; d < 4; d++)
    if (g->grid[d] > 1) cut.push_back(d);
  OverlapRes r;
  r.ms = g->stream;
  r.buf.assign(2 * cut.size(), nullptr);
  std::vector<uint64_t> fb(cut.size());
  void *send[8] = {}, *recv[8] = {};
  for (size_t c = 0; c < cut.size(); c++) {
    const int d = cut[c];
    fb[c] = g->face[d] * slot_units(in) * (proj ? 8 : 16);
    for (int k = 0; k < 2; k++) {
      if (hipMalloc(&r.buf[2 * c + k], fb[c]) != hipSuccess) {
        (void)hipGetLastError();
        r.buf[2 * c + k] = nullptr;
        return gbx_fail(GBX_ENOMEM, "%s: face buffers", who);
      }
      send[2 * d + k] = r.buf[2 * c + k];  // k = 0 bottom face, 1 top face
      // the receive lands in place: whole faces in the ghost slots, projected
      // ones in the half-slots the face launch reads
      GBX_TRY(proj ? gbx_halo_ghost_proj(in, d, k, &recv[2 * d + k]) : gbx_halo_ghost(in, d, k, &recv[2 * d + k]));
    }
  }
  GBX_HIP(hipStreamCreate(&r.cs));
  for (hipEvent_t &e : r.ev) GBX_HIP(hipEventCreate(&e));
  hipStream_t ms = g->stream;
  const gbx_kcfg k = dhop_range_cfg(u, in);
  SlotRange interior, faces;
  interior.end = g->ninterior;
  faces.begin = g->ninterior;
  // faces out (before anything else: the pack kernels need CUs, which the
  // interior kernel would otherwise hold), then moved and into the ghosts:
  // the top face feeds the minus side
  auto pack = [&](hipStream_t st) -> int { return pack_all_impl(in, dag, proj != 0, send, st, who); };
  // the minus side receives the previous part's top face, the plus side the next part's bottom face
  auto move = [&](hipStream_t st) -> int {
    for (size_t c = 0; c < cut.size(); c++)
      for (int side = 0; side < 2; side++) {
        const int d = cut[c];
        GBX_HIP(hipMemcpyAsync(recv[2 * d + side], send[2 * d + 1 - side], fb[c], hipMemcpyDeviceToDevice, st));
      }
    return GBX_OK;
  };
  // the face sites: projected ghosts are read as they are, no unpack
  auto face_sites = [&]() -> int {
    return proj ? gbx_dhop_faces_projected(u, out, in, dag) : dhop_run(u, out, in, dag, k, ms, faces);
  };
  auto serial = [&]() -> int {
    GBX_TRY(pack(ms));
    GBX_TRY(move(ms));
    GBX_TRY(dhop_run(u, out, in, dag, k, ms, interior));
    return face_sites();
  };
  auto overlapped = [&]() -> int {
    GBX_TRY(pack(ms));
    GBX_HIP(hipEventRecord(r.ev[2], ms));  // faces packed, previous face sites done
    GBX_HIP(hipStreamWaitEvent(r.cs, r.ev[2], 0));
    GBX_TRY(move(r.cs));
    GBX_HIP(hipEventRecord(r.ev[3], r.cs));
    GBX_TRY(dhop_run(u, out, in, dag, k, ms, interior));
    GBX_HIP(hipStreamWaitEvent(ms, r.ev[3], 0));
    return face_sites();
  };
  auto timed = [&](const std::function<int()> &f, double *us) -> int {
    GBX_TRY(f());  // warm-up
    GBX_HIP(hipEventRecord(r.ev[0], ms));
    for (int i = 0; i < nrep; i++) GBX_TRY(f());
    GBX_HIP(hipEventRecord(r.ev[1], ms));
    GBX_HIP(hipEventSynchronize(r.ev[1]));
    float t = 0.f;
    GBX_HIP(hipEventElapsedTime(&t, r.ev[0], r.ev[1]));
    *us = 1000.0 * (double)t / nrep;
    return GBX_OK;
  };
  GBX_TRY(timed(serial, us_serial));
  GBX_TRY(timed(overlapped, us_overlap));
  return GBX_OK;
}

extern "C" int gbx_buffer_create(uint64_t bytes, void **dptr)
{
  GBX_CHECK(dptr, GBX_ENULL, "gbx_buffer_create: NULL dptr");
  *dptr = nullptr;
  GBX_CHECK(bytes >= 1, GBX_EINVAL, "gbx_buffer_create: %llu bytes", (unsigned long long)bytes);
  GBX_TRY(gbx_require_device());
  if (hipMalloc(dptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return gbx_fail(GBX_ENOMEM, "gbx_buffer_create: cannot allocate %llu bytes", (unsigned long long)bytes);
  }
  return GBX_OK;
}

extern "C" int gbx_buffer_destroy(void *dptr)
{
  if (!dptr) return GBX_OK;
  GBX_HIP(hipDeviceSynchronize());
  GBX_HIP(hipFree(dptr));
  return GBX_OK;
}

extern "C" int gbx_buffer_download(const void *dptr, uint64_t bytes, void *host)
{
  GBX_CHECK(dptr && host, GBX_ENULL, "gbx_buffer_download: NULL argument");
  GBX_HIP(hipMemcpy(host, dptr, bytes, hipMemcpyDeviceToHost));
  return GBX_OK;
}
