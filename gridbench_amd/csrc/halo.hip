// Halo exchange of a domain-decomposed lattice: the face pack and the ghost
// slots of a domain part (gbx_geom_create_part).
//
// A part holds its vol local slots followed by 2*face ghost slots.  Its
// neighbour table sends every leg that crosses the cut to a ghost slot, so
// the Dhop kernels run on a part unchanged once the ghosts hold the
// neighbouring parts' faces:
//
//   part p's top face    (coordinate ld-1) -> part p+1's minus ghosts (side 0)
//   part p's bottom face (coordinate 0)    -> part p-1's plus ghosts  (side 1)
//
// with p+1 and p-1 taken modulo nparts (the lattice is periodic).  Ghost
// slot f of a side is face site f, f = lexicographic index over the three
// dimensions other than the split one, so a packed face is exactly the
// receiving side's ghost range and a receive can land there directly.
//
// A face element is one site's whole spinor over all (s, RHS) minor indices:
// nch chunks of M 16-byte units, the same bytes as the site's slot.

#include <algorithm>

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

int part_ok(gbx_spinor_s *x, const char *who)
{
  GBX_CHECK(x, GBX_ENULL, "%s: NULL field", who);
  GBX_CHECK(x->geom->nghost != 0, GBX_EINVAL, "%s: the field's geometry is not a domain part", who);
  return GBX_OK;
}

uint64_t slot_units(const gbx_spinor_s *x) { return (uint64_t)x->M * 24 * prec_bytes(x->prec) / 16; }

}  // namespace

extern "C" int gbx_halo_bytes(gbx_spinor x, int proj, uint64_t *bytes)
{
  GBX_CHECK(bytes, GBX_ENULL, "gbx_halo_bytes: NULL bytes");
  GBX_TRY(part_ok(x, "gbx_halo_bytes"));
  GBX_CHECK(proj == 0 || proj == 1, GBX_EINVAL, "gbx_halo_bytes: proj=%d", proj);
  *bytes = x->geom->face * slot_units(x) * 16 / (proj ? 2 : 1);
  return GBX_OK;
}

extern "C" int gbx_halo_ghost(gbx_spinor x, int side, void **dptr)
{
  GBX_CHECK(dptr, GBX_ENULL, "gbx_halo_ghost: NULL dptr");
  GBX_TRY(part_ok(x, "gbx_halo_ghost"));
  GBX_CHECK(side == 0 || side == 1, GBX_EINVAL, "gbx_halo_ghost: side=%d (0 = minus, 1 = plus)", side);
  const gbx_geom_s *g = x->geom;
  *dptr = (char *)x->d + (g->vol + (uint64_t)side * g->face) * slot_units(x) * 16;
  return GBX_OK;
}

extern "C" int gbx_halo_pack(gbx_spinor x, int face, void *dbuf)
{
  GBX_CHECK(dbuf, GBX_ENULL, "gbx_halo_pack: NULL buffer");
  GBX_TRY(part_ok(x, "gbx_halo_pack"));
  GBX_CHECK(face == 0 || face == 1, GBX_EINVAL, "gbx_halo_pack: face=%d (0 = bottom, 1 = top)", face);
  GBX_CHECK(((uintptr_t)dbuf & 15) == 0, GBX_EINVAL, "gbx_halo_pack: buffer not 16-byte aligned");
  const gbx_geom_s *g = x->geom;
  const uint64_t units = slot_units(x), total = g->face * units;
  const uint64_t bytes = total * 16;
  const char *b = (const char *)dbuf, *d = (const char *)x->d;
  GBX_CHECK(b + bytes <= d || b >= d + x->bytes, GBX_EINVAL, "gbx_halo_pack: buffer overlaps the field");
  hipLaunchKernelGGL(k_halo_pack, dim3(grid_for(total)), dim3(256), 0, g->stream, (const uint4 *)x->d,
                     g->face_slots + (uint64_t)face * g->face, (uint4 *)dbuf, g->face, units);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

// Projector of the leg that reads the face on the receiving part: the top
// face feeds the next part's backward leg (direction split_dim + 4), the
// bottom face the previous part's forward leg (direction split_dim); the
// non-dagger operator uses the opposite projector type on each leg, the
// dagger the same.
static int face_projector(int split_dim, int face, int dag)
{
  const int dir = face ? split_dim + 4 : split_dim;
  return dag ? dir : (dir ^ 4);
}

extern "C" int gbx_halo_pack_proj(gbx_spinor x, int face, int dag, void *dbuf)
{
  GBX_CHECK(dbuf, GBX_ENULL, "gbx_halo_pack_proj: NULL buffer");
  GBX_TRY(part_ok(x, "gbx_halo_pack_proj"));
  GBX_CHECK(face == 0 || face == 1, GBX_EINVAL, "gbx_halo_pack_proj: face=%d (0 = bottom, 1 = top)", face);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_halo_pack_proj: dag=%d", dag);
  GBX_CHECK(x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED,
            "gbx_halo_pack_proj: FP64 / FP32 fields (16-bit storage is projected in FP32 by the kernels: a projected "
            "FP32 face is no smaller than the full 16-bit one, use gbx_halo_pack)");
  GBX_CHECK(((uintptr_t)dbuf & 15) == 0, GBX_EINVAL, "gbx_halo_pack_proj: buffer not 16-byte aligned");
  const gbx_geom_s *g = x->geom;
  const uint64_t bytes = g->face * slot_units(x) * 8;
  const char *b = (const char *)dbuf, *d = (const char *)x->d;
  GBX_CHECK(b + bytes <= d || b >= d + x->bytes, GBX_EINVAL, "gbx_halo_pack_proj: buffer overlaps the field");
  const int K = face_projector(g->split_dim, face, dag);
  const dim3 grid(grid_for(g->face * x->M));
  const uint32_t *fs = g->face_slots + (uint64_t)face * g->face;
  if (x->prec == GBX_F64) launch_pack_proj<GBX_F64>(K, grid, g->stream, x->d, fs, dbuf, g->face, x->M);
  else launch_pack_proj<GBX_F32>(K, grid, g->stream, x->d, fs, dbuf, g->face, x->M);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

extern "C" int gbx_halo_unpack_proj(gbx_spinor x, int side, const void *dbuf)
{
  GBX_CHECK(dbuf, GBX_ENULL, "gbx_halo_unpack_proj: NULL buffer");
  void *ghost;
  GBX_TRY(gbx_halo_ghost(x, side, &ghost));
  GBX_CHECK(x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED, "gbx_halo_unpack_proj: FP64 / FP32 fields");
  GBX_CHECK(((uintptr_t)dbuf & 15) == 0, GBX_EINVAL, "gbx_halo_unpack_proj: buffer not 16-byte aligned");
  const gbx_geom_s *g = x->geom;
  const uint32_t nch = (uint32_t)spinor_nchunk(x->prec);
  hipLaunchKernelGGL(k_halo_unpack_proj, dim3(grid_for(g->face * nch * x->M)), dim3(256), 0, g->stream,
                     (const uint4 *)dbuf, (uint4 *)ghost, g->face, nch, x->M);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

extern "C" int gbx_halo_unpack(gbx_spinor x, int side, const void *dbuf)
{
  GBX_CHECK(dbuf, GBX_ENULL, "gbx_halo_unpack: NULL buffer");
  void *ghost;
  GBX_TRY(gbx_halo_ghost(x, side, &ghost));
  uint64_t bytes;
  GBX_TRY(gbx_halo_bytes(x, 0, &bytes));
  GBX_HIP(hipMemcpyAsync(ghost, dbuf, bytes, hipMemcpyDeviceToDevice, x->geom->stream));
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
