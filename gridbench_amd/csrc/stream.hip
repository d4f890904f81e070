// Achievable-HBM-bandwidth probes: a grid-stride 16-byte copy kernel (the
// "STREAM copy", also a plain copy of known size for calibrating byte
// counters) and read-weighted variants (R reads per write, or reads only),
// since the hopping kernels read several times what they write; and a
// gather probe of the hopping kernels' access shape.
#include "gbx_internal.h"

namespace {

typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// Each thread moves 4 x 16 B per iteration (4 loads issued before the
// stores) so enough bytes are in flight per CU to cover HBM latency.
__global__ void __launch_bounds__(256) k_copy16(const u4 *__restrict__ src, u4 *__restrict__ dst, uint64_t n)
{
  constexpr int U = 4;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    u4 v[U];
#pragma unroll
    for (int k = 0; k < U; k++) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
    for (int k = 0; k < U; k++) __builtin_nontemporal_store(v[k], dst + i + k * stride);
  }
  for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Read-weighted probes.  R > 0: every written 16-byte element is the XOR of
// R elements read from R separate planes of the source, all R loads issued
// before the store (R : 1 read : write bytes).  R == 0: read only; each
// thread folds its loads into one value and writes it once at the end
// (gridDim * blockDim * 16 bytes of writes, negligible).
template <int R>
__global__ void __launch_bounds__(256) k_readmix(const u4 *__restrict__ src, u4 *__restrict__ dst, uint64_t nout)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if constexpr (R == 0) {
    constexpr int U = 8;
    u4 acc = {0u, 0u, 0u, 0u};
    uint64_t i = t0;
    for (; i + (U - 1) * stride < nout; i += U * stride) {
      u4 v[U];
#pragma unroll
      for (int k = 0; k < U; k++) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
      for (int k = 0; k < U; k++) acc ^= v[k];
    }
    for (; i < nout; i += stride) acc ^= __builtin_nontemporal_load(src + i);
    __builtin_nontemporal_store(acc, dst + t0);
  } else {
    for (uint64_t i = t0; i < nout; i += stride) {
      u4 v[R];
#pragma unroll
      for (int k = 0; k < R; k++) v[k] = __builtin_nontemporal_load(src + i + (uint64_t)k * nout);
      u4 acc = v[0];
#pragma unroll
      for (int k = 1; k < R; k++) acc ^= v[k];
      __builtin_nontemporal_store(acc, dst + i);
    }
  }
}

// Gather probe: each wavefront reads 4 runs of 256 contiguous bytes (16
// lanes x 16 B) at pseudo-random 256-byte-aligned places of a buffer of
// power-of-two size -- the access shape of a hopping kernel's neighbour
// gather (one chunk of 4 sites x 16 minor indices) without its arithmetic.
// U loads in flight per thread; each thread folds its loads into one value,
// written once.  Every index is masked into the buffer.
template <int U>
__global__ void __launch_bounds__(256) k_gather_runs(const u4 *__restrict__ src, u4 *__restrict__ dst, uint64_t mask,
                                                     int iters)
{
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t h = (uint32_t)((gid >> 6) * 2654435761u) ^ 0x9e3779b9u;  // the same in every lane of a wavefront
  u4 acc = {0u, 0u, 0u, 0u};
  for (int it = 0; it < iters; it++) {
    u4 v[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
      h = h * 1664525u + 1013904223u;
      const uint64_t run = (uint64_t)(h >> 4) + (uint64_t)(lane >> 4) * 0x9E37u;
      v[k] = src[(run * 16 + (lane & 15)) & mask];
    }
#pragma unroll
    for (int k = 0; k < U; k++) acc ^= v[k];
  }
  dst[gid] = acc;
}

void launch_readmix(int mix, dim3 grid, hipStream_t st, const u4 *src, u4 *dst, uint64_t nout)
{
  switch (mix) {
  case 0: hipLaunchKernelGGL(k_readmix<0>, grid, dim3(256), 0, st, src, dst, nout); break;
  case 1: hipLaunchKernelGGL(k_readmix<1>, grid, dim3(256), 0, st, src, dst, nout); break;
  case 2: hipLaunchKernelGGL(k_readmix<2>, grid, dim3(256), 0, st, src, dst, nout); break;
  case 3: hipLaunchKernelGGL(k_readmix<3>, grid, dim3(256), 0, st, src, dst, nout); break;
  case 4: hipLaunchKernelGGL(k_readmix<4>, grid, dim3(256), 0, st, src, dst, nout); break;
  default: hipLaunchKernelGGL(k_readmix<8>, grid, dim3(256), 0, st, src, dst, nout); break;
  }
}

}  // namespace

extern "C" int gbx_stream_bw(uint64_t read_bytes, int mix, int nrep, double *gbps)
{
  GBX_CHECK(gbps, GBX_ENULL, "gbx_stream_bw: NULL result");
  GBX_CHECK(mix == 0 || mix == 1 || mix == 2 || mix == 3 || mix == 4 || mix == 8, GBX_EINVAL,
            "gbx_stream_bw: mix=%d (reads per write: 1, 2, 3, 4 or 8; 0 = read only)", mix);
  const uint64_t unit = 16ull * (mix ? (uint64_t)mix : 1ull);
  GBX_CHECK(read_bytes >= 64 * 1024 && read_bytes % unit == 0, GBX_EINVAL, "gbx_stream_bw: read_bytes=%llu",
            (unsigned long long)read_bytes);
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "gbx_stream_bw: nrep=%d", nrep);
  GBX_TRY(gbx_require_device());
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const dim3 grid((unsigned)(ncu * 16));
  // R > 0: nout = read_bytes / (16 R) written elements; R == 0: one per thread
  const uint64_t nread = read_bytes / 16;
  const uint64_t nout = mix ? nread / (uint64_t)mix : nread;
  const uint64_t out_bytes = mix ? nout * 16 : (uint64_t)grid.x * 256 * 16;
  void *a = nullptr, *b = nullptr;
  GBX_HIP(hipMalloc(&a, read_bytes));
  if (hipMalloc(&b, out_bytes) != hipSuccess) {
    (void)hipFree(a);
    return gbx_fail(GBX_ENOMEM, "gbx_stream_bw: cannot allocate %llu bytes", (unsigned long long)out_bytes);
  }
  int rc = GBX_OK;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipStreamCreate failed");
  if (rc == GBX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = gbx_fail(GBX_EHIP, "hipEventCreate failed");
  if (rc == GBX_OK) {
    if (hipMemsetAsync(a, 0x3c, read_bytes, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipMemsetAsync failed");
    if (rc == GBX_OK) launch_readmix(mix, grid, st, (const u4 *)a, (u4 *)b, nout);  // warm-up
    if (rc == GBX_OK && hipEventRecord(e0, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    for (int r = 0; r < nrep && rc == GBX_OK; r++) launch_readmix(mix, grid, st, (const u4 *)a, (u4 *)b, nout);
    if (rc == GBX_OK && hipEventRecord(e1, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "read-mix kernel failed");
    float ms = 0.f;
    if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventElapsedTime failed");
    if (rc == GBX_OK) *gbps = (double)(read_bytes + out_bytes) * nrep / (ms * 1e-3) / 1e9;
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

extern "C" int gbx_gather_bw(uint64_t buffer_bytes, int nrep, double *gbps)
{
  GBX_CHECK(gbps, GBX_ENULL, "gbx_gather_bw: NULL result");
  GBX_CHECK(buffer_bytes >= (1ull << 20) && (buffer_bytes & (buffer_bytes - 1)) == 0, GBX_EINVAL,
            "gbx_gather_bw: buffer_bytes=%llu (a power of two, at least 1 MiB)", (unsigned long long)buffer_bytes);
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "gbx_gather_bw: nrep=%d", nrep);
  GBX_TRY(gbx_require_device());
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  constexpr int U = 8, iters = 64;
  const dim3 grid((unsigned)(ncu * 8));
  const uint64_t nthreads = (uint64_t)grid.x * 256, per_launch = nthreads * iters * U * 16;
  void *a = nullptr, *b = nullptr;
  GBX_HIP(hipMalloc(&a, buffer_bytes));
  if (hipMalloc(&b, nthreads * 16) != hipSuccess) {
    (void)hipFree(a);
    return gbx_fail(GBX_ENOMEM, "gbx_gather_bw: cannot allocate the result array");
  }
  int rc = GBX_OK;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipStreamCreate failed");
  if (rc == GBX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = gbx_fail(GBX_EHIP, "hipEventCreate failed");
  if (rc == GBX_OK) {
    const uint64_t mask = buffer_bytes / 16 - 1;
    auto launch = [&]() {
      hipLaunchKernelGGL((k_gather_runs<U>), grid, dim3(256), 0, st, (const u4 *)a, (u4 *)b, mask, iters);
    };
    if (hipMemsetAsync(a, 0x3c, buffer_bytes, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipMemsetAsync failed");
    if (rc == GBX_OK) launch();  // warm-up
    if (rc == GBX_OK && hipEventRecord(e0, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    for (int r = 0; r < nrep && rc == GBX_OK; r++) launch();
    if (rc == GBX_OK && hipEventRecord(e1, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "gather kernel failed");
    float ms = 0.f;
    if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventElapsedTime failed");
    if (rc == GBX_OK) *gbps = (double)per_launch * nrep / (ms * 1e-3) / 1e9;
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

extern "C" int gbx_stream_copy_bw(uint64_t bytes, int nrep, double *gbps)
{
  GBX_CHECK(gbps, GBX_ENULL, "gbx_stream_copy_bw: NULL result");
  GBX_CHECK(bytes >= 16 * 1024 && bytes % 16 == 0, GBX_EINVAL, "gbx_stream_copy_bw: bytes=%llu",
            (unsigned long long)bytes);
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "gbx_stream_copy_bw: nrep=%d", nrep);
  GBX_TRY(gbx_require_device());
  void *a = nullptr, *b = nullptr;
  GBX_HIP(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return gbx_fail(GBX_ENOMEM, "gbx_stream_copy_bw: cannot allocate %llu bytes", (unsigned long long)bytes);
  }
  int rc = GBX_OK;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipStreamCreate failed");
  if (rc == GBX_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = gbx_fail(GBX_EHIP, "hipEventCreate failed");
  if (rc == GBX_OK) {
    const uint64_t n = bytes / 16;
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const dim3 grid((unsigned)(ncu * 16)), block(256);
    if (hipMemsetAsync(a, 0x3c, bytes, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipMemsetAsync failed");
    if (rc == GBX_OK) hipLaunchKernelGGL(k_copy16, grid, block, 0, st, (const u4 *)a, (u4 *)b, n);  // warm-up
    if (rc == GBX_OK && hipEventRecord(e0, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    for (int r = 0; r < nrep && rc == GBX_OK; r++) hipLaunchKernelGGL(k_copy16, grid, block, 0, st, (const u4 *)a, (u4 *)b, n);
    if (rc == GBX_OK && hipEventRecord(e1, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "copy kernel failed");
    float ms = 0.f;
    if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventElapsedTime failed");
    if (rc == GBX_OK) *gbps = 2.0 * (double)bytes * nrep / (ms * 1e-3) / 1e9;
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
