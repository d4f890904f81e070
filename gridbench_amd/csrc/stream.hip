// Achievable-HBM-bandwidth probe: a grid-stride 16-byte copy kernel, the
// "STREAM copy" against which the hopping kernels' roofline fraction is
// read, and a plain copy of known size for calibrating byte counters.
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

}  // namespace

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
