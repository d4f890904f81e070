// Vector algebra on device spinors, used by the solver.
//
// Reductions are deterministic: a fixed grid (kRedBlocks x 256 threads) walks
// the field in a fixed order, each block reduces in a fixed LDS tree, and a
// single block combines the partials in a fixed tree -- the same bits on
// every run, accumulated in FP64 whatever the storage precision.
// Zeroing uses a fill kernel (never hipMemset), so the work can be captured
// in a HIP graph and replayed.
#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

template <int P>
__device__ __forceinline__ void load_cplx_chunk(const void *b, uint64_t t, double *v)
{
  dev::load_chunk<P, double>(b, t, v);
}

__device__ __forceinline__ void block_reduce2(double &a, double &b, double *sh)
{
  sh[threadIdx.x] = a;
  sh[kRedThreads + threadIdx.x] = b;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[threadIdx.x] += sh[threadIdx.x + s];
      sh[kRedThreads + threadIdx.x] += sh[kRedThreads + threadIdx.x + s];
    }
    __syncthreads();
  }
  a = sh[0];
  b = sh[kRedThreads];
}

// OP 0: (|x|^2, 0); OP 1: (Re<x,y>, Im<x,y>) with <x,y> = sum conj(x) y.
template <int P, int OP>
__global__ void __launch_bounds__(kRedThreads) k_reduce(const void *__restrict__ x, const void *__restrict__ y,
                                                        uint64_t nchunk, double *__restrict__ partial)
{
  __shared__ double sh[2 * kRedThreads];
  constexpr int R = dev::Traits<P>::R;
  double a = 0.0, b = 0.0;
  for (uint64_t t = blockIdx.x * (uint64_t)kRedThreads + threadIdx.x; t < nchunk; t += (uint64_t)gridDim.x * kRedThreads) {
    double xv[R];
    load_cplx_chunk<P>(x, t, xv);
    if (OP == 0) {
#pragma unroll
      for (int j = 0; j < R; j++) a += xv[j] * xv[j];
    } else {
      double yv[R];
      load_cplx_chunk<P>(y, t, yv);
#pragma unroll
      for (int j = 0; j < R; j += 2) {
        a += xv[j] * yv[j] + xv[j + 1] * yv[j + 1];
        b += xv[j] * yv[j + 1] - xv[j + 1] * yv[j];
      }
    }
  }
  block_reduce2(a, b, sh);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

__global__ void __launch_bounds__(kRedThreads) k_reduce_final(const double *__restrict__ partial, int n, double *__restrict__ out)
{
  __shared__ double sh[2 * kRedThreads];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedThreads) {
    a += partial[2 * i];
    b += partial[2 * i + 1];
  }
  block_reduce2(a, b, sh);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

// y = a*x + y, x = a*x, x = 0 on chunks.
template <int P, int OP>
__global__ void __launch_bounds__(256) k_axpy(double a, const void *__restrict__ x, void *__restrict__ y, uint64_t nchunk)
{
  constexpr int R = dev::Traits<P>::R;
  typedef typename dev::Traits<P>::acc A;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nchunk; t += (uint64_t)gridDim.x * blockDim.x) {
    A v[R];
    if (OP == 0) {  // axpy
      A xv[R];
      dev::load_chunk<P, A>(x, t, xv);
      dev::load_chunk<P, A>(y, t, v);
#pragma unroll
      for (int j = 0; j < R; j++) v[j] = (A)a * xv[j] + v[j];
    } else if (OP == 1) {  // scale (y is the field)
      dev::load_chunk<P, A>(y, t, v);
#pragma unroll
      for (int j = 0; j < R; j++) v[j] = (A)a * v[j];
    } else {  // zero
#pragma unroll
      for (int j = 0; j < R; j++) v[j] = (A)0;
    }
    dev::store_chunk<P, A>(y, t, v, false);
  }
}

// Precision conversion between SoA layouts: thread = (slot, m), 24 reals.
template <int PD, int PS>
__global__ void __launch_bounds__(256) k_convert_spinor(const void *__restrict__ src, void *__restrict__ dst, uint64_t npts,
                                                        uint32_t M, double scale)
{
  constexpr int RS = dev::Traits<PS>::R, NS = dev::Traits<PS>::NCH;
  constexpr int RD = dev::Traits<PD>::R, ND = dev::Traits<PD>::NCH;
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < npts; p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = p / M, m = p % M;
    double v[24];
#pragma unroll
    for (int k = 0; k < NS; k++) dev::load_chunk<PS, double>(src, (q * NS + k) * M + m, v + k * RS);
#pragma unroll
    for (int j = 0; j < 24; j++) v[j] *= scale;
#pragma unroll
    for (int k = 0; k < ND; k++) dev::store_chunk<PD, double>(dst, (q * ND + k) * M + m, v + k * RD, false);
  }
}

__device__ __forceinline__ uint64_t dmix64(uint64_t z)
{
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Gaussian fill: thread = (slot, m); the 12 pairs of its spinor are keyed by
// the lexicographic 5D index, so the field does not depend on the site order.
template <int P>
__global__ void __launch_bounds__(256) k_fill_gaussian(void *__restrict__ dst, const uint32_t *__restrict__ site_of_slot,
                                                       uint64_t npts, uint32_t M, uint32_t Ls, uint64_t V, uint64_t key)
{
  constexpr int R = dev::Traits<P>::R, NCH = dev::Traits<P>::NCH;
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < npts; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = t / M;
    const uint32_t m = (uint32_t)(t % M);
    const uint64_t base = (((uint64_t)(m / Ls) * V + site_of_slot[q]) * Ls + m % Ls) * 12;
    double v[24];
#pragma unroll
    for (int j = 0; j < 12; j++) {
      const uint64_t h = dmix64(key ^ dmix64(base + (uint64_t)j));
      const double u1 = ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      const double u2 = (double)(dmix64(h) >> 11) * (1.0 / 9007199254740992.0);
      const double r = sqrt(-2.0 * log(u1));
      double s, c;
      sincos(6.283185307179586 * u2, &s, &c);
      v[2 * j] = r * c;
      v[2 * j + 1] = r * s;
    }
#pragma unroll
    for (int k = 0; k < NCH; k++) dev::store_chunk<P, double>(dst, (q * NCH + k) * M + m, v + k * R, false);
  }
}

uint64_t dmix64_host(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint32_t grid_for(uint64_t n, uint32_t threads = 256)
{
  uint64_t b = (n + threads - 1) / threads;
  if (b > 8192) b = 8192;
  return (uint32_t)(b ? b : 1);
}

uint64_t nchunk(const gbx_spinor_s *x) { return x->geom->vol * (uint64_t)x->M * (uint64_t)spinor_nchunk(x->prec); }

// Per-device scratch for reduction partials and the pinned result.
struct RedScratch {
  double *partial = nullptr;
  double *result = nullptr;  // device
  double *host = nullptr;    // pinned host
};
thread_local RedScratch g_red;

int red_scratch()
{
  if (!g_red.partial) {
    GBX_HIP(hipMalloc(&g_red.partial, sizeof(double) * 2 * kRedBlocks));
    GBX_HIP(hipMalloc(&g_red.result, sizeof(double) * 2));
    GBX_HIP(hipHostMalloc(&g_red.host, sizeof(double) * 2, hipHostMallocDefault));
  }
  return GBX_OK;
}

template <int OP>
int reduce(const gbx_spinor_s *x, const gbx_spinor_s *y, double *a, double *b, hipStream_t st)
{
  GBX_TRY(red_scratch());
  const uint64_t n = nchunk(x);
  switch (x->prec) {
  case GBX_F64: hipLaunchKernelGGL((k_reduce<GBX_F64, OP>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, x->d, y ? y->d : nullptr, n, g_red.partial); break;
  case GBX_F32: hipLaunchKernelGGL((k_reduce<GBX_F32, OP>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, x->d, y ? y->d : nullptr, n, g_red.partial); break;
  case GBX_F16: hipLaunchKernelGGL((k_reduce<GBX_F16, OP>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, x->d, y ? y->d : nullptr, n, g_red.partial); break;
  default: hipLaunchKernelGGL((k_reduce<GBX_BF16, OP>), dim3(kRedBlocks), dim3(kRedThreads), 0, st, x->d, y ? y->d : nullptr, n, g_red.partial); break;
  }
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRedThreads), 0, st, (const double *)g_red.partial, kRedBlocks, g_red.result);
  GBX_HIP(hipGetLastError());
  GBX_HIP(hipMemcpyAsync(g_red.host, g_red.result, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  GBX_HIP(hipStreamSynchronize(st));
  *a = g_red.host[0];
  if (b) *b = g_red.host[1];
  return GBX_OK;
}

template <int OP>
int axpy_like(double a, const gbx_spinor_s *x, gbx_spinor_s *y, hipStream_t st)
{
  const uint64_t n = nchunk(y);
  const dim3 grid(grid_for(n)), block(256);
  const void *xd = x ? x->d : nullptr;
  switch (y->prec) {
  case GBX_F64: hipLaunchKernelGGL((k_axpy<GBX_F64, OP>), grid, block, 0, st, a, xd, y->d, n); break;
  case GBX_F32: hipLaunchKernelGGL((k_axpy<GBX_F32, OP>), grid, block, 0, st, a, xd, y->d, n); break;
  case GBX_F16: hipLaunchKernelGGL((k_axpy<GBX_F16, OP>), grid, block, 0, st, a, xd, y->d, n); break;
  default: hipLaunchKernelGGL((k_axpy<GBX_BF16, OP>), grid, block, 0, st, a, xd, y->d, n); break;
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

template <int PD>
void convert_to(int ps, const void *src, void *dst, uint64_t npts, uint32_t M, double scale, hipStream_t st)
{
  const dim3 grid(grid_for(npts)), block(256);
  switch (ps) {
  case GBX_F64: hipLaunchKernelGGL((k_convert_spinor<PD, GBX_F64>), grid, block, 0, st, src, dst, npts, M, scale); break;
  case GBX_F32: hipLaunchKernelGGL((k_convert_spinor<PD, GBX_F32>), grid, block, 0, st, src, dst, npts, M, scale); break;
  case GBX_F16: hipLaunchKernelGGL((k_convert_spinor<PD, GBX_F16>), grid, block, 0, st, src, dst, npts, M, scale); break;
  default: hipLaunchKernelGGL((k_convert_spinor<PD, GBX_BF16>), grid, block, 0, st, src, dst, npts, M, scale); break;
  }
}

int same_shape(const gbx_spinor_s *x, const gbx_spinor_s *y, const char *who)
{
  GBX_CHECK(x && y && x->d && y->d, GBX_ENULL, "%s: NULL spinor", who);
  GBX_CHECK(x->geom == y->geom && x->M == y->M, GBX_EMISMATCH, "%s: different geometry or Ls*nrhs", who);
  return GBX_OK;
}

}  // namespace

// -------------------------------------------------------- internal (solver)
int spinor_convert_scaled(gbx_spinor_s *dst, const gbx_spinor_s *src, double scale, hipStream_t st)
{
  GBX_TRY(same_shape(dst, src, "spinor_convert"));
  const uint64_t npts = dst->geom->vol * (uint64_t)dst->M;
  switch (dst->prec) {
  case GBX_F64: convert_to<GBX_F64>(src->prec, src->d, dst->d, npts, dst->M, scale, st); break;
  case GBX_F32: convert_to<GBX_F32>(src->prec, src->d, dst->d, npts, dst->M, scale, st); break;
  case GBX_F16: convert_to<GBX_F16>(src->prec, src->d, dst->d, npts, dst->M, scale, st); break;
  default: convert_to<GBX_BF16>(src->prec, src->d, dst->d, npts, dst->M, scale, st); break;
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

int spinor_norm2_on(const gbx_spinor_s *x, double *n2, hipStream_t st) { return reduce<0>(x, nullptr, n2, nullptr, st); }
int spinor_dot_on(const gbx_spinor_s *x, const gbx_spinor_s *y, double *re, double *im, hipStream_t st)
{
  return reduce<1>(x, y, re, im, st);
}
int spinor_axpy_on(double a, const gbx_spinor_s *x, gbx_spinor_s *y, hipStream_t st) { return axpy_like<0>(a, x, y, st); }
int spinor_scale_on(gbx_spinor_s *x, double a, hipStream_t st) { return axpy_like<1>(a, nullptr, x, st); }
int spinor_zero_on(gbx_spinor_s *x, hipStream_t st) { return axpy_like<2>(0.0, nullptr, x, st); }

// --------------------------------------------------------------------- C-ABI
extern "C" int gbx_spinor_norm2(gbx_spinor x, double *n2)
{
  GBX_CHECK(x && x->d && n2, GBX_ENULL, "gbx_spinor_norm2: NULL argument");
  return spinor_norm2_on(x, n2, x->geom->stream);
}

extern "C" int gbx_spinor_dot(gbx_spinor x, gbx_spinor y, double *re, double *im)
{
  GBX_CHECK(re && im, GBX_ENULL, "gbx_spinor_dot: NULL result");
  GBX_TRY(same_shape(x, y, "gbx_spinor_dot"));
  GBX_CHECK(x->prec == y->prec, GBX_EMISMATCH, "gbx_spinor_dot: precisions differ");
  GBX_CHECK(x->basis == y->basis, GBX_EMISMATCH, "gbx_spinor_dot: spin bases differ");
  return spinor_dot_on(x, y, re, im, x->geom->stream);
}

extern "C" int gbx_spinor_axpy(double a, gbx_spinor x, gbx_spinor y)
{
  GBX_TRY(same_shape(x, y, "gbx_spinor_axpy"));
  GBX_CHECK(x->prec == y->prec, GBX_EMISMATCH, "gbx_spinor_axpy: precisions differ");
  GBX_CHECK(x->basis == y->basis, GBX_EMISMATCH, "gbx_spinor_axpy: spin bases differ");
  return spinor_axpy_on(a, x, y, x->geom->stream);
}

extern "C" int gbx_spinor_scale(gbx_spinor x, double a)
{
  GBX_CHECK(x && x->d, GBX_ENULL, "gbx_spinor_scale: NULL spinor");
  return spinor_scale_on(x, a, x->geom->stream);
}

extern "C" int gbx_spinor_fill_gaussian(gbx_spinor x, uint64_t seed)
{
  GBX_CHECK(x && x->d, GBX_ENULL, "gbx_spinor_fill_gaussian: NULL spinor");
  const uint64_t npts = x->geom->vol * (uint64_t)x->M;
  const uint64_t key = dmix64_host(seed);
  const dim3 grid(grid_for(npts)), block(256);
  hipStream_t st = x->geom->stream;
  const uint32_t *sos = x->geom->site_of_slot;
  switch (x->prec) {
  case GBX_F64: hipLaunchKernelGGL(k_fill_gaussian<GBX_F64>, grid, block, 0, st, x->d, sos, npts, x->M, (uint32_t)x->Ls, x->geom->vol, key); break;
  case GBX_F32: hipLaunchKernelGGL(k_fill_gaussian<GBX_F32>, grid, block, 0, st, x->d, sos, npts, x->M, (uint32_t)x->Ls, x->geom->vol, key); break;
  case GBX_F16: hipLaunchKernelGGL(k_fill_gaussian<GBX_F16>, grid, block, 0, st, x->d, sos, npts, x->M, (uint32_t)x->Ls, x->geom->vol, key); break;
  default: hipLaunchKernelGGL(k_fill_gaussian<GBX_BF16>, grid, block, 0, st, x->d, sos, npts, x->M, (uint32_t)x->Ls, x->geom->vol, key); break;
  }
  GBX_HIP(hipGetLastError());
  x->basis = GBX_BASIS_CHIRAL;  // the values are those of a chiral-basis field
  return GBX_OK;
}

extern "C" int gbx_spinor_zero(gbx_spinor x)
{
  GBX_CHECK(x && x->d, GBX_ENULL, "gbx_spinor_zero: NULL spinor");
  return spinor_zero_on(x, x->geom->stream);
}
