// Drop-in replacement for the reference entry point
//   template<class Simd> double dslash_kernel(int nrep, Simd *U, Simd *out,
//       Simd *in, uint64_t *nbr, uint64_t nsite, uint64_t Ls, uint8_t *prm)
// (WilsonKernelsHand.h:17-29), on host arrays in the reference layout.
//
// Both device paths compute with the reference's rounding (no fused
// multiply-adds, its order of operations), so the result is the reference
// CPU kernel's to the bit:
//  * canonical tables (recognised by gbx_table_recognise): the fields are
//    repacked into the SoA layout and the VALU kernel (dhop.hip) runs;
//  * any other valid tables: a table-driven kernel on the reference layout
//    honours nbr/prm exactly: thread = (5D outer site, lane), neighbour lane
//    = lane ^ (DATA_SIMD >> (ptype+1)) when prm is set (the reference GPU
//    kernel's gather form, WilsonKernelsHandGpu.h:33-41).
// Host arrays may be RIRI (re,im per lane) or RRII (all real parts, then
// all imaginary parts, per vector; bench.cc:105-109), and vectors may be
// expanded: nsimd lanes holding nsimd/dsimd replicas of a dsimd-lane lattice
// (bench.cc:1-3, 98-138), run as independent problems.
// Device buffers and objects live in a context (gbx_dslash_ctx): one per
// caller that wants its own stream, or the process-wide one behind the
// context-free entries.
// Timing follows the reference: nrep applications, the clock covering
// applications 1..nrep-1 (WilsonKernelsHandCpu.h:407-450), measured here
// with HIP events on the stream the kernels run on.
#include <cstring>
#include <mutex>

#include "dhop_dev.h"
#include "gbx_internal.h"

int64_t table_first_invalid(const uint64_t *nbr, const uint8_t *prm, uint64_t n, uint64_t n5);

namespace {

thread_local int g_last_path = 0;

// Real ri of lane `lane` of complex element c of `width`-lane vectors: RIRI
// [c][lane][ri] or RRII [c][ri][lane] (bench.cc:105-109).
template <bool RRII>
__device__ __forceinline__ uint64_t rpos(uint64_t c, int lane, int ri, int width)
{
  return c * 2 * (uint64_t)width + (RRII ? (uint64_t)ri * width + lane : (uint64_t)lane * 2 + ri);
}

// Thread = (5D outer site, lane) over `width`-lane vectors.  The permute
// mask comes from the data SIMD width dsimd <= width, so an expanded vector
// is width/dsimd independent copies of the dsimd-lane problem
// (Simd_riri_vec.h:466, "keep the permutes fixed as SIMD expanded").
template <typename T, bool DAG, bool RRII>
__global__ void __launch_bounds__(256) k_dhop_table(const T *__restrict__ U, T *__restrict__ out, const T *__restrict__ in,
                                                    const uint64_t *__restrict__ nbr, const uint8_t *__restrict__ prm,
                                                    uint64_t nsite, uint64_t Ls, int nsimd, int dsimd)
{
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (t >= nsite * Ls * (uint64_t)nsimd) return;
  const int lane = (int)(t % (uint64_t)nsimd);
  const uint64_t ss = t / (uint64_t)nsimd;
  const uint64_t sU = ss / Ls;
  T rr[12], ri[12];
  // leg order and projector pairs as dhop_site (WilsonKernelsHandCpu.h:435-442)
#define GBX_TLEG(DIR, PTYPE, FIRST)                                                                                      \
  {                                                                                                                      \
    const uint64_t off = nbr[ss * 8 + DIR];                                                                              \
    const int plane = prm[ss * 8 + DIR] ? (lane ^ (dsimd >> (PTYPE + 1))) : lane;                                        \
    T v[24];                                                                                                             \
    _Pragma("unroll") for (int sc = 0; sc < 12; sc++)                                                                    \
    {                                                                                                                    \
      v[2 * sc] = in[rpos<RRII>(off * 12 + sc, plane, 0, nsimd)];                                                        \
      v[2 * sc + 1] = in[rpos<RRII>(off * 12 + sc, plane, 1, nsimd)];                                                    \
    }                                                                                                                    \
    T Ul[18];                                                                                                            \
    _Pragma("unroll") for (int e = 0; e < 9; e++)                                                                        \
    {                                                                                                                    \
      Ul[2 * e] = U[rpos<RRII>((sU * 8 + DIR) * 9 + e, lane, 0, nsimd)];                                                 \
      Ul[2 * e + 1] = U[rpos<RRII>((sU * 8 + DIR) * 9 + e, lane, 1, nsimd)];                                             \
    }                                                                                                                    \
    T hr[6], hi[6], ur[6], ui[6];                                                                                        \
    dev::project<dev::LegProj<DAG, DIR>::value>(v, hr, hi);                                                              \
    dev::su3_mul<dev::MUL_EXACT>(Ul, hr, hi, ur, ui);                                                                        \
    dev::recon<dev::LegProj<DAG, DIR>::value, FIRST>(ur, ui, rr, ri);                                                    \
  }
  GBX_TLEG(0, 3, true)
  GBX_TLEG(1, 2, false)
  GBX_TLEG(2, 1, false)
  GBX_TLEG(3, 0, false)
  GBX_TLEG(4, 3, false)
  GBX_TLEG(5, 2, false)
  GBX_TLEG(6, 1, false)
  GBX_TLEG(7, 0, false)
#undef GBX_TLEG
#pragma unroll
  for (int sc = 0; sc < 12; sc++) {
    out[rpos<RRII>(ss * 12 + sc, lane, 0, nsimd)] = rr[sc];
    out[rpos<RRII>(ss * 12 + sc, lane, 1, nsimd)] = ri[sc];
  }
}

bool overlaps(const void *a, size_t na, const void *b, size_t nb)
{
  const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
  return a0 < b0 + nb && b0 < a0 + na;
}

}  // namespace

// A drop-in context: the device buffers of the table path and the SoA
// objects of the fast path, cached across calls with the same shape (the
// reference benchmark calls once with nrep = 1000; tests call many times with
// small shapes), and the stream its work runs on.  Contexts share nothing,
// so calls on different contexts may run from different host threads.
struct gbx_dslash_ctx_s {
  std::mutex mu;
  void *dU = nullptr, *dIn = nullptr, *dOut = nullptr, *dNbr = nullptr, *dPrm = nullptr;
  size_t nU = 0, nIn = 0, nOut = 0, nNbr = 0, nPrm = 0;
  gbx_geom geom = nullptr;
  gbx_gauge gauge = nullptr;
  gbx_spinor sin = nullptr, sout = nullptr;
  int dims[4] = {0, 0, 0, 0}, Ls = 0, prec = -1;
  hipStream_t st = nullptr;  // table path; the fast path runs on its geometry's stream
  int last_path = 0;
};

namespace {

int ensure(void **p, size_t *have, size_t need)
{
  if (*have >= need && *p) return GBX_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  GBX_HIP(hipMalloc(p, need));
  *have = need;
  return GBX_OK;
}

int fast_objects(gbx_dslash_ctx_s &c, const int dims[4], int Ls, int prec)
{
  if (c.geom && c.prec == prec && c.Ls == Ls && !memcmp(c.dims, dims, sizeof c.dims)) return GBX_OK;
  gbx_spinor_destroy(c.sin);
  gbx_spinor_destroy(c.sout);
  gbx_gauge_destroy(c.gauge);
  gbx_geom_destroy(c.geom);
  c.sin = c.sout = nullptr;
  c.gauge = nullptr;
  c.geom = nullptr;
  c.prec = -1;
  GBX_TRY(gbx_geom_create(dims, nullptr, &c.geom));
  GBX_TRY(gbx_gauge_create(c.geom, prec, 0, &c.gauge));
  GBX_TRY(gbx_spinor_create(c.geom, Ls, 1, prec, &c.sin));
  GBX_TRY(gbx_spinor_create(c.geom, Ls, 1, prec, &c.sout));
  memcpy(c.dims, dims, sizeof c.dims);
  c.Ls = Ls;
  c.prec = prec;
  return GBX_OK;
}

template <typename T> struct RefPrec;
template <> struct RefPrec<double> {
  static constexpr int prec = GBX_F64;
};
template <> struct RefPrec<float> {
  static constexpr int prec = GBX_F32;
};

// Time nrep launches of `launch` on stream st, reference convention.
template <typename F> int timed_reps(int nrep, hipStream_t st, F launch, double *usec)
{
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GBX_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = gbx_fail(GBX_EHIP, "hipEventCreate failed");
  for (int rep = 0; rep < nrep && rc == GBX_OK; rep++) {
    if (rep == (nrep > 1 ? 1 : 0) && hipEventRecord(e0, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
    if (rc == GBX_OK) rc = launch();
  }
  if (rc == GBX_OK && hipEventRecord(e1, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
  if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "kernel failed: %s",
                                                                         hipGetErrorString(hipGetLastError()));
  float ms = 0.f;
  if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventElapsedTime failed");
  if (rc != GBX_OK) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == GBX_OK && usec) *usec = 1000.0 * (double)ms;
  return rc;
}

bool pow2(int n) { return n >= 1 && (n & (n - 1)) == 0; }

// nsimd = lanes per host vector (EXPAND_SIMD), dsimd = lanes the tables and
// permutes are for (DATA_SIMD, bench.cc:1-3); rrii = the RRII lane layout.
template <typename T>
int dslash_kernel_impl(gbx_dslash_ctx_s &c, int nrep, int nsimd, int dsimd, int rrii, int dag, const T *U, T *out,
                       const T *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec)
{
  GBX_CHECK(U && out && in && nbr && prm, GBX_ENULL, "dslash_kernel: NULL array");
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "dslash_kernel: nrep=%d", nrep);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "dslash_kernel: dag=%d", dag);
  GBX_CHECK(rrii == 0 || rrii == 1, GBX_EINVAL, "dslash_kernel: rrii=%d", rrii);
  GBX_CHECK(pow2(dsimd) && dsimd <= 16, GBX_EINVAL, "dslash_kernel: data nsimd=%d not in {1,2,4,8,16}", dsimd);
  GBX_CHECK(pow2(nsimd) && nsimd >= dsimd && nsimd <= 64, GBX_EINVAL,
            "dslash_kernel: vector nsimd=%d must be a power of two in [%d, 64]", nsimd, dsimd);
  GBX_CHECK(nsite >= 1 && Ls >= 1 && Ls <= 4096 && nsite * Ls <= (1ull << 34), GBX_EINVAL,
            "dslash_kernel: nsite=%llu Ls=%llu", (unsigned long long)nsite, (unsigned long long)Ls);
  const uint64_t n5 = nsite * Ls;
  const size_t nf = (size_t)n5 * 24 * nsimd * sizeof(T), nu = (size_t)nsite * 144 * nsimd * sizeof(T);
  const size_t nt = (size_t)n5 * 8;
  GBX_CHECK(!overlaps(out, nf, in, nf) && !overlaps(out, nf, U, nu) && !overlaps(out, nf, nbr, nt * 8) &&
                !overlaps(out, nf, prm, nt),
            GBX_EALIAS, "dslash_kernel: out overlaps an input array");
  // Validate the tables (the reference trusts them blindly).
  const int64_t bad = table_first_invalid(nbr, prm, nt, n5);
  GBX_CHECK(bad < 0, GBX_ETABLE, "dslash_kernel: table entry %lld invalid (nbr=%llu, prm=%u, nsite*Ls=%llu)",
            (long long)bad, (unsigned long long)nbr[bad < 0 ? 0 : bad], (unsigned)prm[bad < 0 ? 0 : bad],
            (unsigned long long)n5);
  GBX_TRY(gbx_require_device());
  if (!c.st) GBX_HIP(hipStreamCreate(&c.st));
  GBX_TRY(ensure(&c.dU, &c.nU, nu));
  GBX_TRY(ensure(&c.dIn, &c.nIn, nf));
  GBX_TRY(ensure(&c.dOut, &c.nOut, nf));
  int dims[4];
  const int canon = gbx_table_recognise(dsimd, nsite, Ls, nbr, prm, dims);
  gbx_clear_error();
  if (canon == GBX_OK) {
    // The SoA kernel, once per replica of the dsimd-lane lattice.
    GBX_TRY(fast_objects(c, dims, (int)Ls, RefPrec<T>::prec));
    hipStream_t st = c.geom->stream;
    GBX_HIP(hipMemcpyAsync(c.dU, U, nu, hipMemcpyHostToDevice, st));
    GBX_HIP(hipMemcpyAsync(c.dIn, in, nf, hipMemcpyHostToDevice, st));
    double total = 0.0;
    for (int nn = 0; nn < nsimd / dsimd; nn++) {
      gbx_ref_layout lay;
      lay.width = nsimd;
      lay.lane0 = nn * dsimd;
      lay.rrii = rrii;
      GBX_TRY(ref_pack_gauge_dev(c.gauge, (const T *)c.dU, dsimd, lay));
      GBX_TRY(ref_pack_spinor_dev(c.sin, (const T *)c.dIn, dsimd, lay));
      double us = 0.0;
      GBX_TRY(timed_reps(nrep, st, [&]() { return dhop_run_exact(c.gauge, c.sout, c.sin, dag, st); }, &us));
      total += us;
      GBX_TRY(ref_unpack_spinor_dev(c.sout, (T *)c.dOut, dsimd, lay));
    }
    GBX_HIP(hipMemcpyAsync(out, c.dOut, nf, hipMemcpyDeviceToHost, st));
    GBX_HIP(hipStreamSynchronize(st));
    if (usec) *usec = total;
    c.last_path = 1;
    return GBX_OK;
  }
  // Generic table-driven path.
  GBX_TRY(ensure(&c.dNbr, &c.nNbr, nt * 8));
  GBX_TRY(ensure(&c.dPrm, &c.nPrm, nt));
  hipStream_t st = c.st;
  GBX_HIP(hipMemcpyAsync(c.dU, U, nu, hipMemcpyHostToDevice, st));
  GBX_HIP(hipMemcpyAsync(c.dIn, in, nf, hipMemcpyHostToDevice, st));
  GBX_HIP(hipMemcpyAsync(c.dNbr, nbr, nt * 8, hipMemcpyHostToDevice, st));
  GBX_HIP(hipMemcpyAsync(c.dPrm, prm, nt, hipMemcpyHostToDevice, st));
  const uint64_t nthr = n5 * (uint64_t)nsimd;
  const dim3 grid((unsigned)((nthr + 255) / 256)), block(256);
  auto launch = [&]() -> int {
#define GBX_TABLE(D, R)                                                                                                   \
  hipLaunchKernelGGL((k_dhop_table<T, D, R>), grid, block, 0, st, (const T *)c.dU, (T *)c.dOut, (const T *)c.dIn,         \
                     (const uint64_t *)c.dNbr, (const uint8_t *)c.dPrm, nsite, Ls, nsimd, dsimd)
    if (dag && rrii) GBX_TABLE(true, true);
    else if (dag) GBX_TABLE(true, false);
    else if (rrii) GBX_TABLE(false, true);
    else GBX_TABLE(false, false);
#undef GBX_TABLE
    GBX_HIP(hipGetLastError());
    return GBX_OK;
  };
  GBX_TRY(timed_reps(nrep, st, launch, usec));
  GBX_HIP(hipMemcpyAsync(out, c.dOut, nf, hipMemcpyDeviceToHost, st));
  GBX_HIP(hipStreamSynchronize(st));
  c.last_path = 2;
  return GBX_OK;
}

void ctx_release(gbx_dslash_ctx_s &c)
{
  if (c.st) (void)hipStreamSynchronize(c.st);
  for (void *p : {c.dU, c.dIn, c.dOut, c.dNbr, c.dPrm}) (void)hipFree(p);
  c.dU = c.dIn = c.dOut = c.dNbr = c.dPrm = nullptr;
  c.nU = c.nIn = c.nOut = c.nNbr = c.nPrm = 0;
  gbx_spinor_destroy(c.sin);
  gbx_spinor_destroy(c.sout);
  gbx_gauge_destroy(c.gauge);
  gbx_geom_destroy(c.geom);
  c.sin = c.sout = nullptr;
  c.gauge = nullptr;
  c.geom = nullptr;
  c.prec = -1;
  if (c.st) (void)hipStreamDestroy(c.st);
  c.st = nullptr;
}

// One call on a context: serialised on the context's lock; the path taken is
// kept on the context and in this thread's gbx_dslash_last_path().
template <typename T>
int ctx_call(gbx_dslash_ctx_s *c, int nrep, int nsimd, int dsimd, int rrii, int dag, const T *U, T *out, const T *in,
             const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec)
{
  g_last_path = 0;
  GBX_CHECK(c, GBX_ENULL, "dslash_kernel: NULL context");
  std::lock_guard<std::mutex> lk(c->mu);
  c->last_path = 0;
  const int rc = dslash_kernel_impl<T>(*c, nrep, nsimd, dsimd, rrii, dag, U, out, in, nbr, nsite, Ls, prm, usec);
  g_last_path = c->last_path;
  return rc;
}

// The context behind the context-free entries (created on first use, kept
// for the life of the process).
gbx_dslash_ctx_s *default_ctx()
{
  static gbx_dslash_ctx_s *c = new gbx_dslash_ctx_s();
  return c;
}

}  // namespace

extern "C" int gbx_dslash_kernel_f64(int nrep, int nsimd, int dag, const double *U, double *out, const double *in,
                                     const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec)
{
  return ctx_call<double>(default_ctx(), nrep, nsimd, nsimd, 0, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_kernel_f32(int nrep, int nsimd, int dag, const float *U, float *out, const float *in,
                                     const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec)
{
  return ctx_call<float>(default_ctx(), nrep, nsimd, nsimd, 0, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_kernel_rrii_f64(int nrep, int nsimd, int dag, const double *U, double *out, const double *in,
                                          const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec)
{
  return ctx_call<double>(default_ctx(), nrep, nsimd, nsimd, 1, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_kernel_rrii_f32(int nrep, int nsimd, int dag, const float *U, float *out, const float *in,
                                          const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec)
{
  return ctx_call<float>(default_ctx(), nrep, nsimd, nsimd, 1, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_kernel_expand_f64(int nrep, int nsimd, int data_nsimd, int rrii, int dag, const double *U,
                                            double *out, const double *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                                            const uint8_t *prm, double *usec)
{
  return ctx_call<double>(default_ctx(), nrep, nsimd, data_nsimd, rrii, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_kernel_expand_f32(int nrep, int nsimd, int data_nsimd, int rrii, int dag, const float *U,
                                            float *out, const float *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                                            const uint8_t *prm, double *usec)
{
  return ctx_call<float>(default_ctx(), nrep, nsimd, data_nsimd, rrii, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_last_path(void) { return g_last_path; }

extern "C" int gbx_dslash_ctx_create(gbx_dslash_ctx *ctx)
{
  GBX_CHECK(ctx, GBX_ENULL, "gbx_dslash_ctx_create: NULL argument");
  *ctx = new gbx_dslash_ctx_s();
  return GBX_OK;
}

extern "C" int gbx_dslash_ctx_destroy(gbx_dslash_ctx ctx)
{
  if (!ctx) return GBX_OK;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx_release(*ctx);
  }
  delete ctx;
  return GBX_OK;
}

extern "C" int gbx_dslash_ctx_kernel_f64(gbx_dslash_ctx ctx, int nrep, int nsimd, int data_nsimd, int rrii, int dag,
                                         const double *U, double *out, const double *in, const uint64_t *nbr, uint64_t nsite,
                                         uint64_t Ls, const uint8_t *prm, double *usec)
{
  return ctx_call<double>(ctx, nrep, nsimd, data_nsimd, rrii, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_ctx_kernel_f32(gbx_dslash_ctx ctx, int nrep, int nsimd, int data_nsimd, int rrii, int dag,
                                         const float *U, float *out, const float *in, const uint64_t *nbr, uint64_t nsite,
                                         uint64_t Ls, const uint8_t *prm, double *usec)
{
  return ctx_call<float>(ctx, nrep, nsimd, data_nsimd, rrii, dag, U, out, in, nbr, nsite, Ls, prm, usec);
}

extern "C" int gbx_dslash_ctx_last_path(gbx_dslash_ctx ctx)
{
  if (!ctx) return 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  return ctx->last_path;
}
