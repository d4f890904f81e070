// Defect-correction conjugate gradient for  M^dagger M x = b,
// M = 1 - kappa Dhop  (M^dagger = 1 - kappa Dhop^dagger), and the even-odd
// (Schur complement) solve of M x = b on a red-black pair: with
// Mpc = 1 - kappa^2 DhopEO DhopOE on the even sites the outer loop works on
// r = b'_e - Mpc x_e and the inner CG on Mpc^dag Mpc (CGNR).
//
// Outer loop, FP64 on the GPU:
//   r = b - M^dag M x  (FP64 hopping kernels; the "true" residual)
//   stop when |r| / |b| <= tol
//   r_in = s * r with s = sqrt(N) / |r|  (unit-RMS entries: safe for FP16)
//   e = inner_solve(M_lo^dag M_lo e = r_in) to |r_in - ...| <= inner_tol |r_in|
//   x += e / s          (reliable update of x in FP64)
// Inner CG: vectors e, r in FP32; p and A p in the low storage precision of
// the inner gauge field (FP32, FP16 or BF16), which is what the low-precision
// hopping kernel reads and writes; every arithmetic operation in FP32; dots
// and norms in FP64 with deterministic reductions.  One iteration is two
// fused M applications (out = in - kappa Dhop in), one dot, one fused
// update-and-norm pass and one fused direction update.
#include <chrono>
#include <cmath>

#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

constexpr int kRB = 1024, kRT = 256;

template <int P> struct Pt {
  typedef typename dev::Traits<P>::acc A;
  static constexpr int R = dev::Traits<P>::R, NCH = dev::Traits<P>::NCH;
};

template <int P, typename A>
__device__ __forceinline__ void load_pt(const void *b, uint64_t q, uint32_t M, uint32_t m, A *v)
{
#pragma unroll
  for (int k = 0; k < Pt<P>::NCH; k++) dev::load_chunk<P, A>(b, (q * Pt<P>::NCH + k) * M + m, v + k * Pt<P>::R);
}

template <int P, typename A>
__device__ __forceinline__ void store_pt(void *b, uint64_t q, uint32_t M, uint32_t m, const A *v)
{
#pragma unroll
  for (int k = 0; k < Pt<P>::NCH; k++) dev::store_chunk<P, A>(b, (q * Pt<P>::NCH + k) * M + m, v + k * Pt<P>::R, false);
}

__device__ __forceinline__ void bred(double &a, double *sh)
{
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int s = kRT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  a = sh[0];
}

// e += a p; r -= a Ap; partial |r|^2.   (e, r: precision PE; p, Ap: PP)
template <int PE, int PP>
__global__ void __launch_bounds__(kRT) k_cg_update(double a, const void *__restrict__ p, const void *__restrict__ Ap,
                                                   void *__restrict__ e, void *__restrict__ r, uint64_t npts, uint32_t M,
                                                   double *__restrict__ partial)
{
  __shared__ double sh[kRT];
  typedef typename Pt<PE>::A A;
  double acc = 0.0;
  for (uint64_t t = blockIdx.x * (uint64_t)kRT + threadIdx.x; t < npts; t += (uint64_t)gridDim.x * kRT) {
    const uint64_t q = t / M;
    const uint32_t m = (uint32_t)(t % M);
    A pv[24], apv[24], ev[24], rv[24];
    load_pt<PP, A>(p, q, M, m, pv);
    load_pt<PP, A>(Ap, q, M, m, apv);
    load_pt<PE, A>(e, q, M, m, ev);
    load_pt<PE, A>(r, q, M, m, rv);
#pragma unroll
    for (int j = 0; j < 24; j++) {
      ev[j] += (A)a * pv[j];
      rv[j] -= (A)a * apv[j];
    }
    store_pt<PE, A>(e, q, M, m, ev);
    store_pt<PE, A>(r, q, M, m, rv);
    // norm of the stored (rounded) residual
    load_pt<PE, A>(r, q, M, m, rv);
#pragma unroll
    for (int j = 0; j < 24; j++) acc += (double)rv[j] * (double)rv[j];
  }
  bred(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// p = r + b p
template <int PE, int PP>
__global__ void __launch_bounds__(kRT) k_cg_xpay(double b, const void *__restrict__ r, void *__restrict__ p, uint64_t npts,
                                                 uint32_t M)
{
  typedef typename Pt<PE>::A A;
  for (uint64_t t = blockIdx.x * (uint64_t)kRT + threadIdx.x; t < npts; t += (uint64_t)gridDim.x * kRT) {
    const uint64_t q = t / M;
    const uint32_t m = (uint32_t)(t % M);
    A rv[24], pv[24];
    load_pt<PE, A>(r, q, M, m, rv);
    load_pt<PP, A>(p, q, M, m, pv);
#pragma unroll
    for (int j = 0; j < 24; j++) pv[j] = rv[j] + (A)b * pv[j];
    store_pt<PP, A>(p, q, M, m, pv);
  }
}

// z = a x + b y, all FP64 or same precision P (used for r = b - t)
template <int P>
__global__ void __launch_bounds__(kRT) k_axpby(double a, const void *__restrict__ x, double b, const void *__restrict__ y,
                                               void *__restrict__ z, uint64_t nchunk)
{
  typedef typename Pt<P>::A A;
  constexpr int R = Pt<P>::R;
  for (uint64_t t = blockIdx.x * (uint64_t)kRT + threadIdx.x; t < nchunk; t += (uint64_t)gridDim.x * kRT) {
    A xv[R], yv[R];
    dev::load_chunk<P, A>(x, t, xv);
    dev::load_chunk<P, A>(y, t, yv);
#pragma unroll
    for (int j = 0; j < R; j++) xv[j] = (A)a * xv[j] + (A)b * yv[j];
    dev::store_chunk<P, A>(z, t, xv, false);
  }
}

// x (FP64) += a e (precision PS)
template <int PS>
__global__ void __launch_bounds__(kRT) k_axpy_into_f64(double a, const void *__restrict__ e, void *__restrict__ x, uint64_t npts,
                                                       uint32_t M)
{
  for (uint64_t t = blockIdx.x * (uint64_t)kRT + threadIdx.x; t < npts; t += (uint64_t)gridDim.x * kRT) {
    const uint64_t q = t / M;
    const uint32_t m = (uint32_t)(t % M);
    double ev[24], xv[24];
    load_pt<PS, double>(e, q, M, m, ev);
    load_pt<GBX_F64, double>(x, q, M, m, xv);
#pragma unroll
    for (int j = 0; j < 24; j++) xv[j] += a * ev[j];
    store_pt<GBX_F64, double>(x, q, M, m, xv);
  }
}

__global__ void __launch_bounds__(kRT) k_final(const double *__restrict__ partial, int n, double *__restrict__ out)
{
  __shared__ double sh[kRT];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += kRT) a += partial[i];
  bred(a, sh);
  if (threadIdx.x == 0) out[0] = a;
}

struct Scratch {
  double *partial = nullptr, *res = nullptr, *host = nullptr;
  ~Scratch()
  {
    (void)hipFree(partial);
    (void)hipFree(res);
    (void)hipHostFree(host);
  }
  int init()
  {
    GBX_HIP(hipMalloc(&partial, sizeof(double) * kRB));
    GBX_HIP(hipMalloc(&res, sizeof(double)));
    GBX_HIP(hipHostMalloc(&host, sizeof(double), hipHostMallocDefault));
    return GBX_OK;
  }
  int fetch(hipStream_t st, double *v)
  {
    hipLaunchKernelGGL(k_final, dim3(1), dim3(kRT), 0, st, (const double *)partial, kRB, res);
    GBX_HIP(hipMemcpyAsync(host, res, sizeof(double), hipMemcpyDeviceToHost, st));
    GBX_HIP(hipStreamSynchronize(st));
    *v = *host;
    return GBX_OK;
  }
};

// Owning handle for temporaries.
struct Tmp {
  gbx_spinor_s *s = nullptr;
  ~Tmp() { gbx_spinor_destroy(s); }
};

int make(gbx_geom_s *g, int Ls, int prec, Tmp &t) { return gbx_spinor_create(g, Ls, 1, prec, &t.s); }

template <int PE, int PP>
void launch_update(double a, const gbx_spinor_s *p, const gbx_spinor_s *Ap, gbx_spinor_s *e, gbx_spinor_s *r, Scratch &sc,
                   hipStream_t st)
{
  const uint64_t npts = r->geom->vol * (uint64_t)r->M;
  hipLaunchKernelGGL((k_cg_update<PE, PP>), dim3(kRB), dim3(kRT), 0, st, a, p->d, Ap->d, e->d, r->d, npts, r->M, sc.partial);
}

template <int PE, int PP>
void launch_xpay(double b, const gbx_spinor_s *r, gbx_spinor_s *p, hipStream_t st)
{
  const uint64_t npts = r->geom->vol * (uint64_t)r->M;
  hipLaunchKernelGGL((k_cg_xpay<PE, PP>), dim3(kRB), dim3(kRT), 0, st, b, r->d, p->d, npts, r->M);
}

// Launch configuration of the solver's hopping kernels: the tuned VALU
// configuration of this signature (the fused forms run on the VALU kernel)
// with cached output stores, since the solver reads its outputs back at once.
dhop_cfg solver_cfg(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  gbx_kcfg k;
  gbx_dhop_get_cfg((gbx_gauge)u, (gbx_spinor)in, &k);
  dhop_cfg c = dhop_default_cfg(u, in);
  if (k.kind == 0) c = dhop_cfg{k.spw, k.threads, k.xcd, 0, k.mc, k.pf, k.wgpc, k.cblk};
  c.nt = 0;
  return c;
}

// The positive operator A the CG inverts, with its temporaries:
//   whole lattice:  A = M^dag M        (u = the lattice's gauge field, t)
//   red-black pair: A = Mpc^dag Mpc     (ue / uo = even / odd halves' fields,
//                   t on the even half, to on the odd half)
enum { OP_WILSON = 0, OP_DWF = 1 };

struct NormalOp {
  bool eo;
  const gbx_gauge_s *ue, *uo;
  double kappa;
  gbx_spinor_s *t, *to;
  int kind = OP_WILSON;       // red-black only: M = 1 - kappa Dhop, or the DWF operator
  gbx_dwf_params dwf{0, 0};
  gbx_spinor_s *te = nullptr;  // DWF: a second even-half temporary
};

// y = Mpc x (dag = 0) or Mpc^dag x (dag = 1), with dagger kernels for dag
// ((DhopEO DhopOE)^dag = (Dhop^dag)_EO (Dhop^dag)_OE):
//   Wilson:  to = DhopOE x;  y = x - kappa^2 DhopEO to
//   DWF:     to = Mooee^-1 DhopOE x;  te = DhopEO to;  y = Mooee x - te / 4
int apply_schur(const NormalOp &op, gbx_spinor_s *y, const gbx_spinor_s *x, int dag, hipStream_t st)
{
  GBX_TRY(dhop_launch(op.uo, op.to, x, dag, solver_cfg(op.uo, x), 0.0, 1.0, nullptr, st));
  if (op.kind == OP_DWF) {
    GBX_TRY(dwf_s5_launch(op.to, op.to, nullptr, &op.dwf, 1, dag, 1.0, 0.0, st));
    GBX_TRY(dhop_launch(op.ue, op.te, op.to, dag, solver_cfg(op.ue, op.to), 0.0, 1.0, nullptr, st));
    return dwf_s5_launch(y, x, op.te, &op.dwf, 0, dag, 1.0, -0.25, st);
  }
  GBX_TRY(dhop_launch(op.ue, y, op.to, dag, solver_cfg(op.ue, op.to), 1.0, -op.kappa * op.kappa, x, st));
  return GBX_OK;
}

// y = A x.  Whole lattice: t = x - kappa Dhop x; y = t - kappa Dhop^dag t.
int apply_normal(const NormalOp &op, gbx_spinor_s *y, const gbx_spinor_s *x, hipStream_t st)
{
  if (op.eo) {
    GBX_TRY(apply_schur(op, op.t, x, 0, st));
    return apply_schur(op, y, op.t, 1, st);
  }
  const dhop_cfg c = solver_cfg(op.ue, x);
  GBX_TRY(dhop_launch(op.ue, op.t, x, 0, c, 1.0, -op.kappa, x, st));
  GBX_TRY(dhop_launch(op.ue, y, op.t, 1, c, 1.0, -op.kappa, op.t, st));
  return GBX_OK;
}

// Inner CG on A e = r (e zero on entry, r overwritten).  Vectors e, r
// precision PE; p, Ap and the operator's temporaries precision PP (the gauge
// field's).
template <int PE, int PP>
int inner_cg(const NormalOp &op, gbx_spinor_s *e, gbx_spinor_s *r, gbx_spinor_s *p, gbx_spinor_s *Ap, double rel_tol,
             int max_iter, int *iters, int *applies, Scratch &sc, hipStream_t st)
{
  double rr;
  GBX_TRY(spinor_norm2_on(r, &rr, st));
  const double stop = rel_tol * rel_tol * rr;
  GBX_TRY(spinor_zero_on(e, st));
  GBX_TRY(spinor_convert_scaled(p, r, 1.0, st));
  int it = 0;
  while (it < max_iter && rr > stop) {
    GBX_TRY(apply_normal(op, Ap, p, st));
    (*applies)++;
    double pap, dummy;
    GBX_TRY(spinor_dot_on(p, Ap, &pap, &dummy, st));
    if (!(pap > 0.0)) break;  // breakdown (rounding at the accuracy floor)
    const double a = rr / pap;
    launch_update<PE, PP>(a, p, Ap, e, r, sc, st);
    GBX_HIP(hipGetLastError());
    double rr_new;
    GBX_TRY(sc.fetch(st, &rr_new));
    const double b = rr_new / rr;
    rr = rr_new;
    launch_xpay<PE, PP>(b, r, p, st);
    GBX_HIP(hipGetLastError());
    it++;
  }
  *iters += it;
  return GBX_OK;
}

int inner_dispatch(const NormalOp &op, gbx_spinor_s *e, gbx_spinor_s *r, gbx_spinor_s *p, gbx_spinor_s *Ap, double tol,
                   int maxit, int *iters, int *applies, Scratch &sc, hipStream_t st)
{
  if (e->prec == GBX_F64) return inner_cg<GBX_F64, GBX_F64>(op, e, r, p, Ap, tol, maxit, iters, applies, sc, st);
  switch (p->prec) {
  case GBX_F32: return inner_cg<GBX_F32, GBX_F32>(op, e, r, p, Ap, tol, maxit, iters, applies, sc, st);
  case GBX_F16: return inner_cg<GBX_F32, GBX_F16>(op, e, r, p, Ap, tol, maxit, iters, applies, sc, st);
  default: return inner_cg<GBX_F32, GBX_BF16>(op, e, r, p, Ap, tol, maxit, iters, applies, sc, st);
  }
}

int axpby_f64(double a, const gbx_spinor_s *x, double b, const gbx_spinor_s *y, gbx_spinor_s *z, hipStream_t st)
{
  const uint64_t n = z->geom->vol * (uint64_t)z->M * 12;
  hipLaunchKernelGGL(k_axpby<GBX_F64>, dim3(kRB), dim3(kRT), 0, st, a, x->d, b, y->d, z->d, n);
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

int axpy_into_f64(double a, const gbx_spinor_s *e, gbx_spinor_s *x, hipStream_t st)
{
  const uint64_t npts = x->geom->vol * (uint64_t)x->M;
  switch (e->prec) {
  case GBX_F64: hipLaunchKernelGGL(k_axpy_into_f64<GBX_F64>, dim3(kRB), dim3(kRT), 0, st, a, e->d, x->d, npts, x->M); break;
  case GBX_F32: hipLaunchKernelGGL(k_axpy_into_f64<GBX_F32>, dim3(kRB), dim3(kRT), 0, st, a, e->d, x->d, npts, x->M); break;
  case GBX_F16: hipLaunchKernelGGL(k_axpy_into_f64<GBX_F16>, dim3(kRB), dim3(kRT), 0, st, a, e->d, x->d, npts, x->M); break;
  default: hipLaunchKernelGGL(k_axpy_into_f64<GBX_BF16>, dim3(kRB), dim3(kRT), 0, st, a, e->d, x->d, npts, x->M); break;
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

int check_params(const gbx_cg_params *p)
{
  GBX_CHECK(std::fabs(p->kappa) < 0.125, GBX_EINVAL, "cg: |kappa| = %g must be < 1/8", p->kappa);
  GBX_CHECK(p->tol > 0 && p->inner_tol > 0 && p->inner_tol < 1 && p->max_outer >= 1 && p->max_inner >= 1, GBX_EINVAL,
            "cg: bad tolerances or iteration limits");
  return GBX_OK;
}

int check_args(const gbx_gauge_s *u_hi, const gbx_gauge_s *u_lo, const gbx_spinor_s *x, const gbx_spinor_s *b,
               const gbx_cg_params *p)
{
  GBX_CHECK(u_hi && x && b && p, GBX_ENULL, "cg: NULL argument");
  GBX_CHECK(u_hi->prec == GBX_F64, GBX_EINVAL, "cg: the outer gauge field must be FP64");
  GBX_CHECK(x->prec == GBX_F64 && b->prec == GBX_F64, GBX_EINVAL, "cg: x and b must be FP64");
  GBX_CHECK(x->nrhs == 1 && b->nrhs == 1, GBX_EINVAL, "cg: one right-hand side");
  GBX_CHECK(x->geom == u_hi->geom && b->geom == u_hi->geom && x->Ls == b->Ls, GBX_EMISMATCH, "cg: shapes differ");
  GBX_CHECK(x->d != b->d, GBX_EALIAS, "cg: x aliases b");
  if (u_lo) {
    GBX_CHECK(u_lo->geom == u_hi->geom, GBX_EMISMATCH, "cg: inner gauge field on another geometry");
    GBX_CHECK(u_lo->prec != GBX_F64, GBX_EINVAL, "cg: the inner gauge field must be FP32, FP16 or BF16");
  }
  GBX_CHECK(whole_lattice(u_hi->geom), GBX_EINVAL,
            "cg: gbx_cg_solve works on a whole lattice; use gbx_cg_solve_eo on a red-black pair");
  return check_params(p);
}

int solve(const gbx_gauge_s *u_hi, const gbx_gauge_s *u_lo, gbx_spinor_s *x, const gbx_spinor_s *b, const gbx_cg_params *prm,
          gbx_cg_stats *stats)
{
  GBX_TRY(check_args(u_hi, u_lo, x, b, prm));
  // TDIAG fields: the solve runs in that basis (S is orthogonal, so norms
  // and residuals are the chiral ones); FP64 or FP32 inner solves only.
  GBX_CHECK(x->basis == b->basis, GBX_EMISMATCH, "cg: x and b are in different spin bases");
  GBX_CHECK(b->basis == GBX_BASIS_CHIRAL || !u_lo || u_lo->prec == GBX_F32, GBX_EUNSUPPORTED,
            "cg: TDIAG fields need an FP64 or FP32 inner gauge field");
  const auto t0 = std::chrono::steady_clock::now();
  gbx_cg_stats st{};
  gbx_geom_s *g = u_hi->geom;
  hipStream_t s = g->stream;
  const int Ls = x->Ls;
  const int plo = u_lo ? u_lo->prec : GBX_F64;
  const int pe = u_lo ? GBX_F32 : GBX_F64;
  Tmp r, t, y, e, ri, p, Ap, tl;
  Scratch sc;
  GBX_TRY(sc.init());
  GBX_TRY(make(g, Ls, GBX_F64, r));
  GBX_TRY(make(g, Ls, GBX_F64, t));
  GBX_TRY(make(g, Ls, GBX_F64, y));
  GBX_TRY(make(g, Ls, pe, e));
  GBX_TRY(make(g, Ls, pe, ri));
  GBX_TRY(make(g, Ls, plo, p));
  GBX_TRY(make(g, Ls, plo, Ap));
  GBX_TRY(make(g, Ls, plo, tl));
  for (Tmp *f : {&r, &t, &y, &e, &ri, &p, &Ap, &tl}) f->s->basis = b->basis;
  const gbx_gauge_s *uin = u_lo ? u_lo : u_hi;
  const NormalOp op_hi{false, u_hi, nullptr, prm->kappa, t.s, nullptr};
  const NormalOp op_lo{false, uin, nullptr, prm->kappa, tl.s, nullptr};
  double bb;
  GBX_TRY(spinor_norm2_on(b, &bb, s));
  GBX_CHECK(bb > 0.0, GBX_EINVAL, "cg: b is zero");
  const double bnorm = std::sqrt(bb);
  const double N = (double)g->vol * Ls * 24;
  double rel = 1.0;
  for (;;) {
    // true residual r = b - M^dag M x in FP64
    GBX_TRY(apply_normal(op_hi, y.s, x, s));
    st.hi_applies++;
    GBX_TRY(axpby_f64(1.0, b, -1.0, y.s, r.s, s));
    double rr;
    GBX_TRY(spinor_norm2_on(r.s, &rr, s));
    rel = std::sqrt(rr) / bnorm;
    if (rel <= prm->tol || st.outer_iters >= prm->max_outer || !(rr > 0.0)) break;
    const double scale = std::sqrt(N) / std::sqrt(rr);
    GBX_TRY(spinor_convert_scaled(ri.s, r.s, scale, s));
    GBX_TRY(inner_dispatch(op_lo, e.s, ri.s, p.s, Ap.s, prm->inner_tol, prm->max_inner, &st.inner_iters, &st.lo_applies, sc,
                           s));
    GBX_TRY(axpy_into_f64(1.0 / scale, e.s, x, s));
    st.outer_iters++;
  }
  GBX_HIP(hipStreamSynchronize(s));
  st.true_resid = rel;
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  if (rel > prm->tol) return gbx_fail(GBX_ENOCONV, "cg: relative residual %.3e after %d outer iterations", rel, st.outer_iters);
  return GBX_OK;
}

// out = alpha * c + beta * Dhop(in) with the solver's configuration.
int xpay(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *c, const gbx_spinor_s *in, int dag, double alpha,
         double beta, hipStream_t st)
{
  return dhop_launch(u, out, in, dag, solver_cfg(u, in), alpha, beta, c, st);
}

// Even-odd Schur-complement solve on a red-black pair (kind = OP_WILSON:
// M = 1 - kappa Dhop; OP_DWF: D = Mooee - Dhop / 2).
int solve_schur(int kind, const gbx_dwf_params *dwf, const gbx_gauge_s *const u_hi[2], const gbx_gauge_s *const *u_lo,
                gbx_spinor_s *const x[2], const gbx_spinor_s *const b[2], const gbx_cg_params *prm, gbx_cg_stats *stats)
{
  const char *who = kind == OP_DWF ? "dwf_eo" : "cg_eo";
  GBX_CHECK(u_hi && x && b && prm && u_hi[0] && u_hi[1] && x[0] && x[1] && b[0] && b[1], GBX_ENULL, "%s: NULL argument", who);
  const gbx_geom_s *ge = u_hi[0]->geom, *go = u_hi[1]->geom;
  GBX_CHECK(ge->parity == 0 && go->parity == 1 && ge->pair_id == go->pair_id, GBX_EMISMATCH,
            "%s: u_hi must be the even then the odd half of one red-black pair", who);
  GBX_CHECK(ge->nghost == 0, GBX_EINVAL, "%s: the solvers take a whole lattice's halves, not a domain part's", who);
  const int basis = b[0]->basis;
  GBX_CHECK(x[0]->basis == basis && x[1]->basis == basis && b[1]->basis == basis, GBX_EMISMATCH,
            "%s: x and b are in different spin bases", who);
  for (int p = 0; p < 2; p++) {
    const gbx_geom_s *g = p ? go : ge;
    GBX_CHECK(u_hi[p]->prec == GBX_F64, GBX_EINVAL, "%s: the outer gauge fields must be FP64", who);
    GBX_CHECK(x[p]->prec == GBX_F64 && b[p]->prec == GBX_F64, GBX_EINVAL, "%s: x and b must be FP64", who);
    GBX_CHECK(x[p]->geom == g && b[p]->geom == g, GBX_EMISMATCH, "%s: x[%d] / b[%d] not on the %s half", who, p, p,
              p ? "odd" : "even");
    GBX_CHECK(x[p]->nrhs == 1 && b[p]->nrhs == 1 && x[p]->Ls == x[0]->Ls && b[p]->Ls == x[0]->Ls, GBX_EMISMATCH,
              "%s: one right-hand side, equal Ls", who);
    GBX_CHECK(x[p]->d != b[p]->d, GBX_EALIAS, "%s: x aliases b", who);
    if (u_lo) {
      GBX_CHECK(u_lo[p] && u_lo[p]->geom == g, GBX_EMISMATCH, "%s: inner gauge field %d not on the %s half", who, p,
                p ? "odd" : "even");
      GBX_CHECK(u_lo[p]->prec != GBX_F64 && u_lo[p]->prec == u_lo[0]->prec, GBX_EINVAL,
                "%s: inner gauge fields must share one of FP32, FP16, BF16", who);
    }
  }
  // TDIAG fields: the Wilson operator only (the DWF terms' chirality
  // projectors are written for chiral values), FP64 or FP32 inner solves.
  GBX_CHECK(basis == GBX_BASIS_CHIRAL || (kind == OP_WILSON && (!u_lo || u_lo[0]->prec == GBX_F32)), GBX_EUNSUPPORTED,
            "%s: TDIAG fields: the Wilson operator with an FP64 or FP32 inner solve only", who);
  if (kind == OP_DWF) {
    GBX_CHECK(prm->tol > 0 && prm->inner_tol > 0 && prm->inner_tol < 1 && prm->max_outer >= 1 && prm->max_inner >= 1,
              GBX_EINVAL, "%s: bad tolerances or iteration limits", who);
  } else {
    GBX_TRY(check_params(prm));
  }
  const auto t0 = std::chrono::steady_clock::now();
  gbx_cg_stats st{};
  gbx_geom_s *g = u_hi[0]->geom, *g1 = u_hi[1]->geom;
  hipStream_t s = g->stream;  // the pair's one stream
  const int Ls = x[0]->Ls;
  const double kappa = prm->kappa;
  const int plo = u_lo ? u_lo[0]->prec : GBX_F64;
  const int pe = u_lo ? GBX_F32 : GBX_F64;
  Tmp bp, r, y, t, te, to, e, ri, p, Ap, tl, tel, tol_, ro;
  Scratch sc;
  GBX_TRY(sc.init());
  GBX_TRY(make(g, Ls, GBX_F64, bp));
  GBX_TRY(make(g, Ls, GBX_F64, r));
  GBX_TRY(make(g, Ls, GBX_F64, y));
  GBX_TRY(make(g, Ls, GBX_F64, t));
  GBX_TRY(make(g, Ls, GBX_F64, te));
  GBX_TRY(make(g1, Ls, GBX_F64, to));
  GBX_TRY(make(g1, Ls, GBX_F64, ro));
  GBX_TRY(make(g, Ls, pe, e));
  GBX_TRY(make(g, Ls, pe, ri));
  GBX_TRY(make(g, Ls, plo, p));
  GBX_TRY(make(g, Ls, plo, Ap));
  GBX_TRY(make(g, Ls, plo, tl));
  GBX_TRY(make(g, Ls, plo, tel));
  GBX_TRY(make(g1, Ls, plo, tol_));
  for (Tmp *f : {&bp, &r, &y, &t, &te, &to, &ro, &e, &ri, &p, &Ap, &tl, &tel, &tol_}) f->s->basis = basis;
  NormalOp op_hi{true, u_hi[0], u_hi[1], kappa, t.s, to.s};
  NormalOp op_lo = u_lo ? NormalOp{true, u_lo[0], u_lo[1], kappa, tl.s, tol_.s} : NormalOp{true, u_hi[0], u_hi[1], kappa, tl.s,
                                                                                           tol_.s};
  op_hi.kind = op_lo.kind = kind;
  if (kind == OP_DWF) {
    op_hi.dwf = op_lo.dwf = *dwf;
    op_hi.te = te.s;
    op_lo.te = tel.s;
  }
  double bbe, bbo;
  GBX_TRY(spinor_norm2_on(b[0], &bbe, s));
  GBX_TRY(spinor_norm2_on(b[1], &bbo, s));
  GBX_CHECK(bbe + bbo > 0.0, GBX_EINVAL, "%s: b is zero", who);
  const double bnorm = std::sqrt(bbe + bbo);
  const double N = (double)g->vol * Ls * 24;
  // b'_e:  Wilson b_e + kappa DhopEO b_o;  DWF b_e + DhopEO Mooee^-1 b_o / 2
  if (kind == OP_DWF) {
    GBX_TRY(dwf_s5_launch(to.s, b[1], nullptr, dwf, 1, 0, 1.0, 0.0, s));
    GBX_TRY(xpay(u_hi[0], bp.s, b[0], to.s, 0, 1.0, 0.5, s));
  } else {
    GBX_TRY(xpay(u_hi[0], bp.s, b[0], b[1], 0, 1.0, kappa, s));
  }
  double rel = 1.0;
  for (;;) {
    // r = b'_e - Mpc x_e in FP64; |r| is |b - M x| once x_o is rebuilt
    GBX_TRY(apply_schur(op_hi, y.s, x[0], 0, s));
    st.hi_applies++;
    GBX_TRY(axpby_f64(1.0, bp.s, -1.0, y.s, r.s, s));
    double rr;
    GBX_TRY(spinor_norm2_on(r.s, &rr, s));
    rel = std::sqrt(rr) / bnorm;
    if (rel <= prm->tol || st.outer_iters >= prm->max_outer || !(rr > 0.0)) break;
    // inner right-hand side Mpc^dag r, scaled to unit-RMS entries
    GBX_TRY(apply_schur(op_hi, y.s, r.s, 1, s));
    double yy;
    GBX_TRY(spinor_norm2_on(y.s, &yy, s));
    if (!(yy > 0.0)) break;
    const double scale = std::sqrt(N) / std::sqrt(yy);
    GBX_TRY(spinor_convert_scaled(ri.s, y.s, scale, s));
    GBX_TRY(inner_dispatch(op_lo, e.s, ri.s, p.s, Ap.s, prm->inner_tol, prm->max_inner, &st.inner_iters, &st.lo_applies, sc,
                           s));
    GBX_TRY(axpy_into_f64(1.0 / scale, e.s, x[0], s));
    st.outer_iters++;
  }
  if (kind == OP_DWF) {
    // x_o = Mooee^-1 (b_o + DhopOE x_e / 2); then b - D x with D = Mooee - Dhop / 2
    GBX_TRY(xpay(u_hi[1], to.s, b[1], x[0], 0, 1.0, 0.5, s));
    GBX_TRY(dwf_s5_launch(x[1], to.s, nullptr, dwf, 1, 0, 1.0, 0.0, s));
    GBX_TRY(dhop_launch(u_hi[0], te.s, x[1], 0, solver_cfg(u_hi[0], x[1]), 0.0, 1.0, nullptr, s));
    GBX_TRY(dwf_s5_launch(y.s, x[0], te.s, dwf, 0, 0, 1.0, -0.5, s));  // (D x)_e
    GBX_TRY(dhop_launch(u_hi[1], to.s, x[0], 0, solver_cfg(u_hi[1], x[0]), 0.0, 1.0, nullptr, s));
    GBX_TRY(dwf_s5_launch(to.s, x[1], to.s, dwf, 0, 0, 1.0, -0.5, s));  // (D x)_o
  } else {
    // x_o = b_o + kappa DhopOE x_e; then b - M x
    GBX_TRY(xpay(u_hi[1], x[1], b[1], x[0], 0, 1.0, kappa, s));
    GBX_TRY(xpay(u_hi[0], y.s, x[0], x[1], 0, 1.0, -kappa, s));   // (M x)_e
    GBX_TRY(xpay(u_hi[1], to.s, x[1], x[0], 0, 1.0, -kappa, s));  // (M x)_o
  }
  GBX_TRY(axpby_f64(1.0, b[0], -1.0, y.s, r.s, s));
  GBX_TRY(axpby_f64(1.0, b[1], -1.0, to.s, ro.s, s));
  double re2, ro2;
  GBX_TRY(spinor_norm2_on(r.s, &re2, s));
  GBX_TRY(spinor_norm2_on(ro.s, &ro2, s));
  GBX_HIP(hipStreamSynchronize(s));
  const double full = std::sqrt(re2 + ro2) / bnorm;
  st.true_resid = full;
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  if (rel > prm->tol)
    return gbx_fail(GBX_ENOCONV, "%s: relative residual %.3e after %d outer iterations", who, rel, st.outer_iters);
  return GBX_OK;
}

}  // namespace

extern "C" int gbx_cg_solve_eo(const gbx_gauge u_hi[2], const gbx_gauge *u_lo, gbx_spinor x[2], const gbx_spinor b[2],
                               const gbx_cg_params *p, gbx_cg_stats *st)
{
  return solve_schur(OP_WILSON, nullptr, u_hi, u_lo, x, b, p, st);
}

extern "C" int gbx_dwf_solve_eo(const gbx_gauge u_hi[2], const gbx_gauge *u_lo, gbx_spinor x[2], const gbx_spinor b[2],
                                const gbx_dwf_params *dwf, const gbx_cg_params *p, gbx_cg_stats *st)
{
  GBX_CHECK(dwf, GBX_ENULL, "dwf_eo: NULL DWF parameters");
  GBX_CHECK(x && x[0], GBX_ENULL, "dwf_eo: NULL argument");
  GBX_CHECK(5.0 - dwf->M5 > 0.0 && x[0]->Ls >= 2 && x[0]->Ls <= 256, GBX_EINVAL, "dwf_eo: M5 = %g, Ls = %d", dwf->M5,
            x[0]->Ls);
  return solve_schur(OP_DWF, dwf, u_hi, u_lo, x, b, p, st);
}

extern "C" int gbx_cg_solve(gbx_gauge u_hi, gbx_gauge u_lo, gbx_spinor x, gbx_spinor b, const gbx_cg_params *p,
                            gbx_cg_stats *st)
{
  GBX_CHECK(u_lo, GBX_ENULL, "gbx_cg_solve: NULL inner gauge field (use gbx_cg_solve_f64)");
  return solve(u_hi, u_lo, x, b, p, st);
}

extern "C" int gbx_cg_solve_f64(gbx_gauge u, gbx_spinor x, gbx_spinor b, const gbx_cg_params *p, gbx_cg_stats *st)
{
  return solve(u, nullptr, x, b, p, st);
}
