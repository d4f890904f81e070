// Domain-wall (Shamir) fifth-dimension terms and the DWF operator.
//
// On a field of Ls s-slices (chirality P+ = spins 0,1 and P- = spins 2,3 of
// the chiral basis, gamma5 = diag(1,1,-1,-1)):
//
//   Mooee psi(s) = d psi(s) - P- psi(s+1) - P+ psi(s-1),   d = 5 - M5,
//   walls:  -P- psi(Ls)  ->  +mass P- psi(0),   -P+ psi(-1)  ->  +mass P+ psi(Ls-1)
//
// and the DWF operator D = Mooee - (1/2) Dhop (Dhop on every s-slice).  In
// each chirality Mooee is an Ls x Ls matrix A = d I - L with L a shift by
// one slice that picks up a factor -mass when it wraps, so L^Ls = -mass I and
//
//   A^-1 = (1 / (1 + mass d^-Ls)) sum_{k=0}^{Ls-1} L^k / d^(k+1),
//
// applied here in closed form.  The transposes swap the chiralities' shift
// directions: Mooee^dag is Mooee with the roles of P+ and P- exchanged, and
// likewise for the inverse.  The reference creates a DomainWallFermion with
// mass 0.1 and M5 1.0 (TableGenerate.cc:47-50) and writes out only its Dhop;
// the terms here are the Shamir form of that operator.
//
// Kernel: one thread per 16-byte chunk of the SoA layout.  Chunks are stored
// ((q*nch + k)*M + m) with m = rhs*Ls + s, so every Ls consecutive chunks are
// one (site, chunk, right-hand side) group holding all s-slices.  A
// workgroup loads whole groups into LDS (coalesced), then each thread forms
// its slice's output from its group's Ls slices.  In place (out == x or
// out == z) is safe: a workgroup reads its groups completely before writing.
#include <cmath>

#include "gbx_internal.h"
#include "layout_dev.h"

namespace {

// out = alpha * S x + beta * z, S = Mooee (INV false) or Mooee^-1 (INV true),
// their adjoints for dag (chiralities exchanged).  z may be null (beta = 0).
template <int P, bool INV>
__global__ void __launch_bounds__(256) k_dwf_s5(const void *x, void *out, const void *z, uint64_t ngroups,
                                                uint32_t Ls, uint32_t nrhs, uint32_t gpw, double d_, double mass_,
                                                double alpha_, double beta_, int dag)
{
  typedef typename dev::Traits<P>::acc A;
  constexpr int R = dev::Traits<P>::R, NCH = dev::Traits<P>::NCH;
  extern __shared__ __align__(16) unsigned char smem[];
  A *sx = (A *)smem;  // [gpw * Ls][R]
  const A d = (A)d_, mass = (A)mass_, alpha = (A)alpha_, beta = (A)beta_;
  const uint32_t lt = threadIdx.x, lg = lt / Ls, s = lt - lg * Ls;
  const bool in_group = lt < gpw * Ls;
  A dinv = (A)1 / d, dls = (A)1;
  for (uint32_t k = 0; k < Ls; k++) dls *= dinv;
  const A gnorm = (A)1 / ((A)1 + mass * dls);  // 1 / (1 + mass d^-Ls)
  for (uint64_t g0 = (uint64_t)blockIdx.x * gpw; g0 < ngroups; g0 += (uint64_t)gridDim.x * gpw) {
    const uint64_t grp = g0 + lg;
    const bool active = in_group && grp < ngroups;
    const uint64_t t = grp * Ls + s;  // chunk index in memory order
    if (active) dev::load_chunk<P, A>(x, t, sx + (uint64_t)lt * R);
    __syncthreads();
    if (active) {
      const uint32_t kch = (uint32_t)((grp / nrhs) % NCH);  // chunk of the spinor this group holds
      const A *gx = sx + (uint64_t)lg * Ls * R;
      A v[R];
#pragma unroll
      for (int j = 0; j < R; j++) {
        // P+ (reals 0..11) hops down in s (psi(s-1)), P- (12..23) up; dag exchanges them
        const bool down = ((kch * R + j) < 12) != (dag != 0);
        A acc;
        if (!INV) {
          // d x(s) - x(s -/+ 1), the wrap replaced by +mass x(other end)
          uint32_t sn;
          A c;
          if (down) {
            sn = s == 0 ? Ls - 1 : s - 1;
            c = s == 0 ? mass : (A)-1;
          } else {
            sn = s == Ls - 1 ? 0 : s + 1;
            c = s == Ls - 1 ? mass : (A)-1;
          }
          acc = d * gx[(uint64_t)s * R + j] + c * gx[(uint64_t)sn * R + j];
        } else {
          // sum_k d^-(k+1) L^k x: L^k x(s) = x(s -/+ k), times -mass once it wraps
          acc = (A)0;
          A w = dinv;
          for (uint32_t k = 0; k < Ls; k++) {
            uint32_t sk;
            bool wrap;
            if (down) {
              wrap = k > s;
              sk = wrap ? s + Ls - k : s - k;
            } else {
              wrap = s + k >= Ls;
              sk = wrap ? s + k - Ls : s + k;
            }
            const A c = wrap ? -mass * w : w;
            acc += c * gx[(uint64_t)sk * R + j];
            w *= dinv;
          }
          acc *= gnorm;
        }
        v[j] = alpha * acc;
      }
      if (z) {
        A zv[R];
        dev::load_chunk<P, A>(z, t, zv);
#pragma unroll
        for (int j = 0; j < R; j++) v[j] += beta * zv[j];
      }
      dev::store_chunk<P, A>(out, t, v, false);
    }
    __syncthreads();
  }
}

template <int P>
void launch_s5(bool inv, const void *x, void *out, const void *z, uint64_t ngroups, uint32_t Ls, uint32_t nrhs, double d,
               double mass, double alpha, double beta, int dag, hipStream_t st)
{
  typedef typename dev::Traits<P>::acc A;
  constexpr int R = dev::Traits<P>::R;
  const uint32_t gpw = 256 / Ls;
  const size_t lds = (size_t)gpw * Ls * R * sizeof(A);
  uint64_t nb = (ngroups + gpw - 1) / gpw;
  if (nb > (1u << 20)) nb = 1u << 20;
  const dim3 grid((unsigned)nb), block(gpw * Ls);
  if (inv)
    hipLaunchKernelGGL((k_dwf_s5<P, true>), grid, block, lds, st, x, out, z, ngroups, Ls, nrhs, gpw, d, mass, alpha, beta, dag);
  else
    hipLaunchKernelGGL((k_dwf_s5<P, false>), grid, block, lds, st, x, out, z, ngroups, Ls, nrhs, gpw, d, mass, alpha, beta,
                       dag);
}

int check_params(const gbx_dwf_params *p, int Ls)
{
  GBX_CHECK(p, GBX_ENULL, "dwf: NULL parameters");
  const double d = 5.0 - p->M5;
  GBX_CHECK(std::isfinite(p->M5) && std::isfinite(p->mass) && d > 0.0, GBX_EINVAL, "dwf: M5 = %g must be below 5",
            p->M5);
  GBX_CHECK(Ls >= 2 && Ls <= 256, GBX_EINVAL, "dwf: Ls = %d must be in 2..256", Ls);
  GBX_CHECK(std::fabs(1.0 + p->mass * std::pow(d, -(double)Ls)) > 1e-12, GBX_EINVAL,
            "dwf: Mooee is singular for M5 = %g, mass = %g, Ls = %d", p->M5, p->mass, Ls);
  return GBX_OK;
}

}  // namespace

// out = alpha * S x + beta * z on any geometry (site-local), S = Mooee or
// its inverse (inv), adjoint for dag.  z may be null when beta == 0.
int dwf_s5_launch(gbx_spinor_s *out, const gbx_spinor_s *x, const gbx_spinor_s *z, const gbx_dwf_params *p, int inv, int dag,
                  double alpha, double beta, hipStream_t st)
{
  GBX_CHECK(out && x && out->d && x->d, GBX_ENULL, "dwf_mooee: NULL field");
  GBX_TRY(require_chiral("dwf_mooee", x, z));
  GBX_TRY(check_params(p, x->Ls));
  GBX_CHECK(out->geom == x->geom && out->Ls == x->Ls && out->nrhs == x->nrhs && out->prec == x->prec, GBX_EMISMATCH,
            "dwf_mooee: out and in differ in geometry, shape or precision");
  if (z) {
    GBX_CHECK(z->geom == x->geom && z->Ls == x->Ls && z->nrhs == x->nrhs && z->prec == x->prec, GBX_EMISMATCH,
              "dwf_mooee: the added field differs in geometry, shape or precision");
  } else {
    GBX_CHECK(beta == 0.0, GBX_EINVAL, "dwf_mooee: beta != 0 without a field to add");
  }
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "dwf_mooee: dag must be 0 or 1");
  const uint64_t ngroups = x->geom->vol * (uint64_t)spinor_nchunk(x->prec) * (uint64_t)x->nrhs;
  const double d = 5.0 - p->M5;
  const void *zd = z ? z->d : nullptr;
  switch (x->prec) {
  case GBX_F64: launch_s5<GBX_F64>(inv != 0, x->d, out->d, zd, ngroups, (uint32_t)x->Ls, (uint32_t)x->nrhs, d, p->mass, alpha, beta, dag, st); break;
  case GBX_F32: launch_s5<GBX_F32>(inv != 0, x->d, out->d, zd, ngroups, (uint32_t)x->Ls, (uint32_t)x->nrhs, d, p->mass, alpha, beta, dag, st); break;
  case GBX_F16: launch_s5<GBX_F16>(inv != 0, x->d, out->d, zd, ngroups, (uint32_t)x->Ls, (uint32_t)x->nrhs, d, p->mass, alpha, beta, dag, st); break;
  default: launch_s5<GBX_BF16>(inv != 0, x->d, out->d, zd, ngroups, (uint32_t)x->Ls, (uint32_t)x->nrhs, d, p->mass, alpha, beta, dag, st); break;
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

extern "C" int gbx_dwf_mooee(gbx_spinor out, gbx_spinor in, const gbx_dwf_params *p, int inv, int dag)
{
  GBX_CHECK(out && in, GBX_ENULL, "gbx_dwf_mooee: NULL field");
  GBX_CHECK(inv == 0 || inv == 1, GBX_EINVAL, "gbx_dwf_mooee: inv must be 0 or 1");
  return dwf_s5_launch(out, in, nullptr, p, inv, dag, 1.0, 0.0, in->geom->stream);
}

// out = D in = Mooee in - (1/2) Dhop in on a whole lattice: out = Dhop in,
// then out = Mooee in - (1/2) out (the 5th-dimension kernel reads its own
// output element of z before writing it).
extern "C" int gbx_dwf_apply(gbx_gauge u, gbx_spinor out, gbx_spinor in, const gbx_dwf_params *p, int dag)
{
  GBX_CHECK(u && out && in, GBX_ENULL, "gbx_dwf_apply: NULL argument");
  GBX_CHECK(whole_lattice(u->geom), GBX_EINVAL, "gbx_dwf_apply: whole lattices only (red-black halves: gbx_dwf_solve_eo)");
  GBX_TRY(check_params(p, in->Ls));
  GBX_TRY(require_chiral("gbx_dwf_apply", in));
  GBX_TRY(gbx_dhop(u, out, in, dag));
  return dwf_s5_launch(out, in, out, p, 0, dag, 1.0, -0.5, u->geom->stream);
}
