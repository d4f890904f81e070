// Spin basis of stored spinor values (GBX_BASIS_CHIRAL / GBX_BASIS_TDIAG).
//
// TDIAG stores psi' = S psi, S = [[1, 1], [1, -1]] / sqrt(2) acting on each
// (upper spin, lower spin) pair of the same colour: (0, 2) and (1, 3).  In
// the SoA layout spins 0, 1 fill chunks 0 .. NCH/2-1 and spins 2, 3 chunks
// NCH/2 .. NCH-1 (FP64: one complex per chunk, FP32: two), and real e of
// chunk k pairs with real e of chunk k + NCH/2.  S is symmetric and
// orthogonal, S = S^-1, so one kernel converts either way and norms and
// inner products are those of the chiral values.
#include <algorithm>

#include "gbx_internal.h"
#include "layout_dev.h"

namespace {

template <int P>
__global__ void __launch_bounds__(256) k_spin_rotate(void *x, uint64_t nslot, uint32_t M)
{
  typedef typename dev::Traits<P>::vec V;
  typedef typename dev::Traits<P>::st S;
  constexpr int NCH = dev::Traits<P>::NCH, H = NCH / 2, R = dev::Traits<P>::R;
  const S c = (S)0.70710678118654752440;  // 1/sqrt(2), rounded to the storage precision
  V *v = (V *)x;
  const uint64_t n = nslot * H * M;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t m = i % M, qk = i / M, q = qk / H, k = qk - q * H;
    const uint64_t lo = (q * NCH + k) * M + m, hi = lo + (uint64_t)H * M;
    const V a = v[lo], b = v[hi];
    V s, d;
#pragma unroll
    for (int e = 0; e < R; e++) {
      s[e] = (a[e] + b[e]) * c;
      d[e] = (a[e] - b[e]) * c;
    }
    v[lo] = s;
    v[hi] = d;
  }
}

}  // namespace

int require_chiral(const char *what, const gbx_spinor_s *a, const gbx_spinor_s *b, const gbx_spinor_s *c)
{
  const gbx_spinor_s *f[3] = {a, b, c};
  for (const gbx_spinor_s *x : f)
    GBX_CHECK(!x || x->basis == GBX_BASIS_CHIRAL, GBX_EUNSUPPORTED,
              "%s: takes chiral-basis fields only (gbx_spinor_set_basis(x, GBX_BASIS_CHIRAL) first)", what);
  return GBX_OK;
}

extern "C" int gbx_spinor_set_basis(gbx_spinor x, int basis)
{
  GBX_CHECK(x && x->d, GBX_ENULL, "gbx_spinor_set_basis: NULL field");
  GBX_CHECK(basis == GBX_BASIS_CHIRAL || basis == GBX_BASIS_TDIAG, GBX_EINVAL, "gbx_spinor_set_basis: basis %d", basis);
  GBX_CHECK(x->prec == GBX_F64 || x->prec == GBX_F32, GBX_EUNSUPPORTED,
            "gbx_spinor_set_basis: FP64 / FP32 fields only (16-bit chunks straddle the spin halves)");
  if (basis == x->basis) return GBX_OK;
  const uint64_t nslot = x->geom->vol + x->geom->nghost;
  const uint64_t n = nslot * (uint64_t)(spinor_nchunk(x->prec) / 2) * x->M;
  const unsigned nb = (unsigned)std::min<uint64_t>((n + 255) / 256, 8192);
  hipStream_t st = x->geom->stream;
  if (x->prec == GBX_F64) hipLaunchKernelGGL(k_spin_rotate<GBX_F64>, dim3(nb), dim3(256), 0, st, x->d, nslot, x->M);
  else hipLaunchKernelGGL(k_spin_rotate<GBX_F32>, dim3(nb), dim3(256), 0, st, x->d, nslot, x->M);
  GBX_HIP(hipGetLastError());
  x->basis = basis;
  return GBX_OK;
}

extern "C" int gbx_spinor_get_basis(gbx_spinor x, int *basis)
{
  GBX_CHECK(x && basis, GBX_ENULL, "gbx_spinor_get_basis: NULL argument");
  *basis = x->basis;
  return GBX_OK;
}
