// Stencil tables in the reference's form and their recognition.
//
// The reference's tables come from Grid (TableGenerate.cc:72-91):
//   nbr[mu + 8*(s + Ls*site)] = 5D outer-site index of the neighbour,
//   prm[...]                  = 1 iff the hop crosses the outer boundary of a
//                               dimension that is split over SIMD lanes.
// Grid's virtual-node convention: Nsimd lanes are split by factors of two
// over T, Z, Y, X (in that order); outer dims rd = L/simd; the 5D index has s
// fastest.  This file generates such tables for a periodic lattice and
// recognises them, so the drop-in entry can route canonical tables to the
// SoA kernel that computes neighbours from its own site order.
#include <atomic>

#include "gbx_internal.h"

namespace {

bool simd_split(int nsimd, int simd[4])
{
  for (int d = 0; d < 4; d++) simd[d] = 1;
  if (nsimd != 1 && nsimd != 2 && nsimd != 4 && nsimd != 8 && nsimd != 16) return false;
  int d = 3;
  for (int n = nsimd; n > 1; n >>= 1) {
    simd[d] *= 2;
    d = (d + 3) % 4;  // T, Z, Y, X, T, ...
  }
  return true;
}

struct OuterGeom {
  int rd[4];
  uint64_t stride[4];  // lexicographic strides of the outer lattice
  uint64_t nsite;
};

OuterGeom outer_geom(const int rd[4])
{
  OuterGeom g;
  uint64_t st = 1;
  for (int d = 0; d < 4; d++) {
    g.rd[d] = rd[d];
    g.stride[d] = st;
    st *= (uint64_t)rd[d];
  }
  g.nsite = st;
  return g;
}

// Neighbour outer site of `site` in direction mu and whether the hop wraps.
inline uint64_t hop(const OuterGeom &g, uint64_t site, int mu, bool &wrapped)
{
  const int d = mu & 3;
  const int x = (int)((site / g.stride[d]) % (uint64_t)g.rd[d]);
  wrapped = false;
  if (mu < 4) {
    if (x + 1 == g.rd[d]) {
      wrapped = true;
      return site - (uint64_t)x * g.stride[d];
    }
    return site + g.stride[d];
  }
  if (x == 0) {
    wrapped = true;
    return site + (uint64_t)(g.rd[d] - 1) * g.stride[d];
  }
  return site - g.stride[d];
}

}  // namespace

extern "C" int gbx_table_generate(const int dims[4], int nsimd, uint64_t Ls, uint64_t *nbr, uint8_t *prm)
{
  int simd[4], rd[4];
  GBX_CHECK(dims, GBX_ENULL, "gbx_table_generate: dims is NULL");
  GBX_CHECK(nbr && prm, GBX_ENULL, "gbx_table_generate: NULL table");
  GBX_CHECK(simd_split(nsimd, simd), GBX_EINVAL, "gbx_table_generate: nsimd=%d not in {1,2,4,8,16}", nsimd);
  GBX_CHECK(Ls > 0, GBX_EINVAL, "gbx_table_generate: Ls=0");
  for (int d = 0; d < 4; d++) {
    GBX_CHECK(dims[d] > 0 && dims[d] % simd[d] == 0, GBX_EINVAL,
              "gbx_table_generate: dim %d = %d not divisible by its %d SIMD lanes", d, dims[d], simd[d]);
    rd[d] = dims[d] / simd[d];
  }
  const OuterGeom g = outer_geom(rd);
#pragma omp parallel for schedule(static)
  for (int64_t site = 0; site < (int64_t)g.nsite; site++) {
    for (int mu = 0; mu < 8; mu++) {
      bool w;
      const uint64_t n4 = hop(g, (uint64_t)site, mu, w);
      const uint8_t p = (w && simd[mu & 3] > 1) ? 1 : 0;
      for (uint64_t s = 0; s < Ls; s++) {
        const uint64_t e = 8 * ((uint64_t)site * Ls + s) + (uint64_t)mu;
        nbr[e] = n4 * Ls + s;
        prm[e] = p;
      }
    }
  }
  return GBX_OK;
}

extern "C" int gbx_table_recognise(int nsimd, uint64_t nsite, uint64_t Ls, const uint64_t *nbr, const uint8_t *prm,
                                   int dims[4])
{
  int simd[4];
  GBX_CHECK(nbr && prm && dims, GBX_ENULL, "gbx_table_recognise: NULL argument");
  GBX_CHECK(simd_split(nsimd, simd), GBX_EINVAL, "gbx_table_recognise: nsimd=%d not in {1,2,4,8,16}", nsimd);
  GBX_CHECK(nsite > 0 && Ls > 0, GBX_EINVAL, "gbx_table_recognise: empty lattice");
  const uint64_t n5 = nsite * Ls;
  // Outer extents from the cycle lengths of the +mu hops starting at site 0
  // (s = 0 entries), each hop staying within the lattice.
  int rd[4];
  uint64_t prod = 1;
  for (int d = 0; d < 4; d++) {
    uint64_t cur = 0, len = 0;
    do {
      const uint64_t e = 8 * cur * Ls + (uint64_t)d;
      const uint64_t nx = nbr[e];
      if (nx >= n5 || nx % Ls != 0) return gbx_fail(GBX_ETABLE, "not canonical: entry %llu", (unsigned long long)e);
      cur = nx / Ls;
      len++;
    } while (cur != 0 && len <= nsite);
    if (cur != 0 || len == 0 || len > 65536) return gbx_fail(GBX_ETABLE, "not canonical: no +%d cycle", d);
    rd[d] = (int)len;
    prod *= len;
  }
  if (prod != nsite) return gbx_fail(GBX_ETABLE, "not canonical: outer extents do not multiply to nsite");
  const OuterGeom g = outer_geom(rd);
  // The cycle along +mu from 0 must step by the lexicographic stride.
  std::atomic<uint64_t> bad{~0ull};
#pragma omp parallel for schedule(static)
  for (int64_t site = 0; site < (int64_t)nsite; site++) {
    if (bad.load(std::memory_order_relaxed) != ~0ull) continue;
    for (int mu = 0; mu < 8; mu++) {
      bool w;
      const uint64_t n4 = hop(g, (uint64_t)site, mu, w);
      const uint8_t p = (w && simd[mu & 3] > 1) ? 1 : 0;
      for (uint64_t s = 0; s < Ls; s++) {
        const uint64_t e = 8 * ((uint64_t)site * Ls + s) + (uint64_t)mu;
        if (nbr[e] != n4 * Ls + s || prm[e] != p) {
          uint64_t exp = ~0ull;
          bad.compare_exchange_strong(exp, e);
          break;
        }
      }
    }
  }
  if (bad.load() != ~0ull)
    return gbx_fail(GBX_ETABLE, "not canonical: entry %llu differs", (unsigned long long)bad.load());
  for (int d = 0; d < 4; d++) dims[d] = rd[d] * simd[d];
  return GBX_OK;
}

// Index of the first entry with nbr >= n5 or prm > 1, or -1.
int64_t table_first_invalid(const uint64_t *nbr, const uint8_t *prm, uint64_t n, uint64_t n5)
{
  int64_t bad = INT64_MAX;
#pragma omp parallel for schedule(static) reduction(min : bad)
  for (int64_t i = 0; i < (int64_t)n; i++)
    if ((nbr[i] >= n5 || prm[i] > 1) && i < bad) bad = i;
  return bad == INT64_MAX ? -1 : bad;
}
