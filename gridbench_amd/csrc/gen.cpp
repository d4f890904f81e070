// Seeded synthetic inputs: counter-based Gaussian reals, Haar-like SU(3)
// links and Gaussian spinors, bit-reproducible on any x86 host.
//
// Reproducibility rules: no libm transcendental is called (their last-bit
// results may depend on the host CPU's dispatch); the natural logarithm is
// this file's own (below); sqrt and division are correctly rounded IEEE
// operations; the file is compiled with FP contraction off.
#pragma clang fp contract(off)

#include <cmath>
#include <cstring>

#include "gbx_internal.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t key(uint64_t seed, uint64_t stream) { return mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull)); }

// Uniform in the open interval (-1, 1) for counter c of the keyed stream.
inline double uniform_pm1(uint64_t k, uint64_t c)
{
  const uint64_t r = mix64(k ^ mix64(c));
  const double u = ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0,1)
  return 2.0 * u - 1.0;
}

// Natural logarithm of a positive finite double: x = m 2^e with
// m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(t), t = (m-1)/(m+1), |t| < 0.172,
// series to t^25 (truncation < 1e-20 relative).
double own_log(double x)
{
  int e = 0;
  double m = std::frexp(x, &e);  // exact
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double t = (m - 1.0) / (m + 1.0);
  const double t2 = t * t;
  double p = 1.0 / 25.0;
  for (int k = 11; k >= 0; k--) p = p * t2 + 1.0 / (double)(2 * k + 1);
  const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
  const double ln2_lo = 1.90821492927058770002e-10;
  return (double)e * ln2_hi + ((double)e * ln2_lo + 2.0 * t * p);
}

// Marsaglia polar method on the counter stream: pair p uses counters
// (p << 8) | j for j = 0, 1, ... until a point falls inside the unit disc.
inline void gauss_pair(uint64_t k, uint64_t p, double &g0, double &g1)
{
  for (uint64_t a = 0; a < 128; a++) {
    const double v1 = uniform_pm1(k, (p << 8) | (2 * a));
    const double v2 = uniform_pm1(k, (p << 8) | (2 * a + 1));
    const double s = v1 * v1 + v2 * v2;
    if (s > 0.0 && s < 1.0) {
      const double f = std::sqrt(-2.0 * own_log(s) / s);
      g0 = v1 * f;
      g1 = v2 * f;
      return;
    }
  }
  g0 = g1 = 0.0;  // unreachable in practice: (1 - pi/4)^128 < 1e-85
}

void gaussians(uint64_t seed, uint64_t stream, uint64_t index0, uint64_t n, double *out)
{
  const uint64_t k = key(seed, stream);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const uint64_t idx = index0 + (uint64_t)i;
    double g0, g1;
    gauss_pair(k, idx >> 1, g0, g1);
    out[i] = (idx & 1) ? g1 : g0;
  }
}

struct cplx {
  double re, im;
};
inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cplx cconjmul(cplx a, cplx b) { return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; }  // conj(a) b
inline cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }

bool dims_ok(const int dims[4], uint64_t *vol)
{
  if (!dims) return false;
  uint64_t v = 1;
  for (int d = 0; d < 4; d++) {
    if (dims[d] <= 0 || dims[d] > 4096) return false;
    v *= (uint64_t)dims[d];
  }
  if (v > (1ull << 31)) return false;
  *vol = v;
  return true;
}

}  // namespace

extern "C" int gbx_gen_gaussian(uint64_t seed, uint64_t stream, uint64_t index0, uint64_t n, double *out)
{
  GBX_CHECK(out || n == 0, GBX_ENULL, "gbx_gen_gaussian: out is NULL");
  gaussians(seed, stream, index0, n, out);
  return GBX_OK;
}

extern "C" int gbx_gen_su3_links(const int dims[4], uint64_t seed, double *links)
{
  uint64_t V;
  GBX_CHECK(dims_ok(dims, &V), GBX_EINVAL, "gbx_gen_su3_links: bad dims");
  GBX_CHECK(links, GBX_ENULL, "gbx_gen_su3_links: links is NULL");
  const uint64_t k = key(seed, 0);
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < (int64_t)(V * 4); l++) {
    // 12 Gaussians per link: pairs 6*l .. 6*l+5 give (re, im) of rows 0, 1
    cplx r0[3], r1[3], r2[3];
    for (int j = 0; j < 3; j++) {
      gauss_pair(k, (uint64_t)l * 6 + (uint64_t)j, r0[j].re, r0[j].im);
      gauss_pair(k, (uint64_t)l * 6 + 3 + (uint64_t)j, r1[j].re, r1[j].im);
    }
    double n0 = 0.0;
    for (int j = 0; j < 3; j++) n0 += r0[j].re * r0[j].re + r0[j].im * r0[j].im;
    n0 = std::sqrt(n0);
    for (int j = 0; j < 3; j++) r0[j] = {r0[j].re / n0, r0[j].im / n0};
    cplx c = {0.0, 0.0};
    for (int j = 0; j < 3; j++) {
      const cplx t = cconjmul(r0[j], r1[j]);
      c = {c.re + t.re, c.im + t.im};
    }
    for (int j = 0; j < 3; j++) r1[j] = csub(r1[j], cmul(c, r0[j]));
    double n1 = 0.0;
    for (int j = 0; j < 3; j++) n1 += r1[j].re * r1[j].re + r1[j].im * r1[j].im;
    n1 = std::sqrt(n1);
    for (int j = 0; j < 3; j++) r1[j] = {r1[j].re / n1, r1[j].im / n1};
    // row 2 = conj(r0 x r1)
    for (int j = 0; j < 3; j++) {
      const int a = (j + 1) % 3, b = (j + 2) % 3;
      const cplx x = csub(cmul(r0[a], r1[b]), cmul(r0[b], r1[a]));
      r2[j] = {x.re, -x.im};
    }
    double *o = links + (uint64_t)l * 18;
    const cplx *rows[3] = {r0, r1, r2};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        o[(i * 3 + j) * 2] = rows[i][j].re;
        o[(i * 3 + j) * 2 + 1] = rows[i][j].im;
      }
  }
  return GBX_OK;
}

extern "C" int gbx_double_links(const int dims[4], const double *links, double *dbl)
{
  uint64_t V;
  GBX_CHECK(dims_ok(dims, &V), GBX_EINVAL, "gbx_double_links: bad dims");
  GBX_CHECK(links && dbl, GBX_ENULL, "gbx_double_links: NULL array");
  GBX_CHECK(links != dbl, GBX_EALIAS, "gbx_double_links: output aliases input");
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < (int64_t)V; x++) {
    int c[4];
    uint64_t r = (uint64_t)x;
    for (int d = 0; d < 4; d++) {
      c[d] = (int)(r % (uint64_t)dims[d]);
      r /= (uint64_t)dims[d];
    }
    for (int mu = 0; mu < 4; mu++) {
      int b[4] = {c[0], c[1], c[2], c[3]};
      b[mu] = (c[mu] + dims[mu] - 1) % dims[mu];
      const uint64_t xb = (uint64_t)b[0] + (uint64_t)dims[0] * ((uint64_t)b[1] + (uint64_t)dims[1] * ((uint64_t)b[2] + (uint64_t)dims[2] * (uint64_t)b[3]));
      const double *f = links + ((uint64_t)x * 4 + (uint64_t)mu) * 18;
      const double *g = links + (xb * 4 + (uint64_t)mu) * 18;
      double *of = dbl + ((uint64_t)x * 8 + (uint64_t)mu) * 18;
      double *ob = dbl + ((uint64_t)x * 8 + (uint64_t)mu + 4) * 18;
      std::memcpy(of, f, 18 * sizeof(double));
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          ob[(i * 3 + j) * 2] = g[(j * 3 + i) * 2];
          ob[(i * 3 + j) * 2 + 1] = -g[(j * 3 + i) * 2 + 1];
        }
    }
  }
  return GBX_OK;
}

extern "C" int gbx_gen_spinor(const int dims[4], int Ls, int nrhs, uint64_t seed, double *psi)
{
  uint64_t V;
  GBX_CHECK(dims_ok(dims, &V), GBX_EINVAL, "gbx_gen_spinor: bad dims");
  GBX_CHECK(Ls > 0 && nrhs > 0, GBX_EINVAL, "gbx_gen_spinor: Ls=%d nrhs=%d", Ls, nrhs);
  GBX_CHECK(psi, GBX_ENULL, "gbx_gen_spinor: psi is NULL");
  const uint64_t n = V * (uint64_t)Ls * 24;
  for (int r = 0; r < nrhs; r++) gaussians(seed, (uint64_t)r, 0, n, psi + (uint64_t)r * n);
  return GBX_OK;
}
