/*
 * ORACLE -- test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker / the timed CPU restatement.
 *
 * Plain-C restatement of GridBench's CPU path for the DWF Wilson hopping term:
 *   - oracle_dhop_hand_{f64,f32}: the "Hand" kernel (hand_impl.h; reference
 *     WilsonKernelsHandCpu.h:385-451), arbitrary Nsimd virtual-node layout,
 *     arbitrary nbr/prm tables;
 *   - oracle_tables: the nbr/prm stencil tables in the reference's form
 *     (TableGenerate.cc:72-91: entry mu + 8*(s + Ls*site), 5D index s fastest),
 *     restated from Grid's virtual-node conventions for a periodic lattice;
 *   - oracle_pack_* / oracle_unpack_*: lexicographic <-> reference host layout
 *     (bench.cc:105-138, RIRI: [site][..][lane][re,im]).
 *
 * Parity pins (tests/test_oracle.py): bit-exact against the reference kernel
 * compiled from its own sources (oracle/_ref) on identical inputs for Nsimd in
 * {1,2,4,8,16}, FP64 and FP32; against committed fixtures made from that
 * build (tests/golden); against closed forms (free field plane waves,
 * constant field), gamma5-hermiticity, adjointness, gauge covariance and a
 * brute-force dense-matrix check on a 2^4 lattice.  Tables: invariants plus
 * the fixtures' SHA-256.
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#define OR_XP 0
#define OR_YP 1
#define OR_ZP 2
#define OR_TP 3
#define OR_XM 4
#define OR_YM 5
#define OR_ZM 6
#define OR_TM 7

#define OR_CAT_(a, b) a##b
#define OR_CAT(a, b) OR_CAT_(a, b)

#define OR_T double
#define OR_SFX f64
#include "hand_impl.h"
#undef OR_T
#undef OR_SFX

#define OR_T float
#define OR_SFX f32
#include "hand_impl.h"
#undef OR_T
#undef OR_SFX

/* ------------------------------------------------------------------------ */
/* Virtual-node geometry (Grid conventions behind TableGenerate.cc:72-91).   */
/* ------------------------------------------------------------------------ */

/* Split Nsimd lanes over the 4 dimensions by factors of 2, starting from the
 * last dimension (T) and cycling: Nsimd=8 -> {1,2,2,2}, 16 -> {2,2,2,2}. */
int oracle_simd_layout(int nsimd, int simd[4])
{
  for (int d = 0; d < 4; d++) simd[d] = 1;
  if (nsimd <= 0 || (nsimd & (nsimd - 1))) return -1;
  int k = 0;
  while ((1 << k) < nsimd) k++;
  for (int i = 0; i < k; i++) simd[3 - (i % 4)] *= 2;
  return 0;
}

/* Lexicographic index of the global site for (outer site o, lane l). */
static uint64_t global_site(const int dims[4], const int simd[4], const int rd[4], uint64_t o, int l)
{
  int oc[4], ic[4];
  uint64_t r = o;
  for (int d = 0; d < 4; d++) { oc[d] = (int)(r % (uint64_t)rd[d]); r /= (uint64_t)rd[d]; }
  int q = l;
  for (int d = 0; d < 4; d++) { ic[d] = q % simd[d]; q /= simd[d]; }
  uint64_t g = 0;
  for (int d = 3; d >= 0; d--) g = g * (uint64_t)dims[d] + (uint64_t)(oc[d] + rd[d] * ic[d]);
  return g;
}

static int geom(const int dims[4], int nsimd, int simd[4], int rd[4], uint64_t *nsite)
{
  if (oracle_simd_layout(nsimd, simd)) return -1;
  uint64_t n = 1;
  for (int d = 0; d < 4; d++) {
    if (dims[d] <= 0 || dims[d] % simd[d]) return -1;
    rd[d] = dims[d] / simd[d];
    n *= (uint64_t)rd[d];
  }
  *nsite = n;
  return 0;
}

/* nbr[mu + 8*(s + Ls*site)] = 5D outer index of the neighbour in direction mu
 * (mu<4: +1 hop in dimension mu, mu>=4: -1 hop in dimension mu-4);
 * prm = 1 iff the hop wraps the outer extent of a SIMD-split dimension. */
int oracle_tables(const int dims[4], int nsimd, uint64_t Ls, uint64_t *nbr, uint8_t *prm)
{
  int simd[4], rd[4];
  uint64_t nsite;
  if (geom(dims, nsimd, simd, rd, &nsite)) return -1;
  for (uint64_t site = 0; site < nsite; site++) {
    int oc[4];
    uint64_t r = site;
    for (int d = 0; d < 4; d++) { oc[d] = (int)(r % (uint64_t)rd[d]); r /= (uint64_t)rd[d]; }
    for (int mu = 0; mu < 8; mu++) {
      const int d = mu % 4, disp = mu < 4 ? 1 : -1;
      int nc[4];
      memcpy(nc, oc, sizeof nc);
      int x = oc[d] + disp, wrapped = 0;
      if (x >= rd[d]) { x -= rd[d]; wrapped = 1; }
      if (x < 0) { x += rd[d]; wrapped = 1; }
      nc[d] = x;
      uint64_t nsite4 = 0;
      for (int e = 3; e >= 0; e--) nsite4 = nsite4 * (uint64_t)rd[e] + (uint64_t)nc[e];
      for (uint64_t s = 0; s < Ls; s++) {
        const uint64_t j = (uint64_t)mu + 8 * (s + Ls * site);
        nbr[j] = s + Ls * nsite4;
        prm[j] = (uint8_t)(wrapped && simd[d] > 1);
      }
    }
  }
  return 0;
}

/* Lexicographic spinor [V][Ls][4][3][2] (the Nsimd = 1 order)
 * -> reference layout [nsite][Ls][4][3][nsimd][2]. */
int oracle_pack_spinor_f64(const int dims[4], int nsimd, uint64_t Ls, const double *lex, double *ref)
{
  int simd[4], rd[4];
  uint64_t nsite;
  if (geom(dims, nsimd, simd, rd, &nsite)) return -1;
  for (uint64_t o = 0; o < nsite; o++)
    for (int l = 0; l < nsimd; l++) {
      const uint64_t g = global_site(dims, simd, rd, o, l);
      for (uint64_t s = 0; s < Ls; s++)
        for (int sc = 0; sc < 12; sc++)
          for (int ri = 0; ri < 2; ri++)
            ref[(((o * Ls + s) * 12 + (uint64_t)sc) * nsimd + (uint64_t)l) * 2 + (uint64_t)ri] =
                lex[((g * Ls + s) * 12 + (uint64_t)sc) * 2 + (uint64_t)ri];
    }
  return 0;
}

int oracle_unpack_spinor_f64(const int dims[4], int nsimd, uint64_t Ls, const double *ref, double *lex)
{
  int simd[4], rd[4];
  uint64_t nsite;
  if (geom(dims, nsimd, simd, rd, &nsite)) return -1;
  for (uint64_t o = 0; o < nsite; o++)
    for (int l = 0; l < nsimd; l++) {
      const uint64_t g = global_site(dims, simd, rd, o, l);
      for (uint64_t s = 0; s < Ls; s++)
        for (int sc = 0; sc < 12; sc++)
          for (int ri = 0; ri < 2; ri++)
            lex[((g * Ls + s) * 12 + (uint64_t)sc) * 2 + (uint64_t)ri] =
                ref[(((o * Ls + s) * 12 + (uint64_t)sc) * nsimd + (uint64_t)l) * 2 + (uint64_t)ri];
    }
  return 0;
}

/* Lexicographic doubled gauge field [V][8][3][3][2] -> reference layout
 * [nsite][8][3][3][nsimd][2]. */
int oracle_pack_gauge_f64(const int dims[4], int nsimd, const double *lex, double *ref)
{
  int simd[4], rd[4];
  uint64_t nsite;
  if (geom(dims, nsimd, simd, rd, &nsite)) return -1;
  for (uint64_t o = 0; o < nsite; o++)
    for (int l = 0; l < nsimd; l++) {
      const uint64_t g = global_site(dims, simd, rd, o, l);
      for (int k = 0; k < 72; k++)
        for (int ri = 0; ri < 2; ri++)
          ref[((o * 72 + (uint64_t)k) * nsimd + (uint64_t)l) * 2 + (uint64_t)ri] = lex[(g * 72 + (uint64_t)k) * 2 + (uint64_t)ri];
    }
  return 0;
}
