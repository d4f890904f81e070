// Autotuner for the hopping kernels.
//
// Searches, on the caller's own fields, over the launch parameters that move
// performance on gfx950: sites per workgroup (how many sites share one LDS
// link stage and how many neighbour gathers hit in L1), workgroup size,
// XCD-aware workgroup ordering, non-temporal vs cached output stores, and the
// VALU vs MFMA kernel where the MFMA form exists.  Each candidate is timed
// with HIP events over a few applications after a warm-up; the fastest is
// recorded for the problem signature (precision, compression, Ls, nrhs,
// lattice, traversal block) and used by every later gbx_dhop on that
// signature.  An optional JSON-lines cache file avoids re-tuning.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

#include "gbx_internal.h"

namespace {

int time_cfg(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k, int nrep,
             double *usec)
{
  hipStream_t st = u->geom->stream;
  GBX_TRY(dhop_run(u, out, in, dag, k, st));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GBX_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = gbx_fail(GBX_EHIP, "tune: hipEventCreate failed");
  if (rc == GBX_OK && hipEventRecord(e0, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "tune: event record");
  for (int r = 0; r < nrep && rc == GBX_OK; r++) rc = dhop_run(u, out, in, dag, k, st);
  if (rc == GBX_OK && hipEventRecord(e1, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "tune: event record");
  if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "tune: kernel failed");
  float ms = 0.f;
  if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "tune: elapsed time");
  if (rc != GBX_OK) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == GBX_OK) *usec = 1000.0 * ms / nrep;
  return rc;
}

// Minor-chunk sizes worth trying: all of M, and divisors of M down to 8.
std::vector<int> minor_chunks(uint32_t M)
{
  std::vector<int> v{0};
  for (uint32_t c = M / 2; c >= 8; c /= 2)
    if (M % c == 0) v.push_back((int)c);
  if (M > 16 && M % 12 == 0) v.push_back(12);
  return v;
}

// Stage 1: the work shape -- minor chunk, sites per workgroup, workgroup
// size, gather pipeline -- with XCD ordering and streaming stores on.
std::vector<gbx_kcfg> shape_candidates(const gbx_gauge_s *u, const gbx_spinor_s *in, int kinds)
{
  std::vector<gbx_kcfg> v;
  const uint32_t M = in->M;
  const int spws[] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64};
  for (int mc : (kinds & GBX_KIND_VALU) ? minor_chunks(M) : std::vector<int>{}) {
    const uint32_t mw = mc ? (uint32_t)mc : M;  // minor indices per workgroup
    for (int spw : spws) {
      const uint64_t items = (uint64_t)spw * mw;
      if (items > 4096 || items < 64) continue;
      std::vector<int> thr;
      if (items <= 1024 && items % 64 == 0) thr.push_back((int)items);
      for (int t : {128, 256, 512})
        if ((uint64_t)t < items && items % t == 0 && (uint64_t)t != items) thr.push_back(t);
      if (items < 1024 && items % 64 != 0) thr.push_back((int)((items + 63) / 64 * 64));
      for (int t : thr)
        for (int pf = 0; pf <= 2; pf++) {
          dhop_cfg c{spw, t, 1, 1, mc, pf, 0, 0};
          if (dhop_cfg_check(c, u, in) != GBX_OK) continue;
          v.push_back(gbx_kcfg{spw, t, 1, 1, 0, mc, pf, 0, 0});
        }
    }
  }
  gbx_clear_error();
  if (kinds & GBX_KIND_TILE) {
    // LDS-tiled kernel: compact boxes of 8..64 sites, minor chunks, and
    // workgroups of one or several items per thread.
    static const int shapes[][4] = {{2, 2, 2, 2}, {4, 2, 2, 2}, {2, 2, 2, 4}, {2, 4, 2, 2}, {4, 4, 2, 2}, {2, 2, 4, 4},
                                    {4, 2, 2, 4}, {2, 2, 2, 1}, {2, 2, 1, 2}, {3, 3, 3, 1}, {3, 3, 1, 3}, {3, 2, 2, 2}};
    for (const auto &e : shapes)
      for (int mc : minor_chunks(M)) {
        const uint32_t mw = mc ? (uint32_t)mc : M;
        const uint64_t items = (uint64_t)e[0] * e[1] * e[2] * e[3] * mw;
        for (int t : {256, 512, 768, 1024}) {
          if ((uint64_t)t > (items + 63) / 64 * 64 && t != 256) continue;
          gbx_kcfg k{0, t, 1, 1, 2, mc, 0, 0, 0, pack_tile(e)};
          if (tile_cfg_check(k, u, in) != GBX_OK) continue;
          v.push_back(k);
        }
      }
    // Persistent form with async staging: minor chunks dividing M,
    // workgroups of up to one item per thread.
    for (const auto &e : shapes)
      for (int mc : minor_chunks(M)) {
        const uint32_t mw = mc ? (uint32_t)mc : M;
        const uint64_t items = (uint64_t)e[0] * e[1] * e[2] * e[3] * mw;
        for (int t : {128, 256, 512}) {
          if ((uint64_t)t > (items + 63) / 64 * 64 && t != 128) continue;
          gbx_kcfg k{0, t, 1, 1, 2, mc, 1, 0, 0, pack_tile(e)};
          if (tile_cfg_check(k, u, in) != GBX_OK) continue;
          v.push_back(k);
        }
      }
    gbx_clear_error();
  }
  if (kinds & GBX_KIND_MARCH) {
    // Slab-marching kernel: march along t or z; slabs of 1, 2 or 4 sites;
    // tiles of 1..8 sites over the other three dimensions; whole or half
    // segments; minor chunks that make whole wavefronts per tile site.
    const int *L = u->geom->dims;
    static const int t3[][3] = {{1, 1, 1}, {2, 1, 1}, {1, 2, 1}, {2, 2, 1}, {1, 2, 2}, {2, 1, 2}, {4, 1, 1},
                                {1, 4, 1}, {2, 2, 2}, {4, 2, 1}, {2, 4, 1}};
    std::vector<int> mcs{0};
    for (int mc : minor_chunks(M))
      if (mc >= 16) mcs.push_back(mc);
    for (int md : {3, 2})
      for (int ts : {1, 2, 4})
        for (const auto &t : t3)
          for (int mc : mcs)
            for (int half = 0; half <= 1; half++) {
              int e[4], j = 0;
              for (int d = 0; d < 4; d++) e[d] = d == md ? ts : t[j++];
              const int seg = half ? L[md] / 2 : 0;
              if (half && (L[md] % 2 || (L[md] / 2) % ts)) continue;
              const uint32_t mw = mc ? (uint32_t)mc : M;
              const int thr = t[0] * t[1] * t[2] * ts * (int)mw;
              gbx_kcfg k{0, thr, 1, 1, 3, mc, 0, 0, seg, pack_march(e, md)};
              if (march_cfg_check(k, u, in) != GBX_OK) continue;
              v.push_back(k);
            }
    gbx_clear_error();
  }
  if ((kinds & GBX_KIND_MFMA) && mfma_supported(u, in)) {
    // MFMA form: minor columns per wavefront unit (multiples of 16 dividing
    // M), sites per workgroup, wavefronts per workgroup, next-tile gather.
    for (uint32_t mw = M; mw >= 16; mw /= 2) {
      if (M % mw || mw % 16) continue;
      for (int spw : {1, 2, 4, 8})
        for (int t : {128, 256}) {
          if ((uint64_t)spw * (M / mw) * 64 < (uint64_t)t) continue;  // fewer units than wavefronts
          for (int pf = 0; pf <= 1; pf++)
            for (uint32_t sw = M; sw >= mw && sw * 4 >= M; sw /= 2) {  // columns per lattice sweep
              if (M % sw || sw % mw) continue;
              v.push_back(gbx_kcfg{spw, t, 1, 1, 1, mw == M ? 0 : (int)mw, pf, 0, sw == M ? 0 : (int)sw, 0});
            }
        }
    }
  }
  return v;
}

// Stage 2: around the best shape -- XCD ordering, store kind, and the cap on
// resident workgroups per CU (fewer sites in flight per XCD = a working set
// its 4 MB L2 holds, at the cost of latency hiding).
std::vector<gbx_kcfg> refine_candidates(const gbx_kcfg &b)
{
  std::vector<gbx_kcfg> v;
  for (int xcd = 0; xcd <= 1; xcd++)
    for (int nt = 0; nt <= 1; nt++)
      for (int wgpc : {0, 2, 3, 4, 6, 8, 12, 16}) {
        gbx_kcfg k = b;
        k.xcd = xcd;
        k.nt = nt;
        k.wgpc = wgpc;
        v.push_back(k);
      }
  if (b.kind == 2) {  // tile kernel: XCD order, store kind, residency cap
    for (gbx_kcfg &k : v) k.wgpc = k.wgpc > 4 ? 0 : k.wgpc;
    return v;
  }
  if (b.kind == 3) {  // marching kernel: its LDS sets the residency; XCD order and store kind
    v.clear();
    for (int xcd = 0; xcd <= 1; xcd++)
      for (int nt = 0; nt <= 1; nt++) {
        gbx_kcfg k = b;
        k.xcd = xcd;
        k.nt = nt;
        v.push_back(k);
      }
    return v;
  }
  if (b.mc > 0 && b.kind == 0)  // chunk-inner visiting order: site groups per chunk cycle
    for (int cblk : {16, 64, 256, 1024}) {
      gbx_kcfg k = b;
      k.cblk = cblk;
      v.push_back(k);
    }
  return v;
}

bool parse_line(const std::string &line, std::string &sig, gbx_kcfg &k, double &us)
{
  char s[200];
  k = gbx_kcfg{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (sscanf(line.c_str(),
             "{\"sig\": \"%199[^\"]\", \"spw\": %d, \"threads\": %d, \"xcd\": %d, \"nt\": %d, \"kind\": %d, \"mc\": %d, "
             "\"pf\": %d, \"wgpc\": %d, \"cblk\": %d, ",
             s, &k.spw, &k.threads, &k.xcd, &k.nt, &k.kind, &k.mc, &k.pf, &k.wgpc, &k.cblk) != 10)
    return false;
  const size_t t = line.find("\"tile\": ");
  if (t != std::string::npos && sscanf(line.c_str() + t, "\"tile\": %d", &k.tile) != 1) return false;
  const size_t u = line.find("\"usec\": ");
  if (u == std::string::npos || sscanf(line.c_str() + u, "\"usec\": %lf", &us) != 1) return false;
  sig = s;
  return true;
}

}  // namespace

extern "C" int gbx_tune_kinds(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, int kinds, const char *cache_path,
                              gbx_kcfg *best, double *best_usec)
{
  GBX_CHECK(u && out && in && best && best_usec, GBX_ENULL, "gbx_tune: NULL argument");
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_tune: dag=%d", dag);
  GBX_CHECK(kinds >= 1 && kinds <= GBX_KIND_ALL, GBX_EINVAL, "gbx_tune: kinds=%d", kinds);
  const bool all = kinds == GBX_KIND_ALL;
  std::string sig = dhop_signature(u, in);
  if (!all) {
    sig += "#";
    if (kinds & GBX_KIND_VALU) sig += "valu";
    if (kinds & GBX_KIND_MFMA) sig += "mfma";
    if (kinds & GBX_KIND_TILE) sig += "tile";
    if (kinds & GBX_KIND_MARCH) sig += "march";
  }
  *best = gbx_kcfg{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  *best_usec = 0.0;
  if (cache_path) {
    std::ifstream f(cache_path);
    std::string line, s;
    gbx_kcfg k;
    double us;
    while (std::getline(f, line))
      if (parse_line(line, s, k, us) && s == sig) {
        *best = k;
        *best_usec = us;
      }
    if (best->threads > 0 && (best->spw > 0 || best->kind >= 2)) {
      // validate against this build before trusting it
      double us2;
      if (time_cfg(u, out, in, dag, *best, 2, &us2) == GBX_OK) {
        if (all) dhop_remember_cfg(sig, *best);
        *best_usec = us2;
        return GBX_OK;
      }
      gbx_clear_error();
    }
  }
  const std::vector<gbx_kcfg> cands = shape_candidates(u, in, kinds);
  GBX_CHECK(!cands.empty(), GBX_EUNSUPPORTED, "gbx_tune: no valid configuration of kinds %d for this problem", kinds);
  // Repetitions: about 10 ms of kernel time per candidate, at least 2.
  double us0 = 0.0;
  const gbx_kcfg def = cands.front();
  GBX_TRY(time_cfg(u, out, in, dag, def, 1, &us0));
  int nrep = (int)(10000.0 / (us0 > 1.0 ? us0 : 1.0));
  if (nrep < 2) nrep = 2;
  if (nrep > 50) nrep = 50;
  double bus = 1e300;
  gbx_kcfg bk = def;
  auto search = [&](const std::vector<gbx_kcfg> &list) {
    for (const gbx_kcfg &k : list) {
      double us = 0.0;
      if (time_cfg(u, out, in, dag, k, nrep, &us) != GBX_OK) {
        gbx_clear_error();
        continue;
      }
      if (us < bus) {
        bus = us;
        bk = k;
      }
    }
  };
  search(cands);
  GBX_CHECK(bus < 1e300, GBX_EHIP, "gbx_tune: every candidate failed");
  search(refine_candidates(bk));
  *best = bk;
  *best_usec = bus;
  if (all) dhop_remember_cfg(sig, bk);
  if (cache_path) {
    std::ofstream f(cache_path, std::ios::app);
    if (f)
      f << "{\"sig\": \"" << sig << "\", \"spw\": " << bk.spw << ", \"threads\": " << bk.threads << ", \"xcd\": " << bk.xcd
        << ", \"nt\": " << bk.nt << ", \"kind\": " << bk.kind << ", \"mc\": " << bk.mc << ", \"pf\": " << bk.pf
        << ", \"wgpc\": " << bk.wgpc << ", \"cblk\": " << bk.cblk << ", \"tile\": " << bk.tile << ", \"usec\": " << bus
        << "}\n";
  }
  return GBX_OK;
}

extern "C" int gbx_tune(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const char *cache_path, gbx_kcfg *best,
                        double *best_usec)
{
  return gbx_tune_kinds(u, out, in, dag, GBX_KIND_ALL, cache_path, best, best_usec);
}
