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
  hipEvent_t e0, e1;
  GBX_HIP(hipEventCreate(&e0));
  GBX_HIP(hipEventCreate(&e1));
  int rc = GBX_OK;
  (void)hipEventRecord(e0, st);
  for (int r = 0; r < nrep && rc == GBX_OK; r++) rc = dhop_run(u, out, in, dag, k, st);
  (void)hipEventRecord(e1, st);
  if (hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "tune: kernel failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == GBX_OK) *usec = 1000.0 * ms / nrep;
  return rc;
}

std::vector<gbx_kcfg> candidates(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  std::vector<gbx_kcfg> v;
  const uint32_t M = in->M;
  const int spws[] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64};
  for (int spw : spws) {
    const uint64_t items = (uint64_t)spw * M;
    if (items > 8192) continue;
    std::vector<int> thr;
    if (items <= 1024 && items % 64 == 0) thr.push_back((int)items);
    for (int t : {128, 256, 512}) {
      if ((uint64_t)t < items && (items % t == 0 || items > 4u * (uint64_t)t)) thr.push_back(t);
    }
    if (items < 1024 && items % 64 != 0) thr.push_back((int)((items + 63) / 64 * 64));
    for (int t : thr) {
      dhop_cfg c{spw, t, 1, 1};
      if (dhop_cfg_check(c, u, in) != GBX_OK) continue;
      for (int xcd = 0; xcd <= 1; xcd++)
        for (int nt = 0; nt <= 1; nt++) {
          gbx_kcfg k{spw, t, xcd, nt, 0};
          v.push_back(k);
        }
    }
  }
  gbx_clear_error();
  if (mfma_supported(u, in)) {
    for (int spw : {1, 2, 4})
      for (int xcd = 0; xcd <= 1; xcd++) v.push_back(gbx_kcfg{spw, 256, xcd, 1, 1});
  }
  return v;
}

bool parse_line(const std::string &line, std::string &sig, gbx_kcfg &k, double &us)
{
  char s[200];
  if (sscanf(line.c_str(), "{\"sig\": \"%199[^\"]\", \"spw\": %d, \"threads\": %d, \"xcd\": %d, \"nt\": %d, \"kind\": %d, \"usec\": %lf}",
             s, &k.spw, &k.threads, &k.xcd, &k.nt, &k.kind, &us) != 7)
    return false;
  sig = s;
  return true;
}

}  // namespace

extern "C" int gbx_tune(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const char *cache_path, gbx_kcfg *best,
                        double *best_usec)
{
  GBX_CHECK(u && out && in && best && best_usec, GBX_ENULL, "gbx_tune: NULL argument");
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_tune: dag=%d", dag);
  const std::string sig = dhop_signature(u, in);
  *best = gbx_kcfg{0, 0, 0, 0, 0};
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
    if (best->threads > 0 && best->spw > 0) {
      // validate against this build before trusting it
      double us2;
      if (time_cfg(u, out, in, dag, *best, 2, &us2) == GBX_OK) {
        dhop_remember_cfg(sig, *best);
        *best_usec = us2;
        return GBX_OK;
      }
      gbx_clear_error();
    }
  }
  const std::vector<gbx_kcfg> cands = candidates(u, in);
  GBX_CHECK(!cands.empty(), GBX_EINVAL, "gbx_tune: no valid configuration");
  // Repetitions: about 20 ms of kernel time per candidate, at least 2.
  double us0 = 0.0;
  gbx_kcfg def;
  gbx_dhop_get_cfg(u, in, &def);
  GBX_TRY(time_cfg(u, out, in, dag, def, 1, &us0));
  int nrep = (int)(20000.0 / (us0 > 1.0 ? us0 : 1.0));
  if (nrep < 2) nrep = 2;
  if (nrep > 50) nrep = 50;
  double bus = 1e300;
  gbx_kcfg bk = def;
  for (const gbx_kcfg &k : cands) {
    double us;
    if (time_cfg(u, out, in, dag, k, nrep, &us) != GBX_OK) {
      gbx_clear_error();
      continue;
    }
    if (us < bus) {
      bus = us;
      bk = k;
    }
  }
  GBX_CHECK(bus < 1e300, GBX_EHIP, "gbx_tune: every candidate failed");
  *best = bk;
  *best_usec = bus;
  dhop_remember_cfg(sig, bk);
  if (cache_path) {
    std::ofstream f(cache_path, std::ios::app);
    if (f)
      f << "{\"sig\": \"" << sig << "\", \"spw\": " << bk.spw << ", \"threads\": " << bk.threads << ", \"xcd\": " << bk.xcd
        << ", \"nt\": " << bk.nt << ", \"kind\": " << bk.kind << ", \"usec\": " << bus << "}\n";
  }
  return GBX_OK;
}
