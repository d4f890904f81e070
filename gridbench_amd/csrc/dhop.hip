// The VALU hopping-term kernel over the SoA layout and its C-ABI entries.
//
// Workgroup = `spw` consecutive slots x all M = Ls*nrhs minor indices.  The
// 8 links of each of its sites are read from HBM once, converted to the
// arithmetic precision (and rebuilt from 12 reals when compressed) into LDS,
// and applied from LDS to every s-slice and right-hand side of the site.  A
// thread computes one (site, m) 5D point: 8 neighbour gathers of 16-byte
// chunks (a wavefront of consecutive m reads contiguous 1 KiB runs), 8 x
// (project, SU(3) x 2 colour vectors, reconstruct) in registers, one store.
// With xcd = 1 workgroup b runs logical block (b % 8) * ceil(nwg/8) + b / 8:
// the hardware deals workgroups round-robin over the 8 XCDs, so each XCD
// then walks one contiguous eighth of the slot range and its L2 sees the
// spatially compact neighbourhoods the slot order provides.
#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "dhop_dev.h"
#include "gbx_internal.h"

namespace {

// Logical workgroup L covers sites q0 .. q0+spw-1 (site block L % nwg_sites)
// and minor indices m0 .. m0+mc-1 (minor chunk L / nwg_sites): with mc < M
// the lattice is swept once per minor chunk, so the 8-neighbour working set
// is that of an mc-wide field (the links are re-read once per chunk).
// GP: a domain part whose ghosts hold spin-projected halves from ghost0 on
// (dhop_site_gp).  TD: fields in the GBX_BASIS_TDIAG spin basis.
template <int P, bool DAG, bool C12, bool AXPBY, int PF, bool EXACT = false, bool GP = false, bool TD = false>
__global__ void __launch_bounds__(PF ? 256 : 1024)
    k_dhop(const void *__restrict__ gauge, const uint32_t *__restrict__ nbr, const void *__restrict__ in,
           void *__restrict__ out, const void *__restrict__ center, uint32_t nslot, uint32_t M, uint32_t spw, uint32_t mc,
           uint32_t nwg_sites, uint32_t cblk, uint32_t nchunk, uint32_t nwg, uint32_t per_xcd, int xcd, int nt,
           typename dev::Traits<P>::acc alpha,
           typename dev::Traits<P>::acc beta, uint32_t ghost0)
{
  typedef typename dev::Traits<P>::acc A;
  extern __shared__ __align__(16) unsigned char smem[];
  A *sU = (A *)smem;
  uint32_t *sN = (uint32_t *)(sU + (size_t)spw * 8 * 18);

  const uint32_t L = xcd ? (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (L >= nwg) return;  // uniform over the workgroup, before any barrier
  // cblk site groups are visited for every minor chunk before the next
  // cblk groups (chunk-inner order); cblk == nwg_sites sweeps the lattice
  // once per chunk (chunk-outer order).
  const uint32_t per = cblk * nchunk, grp = L / per, r = L - grp * per;
  const uint32_t chunk = r / cblk, b = grp * cblk + (r - chunk * cblk);
  if (b >= nwg_sites) return;
  const uint32_t q0 = b * spw, m0 = chunk * mc;
  const uint32_t nsl = min(spw, nslot - q0), mcc = min(mc, M - m0);
  dev::stage_links<P, C12>(gauge, nbr, q0, nsl, sU, sN);

  const uint32_t items = nsl * mcc;
  for (uint32_t w = threadIdx.x; w < items; w += blockDim.x) {
    const uint32_t sl = w / mcc, m = m0 + (w - sl * mcc);
    A rr[12], ri[12];
    if constexpr (GP) dev::dhop_site_gp<P, DAG, PF>(in, sN + sl * 8, M, m, ghost0, sU + sl * 8 * 18, rr, ri);
    else dev::dhop_site<P, DAG, PF, EXACT, TD>(in, sN + sl * 8, M, m, sU + sl * 8 * 18, rr, ri);
    dev::store_site<P, AXPBY>(out, center, q0 + sl, M, m, rr, ri, alpha, beta, nt != 0);
  }
}

struct LaunchArgs {
  dim3 grid, block;
  size_t lds;
  hipStream_t st;
  const void *gauge;
  const uint32_t *nbr;
  const void *in;
  void *out;
  const void *center;
  uint32_t nslot, M, spw, mc, nwg_sites, cblk, nchunk, nwg, per_xcd;
  int xcd, nt;
  double alpha, beta;
  uint32_t ghost0;
};

template <int P, bool DAG, bool C12, bool AXPBY, int PF, bool EXACT = false, bool GP = false, bool TD = false>
void launch_t(const LaunchArgs &a)
{
  typedef typename dev::Traits<P>::acc A;
  hipLaunchKernelGGL((k_dhop<P, DAG, C12, AXPBY, PF, EXACT, GP, TD>), a.grid, a.block, a.lds, a.st, a.gauge, a.nbr, a.in, a.out,
                     a.center, a.nslot, a.M, a.spw, a.mc, a.nwg_sites, a.cblk, a.nchunk, a.nwg, a.per_xcd, a.xcd, a.nt,
                     (A)a.alpha, (A)a.beta, a.ghost0);
}

// Projected ghosts (FP64 / FP32, no axpby).
template <int P>
void launch_gp(bool dag, bool c12, int pf, const LaunchArgs &a)
{
#define GBX_GP(D_, C_)                                        \
  if (pf == 2) launch_t<P, D_, C_, false, 2, false, true>(a); \
  else if (pf == 1) launch_t<P, D_, C_, false, 1, false, true>(a); \
  else launch_t<P, D_, C_, false, 0, false, true>(a);
  if (dag) {
    if (c12) { GBX_GP(true, true) } else { GBX_GP(true, false) }
  } else {
    if (c12) { GBX_GP(false, true) } else { GBX_GP(false, false) }
  }
#undef GBX_GP
}

template <int P, bool DAG, bool C12, bool TD = false>
void launch_dc(bool axpby, int pf, const LaunchArgs &a)
{
  if (axpby) {
    if (pf == 2) launch_t<P, DAG, C12, true, 2, false, false, TD>(a);
    else if (pf == 1) launch_t<P, DAG, C12, true, 1, false, false, TD>(a);
    else launch_t<P, DAG, C12, true, 0, false, false, TD>(a);
  } else if (pf == 2) launch_t<P, DAG, C12, false, 2, false, false, TD>(a);
  else if (pf == 1) launch_t<P, DAG, C12, false, 1, false, false, TD>(a);
  else launch_t<P, DAG, C12, false, 0, false, false, TD>(a);
}

template <int P, bool TD = false>
void launch_p(bool dag, bool c12, bool axpby, int pf, const LaunchArgs &a)
{
  if (!dag && !c12) launch_dc<P, false, false, TD>(axpby, pf, a);
  else if (!dag && c12) launch_dc<P, false, true, TD>(axpby, pf, a);
  else if (dag && !c12) launch_dc<P, true, false, TD>(axpby, pf, a);
  else launch_dc<P, true, true, TD>(axpby, pf, a);
}

// The reference's rounding (FP64 / FP32 storage, 18-real links, no axpby).
template <int P>
void launch_exact(bool dag, int pf, const LaunchArgs &a)
{
  if (dag) {
    if (pf == 2) launch_t<P, true, false, false, 2, true>(a);
    else if (pf == 1) launch_t<P, true, false, false, 1, true>(a);
    else launch_t<P, true, false, false, 0, true>(a);
  } else {
    if (pf == 2) launch_t<P, false, false, false, 2, true>(a);
    else if (pf == 1) launch_t<P, false, false, false, 1, true>(a);
    else launch_t<P, false, false, false, 0, true>(a);
  }
}

size_t lds_bytes(int prec, int spw) { return (size_t)spw * 8 * 18 * (prec == GBX_F64 ? 8 : 4) + (size_t)spw * 8 * 4; }

// Tuned configurations keyed by problem signature.
std::mutex g_cfg_mu;
std::unordered_map<std::string, gbx_kcfg> g_cfg;

}  // namespace

std::string dhop_signature(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  char buf[160];
  const gbx_geom_s *g = u->geom;
  snprintf(buf, sizeof buf, "p%d_c%d_Ls%d_n%d_%dx%dx%dx%d", in->prec, u->comp12, in->Ls, in->nrhs, g->dims[0], g->dims[1],
           g->dims[2], g->dims[3]);
  std::string s = buf;
  if (g->parity >= 0) s += g->parity ? "_odd" : "_even";
  if (g->nghost) {  // a domain part: ghost slots, no tiled kernel
    snprintf(buf, sizeof buf, "_part%dx%dx%dx%d", g->grid[0], g->grid[1], g->grid[2], g->grid[3]);
    s += buf;
  }
  if (g->nlevels == 0 && g->nghost == 0 && g->parity < 0) s += "_order";  // a caller-defined site order
  if (in->basis == GBX_BASIS_TDIAG) s += "_td";
  for (int l = 0; l < g->nlevels; l++) {
    snprintf(buf, sizeof buf, "_b%dx%dx%dx%d", g->levels[l * 4], g->levels[l * 4 + 1], g->levels[l * 4 + 2],
             g->levels[l * 4 + 3]);
    s += buf;
  }
  return s;
}

void dhop_remember_cfg(const std::string &sig, const gbx_kcfg &c)
{
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_cfg[sig] = c;
}

bool dhop_lookup_cfg(const std::string &sig, gbx_kcfg *c)
{
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  auto it = g_cfg.find(sig);
  if (it == g_cfg.end()) return false;
  *c = it->second;
  return true;
}

dhop_cfg dhop_default_cfg(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  (void)u;
  dhop_cfg c;
  const uint32_t M = in->M;
  // LDS per workgroup stays <= 40 KB so several workgroups share a CU.
  const int spw_max = in->prec == GBX_F64 ? 32 : 64;
  // Prefer a workgroup of exactly spw*M threads, a multiple of 64, near 256.
  int spw = 0;
  for (int s = 1; s <= spw_max; s++) {
    const uint32_t t = (uint32_t)s * M;
    if (t > 1024) break;
    if (t % 64 == 0 && t >= 256) {
      spw = s;
      break;
    }
  }
  if (spw) {
    c.spw = spw;
    c.threads = (int)(spw * M);
  } else if (M >= 256) {
    c.spw = 1;
    c.threads = 256;
  } else {
    c.spw = (int)((256 + M - 1) / M);
    if (c.spw > spw_max) c.spw = spw_max;
    c.threads = (int)(((uint32_t)c.spw * M + 63) / 64 * 64);
    if (c.threads > 1024) c.threads = 1024;
  }
  c.xcd = 1;
  c.nt = 1;
  c.mc = 0;
  c.pf = 0;
  c.wgpc = 0;
  c.cblk = 0;
  return c;
}

int dhop_cfg_check(const dhop_cfg &c, const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  (void)in;
  GBX_CHECK(c.spw >= 1 && c.spw <= 4096, GBX_EINVAL, "kernel config: spw=%d", c.spw);
  GBX_CHECK(c.pf >= 0 && c.pf <= 2, GBX_EINVAL, "kernel config: pf=%d", c.pf);
  GBX_CHECK(c.threads >= 64 && c.threads <= (c.pf ? 256 : 1024) && c.threads % 64 == 0, GBX_EINVAL,
            "kernel config: threads=%d (pf=%d)", c.threads, c.pf);
  GBX_CHECK(c.mc >= 0, GBX_EINVAL, "kernel config: mc=%d", c.mc);
  GBX_CHECK(c.wgpc >= 0 && c.wgpc <= 64, GBX_EINVAL, "kernel config: wgpc=%d", c.wgpc);
  GBX_CHECK(lds_bytes(u->prec, c.spw) <= 160 * 1024, GBX_EINVAL, "kernel config: spw=%d needs %zu B of LDS", c.spw,
            lds_bytes(u->prec, c.spw));
  return GBX_OK;
}

static int check_operands(const gbx_gauge_s *u, const gbx_spinor_s *out, const gbx_spinor_s *in)
{
  GBX_CHECK(u && out && in, GBX_ENULL, "dhop: NULL handle");
  GBX_CHECK(u->d && out->d && in->d, GBX_ENULL, "dhop: unallocated field");
  GBX_CHECK(out != in && out->d != in->d, GBX_EALIAS, "dhop: out aliases in");
  GBX_CHECK(u->geom == out->geom, GBX_EMISMATCH, "dhop: gauge and output fields on different geometries");
  GBX_CHECK(dhop_input_geom_ok(out->geom, in->geom), GBX_EMISMATCH,
            out->geom->parity < 0 ? "dhop: input field on a different geometry"
                                  : "dhop: on a red-black half the input must live on the partner half of the same pair");
  GBX_CHECK(out->Ls == in->Ls && out->nrhs == in->nrhs, GBX_EMISMATCH, "dhop: in/out shapes differ");
  GBX_CHECK(out->prec == in->prec && u->prec == in->prec, GBX_EMISMATCH,
            "dhop: precisions differ (gauge %d, in %d, out %d)", u->prec, in->prec, out->prec);
  return GBX_OK;
}

int dhop_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const dhop_cfg &c, double alpha,
                double beta, const gbx_spinor_s *center, hipStream_t st, bool exact, SlotRange range)
{
  GBX_TRY(check_operands(u, out, in));
  GBX_TRY(dhop_cfg_check(c, u, in));
  const bool axpby = !(alpha == 0.0 && beta == 1.0);
  GBX_CHECK(!exact || ((u->prec == GBX_F64 || u->prec == GBX_F32) && !u->comp12 && !axpby), GBX_EUNSUPPORTED,
            "dhop: the reference-rounding kernel takes FP64 / FP32 fields with 18-real links");
  const bool td = in->basis == GBX_BASIS_TDIAG;
  GBX_CHECK(!td || (!exact && !c.projected_ghosts), GBX_EUNSUPPORTED,
            "dhop: the reference-rounding and projected-ghost forms take chiral-basis fields only");
  if (axpby) {
    GBX_CHECK(center && center->d, GBX_ENULL, "dhop_axpby: NULL center field");
    GBX_CHECK(center->d != out->d, GBX_EALIAS, "dhop_axpby: out aliases the center field");
    GBX_CHECK(center->prec == in->prec && center->M == in->M && center->geom == out->geom, GBX_EMISMATCH,
              "dhop_axpby: the added field must have the output's geometry, precision and shape");
    GBX_CHECK(center->basis == in->basis, GBX_EMISMATCH, "dhop_axpby: the added field is in another spin basis");
  }
  const uint64_t q0 = range.begin, q1 = std::min<uint64_t>(range.end, u->geom->vol);
  GBX_CHECK(q0 <= q1, GBX_EINVAL, "dhop: slot range %llu .. %llu", (unsigned long long)q0, (unsigned long long)q1);
  // The output takes the input's basis when it is written whole; a part of
  // it only when it is in that basis already.
  const bool whole = q0 == 0 && q1 == u->geom->vol;
  GBX_CHECK(whole || out->basis == in->basis, GBX_EMISMATCH,
            "dhop: a slot range of an output in another spin basis than the input");
  out->basis = in->basis;
  if (q0 == q1) return GBX_OK;
  LaunchArgs a;
  a.nslot = (uint32_t)(q1 - q0);
  a.M = in->M;
  a.spw = (uint32_t)c.spw;
  a.mc = (c.mc <= 0 || (uint32_t)c.mc > in->M) ? in->M : (uint32_t)c.mc;
  a.nwg_sites = (a.nslot + a.spw - 1) / a.spw;
  a.nchunk = (a.M + a.mc - 1) / a.mc;
  a.cblk = (c.cblk <= 0 || (uint32_t)c.cblk >= a.nwg_sites) ? a.nwg_sites : (uint32_t)c.cblk;
  const uint64_t ngrp = (a.nwg_sites + a.cblk - 1) / a.cblk;
  const uint64_t nwg64 = ngrp * a.cblk * a.nchunk;
  GBX_CHECK(nwg64 < (1ull << 31), GBX_EINVAL, "dhop: %llu workgroups", (unsigned long long)nwg64);
  a.nwg = (uint32_t)nwg64;
  a.per_xcd = (a.nwg + 7) / 8;
  a.grid = dim3(c.xcd ? a.per_xcd * 8 : a.nwg);
  a.block = dim3((unsigned)c.threads);
  a.lds = lds_bytes(u->prec, c.spw);
  if (c.wgpc > 0) {
    const size_t pad = ((size_t)160 * 1024 / (size_t)c.wgpc) & ~(size_t)511;
    if (pad > a.lds) a.lds = pad;
  }
  a.st = st;
  const uint64_t pb = (uint64_t)prec_bytes(u->prec), slot_bytes = (uint64_t)in->M * 24 * pb;
  a.gauge = (const char *)u->d + q0 * 8 * (u->comp12 ? 12 : 18) * pb;
  a.nbr = u->geom->nbr + q0 * 8;
  a.in = in->d;
  a.out = (char *)out->d + q0 * slot_bytes;
  a.center = axpby ? (const char *)center->d + q0 * slot_bytes : nullptr;
  a.xcd = c.xcd;
  a.nt = c.nt;
  a.alpha = alpha;
  a.beta = beta;
  a.ghost0 = (uint32_t)u->geom->vol;
  const int pf = (axpby && !c.axpby_pf) ? 0 : c.pf;
  if (c.projected_ghosts) {
    GBX_CHECK((u->prec == GBX_F64 || u->prec == GBX_F32) && !axpby && !exact && u->geom->nghost, GBX_EUNSUPPORTED,
              "dhop: projected ghosts need a domain part, FP64 / FP32 fields and no axpby");
    if (u->prec == GBX_F64) launch_gp<GBX_F64>(dag, u->comp12, pf, a);
    else launch_gp<GBX_F32>(dag, u->comp12, pf, a);
    GBX_HIP(hipGetLastError());
    return GBX_OK;
  }
  if (exact) {
    if (u->prec == GBX_F64) launch_exact<GBX_F64>(dag, pf, a);
    else launch_exact<GBX_F32>(dag, pf, a);
    GBX_HIP(hipGetLastError());
    return GBX_OK;
  }
  if (td) {
    if (u->prec == GBX_F64) launch_p<GBX_F64, true>(dag, u->comp12, axpby, pf, a);
    else launch_p<GBX_F32, true>(dag, u->comp12, axpby, pf, a);
    GBX_HIP(hipGetLastError());
    return GBX_OK;
  }
  switch (u->prec) {
  case GBX_F64: launch_p<GBX_F64>(dag, u->comp12, axpby, pf, a); break;
  case GBX_F32: launch_p<GBX_F32>(dag, u->comp12, axpby, pf, a); break;
  case GBX_F16: launch_p<GBX_F16>(dag, u->comp12, axpby, pf, a); break;
  default: launch_p<GBX_BF16>(dag, u->comp12, axpby, pf, a); break;
  }
  GBX_HIP(hipGetLastError());
  return GBX_OK;
}

// ------------------------------------------------------------------- C-ABI
int mfma_supported(const gbx_gauge_s *u, const gbx_spinor_s *in);
int mfma_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &c, hipStream_t st,
                SlotRange range);

static gbx_kcfg to_kcfg(const dhop_cfg &c)
{
  gbx_kcfg k;
  k.spw = c.spw;
  k.threads = c.threads;
  k.xcd = c.xcd;
  k.nt = c.nt;
  k.kind = 0;
  k.mc = c.mc;
  k.pf = c.pf;
  k.wgpc = c.wgpc;
  k.cblk = c.cblk;
  k.tile = 0;
  return k;
}

static dhop_cfg from_kcfg(const gbx_kcfg &k)
{
  dhop_cfg c;
  c.spw = k.spw;
  c.threads = k.threads;
  c.xcd = k.xcd;
  c.nt = k.nt;
  c.mc = k.mc;
  c.pf = k.pf;
  c.wgpc = k.wgpc;
  c.cblk = k.cblk;
  return c;
}

// The reference-rounding kernel's shape when no tuned one is cached: the
// gather of the next leg issued ahead (pf 1), FP64 on 64-thread and FP32 on
// 256-thread workgroups when Ls*nrhs allows (24^4 x Ls16 drop-in: 832 us
// FP64 / 416 us FP32, against 854 / 424 us for the generic default).
static dhop_cfg exact_default_cfg(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  dhop_cfg c = dhop_default_cfg(u, in);
  const uint32_t M = in->M, target = u->prec == GBX_F64 ? 64u : 256u;
  if (M <= target) {
    c.spw = (int)(target / M);
    c.threads = (int)(((uint32_t)c.spw * M + 63) / 64 * 64);
    c.pf = 1;
  }
  return c;
}

int dhop_run_exact(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, hipStream_t st)
{
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "dhop: dag must be 0 or 1");
  gbx_kcfg k;
  const dhop_cfg c = dhop_lookup_cfg(dhop_signature(u, in), &k) && k.kind == 0 ? from_kcfg(k) : exact_default_cfg(u, in);
  return dhop_launch(u, out, in, dag, c, 0.0, 1.0, nullptr, st, true);
}

int dhop_run(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k, hipStream_t st,
             SlotRange range)
{
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "dhop: dag must be 0 or 1");
  const bool whole = range.begin == 0 && range.end >= u->geom->vol;
  if (k.kind == 1) return mfma_launch(u, out, in, dag, k, st, range);
  if (k.kind == 2) {
    GBX_CHECK(whole, GBX_EUNSUPPORTED, "dhop: the tiled kernel (kind 2) runs on the whole lattice only");
    return tile_launch(u, out, in, dag, k, st);
  }
  if (k.kind == 3) {
    GBX_CHECK(whole, GBX_EUNSUPPORTED, "dhop: the slab-marching kernel (kind 3) runs on the whole lattice only");
    return march_launch(u, out, in, dag, k, st);
  }
  GBX_CHECK(k.kind == 0, GBX_EINVAL, "dhop: kernel kind %d", k.kind);
  return dhop_launch(u, out, in, dag, from_kcfg(k), 0.0, 1.0, nullptr, st, false, range);
}

extern "C" int gbx_dhop_get_cfg(gbx_gauge u, gbx_spinor in, gbx_kcfg *cfg)
{
  GBX_CHECK(u && in && cfg, GBX_ENULL, "gbx_dhop_get_cfg: NULL argument");
  if (!dhop_lookup_cfg(dhop_signature(u, in), cfg)) *cfg = to_kcfg(dhop_default_cfg(u, in));
  return GBX_OK;
}

extern "C" int gbx_dhop_cfg(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const gbx_kcfg *cfg)
{
  GBX_CHECK(cfg, GBX_ENULL, "gbx_dhop_cfg: NULL config");
  GBX_TRY(check_operands(u, out, in));
  return dhop_run(u, out, in, dag, *cfg, u->geom->stream);
}

extern "C" int gbx_dhop(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag)
{
  GBX_TRY(check_operands(u, out, in));
  gbx_kcfg k;
  gbx_dhop_get_cfg(u, in, &k);
  return dhop_run(u, out, in, dag, k, u->geom->stream);
}

gbx_kcfg dhop_range_cfg(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  gbx_kcfg k;
  gbx_dhop_get_cfg(const_cast<gbx_gauge_s *>(u), const_cast<gbx_spinor_s *>(in), &k);
  if (k.kind >= 2) k = to_kcfg(dhop_default_cfg(u, in));  // the tiled and marching kernels cover whole lattices only
  return k;
}

int dhop_check_operands(const gbx_gauge_s *u, const gbx_spinor_s *out, const gbx_spinor_s *in)
{
  return check_operands(u, out, in);
}

extern "C" int gbx_dhop_range(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, uint64_t slot_begin, uint64_t slot_end)
{
  GBX_TRY(check_operands(u, out, in));
  GBX_CHECK(slot_begin <= slot_end && slot_end <= u->geom->vol, GBX_EINVAL, "gbx_dhop_range: slots %llu .. %llu of %llu",
            (unsigned long long)slot_begin, (unsigned long long)slot_end, (unsigned long long)u->geom->vol);
  SlotRange r;
  r.begin = slot_begin;
  r.end = slot_end;
  return dhop_run(u, out, in, dag, dhop_range_cfg(u, in), u->geom->stream, r);
}

extern "C" int gbx_dhop_faces_projected(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag)
{
  GBX_TRY(check_operands(u, out, in));
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_dhop_faces_projected: dag must be 0 or 1");
  const gbx_geom_s *g = in->geom;
  GBX_CHECK(g->nghost != 0, GBX_EINVAL, "gbx_dhop_faces_projected: the field's geometry is not a domain part");
  GBX_CHECK(in->prec == GBX_F64 || in->prec == GBX_F32, GBX_EUNSUPPORTED,
            "gbx_dhop_faces_projected: projected ghosts of FP64 / FP32 fields only");
  const gbx_kcfg k = dhop_range_cfg(u, in);
  dhop_cfg c = k.kind == 0 ? from_kcfg(k) : dhop_default_cfg(u, in);
  c.projected_ghosts = true;
  SlotRange r;
  r.begin = g->ninterior;
  r.end = g->vol;
  return dhop_launch(u, out, in, dag, c, 0.0, 1.0, nullptr, g->stream, false, r);
}

extern "C" int gbx_dhop_xpay(gbx_gauge u, gbx_spinor out, gbx_spinor x, gbx_spinor in, int dag, double alpha, double beta)
{
  GBX_TRY(check_operands(u, out, in));
  GBX_CHECK(x, GBX_ENULL, "gbx_dhop_xpay: NULL field x");
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_dhop_xpay: dag must be 0 or 1");
  gbx_kcfg k;
  gbx_dhop_get_cfg(u, in, &k);
  if (k.kind != 0) k = to_kcfg(dhop_default_cfg(u, in));
  return dhop_launch(u, out, in, dag, from_kcfg(k), alpha, beta, x, u->geom->stream);
}

extern "C" int gbx_dhop_xpay_cfg(gbx_gauge u, gbx_spinor out, gbx_spinor x, gbx_spinor in, int dag, double alpha,
                                 double beta, const gbx_kcfg *cfg)
{
  GBX_CHECK(cfg, GBX_ENULL, "gbx_dhop_xpay_cfg: NULL config");
  GBX_TRY(check_operands(u, out, in));
  GBX_CHECK(x, GBX_ENULL, "gbx_dhop_xpay_cfg: NULL field x");
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_dhop_xpay_cfg: dag must be 0 or 1");
  GBX_CHECK(cfg->kind == 0, GBX_EUNSUPPORTED, "gbx_dhop_xpay_cfg: the fused form runs on the VALU kernel (kind 0)");
  dhop_cfg c = from_kcfg(*cfg);
  c.axpby_pf = true;
  return dhop_launch(u, out, in, dag, c, alpha, beta, x, u->geom->stream);
}

extern "C" int gbx_dhop_axpby(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, double alpha, double beta)
{
  GBX_CHECK(!u || u->geom->parity < 0, GBX_EINVAL,
            "gbx_dhop_axpby: on a red-black half in and out have different parities; use gbx_dhop_xpay");
  return gbx_dhop_xpay(u, out, in, in, dag, alpha, beta);
}

extern "C" int gbx_dhop_time(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const gbx_kcfg *cfg, int nrep, double *usec)
{
  GBX_CHECK(usec, GBX_ENULL, "gbx_dhop_time: NULL usec");
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "gbx_dhop_time: nrep=%d", nrep);
  GBX_TRY(check_operands(u, out, in));
  gbx_kcfg k;
  if (cfg) k = *cfg;
  else gbx_dhop_get_cfg(u, in, &k);
  hipStream_t st = u->geom->stream;
  GBX_TRY(dhop_run(u, out, in, dag, k, st));  // warm-up
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GBX_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = gbx_fail(GBX_EHIP, "hipEventCreate failed");
  if (rc == GBX_OK && hipEventRecord(e0, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
  for (int r = 0; r < nrep && rc == GBX_OK; r++) rc = dhop_run(u, out, in, dag, k, st);
  if (rc == GBX_OK && hipEventRecord(e1, st) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventRecord failed");
  if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventSynchronize failed");
  float ms = 0.f;
  if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "hipEventElapsedTime failed");
  if (rc != GBX_OK) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == GBX_OK) *usec = 1000.0 * (double)ms / nrep;
  return rc;
}
