// Batched hopping term over many independent lattices in one persistent,
// HIP-graph-captured launch.
//
// Work list: one queue per XCD.  Queue x holds, for every lattice in turn,
// the x-th eighth of its slot range cut into items of `spw` sites (spw*M
// close to one workgroup of work), in slot order: each XCD's L2 then sees the
// spatially compact neighbourhoods of the slot order, as the single-lattice
// kernel's XCD mapping gives it (one queue over all eight XCDs spread
// consecutive items over different L2s, and most gathers missed).  A
// persistent grid of workgroups (a few per CU; the hardware deals workgroup b
// to XCD b % 8) pulls items from its own XCD's queue through an atomic
// counter, then from the other queues in turn until every queue is exhausted
// -- every workgroup reaches that exit, and the stealing evens out the tail.
// Per item the workgroup stages the item's links into LDS and computes all
// (site, m) points, exactly like the single-lattice kernel (with its gather
// pipeline PF).  The next item's index is claimed while the current item is
// computed, so the atomic's round trip overlaps the work.  The counters are
// cleared by a kernel node (no memset nodes in the graph).
#include <algorithm>
#include <cstring>
#include <vector>

#include <mutex>

#include "dhop_dev.h"
#include "gbx_internal.h"

// The no-allocation entry's work-queue counter and lattice table, one copy
// per GPU, held until the process ends.  In host code the name is not a
// device address: host code uses the address hipGetSymbolAddress returns for
// the current device (na_scratch_ptr), never the name.
static __device__ __align__(256) unsigned char scratch[GBX_NOALLOC_SCRATCH_BYTES];

namespace {

struct LatDesc {
  const void *gauge;
  const uint32_t *nbr;
  const void *in;
  void *out;
  uint32_t M, nslot, c12, pad;
};

constexpr int kThreads = 256;

// The per-XCD queues' counters, 64 bytes apart (their own cache lines).
constexpr int kQueues = 8, kCtrStride = 16;
__global__ void k_reset_queues(uint32_t *c)
{
  if (threadIdx.x < kQueues) c[threadIdx.x * kCtrStride] = 0u;
}

// Items [start[q], start[q+1]) form queue q.
struct QueueBounds {
  uint32_t start[kQueues + 1];
};

// Next item of queue q, or UINT32_MAX once it is exhausted.
__device__ __forceinline__ uint32_t claim(uint32_t *ctr, const QueueBounds &qb, uint32_t q)
{
  const uint32_t i = qb.start[q] + atomicAdd(ctr + q * kCtrStride, 1u);
  return i < qb.start[q + 1] ? i : 0xffffffffu;
}

// One work item: sites q0 .. q0+nsl-1 of lattice L, every minor index.
template <int P, bool DAG, int PF, bool TD, int T = kThreads>
__device__ __forceinline__ void batch_item(const LatDesc &L, uint32_t q0, uint32_t nsl, typename dev::Traits<P>::acc *sU,
                                           uint32_t *sN, int nt)
{
  typedef typename dev::Traits<P>::acc A;
  if (L.c12) dev::stage_links<P, true>(L.gauge, L.nbr, q0, nsl, sU, sN);
  else dev::stage_links<P, false>(L.gauge, L.nbr, q0, nsl, sU, sN);
  const uint32_t M = L.M, n = nsl * M;
  for (uint32_t w = threadIdx.x; w < n; w += T) {
    const uint32_t sl = w / M, m = w - sl * M;
    A rr[12], ri[12];
    dev::dhop_site<P, DAG, PF, false, TD>(L.in, sN + sl * 8, M, m, sU + sl * 8 * 18, rr, ri);
    dev::store_site<P, false>(L.out, nullptr, q0 + sl, M, m, rr, ri, (A)0, (A)1, nt != 0);
  }
}

template <int P, bool DAG, int PF, bool TD, int T>
__global__ void __launch_bounds__(T) k_dhop_batch(const LatDesc *__restrict__ lat, const uint4 *__restrict__ items,
                                                  QueueBounds qb, uint32_t *__restrict__ counter, uint32_t max_spw, int nt)
{
  typedef typename dev::Traits<P>::acc A;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_next;
  A *sU = (A *)smem;
  uint32_t *sN = (uint32_t *)(sU + (size_t)max_spw * 8 * 18);
  const uint32_t x = blockIdx.x % kQueues;  // the XCD this workgroup runs on
  for (uint32_t k = 0; k < kQueues; k++) {  // its own queue, then the others
    const uint32_t q = (x + k) % kQueues;
    if (threadIdx.x == 0) s_next = claim(counter, qb, q);
    __syncthreads();
    uint32_t idx = s_next;
    while (idx != 0xffffffffu) {  // the same in every thread of the workgroup
      __syncthreads();            // every thread has read s_next
      if (threadIdx.x == 0) s_next = claim(counter, qb, q);  // claimed now, read after this item
      const uint4 it = items[idx];
      const LatDesc L = lat[it.x];
      batch_item<P, DAG, PF, TD, T>(L, it.y, it.z, sU, sN, nt);
      __syncthreads();  // LDS is reused by the next item; s_next is written
      idx = s_next;
    }
    __syncthreads();  // every thread has read the exhausted claim before the next queue's
  }
}

// ------------------------------------------------ the no-allocation variant
// gbx_batch_run_noalloc keeps its work-queue counter and its lattice table
// in `scratch`, one fixed array per GPU (the library allocates nothing per
// call).  The table arrives as the set-up kernel's argument (no host-to-
// device copy); items are not listed: item k belongs to the lattice i with
// item0[i] <= k < item0[i+1], and covers sites (k - item0[i]) * spw[i] on.
struct NaLat {
  LatDesc d;
  uint32_t item0, spw;
};
struct NaTable {
  uint32_t n, nitems, max_spw, pad;
  NaLat lat[GBX_NOALLOC_MAX_LATTICES];
};
constexpr size_t kNaCounterBytes = 256;
static_assert(kNaCounterBytes + sizeof(NaTable) <= GBX_NOALLOC_SCRATCH_BYTES, "scratch too small for the table");
static_assert(sizeof(NaTable) <= 4096 - 64, "the table is passed as a kernel argument");

// One thread: clear the counter, copy the table (a graph-capturable kernel
// node, no memset or copy node).
__global__ void k_na_setup(NaTable t, unsigned char *__restrict__ sp)
{
  *(uint32_t *)sp = 0u;
  NaTable *dst = (NaTable *)(sp + kNaCounterBytes);
  dst->n = t.n;
  dst->nitems = t.nitems;
  dst->max_spw = t.max_spw;
  for (uint32_t i = 0; i < t.n; i++) dst->lat[i] = t.lat[i];
}

template <int P, bool DAG, int PF, bool TD = false>
__global__ void __launch_bounds__(kThreads) k_dhop_batch_na(unsigned char *__restrict__ sp, int nt)
{
  typedef typename dev::Traits<P>::acc A;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_next;
  uint32_t *counter = (uint32_t *)sp;
  const NaTable *tab = (const NaTable *)(sp + kNaCounterBytes);
  const uint32_t n = tab->n, nitems = tab->nitems;
  A *sU = (A *)smem;
  uint32_t *sN = (uint32_t *)(sU + (size_t)tab->max_spw * 8 * 18);
  if (threadIdx.x == 0) s_next = atomicAdd(counter, 1u);
  __syncthreads();
  uint32_t idx = s_next;
  while (idx < nitems) {  // idx is the same in every thread of the workgroup
    __syncthreads();
    if (threadIdx.x == 0) s_next = atomicAdd(counter, 1u);
    uint32_t i = 0;
    while (i + 1 < n && tab->lat[i + 1].item0 <= idx) i++;
    const NaLat &e = tab->lat[i];
    const uint32_t q0 = (idx - e.item0) * e.spw;
    batch_item<P, DAG, PF, TD>(e.d, q0, min(e.spw, e.d.nslot - q0), sU, sN, nt);
    __syncthreads();
    idx = s_next;
  }
}

}  // namespace

// Launch shape of the persistent kernel (gbx_batch_tune picks it): threads
// per workgroup, gather pipeline depth, persistent workgroups per CU,
// streaming stores, and the (site, minor) points one work item aims at.
struct BatchShape {
  int threads = kThreads, pf = 1, wgpc = 0, nt = 1, points = kThreads;
};

struct gbx_batch_s {
  int prec = 0, dag = 0, n = 0;
  bool td = false;  // every field in the GBX_BASIS_TDIAG spin basis
  uint32_t nitems = 0, max_spw = 1, cap_items = 0;
  QueueBounds qb{};
  int grid = 0, ncu = 256;
  BatchShape shape;
  hipStream_t stream = nullptr;
  LatDesc *d_lat = nullptr;
  uint4 *d_items = nullptr;
  uint32_t *d_counter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<gbx_geom_s *> geoms;
  std::vector<LatDesc> lat;
  // The fields the graph reads and writes (they must outlive the batch).
  // Every run checks that the inputs are still in the spin basis the graph
  // was captured for and stamps that basis on the outputs it writes.
  std::vector<gbx_spinor_s *> in, out;
  int basis = GBX_BASIS_CHIRAL;
  void drop_graph()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~gbx_batch_s()
  {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    (void)hipFree(d_lat);
    (void)hipFree(d_items);
    (void)hipFree(d_counter);
    if (stream) (void)hipStreamDestroy(stream);
    for (gbx_geom_s *g : geoms) geom_release(g);
  }
};

namespace {

size_t lds_for(int prec, uint32_t spw) { return (size_t)spw * 8 * 18 * (prec == GBX_F64 ? 8 : 4) + (size_t)spw * 8 * 4; }

// Sites per work item for a lattice of minor extent M: about `points`
// (site, minor) points, at most 64 sites.
uint32_t item_sites(uint32_t M, int points)
{
  uint32_t spw = ((uint32_t)points + M - 1) / M;
  return spw < 1 ? 1 : spw > 64 ? 64 : spw;
}

template <int P, bool DAG, bool TD, int T>
void launch_t(gbx_batch_s *b, hipStream_t st, size_t lds)
{
  if (b->shape.pf == 2)
    hipLaunchKernelGGL((k_dhop_batch<P, DAG, 2, TD, T>), dim3(b->grid), dim3(T), lds, st, (const LatDesc *)b->d_lat,
                       (const uint4 *)b->d_items, b->qb, b->d_counter, b->max_spw, b->shape.nt);
  else
    hipLaunchKernelGGL((k_dhop_batch<P, DAG, 1, TD, T>), dim3(b->grid), dim3(T), lds, st, (const LatDesc *)b->d_lat,
                       (const uint4 *)b->d_items, b->qb, b->d_counter, b->max_spw, b->shape.nt);
}

template <int P, bool TD>
void launch_d(gbx_batch_s *b, hipStream_t st, size_t lds)
{
  if (b->shape.threads == 512) {
    if (b->dag) launch_t<P, true, TD, 512>(b, st, lds);
    else launch_t<P, false, TD, 512>(b, st, lds);
  } else {
    if (b->dag) launch_t<P, true, TD, 256>(b, st, lds);
    else launch_t<P, false, TD, 256>(b, st, lds);
  }
}

template <int P>
void launch(gbx_batch_s *b, hipStream_t st)
{
  const size_t lds = lds_for(P, b->max_spw);
  hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(64), 0, st, b->d_counter);
  if constexpr (P == GBX_F64 || P == GBX_F32) {  // the precisions a TDIAG field can have
    if (b->td) {
      launch_d<P, true>(b, st, lds);
      return;
    }
  }
  launch_d<P, false>(b, st, lds);
}

void launch_any(gbx_batch_s *b, hipStream_t st)
{
  switch (b->prec) {
  case GBX_F64: launch<GBX_F64>(b, st); break;
  case GBX_F32: launch<GBX_F32>(b, st); break;
  case GBX_F16: launch<GBX_F16>(b, st); break;
  default: launch<GBX_BF16>(b, st); break;
  }
}

// Work list for the batch's shape: queue x holds, lattice after lattice,
// the slots [x * nslot / 8, (x + 1) * nslot / 8) of each lattice cut into
// items of item_sites(M, points) sites, in slot order; uploaded into d_items
// (grown as needed), queue bounds in qb.
int build_items(gbx_batch_s *b)
{
  std::vector<uint4> h;
  uint32_t max_spw = 1;
  for (int x = 0; x < kQueues; x++) {
    b->qb.start[x] = (uint32_t)h.size();
    for (int i = 0; i < b->n; i++) {
      const LatDesc &L = b->lat[i];
      const uint32_t spw = item_sites(L.M, b->shape.points);
      max_spw = std::max(max_spw, spw);
      const uint32_t s0 = (uint32_t)((uint64_t)L.nslot * x / kQueues), s1 = (uint32_t)((uint64_t)L.nslot * (x + 1) / kQueues);
      for (uint32_t q0 = s0; q0 < s1; q0 += spw) h.push_back(make_uint4((uint32_t)i, q0, std::min(spw, s1 - q0), 0));
    }
  }
  b->qb.start[kQueues] = (uint32_t)h.size();
  if (h.size() > b->cap_items) {
    (void)hipFree(b->d_items);
    b->d_items = nullptr;
    b->cap_items = 0;
    GBX_HIP(hipMalloc(&b->d_items, sizeof(uint4) * h.size()));
    b->cap_items = (uint32_t)h.size();
  }
  GBX_HIP(hipMemcpy(b->d_items, h.data(), sizeof(uint4) * h.size(), hipMemcpyHostToDevice));
  b->nitems = (uint32_t)h.size();
  b->max_spw = max_spw;
  const int wgpc = b->shape.wgpc > 0 ? b->shape.wgpc : (b->prec == GBX_F64 ? 4 : 6);
  b->grid = b->ncu * wgpc;
  return GBX_OK;
}

// (Re)capture the graph for the current shape: {counter reset, persistent kernel}.
int capture(gbx_batch_s *b)
{
  b->drop_graph();
  GBX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  launch_any(b, b->stream);
  hipGraph_t g = nullptr;
  GBX_HIP(hipStreamEndCapture(b->stream, &g));
  GBX_CHECK(g != nullptr, GBX_EHIP, "batch: empty graph capture");
  b->graph = g;
  GBX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return GBX_OK;
}

int shape_check(const BatchShape &sh)
{
  GBX_CHECK(sh.threads == 256 || sh.threads == 512, GBX_EINVAL, "batch: threads=%d (256 or 512)", sh.threads);
  GBX_CHECK(sh.pf == 1 || sh.pf == 2, GBX_EINVAL, "batch: pf=%d (1 or 2)", sh.pf);
  GBX_CHECK(sh.wgpc >= 0 && sh.wgpc <= 16, GBX_EINVAL, "batch: wgpc=%d", sh.wgpc);
  GBX_CHECK(sh.points >= 64 && sh.points <= 4096, GBX_EINVAL, "batch: points per item %d", sh.points);
  return GBX_OK;
}

}  // namespace

extern "C" int gbx_batch_create(int n, const gbx_gauge *u, const gbx_spinor *out, const gbx_spinor *in, int dag, gbx_batch *bp)
{
  GBX_CHECK(bp && u && out && in, GBX_ENULL, "gbx_batch_create: NULL argument");
  *bp = nullptr;
  GBX_CHECK(n >= 1, GBX_EINVAL, "gbx_batch_create: n=%d", n);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_batch_create: dag=%d", dag);
  const int prec = u[0] ? u[0]->prec : -1;
  std::vector<LatDesc> lat(n);
  for (int i = 0; i < n; i++) {
    GBX_CHECK(u[i] && out[i] && in[i], GBX_ENULL, "gbx_batch_create: NULL handle at %d", i);
    // the single-lattice entries' checks: allocated fields, geometry, equal
    // Ls and nrhs, one precision, no aliasing
    GBX_TRY(dhop_check_operands(u[i], out[i], in[i]));
    GBX_CHECK(u[i]->prec == prec, GBX_EMISMATCH, "gbx_batch_create: lattice %d precision differs from lattice 0", i);
    GBX_CHECK(in[i]->basis == in[0]->basis, GBX_EMISMATCH, "gbx_batch_create: lattice %d is in another spin basis", i);
    for (int j = 0; j < n; j++)
      GBX_CHECK(out[i]->d != in[j]->d && (j == i || out[i]->d != out[j]->d), GBX_EALIAS,
                "gbx_batch_create: output %d aliases another lattice's field", i);
    lat[i] = LatDesc{u[i]->d, u[i]->geom->nbr, in[i]->d, out[i]->d, in[i]->M, (uint32_t)u[i]->geom->vol,
                     (uint32_t)u[i]->comp12, 0};
  }
  GBX_TRY(gbx_require_device());
  gbx_batch_s *b = new gbx_batch_s();
  b->prec = prec;
  b->dag = dag;
  b->td = in[0]->basis == GBX_BASIS_TDIAG;
  b->basis = in[0]->basis;
  b->in.assign(in, in + n);
  b->out.assign(out, out + n);
  b->n = n;
  b->lat = lat;
  for (int i = 0; i < n; i++) {
    geom_retain(u[i]->geom);
    b->geoms.push_back(u[i]->geom);
  }
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  auto fail = [&](int s) {
    delete b;
    return s;
  };
  if (hipStreamCreate(&b->stream) != hipSuccess) return fail(gbx_fail(GBX_EHIP, "gbx_batch_create: stream"));
  if (hipMalloc(&b->d_lat, sizeof(LatDesc) * n) != hipSuccess ||
      hipMalloc(&b->d_counter, sizeof(uint32_t) * kQueues * kCtrStride) != hipSuccess)
    return fail(gbx_fail(GBX_ENOMEM, "gbx_batch_create: device allocation"));
  if (hipMemcpy(b->d_lat, lat.data(), sizeof(LatDesc) * n, hipMemcpyHostToDevice) != hipSuccess)
    return fail(gbx_fail(GBX_EHIP, "gbx_batch_create: upload"));
  int rc = build_items(b);
  if (rc != GBX_OK) return fail(rc);
  // Inputs may still be in flight on their geometries' streams.
  for (gbx_geom_s *g : b->geoms) (void)hipStreamSynchronize(g->stream);
  rc = capture(b);
  if (rc != GBX_OK) return fail(rc);
  for (int i = 0; i < n; i++) out[i]->basis = in[i]->basis;  // what every run writes
  *bp = b;
  return GBX_OK;
}

namespace {

std::mutex g_na_mu;
unsigned char *g_na_ptr[GBX_MAX_DEVICES] = {};

// The device address of `scratch` on device `dev` (made current), looked up
// once per device.
int na_scratch_ptr(int dev, unsigned char **ptr)
{
  GBX_CHECK(dev >= 0 && dev < GBX_MAX_DEVICES, GBX_EINVAL, "gbx_batch_run_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    GBX_HIP(hipSetDevice(dev));
    void *p = nullptr;
    GBX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GBX_CHECK(p != nullptr, GBX_EHIP, "gbx_batch_run_noalloc: no device address for the scratch array");
    g_na_ptr[dev] = (unsigned char *)p;
  }
  *ptr = g_na_ptr[dev];
  return GBX_OK;
}

template <int P, bool DAG, bool TD>
void launch_na_t(dim3 grid, size_t lds, hipStream_t st, unsigned char *sp)
{
  hipLaunchKernelGGL((k_dhop_batch_na<P, DAG, 1, TD>), grid, dim3(kThreads), lds, st, sp, 1);
}

template <int P>
void launch_na(bool dag, bool td, dim3 grid, size_t lds, hipStream_t st, unsigned char *sp)
{
  if constexpr (P == GBX_F64 || P == GBX_F32) {
    if (td) {
      if (dag) launch_na_t<P, true, true>(grid, lds, st, sp);
      else launch_na_t<P, false, true>(grid, lds, st, sp);
      return;
    }
  }
  if (dag) launch_na_t<P, true, false>(grid, lds, st, sp);
  else launch_na_t<P, false, false>(grid, lds, st, sp);
}

}  // namespace

extern "C" int gbx_batch_run_noalloc(int n, const gbx_gauge *u, const gbx_spinor *out, const gbx_spinor *in, int dag)
{
  GBX_CHECK(u && out && in, GBX_ENULL, "gbx_batch_run_noalloc: NULL argument");
  GBX_CHECK(n >= 1, GBX_EINVAL, "gbx_batch_run_noalloc: n=%d", n);
  GBX_CHECK(n <= GBX_NOALLOC_MAX_LATTICES, GBX_EINVAL,
            "gbx_batch_run_noalloc: %d lattices; at most %d fit its fixed lattice table (use gbx_batch_create)", n,
            GBX_NOALLOC_MAX_LATTICES);
  GBX_CHECK(dag == 0 || dag == 1, GBX_EINVAL, "gbx_batch_run_noalloc: dag=%d", dag);
  for (int i = 0; i < n; i++) GBX_CHECK(u[i] && out[i] && in[i], GBX_ENULL, "gbx_batch_run_noalloc: NULL handle at %d", i);
  const int prec = u[0]->prec, dev = u[0]->geom->device, basis = in[0]->basis;
  NaTable t;
  memset(&t, 0, sizeof t);
  uint64_t nitems = 0;
  uint32_t max_spw = 1;
  for (int i = 0; i < n; i++) {
    GBX_TRY(dhop_check_operands(u[i], out[i], in[i]));
    GBX_CHECK(u[i]->prec == prec, GBX_EMISMATCH, "gbx_batch_run_noalloc: lattice %d precision differs from lattice 0", i);
    GBX_CHECK(in[i]->basis == basis, GBX_EMISMATCH, "gbx_batch_run_noalloc: lattice %d is in another spin basis", i);
    GBX_CHECK(u[i]->geom->device == dev, GBX_EMISMATCH, "gbx_batch_run_noalloc: lattice %d is on another device", i);
    for (int j = 0; j < n; j++)
      GBX_CHECK(out[i]->d != in[j]->d && (j == i || out[i]->d != out[j]->d), GBX_EALIAS,
                "gbx_batch_run_noalloc: output %d aliases another lattice's field", i);
    const uint32_t M = in[i]->M, nslot = (uint32_t)u[i]->geom->vol;
    uint32_t spw = (kThreads + M - 1) / M;  // as gbx_batch_create cuts items
    if (spw < 1) spw = 1;
    if (spw > 64) spw = 64;
    max_spw = std::max(max_spw, spw);
    t.lat[i].d = LatDesc{u[i]->d, u[i]->geom->nbr, in[i]->d, out[i]->d, M, nslot, (uint32_t)u[i]->comp12, 0};
    t.lat[i].item0 = (uint32_t)nitems;
    t.lat[i].spw = spw;
    nitems += (nslot + spw - 1) / spw;
    GBX_CHECK(nitems < (1ull << 31), GBX_EINVAL, "gbx_batch_run_noalloc: %llu work items", (unsigned long long)nitems);
  }
  t.n = (uint32_t)n;
  t.nitems = (uint32_t)nitems;
  t.max_spw = max_spw;
  unsigned char *sp = nullptr;
  GBX_TRY(na_scratch_ptr(dev, &sp));
  int ncu = 256;
  GBX_HIP(hipSetDevice(dev));
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const dim3 grid((unsigned)(ncu * (prec == GBX_F64 ? 4 : 6)));
  const size_t lds = lds_for(prec, max_spw);
  // inputs may still be in flight on their geometries' streams
  for (int i = 0; i < n; i++) GBX_HIP(hipStreamSynchronize(u[i]->geom->stream));
  hipStream_t st = u[0]->geom->stream;
  hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(1), 0, st, t, sp);
  const bool td = basis == GBX_BASIS_TDIAG;
  switch (prec) {
  case GBX_F64: launch_na<GBX_F64>(dag, td, grid, lds, st, sp); break;
  case GBX_F32: launch_na<GBX_F32>(dag, td, grid, lds, st, sp); break;
  case GBX_F16: launch_na<GBX_F16>(dag, td, grid, lds, st, sp); break;
  default: launch_na<GBX_BF16>(dag, td, grid, lds, st, sp); break;
  }
  GBX_HIP(hipGetLastError());
  for (int i = 0; i < n; i++) out[i]->basis = basis;
  GBX_HIP(hipStreamSynchronize(st));
  return GBX_OK;
}

extern "C" int gbx_batch_destroy(gbx_batch b)
{
  delete b;
  return GBX_OK;
}

// The captured graph applies the kernel of the inputs' spin basis at capture
// time: refuse to run it on inputs converted since.
static int batch_check_basis(const gbx_batch_s *b, const char *what)
{
  for (int i = 0; i < b->n; i++)
    GBX_CHECK(b->in[i]->basis == b->basis, GBX_EMISMATCH,
              "%s: input %d is in another spin basis than when the batch was created", what, i);
  return GBX_OK;
}

static void batch_stamp_outputs(gbx_batch_s *b)
{
  for (gbx_spinor_s *o : b->out) o->basis = b->basis;
}

extern "C" int gbx_batch_run(gbx_batch b)
{
  GBX_CHECK(b, GBX_ENULL, "gbx_batch_run: NULL batch");
  GBX_TRY(batch_check_basis(b, "gbx_batch_run"));
  for (gbx_geom_s *g : b->geoms) GBX_HIP(hipStreamSynchronize(g->stream));
  GBX_HIP(hipGraphLaunch(b->exec, b->stream));
  batch_stamp_outputs(b);
  GBX_HIP(hipStreamSynchronize(b->stream));
  return GBX_OK;
}

extern "C" int gbx_batch_time(gbx_batch b, int nrep, double *usec)
{
  GBX_CHECK(b && usec, GBX_ENULL, "gbx_batch_time: NULL argument");
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "gbx_batch_time: nrep=%d", nrep);
  GBX_TRY(gbx_batch_run(b));  // warm-up (checks the inputs' basis, stamps the outputs')
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GBX_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = gbx_fail(GBX_EHIP, "gbx_batch_time: hipEventCreate failed");
  if (rc == GBX_OK && hipEventRecord(e0, b->stream) != hipSuccess) rc = gbx_fail(GBX_EHIP, "gbx_batch_time: event record");
  for (int r = 0; r < nrep && rc == GBX_OK; r++)
    if (hipGraphLaunch(b->exec, b->stream) != hipSuccess) rc = gbx_fail(GBX_EHIP, "gbx_batch_time: graph launch");
  if (rc == GBX_OK && hipEventRecord(e1, b->stream) != hipSuccess) rc = gbx_fail(GBX_EHIP, "gbx_batch_time: event record");
  if (rc == GBX_OK && hipEventSynchronize(e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "gbx_batch_time: kernel failed");
  float ms = 0.f;
  if (rc == GBX_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = gbx_fail(GBX_EHIP, "gbx_batch_time: elapsed time");
  if (rc != GBX_OK) (void)hipStreamSynchronize(b->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == GBX_OK) *usec = 1000.0 * ms / nrep;
  return rc;
}

namespace {

gbx_kcfg shape_to_kcfg(const BatchShape &sh)
{
  gbx_kcfg k{};
  k.spw = sh.points;
  k.threads = sh.threads;
  k.pf = sh.pf;
  k.wgpc = sh.wgpc;
  k.nt = sh.nt;
  return k;
}

int apply_shape(gbx_batch_s *b, const BatchShape &sh)
{
  GBX_TRY(shape_check(sh));
  GBX_HIP(hipStreamSynchronize(b->stream));  // no replay may still read the old work list
  b->shape = sh;
  GBX_TRY(build_items(b));
  return capture(b);
}

}  // namespace

extern "C" int gbx_batch_set_cfg(gbx_batch b, const gbx_kcfg *cfg)
{
  GBX_CHECK(b && cfg, GBX_ENULL, "gbx_batch_set_cfg: NULL argument");
  BatchShape sh;
  sh.points = cfg->spw;
  sh.threads = cfg->threads;
  sh.pf = cfg->pf;
  sh.wgpc = cfg->wgpc;
  sh.nt = cfg->nt;
  return apply_shape(b, sh);
}

extern "C" int gbx_batch_get_cfg(gbx_batch b, gbx_kcfg *cfg)
{
  GBX_CHECK(b && cfg, GBX_ENULL, "gbx_batch_get_cfg: NULL argument");
  *cfg = shape_to_kcfg(b->shape);
  return GBX_OK;
}

extern "C" int gbx_batch_tune(gbx_batch b, int nrep, gbx_kcfg *best, double *best_usec)
{
  GBX_CHECK(b && best && best_usec, GBX_ENULL, "gbx_batch_tune: NULL argument");
  GBX_CHECK(nrep >= 1, GBX_EINVAL, "gbx_batch_tune: nrep=%d", nrep);
  GBX_TRY(batch_check_basis(b, "gbx_batch_tune"));
  std::vector<BatchShape> cands;
  for (int threads : {256, 512})
    for (int pf : {1, 2})
      for (int wgpc : {2, 3, 4, 6, 8})
        for (int mult : {1, 2}) {
          BatchShape sh;
          sh.threads = threads;
          sh.pf = pf;
          sh.wgpc = wgpc;
          sh.points = threads * mult;
          sh.nt = 1;
          cands.push_back(sh);
        }
  double bus = 1e300;
  BatchShape bsh = b->shape;
  auto trial = [&](const BatchShape &sh) {
    double us = 0.0;
    if (apply_shape(b, sh) != GBX_OK || gbx_batch_time(b, nrep, &us) != GBX_OK) {
      gbx_clear_error();
      return;
    }
    if (us < bus) {
      bus = us;
      bsh = sh;
    }
  };
  for (const BatchShape &sh : cands) trial(sh);
  GBX_CHECK(bus < 1e300, GBX_EHIP, "gbx_batch_tune: every candidate failed");
  BatchShape alt = bsh;
  alt.nt = 0;
  trial(alt);
  GBX_TRY(apply_shape(b, bsh));
  *best = shape_to_kcfg(bsh);
  *best_usec = bus;
  return GBX_OK;
}

extern "C" int gbx_batch_info(gbx_batch b, uint64_t *items, int *workgroups)
{
  GBX_CHECK(b && items && workgroups, GBX_ENULL, "gbx_batch_info: NULL argument");
  *items = b->nitems;
  *workgroups = b->grid;
  return GBX_OK;
}
