// Internal declarations shared by the libgbx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "gbx.h"

// ---------------------------------------------------------------------------
// Error reporting: every C-ABI entry point returns a status code and leaves a
// thread-local message describing the failure (gbx_last_error).
// ---------------------------------------------------------------------------
int gbx_fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void gbx_clear_error();

#define GBX_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return gbx_fail(e_ == hipErrorOutOfMemory ? GBX_ENOMEM : GBX_EHIP, "%s failed: %s (%s:%d)", \
                      #call, hipGetErrorString(e_), __FILE__, __LINE__);                       \
  } while (0)

#define GBX_CHECK(cond, status, ...)                                                           \
  do {                                                                                         \
    if (!(cond)) return gbx_fail(status, __VA_ARGS__);                                          \
  } while (0)

#define GBX_TRY(expr)                                                                          \
  do {                                                                                         \
    int st_ = (expr);                                                                          \
    if (st_ != GBX_OK) return st_;                                                             \
  } while (0)

int gbx_require_device();  // GBX_ENODEV unless a HIP device is usable

// Devices the library keeps per-device state for (the no-allocation entry's
// scratch address).
#define GBX_MAX_DEVICES 64

// ---------------------------------------------------------------------------
// Storage precisions.  Spinors and links are stored as FP64, FP32, FP16 or
// BF16; arithmetic is FP64 for FP64 storage and FP32 for everything else.
// ---------------------------------------------------------------------------
inline int prec_bytes(int p) { return p == GBX_F64 ? 8 : p == GBX_F32 ? 4 : 2; }
inline bool prec_valid(int p) { return p >= GBX_F64 && p <= GBX_BF16; }

// Device spinor layout ("site-major, Ls/RHS-minor"):
//   chunk k of slot q, minor index m  ->  byte offset ((q*nch + k)*M + m)*16
// where a 16-byte chunk holds 16/b consecutive reals of the 24-real spinor
// (real r = (spin*3+colour)*2 + re/im), nch = 24*b/16 chunks per spinor,
// M = Ls*nrhs and m = r_hs*Ls + s.  One lane loading one chunk therefore moves
// 16 B and a wavefront of consecutive m moves 1 KiB of contiguous memory.
inline int spinor_nchunk(int p) { return 24 * prec_bytes(p) / 16; }

// Device gauge layout: slot-major, [slot][dir 0..7][R] with R = 18 reals
// (3x3 complex, row-major) or 12 reals (rows 0 and 1 only; row 2 rebuilt as
// the conjugated cross product), each real in the gauge storage precision.

// ---------------------------------------------------------------------------
// Objects behind the opaque C handles.
// ---------------------------------------------------------------------------
struct gbx_geom_s {
  int dims[4];
  uint64_t vol;           // 4D sites
  int block[4];           // innermost traversal block shape
  int nlevels = 1;        // traversal blocking levels
  int levels[4 * GBX_MAX_LEVELS] = {0};
  int device;
  hipStream_t stream;
  // slot <-> lexicographic site permutation (host and device copies)
  std::vector<uint32_t> site_of_slot_h, slot_of_site_h;
  uint32_t *site_of_slot = nullptr, *slot_of_site = nullptr;
  uint32_t *nbr = nullptr;  // [slot][8] neighbour slots, direction order Xp Yp Zp Tp Xm Ym Zm Tm
  int refcount = 1;
  // Red-black halves.  parity -1: the whole lattice.  parity 0 / 1: the even /
  // odd sites ((x+y+z+t) mod 2) of a pair made by gbx_geom_create_rb; then vol
  // = V/2, site_of_slot holds checkerboarded lexicographic indices ((x/2)
  // fastest), nbr holds slots of the partner half (same pair_id, other
  // parity), and gsite_of_slot holds each slot's full-lattice lexicographic
  // site (for the gauge field, which is uploaded from the full lattice).
  int parity = -1;
  uint64_t pair_id = 0;
  uint32_t *gsite_of_slot = nullptr;
  // The odd half runs on the even half's stream (work on the two halves of a
  // pair reads across them, so it is ordered on one stream) and holds a
  // reference to it; nullptr when the geometry owns its stream.
  gbx_geom_s *stream_owner = nullptr;
  // Tile tables of the LDS-tiled kernel, by packed tile shape (device arrays).
  std::map<int, uint32_t *> tile_tabs;
  // Sites of the host gauge array an upload reads: the whole lattice (2*vol
  // for a red-black half, the global volume for a domain part).
  uint64_t gauge_vol = 0;
  // Domain part (gbx_geom_create_part): the block at grid coordinate
  // `coord` of the global lattice `gdims` cut into grid[0] x .. x grid[3]
  // equal blocks.  Fields carry nghost ghost slots after the vol local ones:
  // for each cut dimension d (grid[d] > 1), ascending, face[d] slots of the
  // previous part's top face (ghost_base[d] ..) then face[d] slots of the
  // next part's bottom face; nbr points into them across the cuts.
  // face_slots + fs_base[d]: [2][face[d]] local slots of the bottom
  // (coordinate 0) and top (coordinate dims[d]-1) faces of dimension d, face
  // index f = lexicographic over the other three dimensions.
  int grid[4] = {1, 1, 1, 1}, coord[4] = {0, 0, 0, 0};
  int gdims[4] = {0, 0, 0, 0};
  uint64_t face[4] = {0, 0, 0, 0}, ghost_base[4] = {0, 0, 0, 0}, fs_base[4] = {0, 0, 0, 0};
  // Slots 0 .. ninterior-1 are the interior sites (no leg reaches a ghost),
  // slots ninterior .. vol-1 the face sites.
  uint64_t nghost = 0, ninterior = 0;
  uint32_t *face_slots = nullptr;
};

// Full-lattice lexicographic site of each slot (the gauge upload's index).
inline const uint32_t *gauge_sites(const gbx_geom_s *g) { return g->gsite_of_slot ? g->gsite_of_slot : g->site_of_slot; }
// A whole periodic lattice (not a red-black half, not a domain part).
inline bool whole_lattice(const gbx_geom_s *g) { return g->parity < 0 && g->nghost == 0; }
// True when `in` holds the sites Dhop on `out_geom` reads: the same geometry
// for a whole lattice, the partner half for a red-black half.
inline bool dhop_input_geom_ok(const gbx_geom_s *out_geom, const gbx_geom_s *in_geom)
{
  if (out_geom->parity < 0) return in_geom == out_geom;
  return in_geom->pair_id == out_geom->pair_id && in_geom->parity == 1 - out_geom->parity;
}

struct gbx_gauge_s {
  gbx_geom_s *geom;
  int prec;
  int comp12;             // 1: 12 reals per link stored
  void *d = nullptr;      // device storage
  size_t bytes = 0;
};

struct gbx_spinor_s {
  gbx_geom_s *geom;
  int Ls, nrhs, prec;
  uint32_t M;             // Ls * nrhs
  void *d = nullptr;
  size_t bytes = 0;
  int basis = GBX_BASIS_CHIRAL;  // spin basis of the stored values
};

// GBX_EUNSUPPORTED unless every given field stores chiral-basis values
// (entries that have no GBX_BASIS_TDIAG form).
int require_chiral(const char *what, const gbx_spinor_s *a, const gbx_spinor_s *b = nullptr,
                   const gbx_spinor_s *c = nullptr);

void geom_retain(gbx_geom_s *g);
void geom_release(gbx_geom_s *g);

// Scratch buffer reused by uploads/downloads (grows on demand, per geometry
// stream; freed with the library's cached allocations at exit).
int staging_buffer(size_t bytes, void **ptr);

// Lane placement of a reference-layout array: `width` lanes per vector (0 =
// the lattice's own nsimd), the lattice's lanes starting at lane0, RIRI or
// RRII order of the real and imaginary parts.
struct gbx_ref_layout {
  int width = 0, lane0 = 0, rrii = 0;
};
// Reference-layout arrays in device memory <-> device objects, enqueued on
// the geometry's stream (whole lattices only).
int ref_pack_gauge_dev(gbx_gauge u, const double *d, int nsimd, const gbx_ref_layout &l);
int ref_pack_gauge_dev(gbx_gauge u, const float *d, int nsimd, const gbx_ref_layout &l);
int ref_pack_spinor_dev(gbx_spinor x, const double *d, int nsimd, const gbx_ref_layout &l);
int ref_pack_spinor_dev(gbx_spinor x, const float *d, int nsimd, const gbx_ref_layout &l);
int ref_unpack_spinor_dev(gbx_spinor x, double *d, int nsimd, const gbx_ref_layout &l);
int ref_unpack_spinor_dev(gbx_spinor x, float *d, int nsimd, const gbx_ref_layout &l);

// Kernel launch configuration of the VALU hopping kernel.
struct dhop_cfg {
  int spw;       // 4D sites per workgroup (links of these sites staged in LDS)
  int threads;   // workgroup size
  int xcd;       // 1: XCD-aware workgroup order (contiguous slot ranges per XCD)
  int nt;        // 1: non-temporal output stores
  int mc;        // minor indices per workgroup (0 = all Ls*nrhs)
  int pf;        // 1: two-buffer gather pipeline (workgroups of <= 256 threads)
  int wgpc;      // cap on resident workgroups per CU via padded LDS (0 = none)
  int cblk;      // site groups per minor-chunk cycle (0 = whole lattice per chunk)
  // The fused axpby form runs without the gather pipeline (its tuned pf is
  // that of the plain kernel, and the extra centre loads make the pipeline
  // slower there) unless the caller asked for a pf explicitly.
  bool axpby_pf = false;
  // The field's ghosts hold spin-projected halves (a domain part's face
  // slots after gbx_halo_ghost_proj receives).
  bool projected_ghosts = false;
};

// Output slots begin .. end-1 (end clipped to the volume).  The VALU and MFMA
// kernels run on a slot range by offsetting their per-slot pointers (links,
// neighbour table, output): neighbour slots stay absolute.
struct SlotRange {
  uint64_t begin = 0, end = UINT64_MAX;
};

// out = alpha*in_c + beta*Dhop(in) where in_c is `center` (may be null when alpha == 0).
// exact: the reference's rounding (no fused multiply-adds; FP64 / FP32 fields
// with 18-real links, no axpby).
int dhop_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag,
                const dhop_cfg &cfg, double alpha, double beta, const gbx_spinor_s *center,
                hipStream_t stream, bool exact = false, SlotRange range = SlotRange());
// Dhop with the reference's rounding, on the problem's tuned VALU shape.
int dhop_run_exact(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, hipStream_t st);
dhop_cfg dhop_default_cfg(const gbx_gauge_s *u, const gbx_spinor_s *in);
int dhop_cfg_check(const dhop_cfg &cfg, const gbx_gauge_s *u, const gbx_spinor_s *in);
// LDS-tiled kernel (kind 2): out = Dhop(in) on a whole lattice.
int tile_cfg_check(const gbx_kcfg &k, const gbx_gauge_s *u, const gbx_spinor_s *in);
int tile_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k, hipStream_t st);
void tile_tables_free(gbx_geom_s *g);
int pack_tile(const int e[4]);
// Slab-marching kernel (kind 3): out = Dhop(in) on a whole lattice; tile
// extents e (e[md] = slab thickness) and march dimension md packed by pack_march.
int march_cfg_check(const gbx_kcfg &k, const gbx_gauge_s *u, const gbx_spinor_s *in);
int march_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k, hipStream_t st);
int pack_march(const int e[4], int md);
// Any kernel kind (VALU or MFMA) with an explicit configuration.
int dhop_run(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &k, hipStream_t st,
             SlotRange range = SlotRange());
// The configuration gbx_dhop_range runs (the tuned one unless it is the
// whole-lattice tiled kernel, then the VALU default).
gbx_kcfg dhop_range_cfg(const gbx_gauge_s *u, const gbx_spinor_s *in);
int dhop_check_operands(const gbx_gauge_s *u, const gbx_spinor_s *out, const gbx_spinor_s *in);
std::string dhop_signature(const gbx_gauge_s *u, const gbx_spinor_s *in);
void dhop_remember_cfg(const std::string &sig, const gbx_kcfg &c);
bool dhop_lookup_cfg(const std::string &sig, gbx_kcfg *c);
int mfma_supported(const gbx_gauge_s *u, const gbx_spinor_s *in);

// out = alpha * S x + beta * z with S = Mooee / Mooee^-1 (inv) or their
// adjoints (dag), site-local in s (dwf.hip).  z may be null when beta == 0.
int dwf_s5_launch(gbx_spinor_s *out, const gbx_spinor_s *x, const gbx_spinor_s *z, const gbx_dwf_params *p, int inv, int dag,
                  double alpha, double beta, hipStream_t st);

// Vector algebra on a given stream (blas.hip).
int spinor_convert_scaled(gbx_spinor_s *dst, const gbx_spinor_s *src, double scale, hipStream_t st);
int spinor_norm2_on(const gbx_spinor_s *x, double *n2, hipStream_t st);
int spinor_dot_on(const gbx_spinor_s *x, const gbx_spinor_s *y, double *re, double *im, hipStream_t st);
int spinor_axpy_on(double a, const gbx_spinor_s *x, gbx_spinor_s *y, hipStream_t st);
int spinor_scale_on(gbx_spinor_s *x, double a, hipStream_t st);
int spinor_zero_on(gbx_spinor_s *x, hipStream_t st);
