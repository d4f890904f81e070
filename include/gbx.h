/*
 * gbx.h -- C-ABI of libgbx, an MI355X (gfx950) implementation of GridBench's
 * domain-wall-fermion Wilson hopping term (the "Hand" DhopSite kernel applied
 * on every one of Ls fifth-dimension slices that share each gauge link).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (GBX_OK = 0).  No call ever exits the process.  Handles may be used from
 * different host threads as long as one handle is not used by two threads at
 * once.  Reference files are cited as path:line of the GridBench sources.
 */
#ifndef GBX_H
#define GBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Status codes.  The reference has none: a CUDA failure prints and calls  */
/* exit(0) (Macros.h:46-51) and allocation failure asserts                 */
/* (AlignedAllocator.h:47-50).                                             */
/* ---------------------------------------------------------------------- */
#define GBX_OK           0
#define GBX_EINVAL       1  /* argument value out of range / inconsistent      */
#define GBX_ENULL        2  /* required pointer is NULL                        */
#define GBX_EALIAS       3  /* output buffer aliases an input buffer           */
#define GBX_ETABLE       4  /* nbr/prm table entry invalid                     */
#define GBX_ENOMEM       5  /* host or device allocation failed                */
#define GBX_EHIP         6  /* HIP runtime error                               */
#define GBX_ENODEV       7  /* no usable GPU                                   */
#define GBX_EUNSUPPORTED 8  /* valid request this build does not implement     */
#define GBX_EMISMATCH    9  /* objects of different geometry/shape combined    */
#define GBX_ENOCONV     10  /* iterative solver did not reach its tolerance    */

const char *gbx_strerror(int status);
/* Message describing the last failing call made by this thread ("" if none). */
const char *gbx_last_error(void);
int gbx_version(void);
/* Number of usable HIP devices (0 on a host without a GPU; never fails). */
int gbx_device_count(void);

/* Storage precisions of spinors and gauge links.  Arithmetic is FP64 for
 * GBX_F64 storage and FP32 for the others. */
#define GBX_F64  0
#define GBX_F32  1
#define GBX_F16  2
#define GBX_BF16 3

/* ---------------------------------------------------------------------- */
/* Drop-in replacement for the reference kernel entry                      */
/*   template<class Simd> double dslash_kernel(int nrep, Simd *U,          */
/*     Simd *out, Simd *in, uint64_t *nbr, uint64_t nsite, uint64_t Ls,    */
/*     uint8_t *prm)                  (WilsonKernelsHand.h:17-29)          */
/* with the Simd type's lane count (SimdVector.h:89-91) and the dagger     */
/* made explicit.  Host arrays in the reference layout (bench.cc:105-138): */
/*   U   [nsite][8][3][3][nsimd][re,im]   doubled gauge field              */
/*   in, out [nsite*Ls][4][3][nsimd][re,im]                                */
/*   nbr, prm [nsite*Ls*8], entry ss*8+dir (TableGenerate.cc:72-91)        */
/* U, in, nbr, prm are read-only; out is fully overwritten; out == in (or  */
/* any overlap) is rejected with GBX_EALIAS.  nrep applications are run;   */
/* *usec receives the device time of applications 1..nrep-1 in            */
/* microseconds (the reference's timer also starts at rep 1,               */
/* WilsonKernelsHandCpu.h:408-409), or of the single one when nrep == 1.   */
/* Every nbr entry must be < nsite*Ls and every prm entry 0 or 1.          */
/* If the tables are the canonical periodic tables of some lattice (as     */
/* TableGenerate.cc writes them) the fast SoA kernel runs; otherwise a     */
/* table-driven kernel honours the tables exactly.                         */
/* dag = 0: Dhop; dag = 1: Dhop^dagger = gamma5 Dhop gamma5 (the           */
/* reference has no dagger kernel).                                        */
/* ---------------------------------------------------------------------- */
int gbx_dslash_kernel_f64(int nrep, int nsimd, int dag, const double *U, double *out, const double *in,
                          const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec);
int gbx_dslash_kernel_f32(int nrep, int nsimd, int dag, const float *U, float *out, const float *in,
                          const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec);
/* The same entry for host arrays in the reference's RRII layout: each     */
/* vector of nsimd complex lanes stores its nsimd real parts, then its     */
/* nsimd imaginary parts (bench.cc:105-106 under -DRRII; RIRI interleaves  */
/* re,im per lane, bench.cc:108).  Tables and timing as above.             */
int gbx_dslash_kernel_rrii_f64(int nrep, int nsimd, int dag, const double *U, double *out, const double *in,
                               const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec);
int gbx_dslash_kernel_rrii_f32(int nrep, int nsimd, int dag, const float *U, float *out, const float *in,
                               const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm, double *usec);
/* The general form: host vectors of nsimd lanes (EXPAND_SIMD) holding     */
/* nsimd/data_nsimd replicas of a data_nsimd-lane lattice (DATA_SIMD), the  */
/* replica nn at lanes nn*data_nsimd + n (bench.cc:1-3, 98-138); tables and */
/* permutes are those of data_nsimd, so each replica is an independent      */
/* problem (Simd_riri_vec.h:466).  rrii selects the RRII lane layout.       */
/* data_nsimd == nsimd is the plain entry above.                            */
int gbx_dslash_kernel_expand_f64(int nrep, int nsimd, int data_nsimd, int rrii, int dag, const double *U, double *out,
                                 const double *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm,
                                 double *usec);
int gbx_dslash_kernel_expand_f32(int nrep, int nsimd, int data_nsimd, int rrii, int dag, const float *U, float *out,
                                 const float *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls, const uint8_t *prm,
                                 double *usec);
/* Drop-in contexts.  A context owns the device buffers, the device-layout */
/* objects and the stream of its calls, cached across calls of the same    */
/* shape; the caller keeps owning every host array and no pointer to one   */
/* is kept after a call returns.  Calls on one context are serialised;     */
/* calls on different contexts share no state and may come from different */
/* host threads.  The context-free entries above use one process-wide      */
/* context.  Same arguments and results as gbx_dslash_kernel_expand_*      */
/* (WilsonKernelsHand.h:17-29 with Nsimd, lane layout and dagger explicit).*/
typedef struct gbx_dslash_ctx_s *gbx_dslash_ctx;
int gbx_dslash_ctx_create(gbx_dslash_ctx *ctx);
int gbx_dslash_ctx_destroy(gbx_dslash_ctx ctx);
int gbx_dslash_ctx_kernel_f64(gbx_dslash_ctx ctx, int nrep, int nsimd, int data_nsimd, int rrii, int dag, const double *U,
                              double *out, const double *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                              const uint8_t *prm, double *usec);
int gbx_dslash_ctx_kernel_f32(gbx_dslash_ctx ctx, int nrep, int nsimd, int data_nsimd, int rrii, int dag, const float *U,
                              float *out, const float *in, const uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                              const uint8_t *prm, double *usec);
/* Path the context's last call took (1 SoA kernel, 2 table-driven, 0 none). */
int gbx_dslash_ctx_last_path(gbx_dslash_ctx ctx);
/* Which path the last gbx_dslash_kernel_* call of this thread took:
 * 1 = fast SoA kernel (tables recognised), 2 = table-driven kernel. */
int gbx_dslash_last_path(void);

/* ---------------------------------------------------------------------- */
/* Stencil tables in the reference's form (TableGenerate.cc:72-91, which  */
/* needs Grid): nbr[mu + 8*(s + Ls*site)] = 5D outer index of the          */
/* neighbour, prm = 1 when the hop wraps a SIMD-split dimension.  Periodic */
/* lattice dims[4] = {Lx,Ly,Lz,Lt}, nsimd in {1,2,4,8,16} lanes split over */
/* T, Z, Y, X in that order.  nbr/prm hold nsite*Ls*8 entries,             */
/* nsite = Lx*Ly*Lz*Lt/nsimd.                                              */
/* ---------------------------------------------------------------------- */
int gbx_table_generate(const int dims[4], int nsimd, uint64_t Ls, uint64_t *nbr, uint8_t *prm);
/* Recognise canonical tables: on success dims[4] receives the lattice.    */
/* Returns GBX_OK if canonical, GBX_ETABLE (with a message) if not.        */
int gbx_table_recognise(int nsimd, uint64_t nsite, uint64_t Ls, const uint64_t *nbr, const uint8_t *prm,
                        int dims[4]);

/* ---------------------------------------------------------------------- */
/* Seeded synthetic inputs (host side; bit-reproducible on any x86 host).  */
/* Counter-based generator: every value is a function of (seed, stream,    */
/* index) only, so any sub-range can be produced independently.            */
/* ---------------------------------------------------------------------- */
/* n i.i.d. N(0,1) reals, values index0 .. index0+n-1 of stream `stream`. */
int gbx_gen_gaussian(uint64_t seed, uint64_t stream, uint64_t index0, uint64_t n, double *out);
/* Forward links U_mu(x), lexicographic [V][4][3][3][re,im]: rows 0 and 1  */
/* complex N(0,1), Gram-Schmidt orthonormalised; row 2 = conj(r0 x r1),   */
/* so U is unitary with det 1.                                             */
int gbx_gen_su3_links(const int dims[4], uint64_t seed, double *links);
/* Doubled field U[x][mu] = U_mu(x), U[x][mu+4] = U_mu(x-mu)^dagger,      */
/* lexicographic [V][8][3][3][re,im] (the field TableGenerate.cc:54 dumps). */
int gbx_double_links(const int dims[4], const double *links, double *doubled);
/* Gaussian spinor, lexicographic [nrhs][V][Ls][4][3][re,im]; right-hand   */
/* side r uses stream r.                                                   */
int gbx_gen_spinor(const int dims[4], int Ls, int nrhs, uint64_t seed, double *psi);

/* ---------------------------------------------------------------------- */
/* Device-resident objects.                                                */
/* ---------------------------------------------------------------------- */
typedef struct gbx_geom_s *gbx_geom;     /* 4D periodic lattice + device site order */
typedef struct gbx_gauge_s *gbx_gauge;   /* doubled gauge field on the GPU           */
typedef struct gbx_spinor_s *gbx_spinor; /* Ls x nrhs 5D spinor field on the GPU     */

/* block (may be NULL = default) is the 4D traversal block: sites are stored */
/* and visited block by block, which keeps stencil neighbours in cache.     */
int gbx_geom_create(const int dims[4], const int block[4], gbx_geom *out);
/* Multi-level traversal order: blocks[l*4 .. l*4+3] is the level-l block,  */
/* innermost first; each level divides the next and the last divides dims.  */
#define GBX_MAX_LEVELS 4
int gbx_geom_create_levels(const int dims[4], int nlevels, const int *blocks, gbx_geom *out);
/* Any site order: slot_of_site[site] (a permutation of 0 .. V-1) for every */
/* lexicographic site.  Slots are what the kernels visit in order.          */
int gbx_geom_create_order(const int dims[4], const uint32_t *slot_of_site, gbx_geom *out);
/* Red-black (even-odd) pair: the sites of parity (x+y+z+t) mod 2 = 0 and 1  */
/* as two half-lattice geometries (every dim even).  The 4D checkerboard of */
/* the reference's unused red-black grids (TableGenerate.cc:26,28), with Ls */
/* not checkerboarded.  On a half:                                          */
/*  - spinor host arrays are [nrhs][V/2][Ls][24] in checkerboarded          */
/*    lexicographic order: index (x/2) + (Lx/2)*(y + Ly*(z + Lz*t));          */
/*  - gbx_gauge_upload reads the FULL-lattice doubled field [V][8][18] and   */
/*    keeps the links of the half's own sites;                              */
/*  - gbx_dhop(u, out, in) with out on one half and in on the other applies */
/*    the hopping term between them: DhopEO (out even) / DhopOE (out odd).  */
int gbx_geom_create_rb(const int dims[4], const int block[4], gbx_geom *even, gbx_geom *odd);
/* -1 for a whole lattice, 0 / 1 for the even / odd half of a pair. */
int gbx_geom_parity(gbx_geom g, int *parity);
/* Domain part: the block at grid coordinate coord (0 <= coord[d] < grid[d]) */
/* of the periodic lattice `dims` cut into grid[0] x grid[1] x grid[2] x    */
/* grid[3] equal blocks (at least one grid[d] > 1), as one GPU of a domain-  */
/* decomposed run holds it.  No counterpart in the reference, which runs on */
/* one rank (TableGenerate.cc:84 asserts every stencil neighbour is local). */
/*  - local dims ld[d] = dims[d] / grid[d]; spinor host arrays are          */
/*    [nrhs][V_local][Ls][24] in local lexicographic order;                 */
/*  - gbx_gauge_upload reads the GLOBAL doubled field [V][8][18] and keeps  */
/*    the part's own sites' links;                                          */
/*  - a spinor field holds, after its V_local sites, two faces of ghost     */
/*    slots for every cut dimension d (face_d = V_local / ld[d]): legs that */
/*    cross a cut read them, so they must hold the neighbouring parts'      */
/*    faces (the halo functions below) before a Dhop;                       */
/*  - slots 0 .. ninterior-1 are the interior sites, which read no ghost;   */
/*    slots ninterior .. V_local-1 the sites on a cut face.                 */
/* Parts run the VALU and MFMA kernels (not the tiled kernel) and gbx_dhop,   */
/* gbx_dhop_range, gbx_dhop_xpay; solvers take whole lattices.              */
int gbx_geom_create_part(const int dims[4], const int grid[4], const int coord[4], const int block[4], gbx_geom *out);
/* Red-black halves of a domain part (every local extent even): the sites   */
/* of global parity 0 / 1 of the block, as gbx_geom_create_rb splits a whole */
/* lattice (checkerboarded local lexicographic host arrays, index           */
/* (x/2) + (lx/2)*(y + ly*(z + lz*t)) in local coordinates).  Each half's   */
/* ghosts hold the neighbouring blocks' face sites of its parity: face_d/2   */
/* per face, face site f of the block at index f/2.  gbx_dhop(u_e, out_e,    */
/* in_o) is DhopEO on the block once in_o's ghosts are filled (the halo     */
/* functions work on halves unchanged).                                     */
int gbx_geom_create_part_rb(const int dims[4], const int grid[4], const int coord[4], const int block[4], gbx_geom *even,
                            gbx_geom *odd);
/* grid, coord, face (0 for uncut dimensions) and ninterior; grid = 1 1 1 1,  */
/* ninterior = V for a geometry that is not a part.                          */
int gbx_geom_part_info(gbx_geom g, int grid[4], int coord[4], uint64_t face[4], uint64_t *ninterior);
/* The part's tables on the host, without a device (what the geometry      */
/* uploads): site_of_slot (local lexicographic site of each slot) and gsite */
/* (global site) [V_local], nbr [V_local][8] (ghost slots from V_local on),  */
/* face_slots: for each cut dimension d ascending, [2][face_d] (the slots of */
/* its bottom and top faces, face site f lexicographic over the other three */
/* dimensions).  The ghosts of cut dimension d follow those of the cut      */
/* dimensions before it: minus side, then plus side, face_d slots each.     */
int gbx_part_tables(const int dims[4], const int grid[4], const int coord[4], const int block[4], uint32_t *site_of_slot,
                    uint32_t *gsite, uint32_t *nbr, uint32_t *face_slots, uint64_t *ninterior);
int gbx_geom_destroy(gbx_geom g);
int gbx_geom_volume(gbx_geom g, uint64_t *vol);
/* The geometry's stream (hipStream_t) on which all its work runs. */
int gbx_geom_stream(gbx_geom g, void **stream);
int gbx_sync(gbx_geom g);

/* comp12 = 1 stores 12 reals per link (two rows) and rebuilds the third   */
/* row in the kernel; the links must then be SU(3).                        */
int gbx_gauge_create(gbx_geom g, int prec, int comp12, gbx_gauge *out);
int gbx_gauge_destroy(gbx_gauge u);
/* Doubled field, lexicographic FP64 [V][8][3][3][re,im]. */
int gbx_gauge_upload(gbx_gauge u, const double *doubled);
/* Doubled field in the reference host layout [nsite][8][3][3][nsimd][2]. */
int gbx_gauge_upload_ref_f64(gbx_gauge u, const double *ref, int nsimd);
int gbx_gauge_upload_ref_f32(gbx_gauge u, const float *ref, int nsimd);

int gbx_spinor_create(gbx_geom g, int Ls, int nrhs, int prec, gbx_spinor *out);
int gbx_spinor_destroy(gbx_spinor x);
int gbx_spinor_bytes(gbx_spinor x, uint64_t *bytes);
/* Lexicographic FP64 [nrhs][V][Ls][4][3][re,im]. */
int gbx_spinor_upload(gbx_spinor x, const double *lex);
int gbx_spinor_download(gbx_spinor x, double *lex);
/* One right-hand side r, lexicographic FP64 [V][Ls][4][3][re,im]. */
int gbx_spinor_upload_rhs(gbx_spinor x, int r, const double *lex);
int gbx_spinor_download_rhs(gbx_spinor x, int r, double *lex);
/* Reference host layout [nsite*Ls][4][3][nsimd][re,im] (nrhs must be 1). */
int gbx_spinor_upload_ref_f64(gbx_spinor x, const double *ref, int nsimd);
int gbx_spinor_download_ref_f64(gbx_spinor x, double *ref, int nsimd);
int gbx_spinor_upload_ref_f32(gbx_spinor x, const float *ref, int nsimd);
int gbx_spinor_download_ref_f32(gbx_spinor x, float *ref, int nsimd);
/* dst = src converted between precisions (same geometry, Ls, nrhs). */
int gbx_spinor_copy(gbx_spinor dst, gbx_spinor src);
/* Fill every real with an N(0,1) value generated on the GPU (counter-based */
/* hash of (seed, rhs, site, s, component), Box-Muller with the device's    */
/* log/sin/cos).  For throughput runs on fields too large to generate on    */
/* the host; the values are not the host generator's bits.                  */
int gbx_spinor_fill_gaussian(gbx_spinor x, uint64_t seed);

/* Spin basis of a field's stored values.  GBX_BASIS_CHIRAL (every field's */
/* basis at creation) is the reference's (WilsonKernelsHandCpu.h:94-303).  */
/* GBX_BASIS_TDIAG is psi' = S psi with S = [[1, 1], [1, -1]] / sqrt(2) on  */
/* (upper, lower) spin pairs: there gamma_t = diag(1, 1, -1, -1) and        */
/* gamma_x,y,z are the chiral ones negated, so (1 -/+ gamma_t) keep one     */
/* spin half and the VALU kernel's +-t legs gather half of each spinor.     */
/* Dhop is basis-covariant (out takes the input's basis); the result, set  */
/* back to the chiral basis, agrees with the chiral kernel's to rounding.   */
/* No counterpart in the reference (its kernel stores chiral spinors).      */
/* TDIAG fields: FP64 / FP32; accepted by gbx_dhop* and gbx_tune* (VALU,  */
/* LDS-tiled, slab-marching and FP32 MFMA forms), gbx_batch_*,             */
/* gbx_cg_solve (FP32 inner),                                              */
/* gbx_cg_solve_f64 and gbx_cg_solve_eo (FP32 inner or none),              */
/* the norm / dot / axpy / scale / zero / copy routines and the whole-face  */
/* halo routines (a face is copied as stored: sending and receiving parts  */
/* must be in one basis); every other entry returns GBX_EUNSUPPORTED for   */
/* them, downloads and one-RHS uploads included (set the basis back        */
/* first).  A whole upload or gbx_spinor_fill_gaussian writes chiral       */
/* values and leaves the field in GBX_BASIS_CHIRAL.                        */
#define GBX_BASIS_CHIRAL 0
#define GBX_BASIS_TDIAG 1
/* Convert the stored values in place (S is its own inverse: a round trip   */
/* returns the values to within a few ulp, not bit for bit).               */
int gbx_spinor_set_basis(gbx_spinor x, int basis);
int gbx_spinor_get_basis(gbx_spinor x, int *basis);

/* ---------------------------------------------------------------------- */
/* The hopping term on device-resident fields: out = Dhop in (dag = 0) or  */
/* Dhop^dagger in (dag = 1).  Same operator as the reference kernel        */
/* (WilsonKernelsHandCpu.h:426-446): on every s-slice and right-hand side, */
/* Dhop psi(x) = sum_mu (1-g_mu) U_mu(x) psi(x+mu)                          */
/*             + (1+g_mu) U_mu(x-mu)^dagger psi(x-mu).                      */
/* The gauge and spinor precisions must match (FP64 with FP64, else any   */
/* of FP32/FP16/BF16 for both).  Asynchronous on the geometry's stream.   */
/* ---------------------------------------------------------------------- */
int gbx_dhop(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag);
/* out = alpha*in + beta*Dhop(in): with alpha = 1, beta = -kappa this is   */
/* M = 1 - kappa*Dhop in one pass.                                         */
int gbx_dhop_axpby(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, double alpha, double beta);
/* out = alpha*x + beta*Dhop(in) with x on out's geometry: on a red-black   */
/* half, x is a field of out's parity (e.g. x - kappa^2 DhopEO DhopOE x).    */
int gbx_dhop_xpay(gbx_gauge u, gbx_spinor out, gbx_spinor x, gbx_spinor in, int dag, double alpha, double beta);

/* Explicit launch configuration of the VALU kernel. */
typedef struct {
  int spw;     /* 4D sites per workgroup (their links are staged in LDS) */
  int threads; /* workgroup size, multiple of 64, <= 1024               */
  int xcd;     /* 1: give each XCD a contiguous range of sites           */
  int nt;      /* 1: non-temporal output stores                          */
  int kind;    /* 0: VALU register kernel, 1: MFMA kernel (multi-RHS;    */
               /* FP16 / BF16 full spinors, FP32 projected half spinors; */
               /* Ls*nrhs a multiple of 16),                             */
               /* 2: LDS-tiled VALU kernel (whole lattices), 3: slab-    */
               /* marching VALU kernel (whole lattices, FP64 / FP32,     */
               /* 18-real links)                                         */
  int mc;      /* minor indices (s, rhs) per workgroup; 0 = all Ls*nrhs  */
  int pf;      /* 1 (2): gather of leg d+1 (d+2) issued before leg d is  */
               /* computed; VALU kernel, workgroups of <= 256 threads.   */
               /* Kind 2: 1 = persistent form, the next (box, minor      */
               /* chunk)'s spinors and links staged by direct global-to- */
               /* LDS loads while the current one computes (FP64 / FP32, */
               /* 18-real links, mc dividing Ls*nrhs, <= 256 / 512       */
               /* threads)                                               */
  int wgpc;    /* cap on resident workgroups per CU (0 = none): the LDS  */
               /* request is padded to 160 KiB / wgpc, which shrinks the */
               /* set of sites in flight per XCD to what its L2 holds    */
  int cblk;    /* with mc < Ls*nrhs: workgroup site-groups visited for   */
               /* every minor chunk before moving on (0 = one sweep of   */
               /* the lattice per chunk); keeps a block's links in L2.   */
               /* Kind 3: sites of the march dimension one workgroup     */
               /* covers (0 = its whole extent)                          */
  int tile;    /* kind 2: 4D box a workgroup owns, extents packed as     */
               /* x | y << 4 | z << 8 | t << 12 (each 1..15, dividing    */
               /* the lattice); legs inside the box read LDS.  Kind 3:   */
               /* the same packing of a tile over three dimensions and  */
               /* the slab thickness along the march dimension md, with  */
               /* md << 16; threads = tile sites x slab x minor chunk,   */
               /* slab x minor chunk a multiple of 64                    */
} gbx_kcfg;
int gbx_dhop_cfg(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const gbx_kcfg *cfg);
/* gbx_dhop_xpay with an explicit VALU-kernel configuration (kind 0).     */
int gbx_dhop_xpay_cfg(gbx_gauge u, gbx_spinor out, gbx_spinor x, gbx_spinor in, int dag, double alpha, double beta,
                      const gbx_kcfg *cfg);
/* Dhop (dag as packed) on the face slots ninterior .. V_local-1 of a      */
/* domain part whose received faces are projected halves placed at         */
/* gbx_halo_ghost_proj: those legs skip the projection and multiply the    */
/* received two-spinor, bit-identical to the whole-spinor route.  FP64 /   */
/* FP32 fields; VALU kernel.                                               */
int gbx_dhop_faces_projected(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag);
/* gbx_dhop on output slots slot_begin .. slot_end-1 only (device order,   */
/* 0 <= begin <= end <= V; other output slots are not written).  On a     */
/* domain part, [0, ninterior) needs no ghost: it can run while the halo   */
/* is in flight, then [ninterior, V) once it has landed.  VALU or MFMA     */
/* kernel (a tuned tiled configuration falls back to the VALU default).    */
int gbx_dhop_range(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, uint64_t slot_begin, uint64_t slot_end);
/* Configuration gbx_dhop would use now (autotuned if gbx_tune ran). */
int gbx_dhop_get_cfg(gbx_gauge u, gbx_spinor in, gbx_kcfg *cfg);
/* Average device time per application over nrep applications after one  */
/* warm-up, from HIP events recorded on the stream the kernels run on.    */
int gbx_dhop_time(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const gbx_kcfg *cfg, int nrep,
                  double *usec);

/* Autotuner: times candidate configurations (kernel kind: VALU, MFMA,    */
/* LDS-tiled, slab-marching; sites per workgroup, workgroup size, minor    */
/* chunk, box, XCD ordering, store kind, residency cap) on these           */
/* fields and records the fastest for this (precision, comp12, Ls, nrhs,   */
/* geometry) signature; gbx_dhop uses it from then on.  Optional cache     */
/* file (JSON lines) is read first and appended to (may be NULL).          */
int gbx_tune(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, const char *cache_path, gbx_kcfg *best,
             double *best_usec);
/* The same search restricted to kernel kinds: bit 0 = VALU (kind 0), bit 1 */
/* = MFMA (kind 1), bit 2 = LDS-tiled (kind 2), bit 3 = slab-marching     */
/* (kind 3); GBX_KIND_ALL is gbx_tune.  A restricted result is cached      */
/* under its own key and does not change what gbx_dhop uses.  For          */
/* comparing the forms on one problem.                                     */
#define GBX_KIND_VALU 1
#define GBX_KIND_MFMA 2
#define GBX_KIND_TILE 4
#define GBX_KIND_MARCH 8
#define GBX_KIND_ALL 15
int gbx_tune_kinds(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, int kinds, const char *cache_path,
                   gbx_kcfg *best, double *best_usec);

/* ---------------------------------------------------------------------- */
/* Halo exchange of a domain part (gbx_geom_create_part).  No counterpart  */
/* in the reference (single rank).  Along a cut dimension dim, the part's   */
/* top face (coordinate ld-1) goes to the next part's minus ghosts (side 0),*/
/* its bottom face (coordinate 0) to the previous part's plus ghosts (side  */
/* 1), neighbours periodic in the grid.  A face is [face][nch][Ls*nrhs]     */
/* 16-byte chunks, face site f ordered lexicographically over the other     */
/* three dimensions: exactly the layout of a side's ghost slots, so a       */
/* receive (RCCL recv, peer copy or a pack into peer memory) can land in    */
/* gbx_halo_ghost's range directly.  All enqueued on the field's geometry  */
/* stream, no synchronisation; device buffers, 16-byte aligned, not         */
/* overlapping the field.                                                   */
/* ---------------------------------------------------------------------- */
/* Bytes of one face of dim: proj = 0 whole spinors, 1 spin-projected.     */
int gbx_halo_bytes(gbx_spinor x, int dim, int proj, uint64_t *bytes);
/* Device address of the ghost slots of dim, side 0 (minus) or 1 (plus).   */
int gbx_halo_ghost(gbx_spinor x, int dim, int side, void **dptr);
/* Gather face 0 (bottom) or 1 (top) of dim into dbuf.                     */
int gbx_halo_pack(gbx_spinor x, int dim, int face, void *dbuf);
/* Copy a received whole-spinor face into the ghosts of (dim, side).        */
int gbx_halo_unpack(gbx_spinor x, int dim, int side, const void *dbuf);
/* Spin-projected face (half the bytes; FP64 / FP32 fields): the two-spinor */
/* the receiving part's leg forms from it for Dhop (dag = 0) or DhopDag    */
/* (dag = 1), in the storage precision.  gbx_halo_unpack_proj expands it   */
/* into ghosts whose projection is that two-spinor exactly, so the result  */
/* is bit-identical to the one from whole-spinor ghosts.                   */
int gbx_halo_pack_proj(gbx_spinor x, int dim, int face, int dag, void *dbuf);
int gbx_halo_unpack_proj(gbx_spinor x, int dim, int side, const void *dbuf);
/* Where a projected face received for (dim, side) can land so that       */
/* gbx_dhop_faces_projected reads it as it is (no unpack): half-slots of   */
/* the ghost region, in the order of the whole-spinor ghosts.  A field's    */
/* ghosts hold either whole spinors or projected halves, not both.         */
int gbx_halo_ghost_proj(gbx_spinor x, int dim, int side, void **dptr);
/* Every face of every cut dimension in one launch (proj 0 / 1 as above):  */
/* pack face f of dimension d into bufs[2*d + f]; unpack the ghosts of     */
/* (d, side) from bufs[2*d + side].  Entries of uncut dimensions are       */
/* ignored.                                                                */
int gbx_halo_pack_all(gbx_spinor x, int proj, int dag, void *const bufs[8]);
int gbx_halo_unpack_all(gbx_spinor x, int proj, const void *const bufs[8]);
/* One rank's application, timed on this GPU (microseconds per application, */
/* HIP events on the geometry's stream, nrep after one warm-up each):       */
/*  serial:  pack the faces of every cut dimension, move each straight to  */
/*           where the face launch reads it (a device-to-device copy        */
/*           standing in for the transfer between GPUs: ghost slots, or the */
/*           projected half-slots), Dhop on the interior slots, then on the */
/*           face slots, all on the geometry's stream;                      */
/*  overlap: pack, then move on a second stream while the interior slots    */
/*           run, then the face slots after an event.                       */
/* The moved faces are the part's own, so afterwards out's face sites hold  */
/* no meaningful values (its interior sites hold Dhop(in)), and the ghosts */
/* of in are overwritten.  proj = 1: spin-projected faces (FP64 / FP32).   */
int gbx_halo_overlap_time(gbx_gauge u, gbx_spinor out, gbx_spinor in, int dag, int proj, int nrep, double *us_overlap,
                          double *us_serial);
/* Device buffers for faces (hipMalloc; destroy synchronises the device),  */
/* and a blocking copy of one to the host.                                  */
int gbx_buffer_create(uint64_t bytes, void **dptr);
int gbx_buffer_destroy(void *dptr);
int gbx_buffer_download(const void *dptr, uint64_t bytes, void *host);

/* ---------------------------------------------------------------------- */
/* Vector algebra on spinors (used by the solver; deterministic order).    */
/* ---------------------------------------------------------------------- */
/* Re<x,y> and |x|^2 accumulated in FP64 with a fixed reduction tree.     */
int gbx_spinor_dot(gbx_spinor x, gbx_spinor y, double *re, double *im);
int gbx_spinor_norm2(gbx_spinor x, double *n2);
/* y = a*x + y */
int gbx_spinor_axpy(double a, gbx_spinor x, gbx_spinor y);
int gbx_spinor_scale(gbx_spinor x, double a);
int gbx_spinor_zero(gbx_spinor x);

/* ---------------------------------------------------------------------- */
/* Defect-correction CG for  M^dagger M x = b,  M = 1 - kappa Dhop.       */
/* The outer loop runs in FP64 on the GPU (true residual, reliable update */
/* of x, global reductions); each inner CG runs with gauge and vectors in */
/* the precision of u_lo (FP32, FP16 or BF16, FP32 arithmetic) on a       */
/* rescaled residual.  u_hi must be FP64; x and b FP64, nrhs = 1.         */
/* ---------------------------------------------------------------------- */
typedef struct {
  double kappa;       /* hopping parameter, |kappa| < 1/8                 */
  double tol;         /* target |b - M^dag M x| / |b|                     */
  double inner_tol;   /* relative reduction per inner solve (e.g. 1e-3)   */
  int max_outer;
  int max_inner;
} gbx_cg_params;
typedef struct {
  int outer_iters;
  int inner_iters;    /* total over all inner solves                      */
  double true_resid;  /* final |b - M^dag M x| / |b| computed in FP64     */
  double seconds;     /* wall time of the solve                           */
  int hi_applies;     /* FP64 M^dag M applications                         */
  int lo_applies;     /* low-precision M^dag M applications                */
} gbx_cg_stats;
int gbx_cg_solve(gbx_gauge u_hi, gbx_gauge u_lo, gbx_spinor x, gbx_spinor b, const gbx_cg_params *p,
                 gbx_cg_stats *st);
/* The same solve in FP64 only (no inner low-precision solves). */
int gbx_cg_solve_f64(gbx_gauge u, gbx_spinor x, gbx_spinor b, const gbx_cg_params *p, gbx_cg_stats *st);
/* Even-odd preconditioned solve of M x = b (not the normal equations),    */
/* M = 1 - kappa*Dhop, on a red-black pair, by the Schur complement on the  */
/* even sites:                                                             */
/*   b'_e = b_e + kappa DhopEO b_o                                          */
/*   (1 - kappa^2 DhopEO DhopOE) x_e = b'_e   defect-correction CGNR: FP64  */
/*       residual and updates outside, FP32/FP16/BF16 inner solves          */
/*   x_o = b_o + kappa DhopOE x_e                                           */
/* u_hi[2] (FP64) and u_lo[2] (NULL array for FP64 only) are the gauge      */
/* fields of the even and odd halves; x[2], b[2] FP64 fields, even first.   */
/* st->true_resid = |b - M x| / |b| over the whole lattice; hi_applies and  */
/* lo_applies count applications of the Schur operator.  Reaches the same  */
/* solution as gbx_cg_solve on M^dag M x = M^dag b in about half the work.  */
int gbx_cg_solve_eo(const gbx_gauge u_hi[2], const gbx_gauge *u_lo, gbx_spinor x[2], const gbx_spinor b[2],
                    const gbx_cg_params *p, gbx_cg_stats *st);

/* ---------------------------------------------------------------------- */
/* The reference's static_data blobs (TableGenerate.cc:93-138 writes them; */
/* bench.cc:25-56 includes the header and links the arrays; the blobs are  */
/* not in the repository, .MISSING_LARGE_BLOBS).  U_static is the doubled  */
/* gauge field, Phi_static the source, Psi_cpp_static Grid's Dhop of it,   */
/* all in the reference host layout for nsimd lanes (bench.cc:105-138),    */
/* plus nbr_static / prm_static in the table form of gbx_table_generate.   */
/* ---------------------------------------------------------------------- */
/* Write static_data.cc / static_data.h in TableGenerate's exact text      */
/* format (%16.8le values, hex tables, a trailing 0 element per array).    */
/* Arrays: U nsite*144*nsimd, Phi and Psi_cpp nsite*24*Ls*nsimd reals, nbr */
/* and prm nsite*Ls*8 entries.                                             */
int gbx_static_data_write(const char *cc_path, const char *h_path, uint64_t nsite, int Ls, int nsimd, const double *U,
                          const double *Phi, const double *Psi_cpp, const uint64_t *nbr, const uint8_t *prm);
int gbx_static_data_read_header(const char *h_path, uint64_t *nsite, int *Ls);
/* Element counts {U, Phi, Psi_cpp, nbr, prm} of the arrays defined in the */
/* given files (one static_data.cc, or a split gauge / fermion pair).      */
int gbx_static_data_sizes(const char *const *paths, int npaths, uint64_t sizes[5]);
/* Read the arrays into caller buffers of exactly those counts. */
int gbx_static_data_read(const char *const *paths, int npaths, double *U, double *Phi, double *Psi_cpp, uint64_t *nbr,
                         uint8_t *prm, const uint64_t sizes[5]);

/* ---------------------------------------------------------------------- */
/* Domain-wall fermions: the fifth-dimension terms around the hop.  The    */
/* reference builds a DomainWallFermion (mass 0.1, M5 1.0,                 */
/* TableGenerate.cc:47-50) and benchmarks only its hopping term; these are */
/* the Shamir terms of that operator, on fields whose Ls slices are s:     */
/*   Mooee psi(s) = (5 - M5) psi(s) - P_- psi(s+1) - P_+ psi(s-1)          */
/*   with the walls  -P_- psi(Ls) -> +mass P_- psi(0)                       */
/*              and  -P_+ psi(-1) -> +mass P_+ psi(Ls-1),                   */
/*   P_+- = (1 +- gamma5)/2, gamma5 = diag(1,1,-1,-1) (spins 0,1 are P_+);  */
/*   D = Mooee - (1/2) Dhop.                                                */
/* On a red-black pair Mee = Moo = Mooee, Meo = -(1/2) DhopEO and          */
/* Moe = -(1/2) DhopOE.  Requires M5 < 5 and 2 <= Ls <= 256.               */
/* ---------------------------------------------------------------------- */
typedef struct {
  double M5;
  double mass;
} gbx_dwf_params;
/* out = Mooee in (inv = 0) or Mooee^-1 in (inv = 1, closed form); dag = 1 */
/* for the adjoints.  Site-local: any geometry, any precision, out may be  */
/* in itself.                                                              */
int gbx_dwf_mooee(gbx_spinor out, gbx_spinor in, const gbx_dwf_params *p, int inv, int dag);
/* out = D in (dag: D^dag = Mooee^dag - (1/2) Dhop^dag) on a whole lattice. */
int gbx_dwf_apply(gbx_gauge u, gbx_spinor out, gbx_spinor in, const gbx_dwf_params *p, int dag);
/* D x = b on a red-black pair by the Schur complement on the even sites:  */
/*   b'_e = b_e + (1/2) DhopEO Mooee^-1 b_o                                 */
/*   (Mooee - (1/4) DhopEO Mooee^-1 DhopOE) x_e = b'_e   (defect-correction */
/*       CGNR as gbx_cg_solve_eo; p->kappa is not used)                     */
/*   x_o = Mooee^-1 (b_o + (1/2) DhopOE x_e)                                */
/* st->true_resid = |b - D x| / |b| over the whole lattice.                */
int gbx_dwf_solve_eo(const gbx_gauge u_hi[2], const gbx_gauge *u_lo, gbx_spinor x[2], const gbx_spinor b[2],
                     const gbx_dwf_params *dwf, const gbx_cg_params *p, gbx_cg_stats *st);

/* ---------------------------------------------------------------------- */
/* Batched application over many independent lattices (each its own      */
/* geometry, gauge field, Ls and nrhs; FP32/FP16/BF16 storage) in one      */
/* persistent launch captured in a HIP graph.  The inputs share one spin  */
/* basis (GBX_BASIS_TDIAG for FP64 / FP32 batches too) and the outputs     */
/* take it.  The fields must outlive the batch.                            */
/* ---------------------------------------------------------------------- */
typedef struct gbx_batch_s *gbx_batch;
int gbx_batch_create(int n, const gbx_gauge *u, const gbx_spinor *out, const gbx_spinor *in, int dag,
                     gbx_batch *b);
int gbx_batch_destroy(gbx_batch b);
/* Replays the graph once on the batch's stream and waits for it.  Every  */
/* run (and gbx_batch_time) returns GBX_EMISMATCH if an input's spin basis */
/* is no longer the one the graph was captured for, and stamps that basis  */
/* on the outputs it writes.                                                */
int gbx_batch_run(gbx_batch b);
int gbx_batch_time(gbx_batch b, int nrep, double *usec);
/* Work items and persistent workgroups of the launch. */
int gbx_batch_info(gbx_batch b, uint64_t *items, int *workgroups);
/* Launch shape of the persistent kernel, in gbx_kcfg fields: threads (256 */
/* or 512 per workgroup), pf (gather pipeline depth 1 or 2), wgpc          */
/* (persistent workgroups per CU; 0 = 4 for FP64, 6 otherwise), nt         */
/* (streaming stores) and spw (the (site, minor) points a work item aims   */
/* at: a lattice of minor extent M is cut into items of ceil(spw / M)      */
/* sites, at most 64).  gbx_batch_create starts at threads 256, pf 1,     */
/* wgpc 0, nt 1, spw 256.  Setting a shape recaptures the graph.           */
int gbx_batch_get_cfg(gbx_batch b, gbx_kcfg *cfg);
int gbx_batch_set_cfg(gbx_batch b, const gbx_kcfg *cfg);
/* Times shapes (threads, pf, wgpc, points per item, then the store kind)  */
/* on the batch's own fields, nrep replays each, and keeps the fastest.    */
int gbx_batch_tune(gbx_batch b, int nrep, gbx_kcfg *best, double *best_usec);

/* No counterpart in the reference.  For callers that cannot allocate     */
/* device memory -- code under HIP graph capture, tight serving loops,     */
/* harnesses that forbid hipMalloc: the same batched application as       */
/* gbx_batch_create + gbx_batch_run (same checks, the same bits), enqueued  */
/* on the stream of lattice 0's geometry and waited for, with no graph and */
/* no allocation.  Its work-queue counter and lattice table live in one    */
/* fixed device array of GBX_NOALLOC_SCRATCH_BYTES per GPU, kept until the */
/* process ends; n is at most GBX_NOALLOC_MAX_LATTICES (GBX_EINVAL above).  */
/* Calls that overlap on one GPU share that array and race.                */
#define GBX_NOALLOC_MAX_LATTICES 32
#define GBX_NOALLOC_SCRATCH_BYTES 4096
int gbx_batch_run_noalloc(int n, const gbx_gauge *u, const gbx_spinor *out, const gbx_spinor *in, int dag);

/* ---------------------------------------------------------------------- */
/* HBM bandwidth probe: device-to-device copy of `bytes` with the          */
/* library's own streaming kernel; returns GB/s (1e9 B/s) counting read + */
/* write bytes.                                                            */
/* ---------------------------------------------------------------------- */
int gbx_stream_copy_bw(uint64_t bytes, int nrep, double *gbps);
/* Read-weighted variant: reads `read_bytes` (a multiple of 16*mix) and     */
/* writes one 16-byte element per `mix` read (mix = 1, 2, 3, 4 or 8), or,  */
/* with mix = 0, only reads (one 16-byte write per thread).  GB/s counts   */
/* read + write bytes.  A kernel that reads several times what it writes  */
/* (Dhop reads about 2.7 B of HBM per byte written) is bounded by these.  */
int gbx_stream_bw(uint64_t read_bytes, int mix, int nrep, double *gbps);
/* Gather variant, the hopping kernels' access shape: every wavefront load */
/* reads 4 runs of 256 contiguous bytes at pseudo-random places of a      */
/* buffer of `buffer_bytes` (a power of two, >= 1 MiB).  A 2 GiB buffer    */
/* measures HBM, a few MiB the L2s.  GB/s counts the bytes read.          */
int gbx_gather_bw(uint64_t buffer_bytes, int nrep, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* GBX_H */
