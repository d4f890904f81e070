"""Python mirror of the GridBench hot-path interface on top of libgbx.

* `dslash_kernel(...)` keeps the reference entry point's argument list and
  meaning (WilsonKernelsHand.h:17-29): host arrays in the reference layout,
  nbr/prm tables, nrep applications, returns microseconds -- with the SIMD
  lane count and the dagger explicit.
* `Geometry`, `GaugeField`, `SpinorField`, `dhop` work on device-resident
  fields in the library's SoA layout (the fast path).
* Seeded input generation and stencil tables in the reference's form.
"""
import ctypes as C

import numpy as np

from ._lib import (lib, check, ptr, ivec, KCfg, CGParams, CGStats, DwfParams, GbxError, PREC_BY_NAME, PREC_NAMES,
                   F64, F32, F16, BF16)

FLOPS_PER_SITE = 1320  # bench.cc:145, per 5D site and application


def _prec(p):
    if isinstance(p, str):
        return PREC_BY_NAME[p]
    return int(p)


def device_count():
    return lib.gbx_device_count()


# ---------------------------------------------------------------- generators
def gen_gaussian(seed, stream, index0, n):
    out = np.empty(n, dtype=np.float64)
    check(lib.gbx_gen_gaussian(seed, stream, index0, n, ptr(out, C.c_double)), "gbx_gen_gaussian")
    return out


def gen_su3_links(dims, seed):
    """Forward links U_mu(x), lexicographic [V][4][3][3] complex128."""
    V = int(np.prod(dims))
    out = np.empty((V, 4, 3, 3, 2), dtype=np.float64)
    check(lib.gbx_gen_su3_links(ivec(dims), seed, ptr(out, C.c_double)), "gbx_gen_su3_links")
    return out


def double_links(dims, links):
    """Doubled field [V][8][3][3][2]: U[x][mu] = U_mu(x), U[x][mu+4] = U_mu(x-mu)^dagger."""
    links = np.ascontiguousarray(links, dtype=np.float64)
    V = int(np.prod(dims))
    out = np.empty((V, 8, 3, 3, 2), dtype=np.float64)
    check(lib.gbx_double_links(ivec(dims), ptr(links, C.c_double), ptr(out, C.c_double)), "gbx_double_links")
    return out


def gen_doubled_gauge(dims, seed):
    return double_links(dims, gen_su3_links(dims, seed))


def gen_spinor(dims, Ls, nrhs, seed):
    """Gaussian spinor, lexicographic [nrhs][V][Ls][4][3][2] float64."""
    V = int(np.prod(dims))
    out = np.empty((nrhs, V, Ls, 4, 3, 2), dtype=np.float64)
    check(lib.gbx_gen_spinor(ivec(dims), Ls, nrhs, seed, ptr(out, C.c_double)), "gbx_gen_spinor")
    return out


# -------------------------------------------------------------------- tables
def table_generate(dims, nsimd, Ls):
    """(nbr, prm) in the reference's form for a periodic lattice."""
    nsite = int(np.prod(dims)) // nsimd
    nbr = np.empty(nsite * Ls * 8, dtype=np.uint64)
    prm = np.empty(nsite * Ls * 8, dtype=np.uint8)
    check(lib.gbx_table_generate(ivec(dims), nsimd, Ls, ptr(nbr, C.c_uint64), ptr(prm, C.c_uint8)),
          "gbx_table_generate")
    return nbr, prm


def table_recognise(nsimd, nsite, Ls, nbr, prm):
    """Lattice dims if (nbr, prm) are canonical periodic tables, else None."""
    dims = (C.c_int * 4)()
    rc = lib.gbx_table_recognise(nsimd, nsite, Ls, ptr(nbr, C.c_uint64), ptr(prm, C.c_uint8), dims)
    return tuple(dims) if rc == 0 else None


# ------------------------------------------------------------------- drop-in
def _dslash_args(U, out, in_, nbr, nsite, Ls, prm):
    usec = C.c_double(0.0)
    if U.dtype == np.float64:
        ct, sfx = C.c_double, "f64"
    elif U.dtype == np.float32:
        ct, sfx = C.c_float, "f32"
    else:
        raise TypeError("U must be float64 or float32")
    for a in (out, in_):
        if a.dtype != U.dtype:
            raise TypeError("U, out, in must share a dtype")
    if nbr.dtype != np.uint64 or prm.dtype != np.uint8:
        raise TypeError("nbr must be uint64 and prm uint8")
    args = (ptr(U, ct), ptr(out, ct), ptr(in_, ct), ptr(nbr, C.c_uint64), int(nsite), int(Ls), ptr(prm, C.c_uint8),
            C.byref(usec))
    return sfx, args, usec


def dslash_kernel(nrep, U, out, in_, nbr, nsite, Ls, prm, nsimd, dag=False, rrii=False, data_nsimd=None):
    """The reference entry dslash_kernel<Simd>(nrep, U, out, in, nbr, nsite, Ls, prm)
    (WilsonKernelsHand.h:17-29) on numpy arrays in the reference layout.
    float64 arrays run the double kernel, float32 the single.  rrii=True
    takes every array in the RRII lane layout (nsimd real parts, then nsimd
    imaginary parts per vector, bench.cc:105-106) instead of RIRI.
    data_nsimd < nsimd: vectors of nsimd lanes expanded from data_nsimd data
    lanes, replica nn at lanes nn*data_nsimd + n, with the tables and
    permutes of data_nsimd (bench.cc:1-3, 98-138).  Returns the microseconds
    of applications 1..nrep-1, like the reference."""
    sfx, args, usec = _dslash_args(U, out, in_, nbr, nsite, Ls, prm)
    if data_nsimd is not None and int(data_nsimd) != int(nsimd):
        fn = getattr(lib, "gbx_dslash_kernel_expand_" + sfx)
        rc = fn(int(nrep), int(nsimd), int(data_nsimd), int(bool(rrii)), int(bool(dag)), *args)
    else:
        fn = getattr(lib, ("gbx_dslash_kernel_rrii_" if rrii else "gbx_dslash_kernel_") + sfx)
        rc = fn(int(nrep), int(nsimd), int(bool(dag)), *args)
    check(rc, "dslash_kernel")
    return usec.value


def dslash_last_path():
    """1 = the last dslash_kernel call ran the SoA kernel, 2 = the table-driven kernel."""
    return lib.gbx_dslash_last_path()


class DslashContext:
    """A drop-in context (gbx_dslash_ctx): its own cached device buffers and
    stream.  Calls on different contexts may come from different threads."""

    def __init__(self):
        h = C.c_void_p()
        check(lib.gbx_dslash_ctx_create(C.byref(h)), "gbx_dslash_ctx_create")
        self.h = h

    def kernel(self, nrep, U, out, in_, nbr, nsite, Ls, prm, nsimd, dag=False, rrii=False, data_nsimd=None):
        """As dslash_kernel(), on this context."""
        sfx, args, usec = _dslash_args(U, out, in_, nbr, nsite, Ls, prm)
        fn = getattr(lib, "gbx_dslash_ctx_kernel_" + sfx)
        dn = int(nsimd if data_nsimd is None else data_nsimd)
        check(fn(self.h, int(nrep), int(nsimd), dn, int(bool(rrii)), int(bool(dag)), *args), "gbx_dslash_ctx_kernel")
        return usec.value

    def last_path(self):
        return lib.gbx_dslash_ctx_last_path(self.h)

    def close(self):
        if self.h:
            lib.gbx_dslash_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------- device objects
class Geometry:
    """Periodic 4D lattice on the GPU.  `block` is the traversal block; or
    `levels` = [innermost block, ..., outermost block] for a multi-level
    (blocks-of-blocks) site order.  Geometry.rb_pair(dims) gives the even and
    odd halves of a red-black pair (parity -1 / 0 / 1; volume = sites held).
    Geometry.part_of(dims, grid, coord) gives one block of a domain
    decomposition (dims = its local dims, gauge_dims = the whole lattice's)."""

    def __init__(self, dims, block=None, levels=None, _handle=None):
        self.dims = tuple(int(d) for d in dims)
        self.gauge_dims = self.dims
        h = C.c_void_p()
        if _handle is not None:
            h = _handle
        elif levels is not None:
            flat = [int(b) for lv in levels for b in lv]
            check(lib.gbx_geom_create_levels(ivec(self.dims), len(levels), ivec(flat), C.byref(h)),
                  "gbx_geom_create_levels")
        else:
            check(lib.gbx_geom_create(ivec(self.dims), ivec(block) if block is not None else None, C.byref(h)),
                  "gbx_geom_create")
        self.h = h
        v = C.c_uint64()
        check(lib.gbx_geom_volume(h, C.byref(v)), "gbx_geom_volume")
        self.volume = int(v.value)
        par = C.c_int()
        check(lib.gbx_geom_parity(h, C.byref(par)), "gbx_geom_parity")
        self.parity = int(par.value)
        grid, coord, face, nint = (C.c_int * 4)(), (C.c_int * 4)(), (C.c_uint64 * 4)(), C.c_uint64()
        check(lib.gbx_geom_part_info(h, grid, coord, face, C.byref(nint)), "gbx_geom_part_info")
        self.grid, self.coord, self.face = tuple(grid), tuple(coord), tuple(int(f) for f in face)
        self.ninterior = int(nint.value)

    @classmethod
    def part_of(cls, dims, grid, coord, block=None):
        """Block `coord` of the lattice `dims` cut into grid[0] x .. x grid[3] blocks."""
        dims, grid, coord = (tuple(int(v) for v in a) for a in (dims, grid, coord))
        h = C.c_void_p()
        check(lib.gbx_geom_create_part(ivec(dims), ivec(grid), ivec(coord), ivec(block) if block is not None else None,
                                       C.byref(h)), "gbx_geom_create_part")
        g = cls(tuple(d // n for d, n in zip(dims, grid)), _handle=h)
        g.gauge_dims = dims
        return g

    @classmethod
    def with_order(cls, dims, slot_of_site):
        """Geometry whose slots follow slot_of_site[lexicographic site]."""
        dims = tuple(int(d) for d in dims)
        a = np.ascontiguousarray(slot_of_site, dtype=np.uint32)
        assert a.size == int(np.prod(dims))
        h = C.c_void_p()
        check(lib.gbx_geom_create_order(ivec(dims), ptr(a, C.c_uint32), C.byref(h)), "gbx_geom_create_order")
        return cls(dims, _handle=h)

    @classmethod
    def part_rb_pair(cls, dims, grid, coord, block=None):
        """(even, odd) halves of block `coord` of the lattice `dims` cut by `grid`."""
        dims, grid, coord = (tuple(int(v) for v in a) for a in (dims, grid, coord))
        he, ho = C.c_void_p(), C.c_void_p()
        check(lib.gbx_geom_create_part_rb(ivec(dims), ivec(grid), ivec(coord), ivec(block) if block is not None else None,
                                          C.byref(he), C.byref(ho)), "gbx_geom_create_part_rb")
        local = tuple(d // n for d, n in zip(dims, grid))
        ge, go = cls(local, _handle=he), cls(local, _handle=ho)
        ge.gauge_dims = go.gauge_dims = dims
        return ge, go

    @classmethod
    def rb_pair(cls, dims, block=None):
        """(even, odd) half-lattice geometries of one red-black pair."""
        dims = tuple(int(d) for d in dims)
        he, ho = C.c_void_p(), C.c_void_p()
        check(lib.gbx_geom_create_rb(ivec(dims), ivec(block) if block is not None else None, C.byref(he),
                                     C.byref(ho)), "gbx_geom_create_rb")
        return cls(dims, _handle=he), cls(dims, _handle=ho)

    def sync(self):
        check(lib.gbx_sync(self.h), "gbx_sync")

    def close(self):
        if self.h:
            lib.gbx_geom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GaugeField:
    def __init__(self, geom, prec="f64", comp12=False):
        self.geom = geom
        self.prec = _prec(prec)
        self.comp12 = bool(comp12)
        h = C.c_void_p()
        check(lib.gbx_gauge_create(geom.h, self.prec, int(self.comp12), C.byref(h)), "gbx_gauge_create")
        self.h = h

    def upload(self, doubled):
        """Doubled field of the WHOLE lattice, lexicographic float64
        [V][8][3][3][2] (a red-black half keeps its own sites' links)."""
        a = np.ascontiguousarray(doubled, dtype=np.float64)
        assert a.size == int(np.prod(self.geom.gauge_dims)) * 144
        check(lib.gbx_gauge_upload(self.h, ptr(a, C.c_double)), "gbx_gauge_upload")
        return self

    def upload_ref(self, ref, nsimd):
        if ref.dtype == np.float64:
            check(lib.gbx_gauge_upload_ref_f64(self.h, ptr(ref, C.c_double), nsimd), "gbx_gauge_upload_ref_f64")
        else:
            check(lib.gbx_gauge_upload_ref_f32(self.h, ptr(ref, C.c_float), nsimd), "gbx_gauge_upload_ref_f32")
        return self

    def close(self):
        if self.h:
            lib.gbx_gauge_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpinorField:
    def __init__(self, geom, Ls, nrhs=1, prec="f64"):
        self.geom = geom
        self.Ls = int(Ls)
        self.nrhs = int(nrhs)
        self.prec = _prec(prec)
        h = C.c_void_p()
        check(lib.gbx_spinor_create(geom.h, self.Ls, self.nrhs, self.prec, C.byref(h)), "gbx_spinor_create")
        self.h = h

    @property
    def shape(self):
        return (self.nrhs, self.geom.volume, self.Ls, 4, 3, 2)

    def upload(self, lex):
        a = np.ascontiguousarray(lex, dtype=np.float64)
        assert a.size == int(np.prod(self.shape)), (a.shape, self.shape)
        check(lib.gbx_spinor_upload(self.h, ptr(a, C.c_double)), "gbx_spinor_upload")
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.float64)
        check(lib.gbx_spinor_download(self.h, ptr(out, C.c_double)), "gbx_spinor_download")
        return out

    def upload_rhs(self, r, lex):
        """Right-hand side r from lexicographic float64 [V][Ls][4][3][2]."""
        a = np.ascontiguousarray(lex, dtype=np.float64)
        assert a.size == self.geom.volume * self.Ls * 24
        check(lib.gbx_spinor_upload_rhs(self.h, int(r), ptr(a, C.c_double)), "gbx_spinor_upload_rhs")
        return self

    def download_rhs(self, r):
        out = np.empty(self.shape[1:], dtype=np.float64)
        check(lib.gbx_spinor_download_rhs(self.h, int(r), ptr(out, C.c_double)), "gbx_spinor_download_rhs")
        return out

    def upload_ref(self, ref, nsimd):
        if ref.dtype == np.float64:
            check(lib.gbx_spinor_upload_ref_f64(self.h, ptr(ref, C.c_double), nsimd), "gbx_spinor_upload_ref_f64")
        else:
            check(lib.gbx_spinor_upload_ref_f32(self.h, ptr(ref, C.c_float), nsimd), "gbx_spinor_upload_ref_f32")
        return self

    def download_ref(self, nsimd, dtype=np.float64):
        out = np.empty(self.geom.volume * self.Ls * 24, dtype=dtype)
        if dtype == np.float64:
            check(lib.gbx_spinor_download_ref_f64(self.h, ptr(out, C.c_double), nsimd), "gbx_spinor_download_ref_f64")
        else:
            check(lib.gbx_spinor_download_ref_f32(self.h, ptr(out, C.c_float), nsimd), "gbx_spinor_download_ref_f32")
        return out

    def copy_from(self, src):
        check(lib.gbx_spinor_copy(self.h, src.h), "gbx_spinor_copy")
        return self

    def fill_gaussian(self, seed):
        """N(0,1) entries generated on the GPU (throughput runs; not the host generator's bits)."""
        check(lib.gbx_spinor_fill_gaussian(self.h, int(seed)), "gbx_spinor_fill_gaussian")
        return self

    def set_basis(self, basis):
        """Convert the stored values to spin basis BASIS_CHIRAL or BASIS_TDIAG
        (gamma_t diagonal: the VALU kernel's t legs gather half a spinor)."""
        check(lib.gbx_spinor_set_basis(self.h, int(basis)), "gbx_spinor_set_basis")
        return self

    @property
    def basis(self):
        b = C.c_int()
        check(lib.gbx_spinor_get_basis(self.h, C.byref(b)), "gbx_spinor_get_basis")
        return b.value

    def norm2(self):
        n = C.c_double()
        check(lib.gbx_spinor_norm2(self.h, C.byref(n)), "gbx_spinor_norm2")
        return n.value

    def dot(self, other):
        re, im = C.c_double(), C.c_double()
        check(lib.gbx_spinor_dot(self.h, other.h, C.byref(re), C.byref(im)), "gbx_spinor_dot")
        return complex(re.value, im.value)

    def axpy(self, a, x):
        """self = a*x + self"""
        check(lib.gbx_spinor_axpy(float(a), x.h, self.h), "gbx_spinor_axpy")
        return self

    def scale(self, a):
        check(lib.gbx_spinor_scale(self.h, float(a)), "gbx_spinor_scale")
        return self

    def zero(self):
        check(lib.gbx_spinor_zero(self.h), "gbx_spinor_zero")
        return self

    @property
    def nbytes(self):
        b = C.c_uint64()
        check(lib.gbx_spinor_bytes(self.h, C.byref(b)), "gbx_spinor_bytes")
        return b.value

    # ---- halo of a domain part: device pointers are plain integers
    def halo_bytes(self, dim, proj=False):
        b = C.c_uint64()
        check(lib.gbx_halo_bytes(self.h, int(dim), int(bool(proj)), C.byref(b)), "gbx_halo_bytes")
        return b.value

    def halo_ghost(self, dim, side):
        p = C.c_void_p()
        check(lib.gbx_halo_ghost(self.h, int(dim), int(side), C.byref(p)), "gbx_halo_ghost")
        return p.value

    def halo_pack(self, dim, face, dptr, proj=False, dag=False):
        if proj:
            check(lib.gbx_halo_pack_proj(self.h, int(dim), int(face), int(bool(dag)), C.c_void_p(dptr)),
                  "gbx_halo_pack_proj")
        else:
            check(lib.gbx_halo_pack(self.h, int(dim), int(face), C.c_void_p(dptr)), "gbx_halo_pack")
        return self

    def halo_ghost_proj(self, dim, side):
        p = C.c_void_p()
        check(lib.gbx_halo_ghost_proj(self.h, int(dim), int(side), C.byref(p)), "gbx_halo_ghost_proj")
        return p.value

    def halo_pack_all(self, bufs, proj=False, dag=False):
        """bufs: {(dim, face): device pointer} for every cut dimension."""
        arr = (C.c_void_p * 8)(*[bufs.get((d, f)) for d in range(4) for f in (0, 1)])
        check(lib.gbx_halo_pack_all(self.h, int(bool(proj)), int(bool(dag)), arr), "gbx_halo_pack_all")
        return self

    def halo_unpack_all(self, bufs, proj=False):
        """bufs: {(dim, side): device pointer} for every cut dimension."""
        arr = (C.c_void_p * 8)(*[bufs.get((d, f)) for d in range(4) for f in (0, 1)])
        check(lib.gbx_halo_unpack_all(self.h, int(bool(proj)), arr), "gbx_halo_unpack_all")
        return self

    def halo_unpack(self, dim, side, dptr, proj=False):
        fn = lib.gbx_halo_unpack_proj if proj else lib.gbx_halo_unpack
        check(fn(self.h, int(dim), int(side), C.c_void_p(dptr)), "gbx_halo_unpack_proj" if proj else "gbx_halo_unpack")
        return self

    def close(self):
        if self.h:
            lib.gbx_spinor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_tile(ext):
    """4D box extents (each 1..15) in gbx_kcfg.tile's packed form."""
    return int(ext[0]) | (int(ext[1]) << 4) | (int(ext[2]) << 8) | (int(ext[3]) << 12)


def pack_march(ext, md):
    """Kind 3: extents of the tile (three dimensions) and of the slab along the
    march dimension md (each 1..15), in gbx_kcfg.tile's packed form."""
    return pack_tile(ext) | (int(md) << 16)


def march_cfg(ext, md, mc, seg=0, xcd=1, nt=1):
    """Kind 3 configuration: threads = tile sites x slab x minor chunk."""
    n = 1
    for d in range(4):
        n *= int(ext[d])
    return KCfg(0, n * int(mc), xcd, nt, 3, int(mc), 0, 0, int(seg), pack_march(ext, md))


def kcfg(spw, threads, xcd=1, nt=1, kind=0, mc=0, pf=0, wgpc=0, cblk=0, tile=0):
    """Kernel configuration; kind 2 (LDS-tiled) takes tile = 4D box extents."""
    if not isinstance(tile, int):
        tile = pack_tile(tile)
    return KCfg(spw, threads, xcd, nt, kind, mc, pf, wgpc, cblk, tile)


def dhop(U, out, inp, dag=False, cfg=None):
    """out = Dhop inp (dag: Dhop^dagger) on device-resident fields."""
    if cfg is None:
        check(lib.gbx_dhop(U.h, out.h, inp.h, int(bool(dag))), "gbx_dhop")
    else:
        check(lib.gbx_dhop_cfg(U.h, out.h, inp.h, int(bool(dag)), C.byref(cfg)), "gbx_dhop_cfg")
    return out


class DeviceBuffer:
    """Device memory for halo faces (`ptr` is the device address)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib.gbx_buffer_create(self.nbytes, C.byref(p)), "gbx_buffer_create")
        self.ptr = p.value

    def download(self, dtype=np.uint8):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        check(lib.gbx_buffer_download(C.c_void_p(self.ptr), self.nbytes, out.ctypes.data_as(C.c_void_p)),
              "gbx_buffer_download")
        return out

    def close(self):
        if self.ptr:
            lib.gbx_buffer_destroy(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dhop_range(U, out, inp, slot_begin, slot_end, dag=False):
    """Dhop on output slots slot_begin .. slot_end-1 (device order)."""
    check(lib.gbx_dhop_range(U.h, out.h, inp.h, int(bool(dag)), int(slot_begin), int(slot_end)), "gbx_dhop_range")
    return out


def dhop_faces_projected(U, out, inp, dag=False):
    """Dhop on a part's face slots, its ghosts holding projected halves."""
    check(lib.gbx_dhop_faces_projected(U.h, out.h, inp.h, int(bool(dag))), "gbx_dhop_faces_projected")
    return out


def halo_overlap_time(U, out, inp, dag=False, proj=True, nrep=20):
    """(us_overlap, us_serial) per application of one part: face exchange
    (faces moved on this GPU) overlapped with the interior slots, or run
    before them on one stream.  Leaves out's face sites and inp's ghosts
    without meaningful values."""
    ov, se = C.c_double(), C.c_double()
    check(lib.gbx_halo_overlap_time(U.h, out.h, inp.h, int(bool(dag)), int(bool(proj)), int(nrep), C.byref(ov),
                                    C.byref(se)), "gbx_halo_overlap_time")
    return ov.value, se.value


def grid_coords(grid):
    """Coordinates of a process grid, lexicographic (x fastest)."""
    return [tuple(int(v) for v in c[::-1]) for c in np.ndindex(*tuple(grid)[::-1])]


def halo_buffers(fields, proj=True):
    """Device buffers for halo_exchange: {(coord, dim): (bottom, top)}."""
    some = next(iter(fields.values()))
    grid = some.geom.grid
    return {(c, d): (DeviceBuffer(x.halo_bytes(d, proj)), DeviceBuffer(x.halo_bytes(d, proj)))
            for c, x in fields.items() for d in range(4) if grid[d] > 1}


def halo_exchange_in_place(fields, dag=False):
    """Projected faces packed straight into the neighbours' half-slots
    (gbx_halo_ghost_proj), one launch per part: after it, every part's face
    slots are gbx_dhop_faces_projected(dag).  Synchronises every part's
    stream before returning."""
    grid = next(iter(fields.values())).geom.grid
    cut = [d for d in range(4) if grid[d] > 1]

    def step(c, d, k):
        c = list(c)
        c[d] = (c[d] + k) % grid[d]
        return tuple(c)

    for c, x in fields.items():
        x.halo_pack_all({(d, f): fields[step(c, d, -1 if f == 0 else 1)].halo_ghost_proj(d, 1 - f)
                         for d in cut for f in (0, 1)}, proj=True, dag=dag)
    for x in fields.values():
        x.geom.sync()
    return fields


def halo_exchange(fields, proj=False, dag=False, bufs=None):
    """Exchange the faces of the parts `fields` ({grid coord: SpinorField},
    every block of one grid), as the ranks of a domain decomposition would:
    along every cut dimension, part c's top face into the next part's minus
    ghosts and its bottom face into the previous part's plus ghosts.  Whole
    spinors are packed straight into the receiving ghost slots; spin-projected
    faces (proj=True) go through the device buffers `bufs` (halo_buffers).
    Synchronises every part's stream before returning."""
    grid = next(iter(fields.values())).geom.grid

    def step(c, d, k):
        c = list(c)
        c[d] = (c[d] + k) % grid[d]
        return tuple(c)

    cut = [d for d in range(4) if grid[d] > 1]
    for c, x in fields.items():  # one pack launch per part
        if proj:
            x.halo_pack_all({(d, f): bufs[(c, d)][f].ptr for d in cut for f in (0, 1)}, proj=True, dag=dag)
        else:
            x.halo_pack_all({(d, f): fields[step(c, d, -1 if f == 0 else 1)].halo_ghost(d, 1 - f)
                             for d in cut for f in (0, 1)})
    for x in fields.values():
        x.geom.sync()
    if proj:
        for c, x in fields.items():  # one unpack launch per part
            x.halo_unpack_all({(d, side): bufs[(step(c, d, -1 if side == 0 else 1), d)][1 - side].ptr
                               for d in cut for side in (0, 1)}, proj=True)
        for x in fields.values():
            x.geom.sync()
    return fields


def dhop_xpay(U, out, x, inp, alpha, beta, dag=False, cfg=None):
    """out = alpha*x + beta*Dhop(inp), x on out's geometry (red-black halves too);
    cfg: an explicit VALU-kernel configuration."""
    if cfg is None:
        check(lib.gbx_dhop_xpay(U.h, out.h, x.h, inp.h, int(bool(dag)), float(alpha), float(beta)), "gbx_dhop_xpay")
    else:
        check(lib.gbx_dhop_xpay_cfg(U.h, out.h, x.h, inp.h, int(bool(dag)), float(alpha), float(beta), C.byref(cfg)),
              "gbx_dhop_xpay_cfg")


def dhop_axpby(U, out, inp, alpha, beta, dag=False):
    """out = alpha*inp + beta*Dhop inp."""
    check(lib.gbx_dhop_axpby(U.h, out.h, inp.h, int(bool(dag)), float(alpha), float(beta)), "gbx_dhop_axpby")
    return out


def dhop_get_cfg(U, inp):
    k = KCfg()
    check(lib.gbx_dhop_get_cfg(U.h, inp.h, C.byref(k)), "gbx_dhop_get_cfg")
    return k


def dhop_time(U, out, inp, dag=False, cfg=None, nrep=10):
    """Average microseconds per application (HIP events on the kernels' stream)."""
    us = C.c_double()
    check(lib.gbx_dhop_time(U.h, out.h, inp.h, int(bool(dag)), C.byref(cfg) if cfg is not None else None, int(nrep),
                            C.byref(us)), "gbx_dhop_time")
    return us.value


def tune(U, out, inp, dag=False, cache_path=None, kinds=None):
    """Autotune; kinds=None searches every kernel form and sets what gbx_dhop
    uses; kinds = "valu" / "mfma" / "tile" (or a list of them) restricts the
    search (result not adopted)."""
    best = KCfg()
    us = C.c_double()
    cp = cache_path.encode() if cache_path else None
    if kinds is None:
        check(lib.gbx_tune(U.h, out.h, inp.h, int(bool(dag)), cp, C.byref(best), C.byref(us)), "gbx_tune")
    else:
        names = [kinds] if isinstance(kinds, str) else list(kinds)
        k = sum({"valu": 1, "mfma": 2, "tile": 4, "march": 8}[n] for n in names)
        check(lib.gbx_tune_kinds(U.h, out.h, inp.h, int(bool(dag)), k, cp, C.byref(best), C.byref(us)),
              "gbx_tune_kinds")
    return best, us.value


def cg_solve(U_hi, U_lo, x, b, kappa, tol=1e-10, inner_tol=1e-3, max_outer=50, max_inner=2000):
    """Defect-correction CG for M^dag M x = b, M = 1 - kappa Dhop; U_lo None = FP64 only."""
    p = CGParams(kappa, tol, inner_tol, max_outer, max_inner)
    st = CGStats()
    if U_lo is None:
        rc = lib.gbx_cg_solve_f64(U_hi.h, x.h, b.h, C.byref(p), C.byref(st))
    else:
        rc = lib.gbx_cg_solve(U_hi.h, U_lo.h, x.h, b.h, C.byref(p), C.byref(st))
    check(rc, "gbx_cg_solve")
    return st.as_dict()


def cg_solve_eo(U_hi, U_lo, x, b, kappa, tol=1e-10, inner_tol=1e-3, max_outer=50, max_inner=2000):
    """Even-odd (Schur complement) solve of M x = b, M = 1 - kappa Dhop, on a
    red-black pair.  U_hi, x, b: (even, odd) pairs of FP64 objects; U_lo an
    (even, odd) pair of FP32/FP16/BF16 gauge fields or None (FP64 only)."""
    p = CGParams(kappa, tol, inner_tol, max_outer, max_inner)
    st = CGStats()
    arr = lambda objs: (C.c_void_p * 2)(*[o.h for o in objs])
    check(lib.gbx_cg_solve_eo(arr(U_hi), arr(U_lo) if U_lo is not None else None, arr(x), arr(b), C.byref(p),
                              C.byref(st)), "gbx_cg_solve_eo")
    return st.as_dict()


# Domain-wall fermions: the reference's DomainWallFermion parameters
# (TableGenerate.cc:47-50).
DWF_M5, DWF_MASS = 1.0, 0.1


def dwf_mooee(out, inp, M5=DWF_M5, mass=DWF_MASS, inv=False, dag=False):
    """out = Mooee inp (or Mooee^-1, or their adjoints): the DWF fifth-dimension term."""
    check(lib.gbx_dwf_mooee(out.h, inp.h, C.byref(DwfParams(M5, mass)), int(bool(inv)), int(bool(dag))),
          "gbx_dwf_mooee")


def dwf_apply(U, out, inp, M5=DWF_M5, mass=DWF_MASS, dag=False):
    """out = D inp, D = Mooee - Dhop / 2 (whole lattice)."""
    check(lib.gbx_dwf_apply(U.h, out.h, inp.h, C.byref(DwfParams(M5, mass)), int(bool(dag))), "gbx_dwf_apply")


def dwf_solve_eo(U_hi, U_lo, x, b, M5=DWF_M5, mass=DWF_MASS, tol=1e-10, inner_tol=1e-3, max_outer=50,
                 max_inner=5000):
    """D x = b on a red-black pair by the even-odd Schur complement."""
    p = CGParams(0.0, tol, inner_tol, max_outer, max_inner)
    st = CGStats()
    arr = lambda objs: (C.c_void_p * 2)(*[o.h for o in objs])
    check(lib.gbx_dwf_solve_eo(arr(U_hi), arr(U_lo) if U_lo is not None else None, arr(x), arr(b),
                               C.byref(DwfParams(M5, mass)), C.byref(p), C.byref(st)), "gbx_dwf_solve_eo")
    return st.as_dict()


# --------------------------------------------------- reference static_data
def static_data_write(cc_path, h_path, nsite, Ls, nsimd, U, Phi, Psi_cpp, nbr, prm):
    """Write static_data.cc / .h in the text format of TableGenerate.cc:93-138."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (U, Phi, Psi_cpp)]
    n = np.ascontiguousarray(nbr, dtype=np.uint64)
    p = np.ascontiguousarray(prm, dtype=np.uint8)
    check(lib.gbx_static_data_write(str(cc_path).encode(), str(h_path).encode(), int(nsite), int(Ls), int(nsimd),
                                    ptr(a[0], C.c_double), ptr(a[1], C.c_double), ptr(a[2], C.c_double),
                                    ptr(n, C.c_uint64), ptr(p, C.c_uint8)), "gbx_static_data_write")


def static_data_read_header(h_path):
    ns, ls = C.c_uint64(), C.c_int()
    check(lib.gbx_static_data_read_header(str(h_path).encode(), C.byref(ns), C.byref(ls)),
          "gbx_static_data_read_header")
    return int(ns.value), int(ls.value)


def static_data_read(paths):
    """Arrays of a static_data blob (one .cc file or several): dict U, Phi, Psi_cpp, nbr, prm."""
    if isinstance(paths, (str, bytes)) or hasattr(paths, "__fspath__"):
        paths = [paths]
    arr = (C.c_char_p * len(paths))(*[str(p).encode() for p in paths])
    sizes = np.zeros(5, dtype=np.uint64)
    check(lib.gbx_static_data_sizes(arr, len(paths), ptr(sizes, C.c_uint64)), "gbx_static_data_sizes")
    out = {"U": np.empty(int(sizes[0])), "Phi": np.empty(int(sizes[1])), "Psi_cpp": np.empty(int(sizes[2])),
           "nbr": np.empty(int(sizes[3]), dtype=np.uint64), "prm": np.empty(int(sizes[4]), dtype=np.uint8)}
    check(lib.gbx_static_data_read(arr, len(paths), ptr(out["U"], C.c_double), ptr(out["Phi"], C.c_double),
                                   ptr(out["Psi_cpp"], C.c_double), ptr(out["nbr"], C.c_uint64),
                                   ptr(out["prm"], C.c_uint8), ptr(sizes, C.c_uint64)), "gbx_static_data_read")
    return out


def part_tables(dims, grid, coord, block=None):
    """Host tables of a domain part (no GPU needed): dict with site_of_slot
    and gsite [V_local], nbr [V_local][8], face_slots {dim: [2][face_dim]},
    ghost_base {dim: first ghost slot of dim}, ninterior, vol, face."""
    dims, grid, coord = (tuple(int(v) for v in a) for a in (dims, grid, coord))
    ok = all(d >= 1 and n >= 1 and d % n == 0 for d, n in zip(dims, grid))
    # invalid arguments: the library rejects them before writing any table
    ld = [d // n if ok else 1 for d, n in zip(dims, grid)]
    vol = int(np.prod(ld))
    face = {d: vol // ld[d] for d in range(4) if ok and grid[d] > 1}
    nfs = max(1, sum(2 * f for f in face.values()))
    sos = np.zeros(vol, np.uint32)
    gs = np.zeros(vol, np.uint32)
    nbr = np.zeros((vol, 8), np.uint32)
    fs = np.zeros(nfs, np.uint32)
    nint = C.c_uint64()
    check(lib.gbx_part_tables(ivec(dims), ivec(grid), ivec(coord), ivec(block) if block is not None else None,
                              ptr(sos, C.c_uint32), ptr(gs, C.c_uint32), ptr(nbr, C.c_uint32), ptr(fs, C.c_uint32),
                              C.byref(nint)), "gbx_part_tables")
    slots, base, ghost = {}, 0, vol
    ghost_base = {}
    for d in sorted(face):
        slots[d] = fs[base:base + 2 * face[d]].reshape(2, face[d])
        ghost_base[d] = ghost
        base += 2 * face[d]
        ghost += 2 * face[d]
    return {"site_of_slot": sos, "gsite": gs, "nbr": nbr, "face_slots": slots, "ghost_base": ghost_base,
            "ninterior": int(nint.value), "vol": vol, "face": face}


def part_split(dims, grid, lex):
    """Lexicographic spinor array [nrhs][V][Ls][...] of the lattice `dims` ->
    {grid coord: [nrhs][V_local][Ls][...]}, each block in its local
    lexicographic order (the layout SpinorField.upload takes on a
    Geometry.part_of geometry)."""
    a = np.asarray(lex)
    nrhs, rest = a.shape[0], a.shape[2:]
    a = a.reshape((nrhs, dims[3], dims[2], dims[1], dims[0]) + rest)
    ld = [d // n for d, n in zip(dims, grid)]
    out = {}
    for c in grid_coords(grid):
        sl = tuple(slice(c[d] * ld[d], (c[d] + 1) * ld[d]) for d in (3, 2, 1, 0))
        out[c] = np.ascontiguousarray(a[(slice(None),) + sl]).reshape((nrhs, -1) + rest)
    return out


def part_merge(dims, grid, parts):
    """Inverse of part_split."""
    some = next(iter(parts.values()))
    nrhs, rest = some.shape[0], some.shape[2:]
    ld = [d // n for d, n in zip(dims, grid)]
    out = np.empty((nrhs, dims[3], dims[2], dims[1], dims[0]) + rest, dtype=some.dtype)
    for c, p in parts.items():
        sl = tuple(slice(c[d] * ld[d], (c[d] + 1) * ld[d]) for d in (3, 2, 1, 0))
        out[(slice(None),) + sl] = np.asarray(p).reshape((nrhs, ld[3], ld[2], ld[1], ld[0]) + rest)
    return out.reshape((nrhs, -1) + rest)


def rb_split(dims, lex):
    """Whole-lattice field [..., V, ...] (site axis -1 of the leading shape is
    the lexicographic site) -> (even, odd) halves in checkerboarded
    lexicographic order ((x/2) fastest).  `lex` shape [nrhs, V, ...]."""
    L0 = int(dims[0])
    V = int(np.prod(dims))
    site = np.arange(V)
    x = [(site // int(np.prod(dims[:d]))) % int(dims[d]) for d in range(4)]
    par = (x[0] + x[1] + x[2] + x[3]) % 2
    hidx = (site // L0) * (L0 // 2) + x[0] // 2
    out = []
    for p in (0, 1):
        sel = np.nonzero(par == p)[0]
        h = np.empty((lex.shape[0], V // 2) + lex.shape[2:], dtype=lex.dtype)
        h[:, hidx[sel]] = lex[:, sel]
        out.append(h)
    return out[0], out[1]


def rb_merge(dims, even, odd):
    """Inverse of rb_split."""
    L0 = int(dims[0])
    V = int(np.prod(dims))
    site = np.arange(V)
    x = [(site // int(np.prod(dims[:d]))) % int(dims[d]) for d in range(4)]
    par = (x[0] + x[1] + x[2] + x[3]) % 2
    hidx = (site // L0) * (L0 // 2) + x[0] // 2
    lex = np.empty((even.shape[0], V) + even.shape[2:], dtype=even.dtype)
    for p, h in ((0, even), (1, odd)):
        sel = np.nonzero(par == p)[0]
        lex[:, sel] = h[:, hidx[sel]]
    return lex


class Batch:
    """Many independent lattices in one persistent, graph-captured launch."""

    def __init__(self, gauges, outs, ins, dag=False):
        n = len(gauges)
        assert n == len(outs) == len(ins)
        arr = lambda objs: (C.c_void_p * n)(*[o.h for o in objs])
        h = C.c_void_p()
        check(lib.gbx_batch_create(n, arr(gauges), arr(outs), arr(ins), int(bool(dag)), C.byref(h)),
              "gbx_batch_create")
        self.h = h
        self._keep = (gauges, outs, ins)

    def run(self):
        check(lib.gbx_batch_run(self.h), "gbx_batch_run")

    def time(self, nrep=10):
        us = C.c_double()
        check(lib.gbx_batch_time(self.h, int(nrep), C.byref(us)), "gbx_batch_time")
        return us.value

    def info(self):
        items, wgs = C.c_uint64(), C.c_int()
        check(lib.gbx_batch_info(self.h, C.byref(items), C.byref(wgs)), "gbx_batch_info")
        return items.value, wgs.value

    def get_cfg(self):
        """Launch shape: threads, pf, wgpc, nt, spw = points per work item."""
        k = KCfg()
        check(lib.gbx_batch_get_cfg(self.h, C.byref(k)), "gbx_batch_get_cfg")
        return k

    def set_cfg(self, cfg):
        check(lib.gbx_batch_set_cfg(self.h, C.byref(cfg)), "gbx_batch_set_cfg")
        return self

    def tune(self, nrep=3):
        """Time launch shapes on these fields and keep the fastest; returns (cfg, us)."""
        k, us = KCfg(), C.c_double()
        check(lib.gbx_batch_tune(self.h, int(nrep), C.byref(k), C.byref(us)), "gbx_batch_tune")
        return k, us.value

    def close(self):
        if self.h:
            lib.gbx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run_noalloc(gauges, outs, ins, dag=False):
    """The batched application of Batch(...).run() without any device
    allocation (gbx_batch_run_noalloc): at most NOALLOC_MAX_LATTICES lattices,
    one fixed scratch array per GPU; calls overlapping on one GPU race."""
    n = len(gauges)
    assert n == len(outs) == len(ins)
    arr = lambda objs: (C.c_void_p * n)(*[o.h for o in objs])
    check(lib.gbx_batch_run_noalloc(n, arr(gauges), arr(outs), arr(ins), int(bool(dag))), "gbx_batch_run_noalloc")


NOALLOC_MAX_LATTICES = 32  # GBX_NOALLOC_MAX_LATTICES


def stream_copy_bw(nbytes=2 << 30, nrep=20):
    """Achievable HBM bandwidth (GB/s, read + write) of a streaming copy."""
    g = C.c_double()
    check(lib.gbx_stream_copy_bw(int(nbytes), int(nrep), C.byref(g)), "gbx_stream_copy_bw")
    return g.value


def stream_bw(read_bytes=4 << 30, mix=3, nrep=20):
    """HBM bandwidth (GB/s, read + write) of a read-weighted stream: one
    16-byte write per `mix` 16-byte reads, or reads only with mix = 0."""
    g = C.c_double()
    check(lib.gbx_stream_bw(int(read_bytes), int(mix), int(nrep), C.byref(g)), "gbx_stream_bw")
    return g.value


def gather_bw(buffer_bytes=2 << 30, nrep=10):
    """Read bandwidth (GB/s) of wavefront gathers of 4 random 256-byte runs
    from a buffer of buffer_bytes (a power of two): 2 GiB measures HBM, a
    few MiB the L2s."""
    g = C.c_double()
    check(lib.gbx_gather_bw(int(buffer_bytes), int(nrep), C.byref(g)), "gbx_gather_bw")
    return g.value


# ------------------------------------------------------------ work models
def bytes_per_site(prec_spinor, prec_gauge, Ls, nrhs, comp12=False):
    """Algorithmic HBM bytes per (5D site, right-hand side): in + out spinor
    once, the doubled gauge field once per 4D site."""
    bs = {F64: 8, F32: 4, F16: 2, BF16: 2}[_prec(prec_spinor)]
    bu = {F64: 8, F32: 4, F16: 2, BF16: 2}[_prec(prec_gauge)]
    RU = 12 if comp12 else 18
    return 48 * bs + 8 * RU * bu / (Ls * nrhs)


# Spin bases of stored spinor values (include/gbx.h).
BASIS_CHIRAL, BASIS_TDIAG = 0, 1

__all__ = [
    "BASIS_CHIRAL", "BASIS_TDIAG", "FLOPS_PER_SITE", "F64", "F32", "F16", "BF16", "PREC_NAMES", "GbxError", "device_count", "gen_gaussian",
    "gen_su3_links", "double_links", "gen_doubled_gauge", "gen_spinor", "table_generate", "table_recognise",
    "dslash_kernel", "dslash_last_path", "DslashContext", "Geometry", "GaugeField", "SpinorField", "kcfg", "pack_tile",
    "pack_march", "march_cfg",
    "dhop", "dhop_axpby",
    "dhop_get_cfg", "dhop_time", "tune", "cg_solve", "Batch", "batch_run_noalloc", "NOALLOC_MAX_LATTICES", "stream_copy_bw", "stream_bw", "gather_bw", "bytes_per_site", "dhop_xpay",
    "cg_solve_eo", "rb_split", "rb_merge", "DWF_M5", "DWF_MASS", "dwf_mooee", "dwf_apply", "dwf_solve_eo",
    "static_data_write", "static_data_read_header", "static_data_read", "dhop_range", "halo_exchange", "part_split",
    "part_merge", "DeviceBuffer", "part_tables", "halo_overlap_time", "grid_coords", "halo_buffers",
    "dhop_faces_projected", "halo_exchange_in_place",
]
