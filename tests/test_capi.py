"""The C-ABI boundary on a host without a GPU: every symbol include/gbx.h
declares is exported, argument validation and error codes (no compute), and
the C++ mirror of the reference entry (include/gbx_dslash.hpp) compiles
against host code written for the reference's own Simd types."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
REFERENCE = "/root/reference"


def declared_functions():
    src = open(os.path.join(INC, "gbx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gbx_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(gbx):
    from gridbench_amd._lib import lib, PROTOTYPES
    names = declared_functions()
    assert len(names) > 40
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(PROTOTYPES), set(names) ^ set(PROTOTYPES)
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "gridbench_amd", "libgbx.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (gbx_\w+)", out))
    assert set(names) <= exported


def test_strerror_and_last_error(gbx):
    from gridbench_amd._lib import lib
    for code in range(11):
        assert lib.gbx_strerror(code)
    assert lib.gbx_strerror(999) == b"unknown status"
    rc = lib.gbx_table_generate((C.c_int * 4)(4, 4, 4, 4), 3, 2, None, None)
    assert rc == 2  # NULL tables
    assert b"NULL" in lib.gbx_last_error()


def drop_in_args(gbx, dims=(4, 4, 4, 4), Ls=2, nsimd=2):
    nsite = int(np.prod(dims)) // nsimd
    U = np.zeros(nsite * 144 * nsimd)
    inp = np.zeros(nsite * Ls * 24 * nsimd)
    out = np.zeros_like(inp)
    nbr, prm = gbx.table_generate(dims, nsimd, Ls)
    return U, out, inp, nbr, nsite, Ls, prm, nsimd


def test_dslash_kernel_argument_errors(gbx):
    from gridbench_amd._lib import lib, ptr
    U, out, inp, nbr, nsite, Ls, prm, nsimd = drop_in_args(gbx)
    d, u64, u8 = C.c_double, C.c_uint64, C.c_uint8
    us = C.c_double()

    def call(nrep=1, ns=nsimd, dag=0, Up=U, o=out, i=inp, nb=nbr, pr=prm, n=nsite, ls=Ls):
        return lib.gbx_dslash_kernel_f64(nrep, ns, dag, ptr(Up, d) if Up is not None else None, ptr(o, d),
                                         ptr(i, d), ptr(nb, u64), n, ls, ptr(pr, u8), C.byref(us))
    assert call(Up=None) == 2                 # NULL
    assert call(nrep=0) == 1                  # nrep
    assert call(ns=3) == 1                    # nsimd
    assert call(dag=2) == 1                   # dag
    assert call(o=inp) == 3                   # out aliases in
    bad = nbr.copy()
    bad[5] = nsite * Ls
    assert call(nb=bad) == 4                  # neighbour index out of range
    badp = prm.copy()
    badp[7] = 9
    assert call(pr=badp) == 4                 # permute flag not 0/1
    assert b"table entry 7" in lib.gbx_last_error()


def test_dslash_context_lifecycle_and_argument_errors(gbx):
    """Contexts are created and destroyed without a device; a call checks its
    arguments before it needs one."""
    from gridbench_amd._lib import lib, ptr
    U, out, inp, nbr, nsite, Ls, prm, nsimd = drop_in_args(gbx)
    d, u64, u8 = C.c_double, C.c_uint64, C.c_uint8
    us = C.c_double()
    h = C.c_void_p()
    assert lib.gbx_dslash_ctx_create(C.byref(h)) == 0 and h.value
    assert lib.gbx_dslash_ctx_create(None) == 2

    def call(ctx=h, ns=nsimd, dn=nsimd, rrii=0, o=out):
        return lib.gbx_dslash_ctx_kernel_f64(ctx, 1, ns, dn, rrii, 0, ptr(U, d), ptr(o, d), ptr(inp, d), ptr(nbr, u64),
                                             nsite, Ls, ptr(prm, u8), C.byref(us))
    assert call(ctx=None) == 2               # NULL context
    assert call(ns=3) == 1                   # nsimd
    assert call(dn=2 * nsimd) == 1           # data lanes wider than the vector
    assert call(rrii=2) == 1
    assert call(o=inp) == 3                  # out aliases in
    assert lib.gbx_dslash_ctx_last_path(h) == 0
    assert lib.gbx_dslash_ctx_destroy(h) == 0
    assert lib.gbx_dslash_ctx_destroy(None) == 0
    with gbx.DslashContext() as ctx:
        with pytest.raises(gbx.GbxError):
            ctx.kernel(1, U, inp, inp, nbr, nsite, Ls, prm, nsimd)


def test_table_and_generator_argument_errors(gbx):
    from gridbench_amd._lib import lib
    nbr = np.zeros(4096, np.uint64)
    prm = np.zeros(4096, np.uint8)
    p = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    assert lib.gbx_table_generate((C.c_int * 4)(4, 4, 4, 4), 3, 2, p(nbr, C.c_uint64), p(prm, C.c_uint8)) == 1
    assert lib.gbx_table_generate((C.c_int * 4)(4, 4, 4, 3), 2, 2, p(nbr, C.c_uint64), p(prm, C.c_uint8)) == 1
    assert lib.gbx_gen_su3_links((C.c_int * 4)(4, 0, 4, 4), 1, p(np.zeros(10), C.c_double)) == 1
    assert lib.gbx_gen_spinor((C.c_int * 4)(2, 2, 2, 2), 0, 1, 1, p(np.zeros(10), C.c_double)) == 1
    assert lib.gbx_gen_gaussian(1, 0, 0, 5, None) == 2
    x = np.zeros(16 * 72)
    assert lib.gbx_double_links((C.c_int * 4)(2, 2, 2, 2), p(x, C.c_double), p(x, C.c_double)) == 3  # aliasing


def test_geometry_argument_errors(gbx):
    from gridbench_amd._lib import lib
    h = C.c_void_p()
    assert lib.gbx_geom_create((C.c_int * 4)(4, 0, 4, 4), None, C.byref(h)) == 1
    assert lib.gbx_geom_create((C.c_int * 4)(4, 4, 4, 4), (C.c_int * 4)(3, 4, 4, 4), C.byref(h)) == 1
    assert lib.gbx_geom_create_levels((C.c_int * 4)(8, 8, 8, 8), 2, (C.c_int * 8)(2, 2, 2, 2, 3, 4, 4, 4),
                                      C.byref(h)) == 1
    assert lib.gbx_geom_create_levels((C.c_int * 4)(8, 8, 8, 8), 5, (C.c_int * 20)(), C.byref(h)) == 1
    # domain parts: the decomposition is checked before any device work
    d4, i4 = (C.c_int * 4)(4, 4, 4, 8), C.c_int * 4
    assert lib.gbx_geom_create_part(d4, i4(1, 1, 1, 3), i4(0, 0, 0, 0), None, C.byref(h)) == 1  # 3 does not divide 8
    assert lib.gbx_geom_create_part(d4, i4(1, 1, 1, 1), i4(0, 0, 0, 0), None, C.byref(h)) == 1  # nothing cut
    assert lib.gbx_geom_create_part(d4, i4(1, 1, 1, 2), i4(0, 0, 0, 2), None, C.byref(h)) == 1  # coord out of range
    assert lib.gbx_geom_create_part(d4, i4(1, 1, 1, 2), i4(0, 0, 0, 0), i4(3, 4, 4, 4), C.byref(h)) == 1  # block
    assert lib.gbx_geom_create_part(None, i4(1, 1, 1, 2), i4(0, 0, 0, 0), None, C.byref(h)) == 2


def test_halo_argument_errors(gbx):
    from gridbench_amd._lib import lib
    b = C.c_uint64()
    p = C.c_void_p()
    assert lib.gbx_halo_bytes(None, 3, 0, C.byref(b)) == 2
    assert lib.gbx_halo_ghost(None, 3, 0, C.byref(p)) == 2
    assert lib.gbx_halo_pack(None, 3, 0, C.c_void_p(64)) == 2
    assert lib.gbx_halo_pack_proj(None, 3, 0, 0, None) == 2
    assert lib.gbx_halo_unpack(None, 3, 1, C.c_void_p(64)) == 2
    assert lib.gbx_halo_unpack_proj(None, 3, 1, None) == 2
    assert lib.gbx_buffer_create(0, C.byref(p)) == 1
    assert lib.gbx_buffer_destroy(None) == 0
    assert lib.gbx_dhop_range(None, None, None, 0, 0, 1) == 2


def test_probe_argument_errors(gbx):
    """The bandwidth probes check their arguments before touching a device."""
    from gridbench_amd._lib import lib
    g = C.c_double()
    assert lib.gbx_gather_bw(1 << 20, 1, None) == 2
    assert lib.gbx_gather_bw(3 << 20, 1, C.byref(g)) == 1      # not a power of two
    assert lib.gbx_gather_bw(1 << 19, 1, C.byref(g)) == 1      # below 1 MiB
    assert lib.gbx_gather_bw(1 << 20, 0, C.byref(g)) == 1
    assert lib.gbx_stream_bw(1 << 20, 5, 1, C.byref(g)) == 1   # 5 reads per write is not offered
    assert lib.gbx_stream_bw(1 << 20, 0, 1, None) == 2


@pytest.mark.skipif(__import__("gridbench_amd").device_count() > 0, reason="host has a GPU")
def test_no_device_is_reported_not_fatal(gbx):
    from gridbench_amd._lib import lib
    h = C.c_void_p()
    assert lib.gbx_geom_create((C.c_int * 4)(4, 4, 4, 4), None, C.byref(h)) == 7
    g = C.c_double()
    assert lib.gbx_stream_copy_bw(1 << 20, 1, C.byref(g)) == 7
    assert lib.gbx_gather_bw(1 << 20, 1, C.byref(g)) == 7
    U, out, inp, nbr, nsite, Ls, prm, nsimd = drop_in_args(gbx)
    with pytest.raises(gbx.GbxError) as e:
        gbx.dslash_kernel(1, U, out, inp, nbr, nsite, Ls, prm, nsimd)
    assert e.value.status == 7


PROGRAM = r"""
#include <cstdio>
%(includes)s
#include "gbx_dslash.hpp"
int main() {
  const uint64_t nsite = 256 / %(nsimd)d, Ls = 2;
  static %(simd)s U[256 * 72 / %(nsimd)d], in[256 * 2 * 12 / %(nsimd)d], out[256 * 2 * 12 / %(nsimd)d];
  static uint64_t nbr[256 * 2 * 8];
  static uint8_t prm[256 * 2 * 8];
  const int dims[4] = {4, 4, 4, 4};
  if (gbx_table_generate(dims, %(nsimd)d, Ls, nbr, prm)) return 10;
  double us = dslash_kernel<%(simd)s>(10, U, out, in, nbr, nsite, Ls, prm);
  std::printf("usec %%g\n", us);
  return us < 0 ? (int)(-us) : 0;
}
"""


def build_and_run(src, extra_flags):
    with tempfile.TemporaryDirectory() as d:
        cpp, exe = os.path.join(d, "t.cc"), os.path.join(d, "t")
        open(cpp, "w").write(src)
        subprocess.run(["g++", "-std=c++14", "-O1", "-w", "-I", INC] + extra_flags +
                       [cpp, "-L", os.path.join(ROOT, "gridbench_amd"), "-lgbx",
                        "-Wl,-rpath," + os.path.join(ROOT, "gridbench_amd"), "-o", exe], check=True)
        return subprocess.run([exe], capture_output=True, text=True)


def test_cpp_mirror_compiles_with_standin_simd_type(gbx):
    r = build_and_run(PROGRAM % {"includes": "", "simd": "gbx::vcomplex<double, 2>", "nsimd": 2}, [])
    expect = 0 if gbx.device_count() > 0 else 7  # GBX_ENODEV without a GPU
    assert r.returncode == expect, r.stderr


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference sources not present")
def test_cpp_mirror_compiles_with_the_references_own_simd_types(gbx):
    """Host code using the reference's vComplexD (SimdVector.h:441) calls the
    library through the same dslash_kernel<Simd>(...) signature."""
    src = PROGRAM % {"includes": '#include <cassert>\n#include "Simd.h"', "simd": "vComplexD", "nsimd": 2}
    r = build_and_run(src, ["-DGEN", "-DGEN_SIMD_WIDTH=32", "-I", REFERENCE])
    expect = 0 if gbx.device_count() > 0 else 7
    assert r.returncode == expect, r.stderr


def test_host_bench_program_is_built():
    exe = os.path.join(ROOT, "gridbench_amd", "dslash_bench")
    assert os.access(exe, os.X_OK)
