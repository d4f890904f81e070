"""ORACLE -- test infrastructure only.

ctypes access to
  * liboracle.so: the plain-C restatement of the reference Hand kernel
    (hand.c / hand_impl.h) and its table/layout helpers;
  * libport_v{3,4}.so: the same source built for timing (multiply-adds may
    be fused), for bench.py's CPU baseline only;
  * _ref/libref_w{16,32,64,128}.so (when built in this checkout): the
    reference's own dslash_kernel_cpu (WilsonKernelsHandCpu.h:385-451) and
    the scalar dslash_simple.h kernel, compiled from the reference sources by
    oracle/Makefile.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker (or the timed CPU restatement).  The
product (gridbench_amd) never does.
"""
import ctypes as C
import hashlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        P = C.POINTER
        for name, T in (("f64", C.c_double), ("f32", C.c_float)):
            f = getattr(L, "oracle_dhop_hand_" + name)
            f.restype = None
            f.argtypes = [C.c_int, C.c_int, P(T), P(T), P(T), P(C.c_uint64), P(C.c_uint8), C.c_uint64, C.c_uint64,
                          C.c_int]
        L.oracle_tables.restype = C.c_int
        L.oracle_tables.argtypes = [P(C.c_int), C.c_int, C.c_uint64, P(C.c_uint64), P(C.c_uint8)]
        L.oracle_simd_layout.restype = C.c_int
        L.oracle_simd_layout.argtypes = [C.c_int, P(C.c_int)]
        for n in ("oracle_pack_spinor_f64", "oracle_unpack_spinor_f64"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [P(C.c_int), C.c_int, C.c_uint64, P(C.c_double), P(C.c_double)]
        L.oracle_pack_gauge_f64.restype = C.c_int
        L.oracle_pack_gauge_f64.argtypes = [P(C.c_int), C.c_int, P(C.c_double), P(C.c_double)]
        _lib = L
    return _lib


_port = None


def host_supports_avx512():
    """x86-64-v4 (AVX-512 F/BW/CD/DQ/VL) on this host, from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= fl
    except OSError:
        pass
    return False


def port_lib():
    """The restatement built for timing (libport_v4.so on AVX-512 hosts, else
    libport_v3.so): same source, multiply-adds may be fused.  Returns
    (library, name).  bench.py's CPU baseline only; never the checker."""
    global _port
    if _port is None:
        name = "libport_v4.so" if host_supports_avx512() else "libport_v3.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise ImportError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        P = C.POINTER
        f = L.oracle_dhop_hand_f64
        f.restype = None
        f.argtypes = [C.c_int, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_uint64), P(C.c_uint8),
                      C.c_uint64, C.c_uint64, C.c_int]
        _port = (L, name)
    return _port


def port_hand(U, inp, nbr, prm, nsite, Ls, nthreads, out=None):
    """One FP64 application of the timing build (Nsimd 1); returns the output."""
    L, _ = port_lib()
    if out is None:
        out = aligned_empty(inp.size, np.float64)
    L.oracle_dhop_hand_f64(1, 0, _p(U, C.c_double), _p(out, C.c_double), _p(inp, C.c_double), _p(nbr, C.c_uint64),
                           _p(prm, C.c_uint8), int(nsite), int(Ls), int(nthreads))
    return out


def _p(a, T):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(T))


def _dims(d):
    return (C.c_int * 4)(*[int(x) for x in d])


def aligned_empty(n, dtype, align=256):
    """1-D array whose data pointer is `align`-byte aligned (the reference's
    generic SIMD types are alignas(GEN_SIMD_WIDTH))."""
    dtype = np.dtype(dtype)
    buf = np.empty(n * dtype.itemsize + align, dtype=np.uint8)
    off = (-buf.ctypes.data) % align
    return buf[off:off + n * dtype.itemsize].view(dtype)


def aligned_copy(a, align=256):
    a = np.ascontiguousarray(a)
    out = aligned_empty(a.size, a.dtype, align)
    out[:] = a.reshape(-1)
    return out


def simd_layout(nsimd):
    s = (C.c_int * 4)()
    if lib().oracle_simd_layout(nsimd, s):
        raise ValueError(nsimd)
    return tuple(s)


def tables(dims, nsimd, Ls):
    nsite = int(np.prod(dims)) // nsimd
    nbr = aligned_empty(nsite * Ls * 8, np.uint64)
    prm = aligned_empty(nsite * Ls * 8, np.uint8)
    if lib().oracle_tables(_dims(dims), nsimd, Ls, _p(nbr, C.c_uint64), _p(prm, C.c_uint8)):
        raise ValueError((dims, nsimd))
    return nbr, prm


def hand(U, inp, nbr, prm, nsite, Ls, nsimd, dag=False, nthreads=1):
    """One application of the oracle Hand kernel; float64 or float32 arrays
    in the reference layout.  Returns the output array."""
    if U.dtype == np.float64:
        f, T = lib().oracle_dhop_hand_f64, C.c_double
    else:
        f, T = lib().oracle_dhop_hand_f32, C.c_float
    out = aligned_empty(inp.size, inp.dtype)
    f(int(nsimd), int(bool(dag)), _p(U, T), _p(out, T), _p(inp, T), _p(nbr, C.c_uint64), _p(prm, C.c_uint8),
      int(nsite), int(Ls), int(nthreads))
    return out


def pack_spinor(dims, nsimd, Ls, lex):
    lex = np.ascontiguousarray(lex, dtype=np.float64).reshape(-1)
    out = aligned_empty(lex.size, np.float64)
    if lib().oracle_pack_spinor_f64(_dims(dims), nsimd, Ls, _p(lex, C.c_double), _p(out, C.c_double)):
        raise ValueError((dims, nsimd))
    return out


def unpack_spinor(dims, nsimd, Ls, ref):
    ref = np.ascontiguousarray(ref, dtype=np.float64).reshape(-1)
    out = np.empty(ref.size, dtype=np.float64)
    if lib().oracle_unpack_spinor_f64(_dims(dims), nsimd, Ls, _p(ref, C.c_double), _p(out, C.c_double)):
        raise ValueError((dims, nsimd))
    return out


def pack_gauge(dims, nsimd, lex_doubled):
    lex = np.ascontiguousarray(lex_doubled, dtype=np.float64).reshape(-1)
    out = aligned_empty(lex.size, np.float64)
    if lib().oracle_pack_gauge_f64(_dims(dims), nsimd, _p(lex, C.c_double), _p(out, C.c_double)):
        raise ValueError((dims, nsimd))
    return out


def gamma5(ref, nsimd):
    """gamma5 = diag(1,1,-1,-1) on a reference-layout spinor array."""
    a = np.array(ref, copy=True).reshape(-1, 4, 3, nsimd, 2)
    a[:, 2:] *= -1
    return aligned_copy(a.reshape(-1))


def to_rrii(ref, width):
    """RIRI -> RRII lane layout of a reference-layout array of `width`-lane
    vectors: each vector's reals first, then its imaginary parts (bench.cc:105-108)."""
    a = np.ascontiguousarray(ref).reshape(-1, width, 2)
    return aligned_copy(a.transpose(0, 2, 1).reshape(-1))


def from_rrii(ref, width):
    a = np.ascontiguousarray(ref).reshape(-1, 2, width)
    return aligned_copy(a.transpose(0, 2, 1).reshape(-1))


def expand_lanes(replicas, data_nsimd):
    """Vectors of len(replicas)*data_nsimd lanes from RIRI arrays of
    data_nsimd lanes: replica nn at lanes nn*data_nsimd + n (bench.cc:110-138)."""
    parts = [np.ascontiguousarray(r).reshape(-1, 1, data_nsimd, 2) for r in replicas]
    return aligned_copy(np.concatenate(parts, axis=1).reshape(-1))


def replica_lanes(ref, width, data_nsimd, nn):
    """Replica nn of an expanded RIRI array, as a data_nsimd-lane array."""
    a = np.ascontiguousarray(ref).reshape(-1, width // data_nsimd, data_nsimd, 2)
    return aligned_copy(a[:, nn].reshape(-1))


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ------------------------------------------------------- the reference build
_ref_libs = {}


def ref_width_for(nsimd, dtype):
    """GEN_SIMD_WIDTH giving `nsimd` complex lanes: 16*nsimd bytes (double), 8*nsimd (float)."""
    return (16 if np.dtype(dtype) == np.float64 else 8) * nsimd


def ref_lib(width):
    """The reference kernel built with GEN_SIMD_WIDTH=width, or None when not built here."""
    if width not in _ref_libs:
        path = os.path.join(REF_DIR, f"libref_w{width}.so")
        if not os.path.exists(path):
            _ref_libs[width] = None
        else:
            L = C.CDLL(path)
            P = C.POINTER
            L.gbxref_dslash_d.restype = C.c_double
            L.gbxref_dslash_d.argtypes = [C.c_int, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_uint64),
                                          C.c_uint64, C.c_uint64, P(C.c_uint8)]
            L.gbxref_dslash_f.restype = C.c_double
            L.gbxref_dslash_f.argtypes = [C.c_int, P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_uint64),
                                          C.c_uint64, C.c_uint64, P(C.c_uint8)]
            L.gbxref_nsimd_d.restype = C.c_int
            L.gbxref_nsimd_f.restype = C.c_int
            if hasattr(L, "gbxref_simple_d"):
                L.gbxref_simple_d.restype = None
                L.gbxref_simple_d.argtypes = [P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_uint64), C.c_uint64,
                                              C.c_uint64, P(C.c_uint8)]
            _ref_libs[width] = L
    return _ref_libs[width]


def ref_hand(U, inp, nbr, prm, nsite, Ls, nsimd, nrep=1):
    """The reference's own dslash_kernel_cpu on the same arrays (None if not built)."""
    L = ref_lib(ref_width_for(nsimd, U.dtype))
    if L is None:
        return None
    out = aligned_empty(inp.size, inp.dtype)
    U, inp = aligned_copy(U), aligned_copy(inp)
    if U.dtype == np.float64:
        assert L.gbxref_nsimd_d() == nsimd
        L.gbxref_dslash_d(nrep, _p(U, C.c_double), _p(out, C.c_double), _p(inp, C.c_double), _p(nbr, C.c_uint64),
                          nsite, Ls, _p(prm, C.c_uint8))
    else:
        assert L.gbxref_nsimd_f() == nsimd
        L.gbxref_dslash_f(nrep, _p(U, C.c_float), _p(out, C.c_float), _p(inp, C.c_float), _p(nbr, C.c_uint64),
                          nsite, Ls, _p(prm, C.c_uint8))
    return out


def ref_time(U, inp, nbr, prm, nsite, Ls, nsimd, nrep):
    """Microseconds the reference's own dslash_kernel_cpu reports for nrep
    applications (its timer covers reps 1..nrep-1), or None if not built."""
    L = ref_lib(ref_width_for(nsimd, U.dtype))
    if L is None:
        return None
    out = aligned_empty(inp.size, inp.dtype)
    U, inp = aligned_copy(U), aligned_copy(inp)
    if U.dtype == np.float64:
        return L.gbxref_dslash_d(int(nrep), _p(U, C.c_double), _p(out, C.c_double), _p(inp, C.c_double),
                                 _p(nbr, C.c_uint64), nsite, Ls, _p(prm, C.c_uint8))
    return L.gbxref_dslash_f(int(nrep), _p(U, C.c_float), _p(out, C.c_float), _p(inp, C.c_float), _p(nbr, C.c_uint64),
                             nsite, Ls, _p(prm, C.c_uint8))


def ref_simple(U, inp, nbr, prm, nsite, Ls):
    """The reference's scalar dslash_simple.h kernel (Nsimd = 1, FP64), or None."""
    L = ref_lib(16)
    if L is None or not hasattr(L, "gbxref_simple_d"):
        return None
    out = aligned_empty(inp.size, np.float64)
    U, inp = aligned_copy(U), aligned_copy(inp)
    L.gbxref_simple_d(_p(U, C.c_double), _p(out, C.c_double), _p(inp, C.c_double), _p(nbr, C.c_uint64), nsite, Ls,
                      _p(prm, C.c_uint8))
    return out
