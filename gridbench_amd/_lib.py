"""ctypes binding of libgbx.so (the C-ABI declared in include/gbx.h).

The library is loaded from this package directory, where
`gridbench_amd/csrc/Makefile` builds it.  There is no fallback: if the
library is missing, importing this module raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgbx.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C gridbench_amd/csrc` "
        "or __graft_entry__.build()")

lib = C.CDLL(LIB_PATH)

OK, EINVAL, ENULL, EALIAS, ETABLE, ENOMEM, EHIP, ENODEV, EUNSUPPORTED, EMISMATCH, ENOCONV = range(11)
F64, F32, F16, BF16 = 0, 1, 2, 3
PREC_NAMES = {F64: "f64", F32: "f32", F16: "f16", BF16: "bf16"}
PREC_BY_NAME = {v: k for k, v in PREC_NAMES.items()}


class GbxError(RuntimeError):
    """A libgbx call returned a non-zero status."""

    def __init__(self, status, where):
        self.status = status
        msg = lib.gbx_last_error().decode()
        super().__init__(f"{where}: {lib.gbx_strerror(status).decode()} ({status}): {msg}")


def check(status, where="libgbx"):
    if status != OK:
        raise GbxError(status, where)


class KCfg(C.Structure):
    _fields_ = [("spw", C.c_int), ("threads", C.c_int), ("xcd", C.c_int), ("nt", C.c_int), ("kind", C.c_int),
                ("mc", C.c_int), ("pf", C.c_int), ("wgpc", C.c_int),
                ("cblk", C.c_int), ("tile", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CGParams(C.Structure):
    _fields_ = [("kappa", C.c_double), ("tol", C.c_double), ("inner_tol", C.c_double),
                ("max_outer", C.c_int), ("max_inner", C.c_int)]


class DwfParams(C.Structure):
    _fields_ = [("M5", C.c_double), ("mass", C.c_double)]


class CGStats(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters", C.c_int), ("true_resid", C.c_double),
                ("seconds", C.c_double), ("hi_applies", C.c_int), ("lo_applies", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


vp = C.c_void_p
i32, u64, dbl = C.c_int, C.c_uint64, C.c_double
pi32, pu64, pdbl, pflt, pu8 = (C.POINTER(C.c_int), C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                               C.POINTER(C.c_float), C.POINTER(C.c_uint8))
pu32 = C.POINTER(C.c_uint32)

# name: (restype, argtypes) -- every function declared in include/gbx.h
PROTOTYPES = {
    "gbx_strerror": (C.c_char_p, [i32]),
    "gbx_last_error": (C.c_char_p, []),
    "gbx_version": (i32, []),
    "gbx_device_count": (i32, []),
    "gbx_dslash_kernel_f64": (i32, [i32, i32, i32, pdbl, pdbl, pdbl, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_kernel_f32": (i32, [i32, i32, i32, pflt, pflt, pflt, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_kernel_rrii_f64": (i32, [i32, i32, i32, pdbl, pdbl, pdbl, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_kernel_rrii_f32": (i32, [i32, i32, i32, pflt, pflt, pflt, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_kernel_expand_f64": (i32, [i32, i32, i32, i32, i32, pdbl, pdbl, pdbl, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_kernel_expand_f32": (i32, [i32, i32, i32, i32, i32, pflt, pflt, pflt, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_last_path": (i32, []),
    "gbx_dslash_ctx_create": (i32, [C.POINTER(C.c_void_p)]),
    "gbx_dslash_ctx_destroy": (i32, [C.c_void_p]),
    "gbx_dslash_ctx_kernel_f64": (i32, [C.c_void_p, i32, i32, i32, i32, i32, pdbl, pdbl, pdbl, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_ctx_kernel_f32": (i32, [C.c_void_p, i32, i32, i32, i32, i32, pflt, pflt, pflt, pu64, u64, u64, pu8, pdbl]),
    "gbx_dslash_ctx_last_path": (i32, [C.c_void_p]),
    "gbx_table_generate": (i32, [pi32, i32, u64, pu64, pu8]),
    "gbx_table_recognise": (i32, [i32, u64, u64, pu64, pu8, pi32]),
    "gbx_gen_gaussian": (i32, [u64, u64, u64, u64, pdbl]),
    "gbx_gen_su3_links": (i32, [pi32, u64, pdbl]),
    "gbx_double_links": (i32, [pi32, pdbl, pdbl]),
    "gbx_gen_spinor": (i32, [pi32, i32, i32, u64, pdbl]),
    "gbx_geom_create": (i32, [pi32, pi32, C.POINTER(vp)]),
    "gbx_geom_create_levels": (i32, [pi32, i32, pi32, C.POINTER(vp)]),
    "gbx_geom_create_rb": (i32, [pi32, pi32, C.POINTER(vp), C.POINTER(vp)]),
    "gbx_geom_create_order": (i32, [pi32, pu32, C.POINTER(vp)]),
    "gbx_geom_create_part_rb": (i32, [pi32, pi32, pi32, pi32, C.POINTER(vp), C.POINTER(vp)]),
    "gbx_geom_parity": (i32, [vp, pi32]),
    "gbx_geom_create_part": (i32, [pi32, pi32, pi32, pi32, C.POINTER(vp)]),
    "gbx_part_tables": (i32, [pi32, pi32, pi32, pi32, pu32, pu32, pu32, pu32, pu64]),
    "gbx_geom_part_info": (i32, [vp, pi32, pi32, pu64, pu64]),
    "gbx_geom_destroy": (i32, [vp]),
    "gbx_geom_volume": (i32, [vp, pu64]),
    "gbx_geom_stream": (i32, [vp, C.POINTER(vp)]),
    "gbx_sync": (i32, [vp]),
    "gbx_gauge_create": (i32, [vp, i32, i32, C.POINTER(vp)]),
    "gbx_gauge_destroy": (i32, [vp]),
    "gbx_gauge_upload": (i32, [vp, pdbl]),
    "gbx_gauge_upload_ref_f64": (i32, [vp, pdbl, i32]),
    "gbx_gauge_upload_ref_f32": (i32, [vp, pflt, i32]),
    "gbx_spinor_create": (i32, [vp, i32, i32, i32, C.POINTER(vp)]),
    "gbx_spinor_destroy": (i32, [vp]),
    "gbx_spinor_bytes": (i32, [vp, pu64]),
    "gbx_spinor_upload": (i32, [vp, pdbl]),
    "gbx_spinor_download": (i32, [vp, pdbl]),
    "gbx_spinor_upload_rhs": (i32, [vp, i32, pdbl]),
    "gbx_spinor_download_rhs": (i32, [vp, i32, pdbl]),
    "gbx_spinor_upload_ref_f64": (i32, [vp, pdbl, i32]),
    "gbx_spinor_download_ref_f64": (i32, [vp, pdbl, i32]),
    "gbx_spinor_upload_ref_f32": (i32, [vp, pflt, i32]),
    "gbx_spinor_download_ref_f32": (i32, [vp, pflt, i32]),
    "gbx_spinor_copy": (i32, [vp, vp]),
    "gbx_spinor_fill_gaussian": (i32, [vp, u64]),
    "gbx_spinor_set_basis": (i32, [vp, i32]),
    "gbx_spinor_get_basis": (i32, [vp, pi32]),
    "gbx_dhop": (i32, [vp, vp, vp, i32]),
    "gbx_dhop_range": (i32, [vp, vp, vp, i32, u64, u64]),
    "gbx_halo_bytes": (i32, [vp, i32, i32, pu64]),
    "gbx_halo_ghost": (i32, [vp, i32, i32, C.POINTER(vp)]),
    "gbx_halo_pack": (i32, [vp, i32, i32, vp]),
    "gbx_halo_unpack": (i32, [vp, i32, i32, vp]),
    "gbx_halo_pack_proj": (i32, [vp, i32, i32, i32, vp]),
    "gbx_halo_unpack_proj": (i32, [vp, i32, i32, vp]),
    "gbx_halo_overlap_time": (i32, [vp, vp, vp, i32, i32, i32, pdbl, pdbl]),
    "gbx_halo_pack_all": (i32, [vp, i32, i32, C.POINTER(vp)]),
    "gbx_halo_ghost_proj": (i32, [vp, i32, i32, C.POINTER(vp)]),
    "gbx_dhop_faces_projected": (i32, [vp, vp, vp, i32]),
    "gbx_halo_unpack_all": (i32, [vp, i32, C.POINTER(vp)]),
    "gbx_buffer_create": (i32, [u64, C.POINTER(vp)]),
    "gbx_buffer_destroy": (i32, [vp]),
    "gbx_buffer_download": (i32, [vp, u64, vp]),
    "gbx_dhop_axpby": (i32, [vp, vp, vp, i32, dbl, dbl]),
    "gbx_dhop_xpay": (i32, [vp, vp, vp, vp, i32, dbl, dbl]),
    "gbx_dhop_xpay_cfg": (i32, [vp, vp, vp, vp, i32, dbl, dbl, C.POINTER(KCfg)]),
    "gbx_dhop_cfg": (i32, [vp, vp, vp, i32, C.POINTER(KCfg)]),
    "gbx_dhop_get_cfg": (i32, [vp, vp, C.POINTER(KCfg)]),
    "gbx_dhop_time": (i32, [vp, vp, vp, i32, C.POINTER(KCfg), i32, pdbl]),
    "gbx_tune": (i32, [vp, vp, vp, i32, C.c_char_p, C.POINTER(KCfg), pdbl]),
    "gbx_tune_kinds": (i32, [vp, vp, vp, i32, i32, C.c_char_p, C.POINTER(KCfg), pdbl]),
    "gbx_spinor_dot": (i32, [vp, vp, pdbl, pdbl]),
    "gbx_spinor_norm2": (i32, [vp, pdbl]),
    "gbx_spinor_axpy": (i32, [dbl, vp, vp]),
    "gbx_spinor_scale": (i32, [vp, dbl]),
    "gbx_spinor_zero": (i32, [vp]),
    "gbx_cg_solve": (i32, [vp, vp, vp, vp, C.POINTER(CGParams), C.POINTER(CGStats)]),
    "gbx_cg_solve_f64": (i32, [vp, vp, vp, C.POINTER(CGParams), C.POINTER(CGStats)]),
    "gbx_cg_solve_eo": (i32, [C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(CGParams),
                              C.POINTER(CGStats)]),
    "gbx_dwf_mooee": (i32, [vp, vp, C.POINTER(DwfParams), i32, i32]),
    "gbx_static_data_write": (i32, [C.c_char_p, C.c_char_p, u64, i32, i32, pdbl, pdbl, pdbl, pu64, pu8]),
    "gbx_static_data_read_header": (i32, [C.c_char_p, pu64, pi32]),
    "gbx_static_data_sizes": (i32, [C.POINTER(C.c_char_p), i32, pu64]),
    "gbx_static_data_read": (i32, [C.POINTER(C.c_char_p), i32, pdbl, pdbl, pdbl, pu64, pu8, pu64]),
    "gbx_dwf_apply": (i32, [vp, vp, vp, C.POINTER(DwfParams), i32]),
    "gbx_dwf_solve_eo": (i32, [C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(DwfParams),
                               C.POINTER(CGParams), C.POINTER(CGStats)]),
    "gbx_batch_create": (i32, [i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), i32, C.POINTER(vp)]),
    "gbx_batch_destroy": (i32, [vp]),
    "gbx_batch_run": (i32, [vp]),
    "gbx_batch_time": (i32, [vp, i32, pdbl]),
    "gbx_batch_info": (i32, [vp, pu64, pi32]),
    "gbx_batch_get_cfg": (i32, [vp, C.POINTER(KCfg)]),
    "gbx_batch_set_cfg": (i32, [vp, C.POINTER(KCfg)]),
    "gbx_batch_tune": (i32, [vp, i32, C.POINTER(KCfg), pdbl]),
    "gbx_batch_run_noalloc": (i32, [i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), i32]),
    "gbx_stream_copy_bw": (i32, [u64, i32, pdbl]),
    "gbx_stream_bw": (i32, [u64, i32, i32, pdbl]),
    "gbx_gather_bw": (i32, [u64, i32, pdbl]),
}

for _name, (_res, _args) in PROTOTYPES.items():
    _f = getattr(lib, _name)  # AttributeError here = the library lacks a declared symbol
    _f.restype = _res
    _f.argtypes = _args


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array's data."""
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def ivec(values):
    return (C.c_int * len(values))(*[int(v) for v in values])
