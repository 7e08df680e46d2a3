"""ctypes bindings to ``libspmm_mi355x.so`` (the HIP C-ABI of
``include/spmm_mi355x.h``) and to ``libspmm_gen.so`` (graph generator).

Loading fails loudly when a library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libspmm_mi355x.so")
GEN_PATH = os.path.join(PKG_DIR, "libspmm_gen.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "spmm_mi355x.h")

SPMM_OK = 0
STATUS = {
    0: "SPMM_OK",
    1: "SPMM_ERR_INVALID_ARG",
    2: "SPMM_ERR_UNSUPPORTED",
    3: "SPMM_ERR_HIP",
    4: "SPMM_ERR_ALLOC",
    5: "SPMM_ERR_MISMATCH",
}


class SpmmError(RuntimeError):
    def __init__(self, status: int, fn: str, msg: str):
        super().__init__(f"{fn} -> {STATUS.get(status, status)}: {msg}")
        self.status = status


class Tune(C.Structure):
    _fields_ = [("lanes_per_row", C.c_int), ("vec", C.c_int), ("rows_scale", C.c_int),
                ("slices", C.c_int), ("variant", C.c_int)]


_vp, _i, _ip = C.c_void_p, C.c_int, C.POINTER(C.c_int)

# name -> argtypes (all return int status)
_SIGS = {
    "spmm_create": [C.POINTER(_vp)],
    "spmm_destroy": [_vp],
    "spmm_version": [],
    "spmm_csr_analyze": [_vp, _i, _i, _vp, C.POINTER(_vp), _vp],
    "spmm_csr_plan_destroy": [_vp],
    "spmm_csr_plan_bins": [_vp, _ip],
    "spmm_csr_f32": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_csr_bf16": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_csr_dropin_geometry": [_i, _i, _i, _i, _i, C.POINTER(Tune)],
    "spmm_csr_f32_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "spmm_csr_bf16_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "spmm_csr_autotune": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _i, C.POINTER(Tune), _vp],
    "spmm_b2sr_autotune": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _i, C.POINTER(Tune), _vp],
    "spmm_b2sr_plan_set": [_vp, _i, _i, C.POINTER(Tune)],
    "csr2b2sr_nnzb": [_vp, _i, _i, _i, _vp, _vp, _i, _vp, _ip, _vp],
    "csr2b2sr": [_vp, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "spmm_b2sr_f32": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_b2sr_bf16": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_b2sr_f32_noalloc": [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_b2sr_analyze": [_vp, _i, _i, _vp, C.POINTER(_vp), _vp],
    "spmm_b2sr_f32_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_b2sr_bf16_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_b2sr_signpacked_i32": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "b2sr_aligned_nnzb": [_vp, _i, _i, _vp, _vp, C.POINTER(_i), _vp],
    "b2sr_aligned_pack": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "spmv_b2sr_aligned_bin": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "spmv_b2sr_aligned_f32": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "pack_sign_rows_u32": [_i, _i, _vp, _i, _vp, _vp],
    "pack_bin_rows_u32": [_i, _i, _vp, _i, _vp, _vp],
    "pack_sign_cols_u32": [_i, _i, _vp, _i, _vp, _vp],
    "unpack_cols_u32_f32": [_i, _i, _vp, _vp, _i, _vp],
    "bmm32_xnor_f32": [_i, _i, _i, _vp, _vp, _vp, _i, _vp],
    "bmm32_xnor_bin": [_i, _i, _i, _vp, _vp, _vp, _vp],
    "csr_transpose": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sddmm_csr_f32": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "sddmm_csr_bf16": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "spmm_csr_t_analyze": [_vp, _i, _i, _i, _vp, _vp, C.POINTER(_vp), _vp],
    "spmm_csr_t_plan_destroy": [_vp],
    "spmm_csr_t_autotune": [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp, _i, C.POINTER(Tune), _vp],
    "spmm_csr_t_f32_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "spmm_csr_t_bf16_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "spmm_csr_t_sddmm_f32_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp,
                                     _i, _vp, _vp, _vp],
    "spmm_csr_t_sddmm_bf16_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp,
                                      _i, _vp, _vp, _vp],
    "spmm_csr_t_f32": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_csr_t_bf16": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "spmm_mtx_read_info": [C.c_char_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)],
    "spmm_mtx_read_csr": [C.c_char_p, _i, _i, _i, _vp, _vp, _vp],
    "sddmm_csr_f32_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "sddmm_csr_bf16_planned": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp],
    "spmm_batch_plan_create": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, C.POINTER(_vp), _vp],
    "spmm_batch_plan_destroy": [_vp],
    "spmm_batch_plan_set": [_vp, _vp, _i, _i],
    "spmm_batch_plan_info": [_vp, C.POINTER(Tune), C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_float)],
    "spmm_batch_fwd_bwd": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
}

_lib = None
_gen = None


def header_functions() -> list[str]:
    """Function names declared in include/spmm_mi355x.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)


def lib():
    """Load libspmm_mi355x.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
            )
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.spmm_status_string.argtypes = [_i]
        L.spmm_status_string.restype = C.c_char_p
        L.spmm_last_error.argtypes = []
        L.spmm_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def call(name: str, *args) -> None:
    """Call a C-ABI function and raise SpmmError on a non-zero status."""
    L = lib()
    st = getattr(L, name)(*args)
    if st != SPMM_OK:
        raise SpmmError(st, name, (L.spmm_last_error() or b"").decode())


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build it first")
        G = C.CDLL(GEN_PATH)
        G.gen_chung_lu.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_uint64]
        G.gen_chung_lu.restype = _vp
        G.gen_graph_nnz.argtypes = [_vp]
        G.gen_graph_nnz.restype = C.c_int64
        G.gen_graph_pairs.argtypes = [_vp]
        G.gen_graph_pairs.restype = C.c_int64
        G.gen_graph_i0.argtypes = [_vp]
        G.gen_graph_i0.restype = C.c_double
        G.gen_graph_copy.argtypes = [_vp, _vp, _vp]
        G.gen_graph_copy.restype = C.c_int
        G.gen_graph_free.argtypes = [_vp]
        G.gen_graph_free.restype = None
        _gen = G
    return _gen
