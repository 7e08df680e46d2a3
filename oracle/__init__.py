"""CPU oracle for the SpMM hot path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings over ``liboracle.so`` (built from ``spmm_oracle.c`` by the
Makefile in this folder). Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, for :func:`csr_spmm_f32_omp`, as the timed CPU baseline). The
product library (``spmm_eval_amd``) never imports it.

Every function restates a piece of the reference; the C source cites the
reference file:line each follows. Arrays are numpy, C-contiguous.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_readmtx.so")
_lib = None
_ref = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i = C.c_int


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-B", "-C", _HERE], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_omp_threads": ([], _i),
            "orc_set_threads": ([_i], None),
            "orc_f32_to_bf16": ([C.c_float], C.c_uint16),
            "orc_f32_to_bf16_array": ([C.c_int64, _f32p, _u16p], None),
            "orc_csr_spmm_f64": ([_i, _i, _i32p, _i32p, _f32p, _f32p, _i, _f64p, _i], None),
            "orc_csr_spmm_f64_bf16": ([_i, _i, _i32p, _i32p, _f32p, _u16p, _i, _f64p, _i], None),
            "orc_csr_spmm_f32_omp": ([_i, _i, _i32p, _i32p, _f32p, _f32p, _i, _f32p, _i], None),
            "orc_csr_spmm_bf16_omp": ([_i, _i, _i32p, _i32p, _f32p, _u16p, _i, _f32p, _i], None),
            "orc_csr_spmm_i64": ([_i, _i, _i32p, _i32p, _i32p, _i, _i64p, _i], None),
            "orc_csr2b2sr_nnzb": ([_i, _i32p, _i32p, _i, _i32p], _i),
            "orc_csr2b2sr": ([_i, _i32p, _i32p, _i, _i32p, _i32p, _vp], None),
            "orc_b2sr_spmm_f64": ([_i, _i, _i, _i32p, _i32p, _vp, _f32p, _i, _f64p, _i], None),
            "orc_b2sr_spmm_sign_i32": ([_i, _i, _i, _i32p, _i32p, _vp, _u32p, _i, _i32p, _i], None),
            "orc_pack_sign_rows": ([_i, _i, _f32p, _i, _u32p], None),
            "orc_pack_bin_rows": ([_i, _i, _f32p, _i, _u32p], None),
            "orc_pack_sign_cols": ([_i, _i, _f32p, _i, _u32p], None),
            "orc_unpack_cols": ([_i, _i, _u32p, _f32p, _i], None),
            "orc_bmm32_xnor_f32": ([_i, _i, _i, _u32p, _u32p, _f32p, _i], None),
            "orc_bmm32_xnor_bin": ([_i, _i, _i, _u32p, _u32p, _u32p], None),
            "orc_csr2csc": ([_i, _i, _i32p, _i32p, _vp, _i32p, _i32p, _vp, _vp], None),
            "orc_sddmm_f64": ([_i, _i, _i32p, _i32p, _f32p, _i, _f32p, _i, _f64p], None),
            "orc_sddmm_f32_omp": ([_i, _i, _i32p, _i32p, _f32p, _i, _f32p, _i, _f32p], None),
            "orc_sddmm_bf16_omp": ([_i, _i, _i32p, _i32p, _u16p, _i, _u16p, _i, _f32p], None),
            "orc_b2sr_aligned_size": ([_i, _i, _i32p], _i),
            "orc_b2sr_aligned_pack": ([_i, _i, _i32p, _i32p, _vp, _i32p, _i32p, _u32p], None),
            "orc_spmv_b2sr_aligned": ([_i, _i, _i32p, _i32p, _u32p, _vp, _i32p, _vp], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def ref():
    """The reference's own readMtx.hpp helpers (oracle/_ref), or None when
    that library was not built (the reference tree is absent)."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        R = C.CDLL(_REF_PATH)
        R.ref_csr2csc.argtypes = [_i, _i, _i32p, _i32p, _f32p, _i32p, _i32p, _f32p]
        R.ref_csr2csc.restype = None
        R.ref_coo2csr.argtypes = [_i, _i, _i, _i32p, _i32p, _f32p, _i32p, _i32p, _f32p]
        R.ref_coo2csr.restype = None
        R.ref_read_mtx_info.argtypes = [C.c_char_p] + [C.POINTER(C.c_int)] * 3
        R.ref_read_mtx_info.restype = C.c_int
        R.ref_read_mtx_csr.argtypes = [C.c_char_p, _i32p, _i32p, _f32p]
        R.ref_read_mtx_csr.restype = C.c_int
        _ref = R
    return _ref


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


WORD_DTYPE = {2: np.uint8, 4: np.uint8, 8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def omp_threads() -> int:
    return lib().orc_omp_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def f32_to_bf16(x) -> np.ndarray:
    x = _c(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    lib().orc_f32_to_bf16_array(x.size, x.reshape(-1), out.reshape(-1))
    return out


def bf16_to_f32(h) -> np.ndarray:
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


def csr_spmm_f64(rowptr, colind, val, B) -> np.ndarray:
    rowptr, colind, val, B = _c(rowptr, np.int32), _c(colind, np.int32), _c(val, np.float32), _c(B, np.float32)
    M, N = len(rowptr) - 1, B.shape[1]
    out = np.empty((M, N), np.float64)
    lib().orc_csr_spmm_f64(M, N, rowptr, colind, val, B, N, out, N)
    return out


def csr_spmm_f64_bf16(rowptr, colind, val, Bh) -> np.ndarray:
    rowptr, colind, val, Bh = _c(rowptr, np.int32), _c(colind, np.int32), _c(val, np.float32), _c(Bh, np.uint16)
    M, N = len(rowptr) - 1, Bh.shape[1]
    out = np.empty((M, N), np.float64)
    lib().orc_csr_spmm_f64_bf16(M, N, rowptr, colind, val, Bh, N, out, N)
    return out


def csr_spmm_f32_omp(rowptr, colind, val, B, out=None) -> np.ndarray:
    M, N = len(rowptr) - 1, B.shape[1]
    if out is None:
        out = np.empty((M, N), np.float32)
    lib().orc_csr_spmm_f32_omp(M, N, rowptr, colind, val, B, N, out, N)
    return out


def csr_spmm_bf16_omp(rowptr, colind, val, Bh, out=None) -> np.ndarray:
    M, N = len(rowptr) - 1, Bh.shape[1]
    if out is None:
        out = np.empty((M, N), np.float32)
    lib().orc_csr_spmm_bf16_omp(M, N, rowptr, colind, val, Bh, N, out, N)
    return out


def csr_spmm_i64(rowptr, colind, B) -> np.ndarray:
    rowptr, colind, B = _c(rowptr, np.int32), _c(colind, np.int32), _c(B, np.int32)
    M, N = len(rowptr) - 1, B.shape[1]
    out = np.empty((M, N), np.int64)
    lib().orc_csr_spmm_i64(M, N, rowptr, colind, B, N, out, N)
    return out


def csr2b2sr(rowptr, colind, T):
    """Returns (bsrRowPtr, bsrColInd, bsrVal words)."""
    rowptr, colind = _c(rowptr, np.int32), _c(colind, np.int32)
    M = len(rowptr) - 1
    MB = (M + T - 1) // T
    brp = np.empty(MB + 1, np.int32)
    nb = lib().orc_csr2b2sr_nnzb(M, rowptr, colind, T, brp)
    if nb < 0:
        raise ValueError(f"unsupported tile size {T}")
    bci = np.empty(max(nb, 1), np.int32)
    words = np.zeros(max(nb, 1) * T, WORD_DTYPE[T])
    lib().orc_csr2b2sr(M, rowptr, colind, T, brp, bci, words.ctypes.data)
    return brp, bci[:nb], words[: nb * T]


def b2sr_spmm_f64(T, M, brp, bci, words, B) -> np.ndarray:
    B = _c(B, np.float32)
    words = _c(words, WORD_DTYPE[T])
    N = B.shape[1]
    out = np.empty((M, N), np.float64)
    lib().orc_b2sr_spmm_f64(T, M, N, _c(brp, np.int32), _c(bci, np.int32) if len(bci) else np.zeros(1, np.int32),
                            words.ctypes.data, B, N, out, N)
    return out


def b2sr_spmm_sign_i32(T, M, N, brp, bci, words, Bw) -> np.ndarray:
    Bw = _c(Bw, np.uint32)
    words = _c(words, WORD_DTYPE[T])
    out = np.empty((M, N), np.int32)
    lib().orc_b2sr_spmm_sign_i32(T, M, N, _c(brp, np.int32), _c(bci, np.int32) if len(bci) else np.zeros(1, np.int32),
                                 words.ctypes.data, Bw, Bw.shape[1], out, N)
    return out


def pack_sign_rows(X) -> np.ndarray:
    X = _c(X, np.float32)
    H, W = X.shape
    out = np.empty((H, (W + 31) // 32), np.uint32)
    lib().orc_pack_sign_rows(H, W, X, W, out)
    return out


def pack_bin_rows(X) -> np.ndarray:
    X = _c(X, np.float32)
    H, W = X.shape
    out = np.empty((H, (W + 31) // 32), np.uint32)
    lib().orc_pack_bin_rows(H, W, X, W, out)
    return out


def pack_sign_cols(X) -> np.ndarray:
    X = _c(X, np.float32)
    H, W = X.shape
    out = np.empty(((W + 31) // 32, 32 * ((H + 31) // 32)), np.uint32)
    lib().orc_pack_sign_cols(H, W, X, W, out)
    return out


def unpack_cols(words, H, W) -> np.ndarray:
    words = _c(words, np.uint32)
    out = np.zeros((H, W), np.float32)
    lib().orc_unpack_cols(H, W, words, out, W)
    return out


def bmm32_xnor_f32(A, B, M, N, Kw) -> np.ndarray:
    out = np.empty((M, N), np.float32)
    lib().orc_bmm32_xnor_f32(M, N, Kw, _c(A, np.uint32), _c(B, np.uint32), out, N)
    return out


def bmm32_xnor_bin(A, B, M, N, Kbits) -> np.ndarray:
    out = np.empty(((N + 31) // 32, M), np.uint32)
    lib().orc_bmm32_xnor_bin(M, N, Kbits, _c(A, np.uint32), _c(B, np.uint32), out)
    return out


def csr2csc(rowptr, colind, val, K):
    """Returns (colptr, rowind, cscval, perm)."""
    rowptr, colind = _c(rowptr, np.int32), _c(colind, np.int32)
    val = _c(val, np.float32)
    M = len(rowptr) - 1
    nnz = int(rowptr[-1])
    cp = np.empty(K + 1, np.int32)
    ri = np.empty(max(nnz, 1), np.int32)
    cv = np.empty(max(nnz, 1), np.float32)
    perm = np.empty(max(nnz, 1), np.int32)
    lib().orc_csr2csc(M, K, rowptr, colind if nnz else np.zeros(1, np.int32), val.ctypes.data, cp, ri,
                      cv.ctypes.data, perm.ctypes.data)
    return cp, ri[:nnz], cv[:nnz], perm[:nnz]


def sddmm_f64(rowptr, colind, dY, X) -> np.ndarray:
    rowptr, colind = _c(rowptr, np.int32), _c(colind, np.int32)
    dY, X = _c(dY, np.float32), _c(X, np.float32)
    M, N = len(rowptr) - 1, dY.shape[1]
    out = np.empty(max(int(rowptr[-1]), 1), np.float64)
    lib().orc_sddmm_f64(M, N, rowptr, colind if len(colind) else np.zeros(1, np.int32), dY, N, X, X.shape[1], out)
    return out[: int(rowptr[-1])]


def sddmm_omp(rowptr, colind, dY, X, out=None) -> np.ndarray:
    """FP32 SDDMM at the stored positions (the timed CPU baseline); BF16
    operands when dY and X are uint16 bit patterns."""
    M, N = len(rowptr) - 1, dY.shape[1]
    if out is None:
        out = np.empty(max(int(rowptr[-1]), 1), np.float32)
    fn = lib().orc_sddmm_bf16_omp if dY.dtype == np.uint16 else lib().orc_sddmm_f32_omp
    fn(M, N, rowptr, colind if len(colind) else np.zeros(1, np.int32), dY, dY.shape[1], X,
       X.shape[1], out)
    return out[: int(rowptr[-1])]


def b2sr_aligned_pack(T, brp, bci, words):
    """Aligned B2SR-8/16 repack -> (new_rowptr, new_colind, new_val uint32)."""
    brp = _c(brp, np.int32)
    MB = len(brp) - 1
    n = lib().orc_b2sr_aligned_size(T, MB, brp)
    if n < 0:
        raise ValueError(f"aligned repack needs T = 8 or 16, got {T}")
    nrp = np.empty(MB + 1, np.int32)
    nci = np.empty(max(n, 1), np.int32)
    nval = np.zeros(max(n // (32 // T) * T, 1), np.uint32)
    words = _c(words, WORD_DTYPE[T])
    lib().orc_b2sr_aligned_pack(T, MB, brp, _c(bci, np.int32) if len(bci) else np.zeros(1, np.int32),
                                words.ctypes.data, nrp, nci, nval)
    return nrp, nci[:n], nval[: n // (32 // T) * T]


def spmv_b2sr_aligned(T, nrp, nci, nval, x_words):
    """Binary SpMV on the aligned layout -> (counts int32 [MB*T], binarised words [MB])."""
    nrp = _c(nrp, np.int32)
    MB = len(nrp) - 1
    xw = _c(x_words, WORD_DTYPE[T])
    cnt = np.empty(MB * T, np.int32)
    yb = np.empty(max(MB, 1), WORD_DTYPE[T])
    lib().orc_spmv_b2sr_aligned(T, MB, nrp, _c(nci, np.int32) if len(nci) else np.zeros(1, np.int32),
                                _c(nval, np.uint32) if len(nval) else np.zeros(1, np.uint32),
                                xw.ctypes.data, cnt, yb.ctypes.data)
    return cnt, yb[:MB]

