"""Python host side over the C-ABI (torch tensors are device-memory and
stream plumbing only; every computation runs in libspmm_mi355x.so).

Names follow the reference's host entry points where one exists:
``spmm_csr`` is the CSR SpMM of evalCSRSpmmFloatCuSPARSE
(test_spmm_full.cu:212-342), ``csr2b2sr`` is CSR2B2SR + csr2bsr_batch_T
(test_spmm_full.cu:127-169, backend/csr2bsr_batch_bsrbmv.cu:207-1070),
``spmm_b2sr`` is the FBF kernel launch of evalB2SRSpmmFull32
(test_spmm_full.cu:344-369), ``bmm32_*`` are the kernels of backend/mm.cu.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import Tune, call

WORD_DTYPE = {2: torch.uint8, 4: torch.uint8, 8: torch.uint8, 16: torch.uint16, 32: torch.uint32,
              64: torch.uint64}


def _p(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def _dev(t: torch.Tensor, dtype=None) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    return t


def _is_bf16(B: torch.Tensor) -> bool:
    return B.dtype in (torch.bfloat16, torch.uint16, torch.int16)


class Handle:
    """spmm_handle_t: owns a device workspace; one per host thread."""

    def __init__(self):
        h = C.c_void_p()
        call("spmm_create", C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            _lib.lib().spmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrPlan:
    """Row schedule of one CSR structure (spmm_csr_analyze)."""

    def __init__(self, h: Handle, rowptr: torch.Tensor, nnz: int, stream=None):
        self.M = rowptr.numel() - 1
        self.nnz = int(nnz)
        p = C.c_void_p()
        call("spmm_csr_analyze", h.h, self.M, self.nnz, _p(rowptr), C.byref(p), _stream(stream))
        self.p = p

    def autotune(self, rowptr, colind, val, B, C_out, K: int | None = None, stream=None) -> tuple:
        """Times the valid launch geometries on these operands, keeps the
        fastest in the plan for (N, dtype) and returns it as
        (lanes_per_row, vec, rows_scale, slices)."""
        K = B.shape[0] if K is None else K
        N = B.shape[1]
        best = Tune()
        call("spmm_csr_autotune", self.p, self.M, K, N, self.nnz, _p(rowptr), _p(colind), _p(val),
             _p(B), int(_is_bf16(B)), B.stride(0), _p(C_out), C_out.stride(0), C.byref(best),
             _stream(stream))
        return (best.lanes_per_row, best.vec, best.rows_scale, best.slices)

    def bins(self) -> list[int]:
        arr = (C.c_int * 32)()
        call("spmm_csr_plan_bins", self.p, arr)
        return list(arr)

    def close(self):
        if self.p:
            _lib.lib().spmm_csr_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrTPlan:
    """Plan of C = A^T B (spmm_csr_t_analyze): A's stable transpose, its entry
    permutation and its row schedule; values are read through the
    permutation at every call."""

    def __init__(self, h: Handle, rowptr: torch.Tensor, colind: torch.Tensor, K: int,
                 stream=None):
        self.M = rowptr.numel() - 1
        self.K = int(K)
        self.nnz = colind.numel()
        p = C.c_void_p()
        call("spmm_csr_t_analyze", h.h, self.M, self.K, self.nnz, _p(rowptr), _p(colind),
             C.byref(p), _stream(stream))
        self.p = p

    def autotune(self, val, B, C_out, stream=None) -> tuple:
        """Times the launch geometries of A^T B (column slices included) and
        keeps the fastest in the plan: (lanes_per_row, vec, rows_scale, slices)."""
        best = Tune()
        call("spmm_csr_t_autotune", self.p, self.M, self.K, B.shape[1], self.nnz, _p(val), _p(B),
             int(_is_bf16(B)), B.stride(0), _p(C_out), C_out.stride(0), C.byref(best),
             _stream(stream))
        return (best.lanes_per_row, best.vec, best.rows_scale, best.slices)

    def close(self):
        if self.p:
            _lib.lib().spmm_csr_t_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spmm_csr_t(h: Handle, rowptr, colind, val, B, K: int, C_out=None,
               plan: CsrTPlan | None = None, tune: tuple | None = None, stream=None):
    """C = A^T B for CSR A [M][K] and row-major B [M][N] (FP32 or BF16);
    FP32 C [K][N]."""
    M = rowptr.numel() - 1
    nnz = colind.numel()
    N = B.shape[1]
    if C_out is None:
        C_out = torch.empty((K, N), dtype=torch.float32, device=B.device)
    bf16 = _is_bf16(B)
    if plan is None:
        call("spmm_csr_t_bf16" if bf16 else "spmm_csr_t_f32", h.h, M, K, N, nnz, _p(rowptr),
             _p(colind), _p(val), _p(B), B.stride(0), _p(C_out), C_out.stride(0), _stream(stream))
    else:
        t = C.byref(Tune(*tune)) if tune else None
        call("spmm_csr_t_bf16_planned" if bf16 else "spmm_csr_t_f32_planned", plan.p, M, K, N, nnz,
             _p(val), _p(B), B.stride(0), _p(C_out), C_out.stride(0), t, _stream(stream))
    return C_out


def spmm_csr(h: Handle, rowptr, colind, val, B, K: int | None = None, C_out=None,
             plan: CsrPlan | None = None, tune: tuple | None = None, stream=None):
    """C = A * B for CSR A (FP32 values) and row-major B (FP32, or BF16
    given as torch.bfloat16 / uint16 bits); FP32 C."""
    M = rowptr.numel() - 1
    nnz = colind.numel()
    K = B.shape[0] if K is None else K
    N = B.shape[1]
    if C_out is None:
        C_out = torch.empty((M, N), dtype=torch.float32, device=B.device)
    bf16 = _is_bf16(B)
    ldb = B.stride(0) if B.dim() == 2 else N
    ldc = C_out.stride(0)
    args = (M, K, N, nnz, _p(rowptr), _p(colind), _p(val), _p(B), ldb, _p(C_out), ldc)
    if plan is None:
        call("spmm_csr_bf16" if bf16 else "spmm_csr_f32", h.h, *args, _stream(stream))
    else:
        t = C.byref(Tune(*tune)) if tune else None
        call("spmm_csr_bf16_planned" if bf16 else "spmm_csr_f32_planned", plan.p, *args, t,
             _stream(stream))
    return C_out


def dropin_geometry(M: int, K: int, N: int, nnz: int, bf16: bool = False) -> tuple:
    """The (lanes_per_row, vec, rows_scale, slices) that the plan-free
    spmm_csr runs at for these sizes with aligned operands; slices 0 is one
    launch, lanes 0 the narrow kernel."""
    t = Tune()
    call("spmm_csr_dropin_geometry", M, K, N, nnz, 2 if bf16 else 4, C.byref(t))
    return (t.lanes_per_row, t.vec, t.rows_scale, t.slices)


def csr2b2sr(h: Handle, rowptr, colind, M: int, K: int, T: int, stream=None):
    """Device CSR -> B2SR-T (bsrRowPtr, bsrColInd, bsrVal words)."""
    MB = (M + T - 1) // T
    dev = rowptr.device
    brp = torch.empty(MB + 1, dtype=torch.int32, device=dev)
    nb = C.c_int()
    call("csr2b2sr_nnzb", h.h, M, K, colind.numel(), _p(rowptr), _p(colind), T, _p(brp),
         C.byref(nb), _stream(stream))
    bci = torch.empty(max(nb.value, 1), dtype=torch.int32, device=dev)
    words = torch.empty(max(nb.value, 1) * T, dtype=WORD_DTYPE[T], device=dev)
    call("csr2b2sr", h.h, M, K, colind.numel(), _p(rowptr), _p(colind), T, _p(brp), _p(bci),
         _p(words), _stream(stream))
    return brp, bci[: nb.value], words[: nb.value * T]


class B2srPlan(CsrPlan):
    """Rows binned by their block row's tile count (spmm_b2sr_analyze)."""

    def __init__(self, h: Handle, T: int, M: int, brp: torch.Tensor, stream=None):
        self.M = M
        self.nnz = 0
        self.T = T
        p = C.c_void_p()
        call("spmm_b2sr_analyze", h.h, T, M, _p(brp), C.byref(p), _stream(stream))
        self.p = p

    def autotune(self, brp, bci, words, B, C_out, stream=None) -> tuple:
        """Times the candidate lane mappings (and, for T <= 8 and B of 8 MB
        or more, column slice counts) on these operands, keeps the fastest
        in the plan for (N, dtype) and returns it as
        (variant, lanes_per_row, rows_scale, slices)."""
        K, N = B.shape
        best = Tune()
        call("spmm_b2sr_autotune", self.p, self.T, self.M, K, N, _p(brp), _p(bci), _p(words), _p(B),
             int(_is_bf16(B)), B.stride(0), _p(C_out), C_out.stride(0), C.byref(best),
             _stream(stream))
        return (best.variant, best.lanes_per_row, best.rows_scale, max(best.slices, 1))

    def set_geometry(self, N: int, bf16: bool, variant: int, lanes_per_row: int = 0,
                     rows_scale: int = 1, slices: int = 0):
        """Pins a candidate for (N, dtype) (spmm_b2sr_plan_set)."""
        call("spmm_b2sr_plan_set", self.p, N, int(bf16),
             C.byref(Tune(lanes_per_row, 1, rows_scale, slices, variant)))


def spmm_b2sr(h: Handle, T: int, M: int, brp, bci, words, B, C_out=None,
              plan: B2srPlan | None = None, stream=None):
    K, N = B.shape
    if C_out is None:
        C_out = torch.empty((M, N), dtype=torch.float32, device=B.device)
    bf16 = _is_bf16(B)
    args = (T, M, K, N, _p(brp), _p(bci), _p(words), _p(B), B.stride(0), _p(C_out),
            C_out.stride(0), _stream(stream))
    if plan is None:
        call("spmm_b2sr_bf16" if bf16 else "spmm_b2sr_f32", h.h, *args)
    else:
        call("spmm_b2sr_bf16_planned" if bf16 else "spmm_b2sr_f32_planned", plan.p, *args)
    return C_out


def spmm_b2sr_sign(h: Handle, T: int, M: int, K: int, N: int, brp, bci, words, Bw,
                   C_out=None, stream=None):
    if C_out is None:
        C_out = torch.empty((M, N), dtype=torch.int32, device=Bw.device)
    call("spmm_b2sr_signpacked_i32", h.h, T, M, K, N, _p(brp), _p(bci), _p(words), _p(Bw),
         Bw.stride(0), _p(C_out), C_out.stride(0), _stream(stream))
    return C_out


def b2sr_aligned(h: Handle, T: int, brp, bci, words, stream=None):
    """B2SR-8/16 -> aligned 32-bit groups (newRowPtr, newColInd, newVal uint32):
    countUcharUnsignedSize + packUchar2AlignedUnsigned and their 16-bit forms
    (backend/csr2bsr_batch_bsrbmv.cu:586-647, 837-888)."""
    MB = brp.numel() - 1
    dev = brp.device
    nrp = torch.empty(MB + 1, dtype=torch.int32, device=dev)
    n = C.c_int()
    call("b2sr_aligned_nnzb", h.h, T, MB, _p(brp), _p(nrp), C.byref(n), _stream(stream))
    nwords = n.value // (32 // T) * T
    nci = torch.empty(max(n.value, 1), dtype=torch.int32, device=dev)
    nval = torch.empty(max(nwords, 1), dtype=torch.uint32, device=dev)
    call("b2sr_aligned_pack", h.h, T, MB, n.value, _p(brp), _p(bci), _p(words), _p(nrp),
         _p(nci), _p(nval), _stream(stream))
    return nrp, nci[: n.value], nval[:nwords]


def spmv_b2sr_aligned(h: Handle, T: int, M: int, K: int, nrp, nci, nval, x, y_count=None,
                      y_bin=None, y=None, stream=None):
    """Binary SpMV on the aligned layout. x of T-bit words (one per block
    column) gives (y_count int32 [M], y_bin T-bit words [ceil(M/T)]); FP32 x
    [K] gives y FP32 [M]."""
    nnzb = nci.numel()
    if x.dtype == torch.float32:
        if y is None:
            y = torch.empty(M, dtype=torch.float32, device=x.device)
        call("spmv_b2sr_aligned_f32", h.h, T, M, K, nnzb, _p(nrp), _p(nci), _p(nval), _p(x),
             _p(y), _stream(stream))
        return y
    _dev(x, WORD_DTYPE[T])
    if y_count is None:
        y_count = torch.empty(M, dtype=torch.int32, device=x.device)
    if y_bin is None:
        y_bin = torch.empty((M + T - 1) // T, dtype=WORD_DTYPE[T], device=x.device)
    call("spmv_b2sr_aligned_bin", h.h, T, M, K, nnzb, _p(nrp), _p(nci), _p(nval), _p(x),
         _p(y_count), _p(y_bin), _stream(stream))
    return y_count, y_bin


def _pack(fn: str, X, cols: bool, stream=None):
    H, W = X.shape
    Wc = (W + 31) // 32
    shape = (Wc, 32 * ((H + 31) // 32)) if cols else (H, Wc)
    out = torch.empty(shape, dtype=torch.uint32, device=X.device)
    call(fn, H, W, _p(X), X.stride(0), _p(out), _stream(stream))
    return out


def pack_sign_rows(X, stream=None):
    return _pack("pack_sign_rows_u32", X, False, stream)


def pack_bin_rows(X, stream=None):
    return _pack("pack_bin_rows_u32", X, False, stream)


def pack_sign_cols(X, stream=None):
    return _pack("pack_sign_cols_u32", X, True, stream)


def unpack_cols(words, H: int, W: int, stream=None):
    X = torch.zeros((H, W), dtype=torch.float32, device=words.device)
    call("unpack_cols_u32_f32", H, W, _p(words), _p(X), W, _stream(stream))
    return X


def bmm32_xnor_f32(A, B, M: int, N: int, Kw: int, out=None, stream=None):
    Cm = torch.empty((M, N), dtype=torch.float32, device=A.device) if out is None else out
    call("bmm32_xnor_f32", M, N, Kw, _p(A), _p(B), _p(Cm), Cm.stride(0), _stream(stream))
    return Cm


def bmm32_xnor_bin(A, B, M: int, N: int, Kbits: int, out=None, stream=None):
    Cw = torch.empty(((N + 31) // 32, M), dtype=torch.uint32, device=A.device) if out is None else out
    call("bmm32_xnor_bin", M, N, Kbits, _p(A), _p(B), _p(Cw), _stream(stream))
    return Cw


def csr_transpose(h: Handle, rowptr, colind, val, K: int, stream=None):
    """Returns (tRowPtr[K+1], tColInd, tVal, perm)."""
    M = rowptr.numel() - 1
    nnz = colind.numel()
    dev = rowptr.device
    trp = torch.empty(K + 1, dtype=torch.int32, device=dev)
    tci = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    tv = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev) if val is not None else None
    perm = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    call("csr_transpose", h.h, M, K, nnz, _p(rowptr), _p(colind), _p(val), _p(trp), _p(tci),
         _p(tv), _p(perm), _stream(stream))
    return trp, tci[:nnz], (tv[:nnz] if tv is not None else None), perm[:nnz]


def spmm_csr_t_sddmm(plan: CsrTPlan, rowptr, colind, val, dY, X, dX_out=None, dval_out=None,
                     tune: tuple | None = None, stream=None):
    """A training step's backward in one pass over A's transposed plan:
    dX = A^T dY [K][N] and dval[e] = dY[row(e)] . X[col(e)] (FP32 outputs;
    dY and X both FP32 or both BF16). Returns (dX, dval)."""
    M, K, nnz = plan.M, plan.K, plan.nnz
    N = dY.shape[1]
    if dX_out is None:
        dX_out = torch.empty((K, N), dtype=torch.float32, device=dY.device)
    if dval_out is None:
        dval_out = torch.empty(max(nnz, 1), dtype=torch.float32, device=dY.device)[:nnz]
    fn = "spmm_csr_t_sddmm_bf16_planned" if _is_bf16(dY) else "spmm_csr_t_sddmm_f32_planned"
    t = C.byref(Tune(*tune)) if tune else None
    call(fn, plan.p, M, K, N, nnz, _p(rowptr), _p(colind), _p(val), _p(dY), dY.stride(0),
         _p(X), X.stride(0), _p(dX_out), dX_out.stride(0), _p(dval_out), t, _stream(stream))
    return dX_out, dval_out


def sddmm_csr(h: Handle, rowptr, colind, dY, X, out=None, plan: CsrPlan | None = None,
              stream=None):
    """dval[e] = dY[row(e)] . X[col(e)] at the stored positions (FP32 out).
    With `plan` (a CsrPlan of rowptr) the rows run on its schedule."""
    M = rowptr.numel() - 1
    K, N = X.shape
    nnz = colind.numel()
    if out is None:
        out = torch.empty(max(nnz, 1), dtype=torch.float32, device=X.device)[:nnz]
    fn = "sddmm_csr_bf16" if _is_bf16(X) else "sddmm_csr_f32"
    first = h.h
    if plan is not None:
        fn += "_planned"
        first = plan.p
    call(fn, first, M, K, N, nnz, _p(rowptr), _p(colind), _p(dY), dY.stride(0), _p(X),
         X.stride(0), _p(out), _stream(stream))
    return out


def read_mtx(path: str):
    """Matrix Market file -> host CSR (numpy), the reference harness's
    ingest (spmm_mtx_read_info / spmm_mtx_read_csr of the C-ABI). Returns
    (rowptr, colind, val, K, kind)."""
    import numpy as np

    M, K, nnz, kind = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    p = str(path).encode()
    call("spmm_mtx_read_info", p, C.byref(M), C.byref(K), C.byref(nnz), C.byref(kind))
    rowptr = np.empty(M.value + 1, np.int32)
    colind = np.empty(max(nnz.value, 1), np.int32)[: nnz.value]
    val = np.empty(max(nnz.value, 1), np.float32)[: nnz.value]
    call("spmm_mtx_read_csr", p, M.value, K.value, nnz.value, rowptr.ctypes.data,
         colind.ctypes.data, val.ctypes.data)
    return rowptr, colind, val, K.value, kind.value


def to_bf16_bits(x: torch.Tensor) -> torch.Tensor:
    """FP32 -> BF16 (round to nearest even) as a uint16 view."""
    return x.to(torch.bfloat16).view(torch.uint16)
