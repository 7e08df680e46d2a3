"""Cross-checks against independent implementations.

On the GPU, through rocSPARSE: torch.sparse CSR @ dense and sampled_addmm
run rocsparse::csrmmnn_row_split_kernel and rocsparse::sddmm_csx_kernel on
ROCm (kernel trace in profiles/rocsparse_xcheck.md); csr2bsr is called in
librocsparse directly (rocsparse_csr2bsr_nnz + rocsparse_scsr2bsr, ctypes,
on torch's stream), its float blocks packed to words on the host. Products
are FP32 on both sides, so their bar is the FP32 tolerance (relative
Frobenius <= 1e-5), not bits; the conversion must match word for word.

On the CPU: torch's to_sparse_bsr (it converts on the host even for device
tensors in this torch build, as its warning says) as a second, independent
CSR -> BSR for the same comparison. A build lacking an operation skips that
case."""
import ctypes as C

import numpy as np
import pytest
import torch

from spmm_eval_amd import api, graphs

pytestmark = pytest.mark.gpu


def rel(a, b):
    n = torch.linalg.norm(b)
    return float(torch.linalg.norm(a - b) / (n if n > 0 else 1.0))


@pytest.fixture(scope="module")
def graph():
    g = graphs.chung_lu(20000, 400000, seed=21)
    g.val = graphs.row_normalised_uniform(g, 21)
    return g


@pytest.mark.parametrize("N", [32, 64, 128])
def test_csr_spmm_vs_rocsparse(handle, dev, graph, N):
    g = graph
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    B = torch.from_numpy(graphs.uniform_pm1((g.K, N), seed=N)).to(dev)
    A = torch.sparse_csr_tensor(rp.long(), ci.long(), v, size=(g.M, g.K))
    try:
        ref = (A @ B).to_dense() if (A @ B).is_sparse else A @ B
    except (RuntimeError, NotImplementedError) as e:
        pytest.skip(f"torch.sparse CSR @ dense unavailable here: {e}")
    C = api.spmm_csr(handle, rp, ci, v, B)
    assert rel(C, ref) <= 1e-5
    Ct = api.spmm_csr_t(handle, rp, ci, v, B[: g.M], g.K)
    reft = A.to_dense().T @ B[: g.M]  # dense transpose product as the A^T reference
    assert rel(Ct, reft) <= 1e-5


def test_sddmm_vs_rocsparse(handle, dev, graph):
    g = graph
    N = 64
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    dY = torch.from_numpy(graphs.normal((g.M, N), 4, 1)).to(dev)
    X = torch.from_numpy(graphs.normal((g.K, N), 4, 2)).to(dev)
    A = torch.sparse_csr_tensor(rp.long(), ci.long(), torch.ones_like(v), size=(g.M, g.K))
    try:
        S = torch.sparse.sampled_addmm(A, dY, X.T.contiguous(), beta=0.0, alpha=1.0)
    except (RuntimeError, NotImplementedError) as e:
        pytest.skip(f"torch.sparse.sampled_addmm unavailable here: {e}")
    ref = S.values()
    plan = api.CsrPlan(handle, rp, g.nnz)
    out = api.sddmm_csr(handle, rp, ci, dY, X, plan=plan)
    assert rel(out, ref) <= 1e-5


def pack_blocks(vals: np.ndarray, T: int) -> np.ndarray:
    """[nb][T][T] block values -> [nb*T] words, bit T-1-c of word r set iff
    block[r][c] != 0 (numpy.packbits, big-endian within each row)."""
    nz = (vals != 0).astype(np.uint8)
    if T < 8:
        return (np.packbits(nz, axis=2)[:, :, 0] >> (8 - T)).reshape(-1)
    b = np.packbits(nz, axis=2)  # [nb][T][T/8] bytes, most significant first
    return b.reshape(-1, T // 8).view(f">u{T // 8}").reshape(-1).astype(f"<u{T // 8}")


def _rocsparse():
    for name in ("librocsparse.so", "/opt/rocm/lib/librocsparse.so"):
        try:
            L = C.CDLL(name)
            break
        except OSError:
            L = None
    if L is None:
        return None
    vp, i = C.c_void_p, C.c_int
    L.rocsparse_create_handle.argtypes = [C.POINTER(vp)]
    L.rocsparse_destroy_handle.argtypes = [vp]
    L.rocsparse_set_stream.argtypes = [vp, vp]
    L.rocsparse_create_mat_descr.argtypes = [C.POINTER(vp)]
    L.rocsparse_destroy_mat_descr.argtypes = [vp]
    L.rocsparse_csr2bsr_nnz.argtypes = [vp, i, i, i, vp, vp, vp, i, vp, vp, C.POINTER(i)]
    L.rocsparse_scsr2bsr.argtypes = [vp, i, i, i, vp, vp, vp, vp, i, vp, vp, vp, vp]
    for f in ("rocsparse_create_handle", "rocsparse_destroy_handle", "rocsparse_set_stream",
              "rocsparse_create_mat_descr", "rocsparse_destroy_mat_descr",
              "rocsparse_csr2bsr_nnz", "rocsparse_scsr2bsr"):
        getattr(L, f).restype = i
    return L


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_csr2b2sr_vs_rocsparse_csr2bsr(handle, dev, T):
    """Our device conversion against rocSPARSE's csr2bsr on the GPU (block
    rows by row, row-major blocks), on the same 0/1 matrix: block row
    pointers, block columns and every tile word equal."""
    L = _rocsparse()
    if L is None:
        pytest.skip("librocsparse not loadable")
    g = graphs.random_csr(700, 900, 0.01, seed=T, long_rows=2, long_len=900)
    rp, ci = torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)
    ones = torch.ones(max(g.nnz, 1), device=dev)
    MB = -(-g.M // T)
    h, da, db = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert L.rocsparse_create_handle(C.byref(h)) == 0
    assert L.rocsparse_set_stream(h, C.c_void_p(torch.cuda.current_stream().cuda_stream)) == 0
    assert L.rocsparse_create_mat_descr(C.byref(da)) == 0
    assert L.rocsparse_create_mat_descr(C.byref(db)) == 0
    try:
        brp = torch.empty(MB + 1, dtype=torch.int32, device=dev)
        nb = C.c_int()
        assert L.rocsparse_csr2bsr_nnz(h, 0, g.M, g.K, da, C.c_void_p(rp.data_ptr()),
                                       C.c_void_p(ci.data_ptr()), T, db, C.c_void_p(brp.data_ptr()),
                                       C.byref(nb)) == 0
        bval = torch.empty(max(nb.value, 1) * T * T, device=dev)
        bci = torch.empty(max(nb.value, 1), dtype=torch.int32, device=dev)
        assert L.rocsparse_scsr2bsr(h, 0, g.M, g.K, da, C.c_void_p(ones.data_ptr()),
                                    C.c_void_p(rp.data_ptr()), C.c_void_p(ci.data_ptr()), T, db,
                                    C.c_void_p(bval.data_ptr()), C.c_void_p(brp.data_ptr()),
                                    C.c_void_p(bci.data_ptr())) == 0
        torch.cuda.synchronize()
    finally:
        L.rocsparse_destroy_mat_descr(da)
        L.rocsparse_destroy_mat_descr(db)
        L.rocsparse_destroy_handle(h)
    ours = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    assert ours[0].cpu().tolist() == brp.cpu().tolist()
    assert ours[1].cpu().tolist() == bci[: nb.value].cpu().tolist()
    ref = pack_blocks(bval[: nb.value * T * T].cpu().numpy().reshape(nb.value, T, T), T)
    signed = {torch.uint16: torch.int16, torch.uint32: torch.int32, torch.uint64: torch.int64}
    w = ours[2].view(signed[ours[2].dtype]) if ours[2].dtype in signed else ours[2]
    np.testing.assert_array_equal(w.cpu().numpy().view(ref.dtype), ref)


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_csr2b2sr_vs_torch_cpu_bsr(handle, dev, T):
    """Our device conversion against torch's to_sparse_bsr (a host
    conversion in this torch build) on the same 0/1 matrix, padded to whole
    blocks: block row pointers, block columns and every tile word equal."""
    g = graphs.random_csr(700, 900, 0.01, seed=T, long_rows=2, long_len=900)
    MB, KB = -(-g.M // T), -(-g.K // T)
    rp = np.concatenate([g.rowptr, np.full(MB * T - g.M, g.rowptr[-1], np.int32)])
    A = torch.sparse_csr_tensor(torch.from_numpy(rp).long().to(dev),
                                torch.from_numpy(g.colind).long().to(dev),
                                torch.ones(g.nnz, device=dev), size=(MB * T, KB * T))
    try:
        bsr = A.to_sparse_bsr((T, T))
        torch.cuda.synchronize()
    except (RuntimeError, NotImplementedError) as e:
        pytest.skip(f"torch.sparse CSR -> BSR unavailable on this device: {e}")
    brp, bci, words = api.csr2b2sr(handle, torch.from_numpy(g.rowptr).to(dev),
                                   torch.from_numpy(g.colind).to(dev), g.M, g.K, T)
    assert bsr.crow_indices().cpu().numpy().tolist() == brp.cpu().numpy().tolist()
    assert bsr.col_indices().cpu().numpy().tolist() == bci.cpu().numpy().tolist()
    ref = pack_blocks(bsr.values().cpu().numpy(), T)
    signed = {torch.uint16: torch.int16, torch.uint32: torch.int32, torch.uint64: torch.int64}
    w = words.view(signed[words.dtype]) if words.dtype in signed else words
    got = w.cpu().numpy().view(ref.dtype)
    np.testing.assert_array_equal(got, ref)
