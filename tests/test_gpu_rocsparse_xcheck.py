"""Cross-check against an independent implementation on the same GPU:
torch.sparse on ROCm runs hipSPARSE/rocSPARSE. Both sides compute in FP32,
so the bar is the FP32 tolerance (relative Frobenius <= 1e-5), not bits.
A torch build whose sparse backend lacks an operation skips that case."""
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


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_csr2b2sr_vs_rocsparse_bsr(handle, dev, T):
    """Our device conversion against rocSPARSE csr2bsr (torch's
    to_sparse_bsr on the GPU) on the same 0/1 matrix, padded to whole
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
