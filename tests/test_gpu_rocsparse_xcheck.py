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
