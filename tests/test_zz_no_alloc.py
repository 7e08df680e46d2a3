"""Allocation-free B2SR FBF SpMM (spmm_b2sr_f32_noalloc, static device
scratch) against the handle form spmm_b2sr_f32 on the same operands: two
calls in a row per tile size, the second on a smaller graph, so the scratch
lists and counters left by the first call must not leak into the second.
Both forms run the same kernels, so the outputs must be bitwise equal; the
FP64 oracle bounds both at 1e-5 relative."""
import numpy as np
import pytest
import torch

import oracle
from spmm_eval_amd import api, graphs
from spmm_eval_amd._lib import SpmmError, call

pytestmark = pytest.mark.gpu


def _noalloc(T, M, brp, bci, words, B, C):
    call("spmm_b2sr_f32_noalloc", T, M, B.shape[0], B.shape[1], api._p(brp), api._p(bci),
         api._p(words), api._p(B), B.stride(0), api._p(C), C.stride(0), api._stream())


@pytest.mark.parametrize("T,N", [(2, 64), (4, 64), (8, 32), (16, 64), (32, 32), (64, 64), (4, 3)])
def test_noalloc_matches_handle_form(handle, dev, T, N):
    for n, nnz, seed in ((20000, 200000, 11), (3000, 24000, 12)):
        print(f"noalloc T={T} N={N} n={n} seed={seed}", flush=True)
        g = graphs.chung_lu(n, nnz, seed=seed)
        rp = torch.from_numpy(g.rowptr).to(dev)
        ci = torch.from_numpy(g.colind).to(dev)
        brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
        Bh = graphs.uniform_pm1((g.K, N), seed=seed)
        B = torch.from_numpy(Bh).to(dev)
        ref = api.spmm_b2sr(handle, T, g.M, brp, bci, words, B)
        C = torch.full((g.M, N), float("nan"), device=dev)
        _noalloc(T, g.M, brp, bci, words, B, C)
        torch.cuda.synchronize()
        assert torch.equal(C, ref), f"T={T} N={N} n={n}: noalloc differs from spmm_b2sr_f32"
        exact = oracle.csr_spmm_f64(g.rowptr, g.colind, np.ones(g.nnz), Bh)
        err = np.linalg.norm(C.cpu().numpy() - exact) / np.linalg.norm(exact)
        assert err <= 1e-5, f"relative error {err}"


def test_noalloc_rejects_too_many_block_rows(dev):
    M = 4 * (1 << 18) + 1  # one block row beyond the static scratch
    with pytest.raises(SpmmError) as e:
        call("spmm_b2sr_f32_noalloc", 4, M, 4, 4, None, None, None, None, 4, None, 4, None)
    assert "262144" in str(e.value)


@pytest.mark.parametrize("T", [4, 16])
def test_noalloc_graph_replay(handle, dev, T):
    """Captured in a hipGraph (the caller this form is for) and replayed
    three times, with an eager allocation-free call on another graph between
    replays sharing the same scratch: every replay equals the handle form."""
    g = graphs.chung_lu(12000, 120000, seed=21)
    other = graphs.random_csr(300, 8192, 0.01, seed=7, long_rows=2, long_len=4000)
    ops = []
    for h in (g, other):
        rp = torch.from_numpy(h.rowptr).to(dev)
        ci = torch.from_numpy(h.colind).to(dev)
        brp, bci, words = api.csr2b2sr(handle, rp, ci, h.M, h.K, T)
        B = torch.from_numpy(graphs.uniform_pm1((h.K, 64), seed=T)).to(dev)
        ops.append((h.M, brp, bci, words, B, api.spmm_b2sr(handle, T, h.M, brp, bci, words, B)))
    (M, brp, bci, words, B, ref), (Mo, obrp, obci, owords, Bo, o_ref) = ops
    C = torch.empty(M, 64, device=dev)
    Co = torch.empty(Mo, 64, device=dev)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        _noalloc(T, M, brp, bci, words, B, C)  # first call on the device outside capture
        s.synchronize()
        with torch.cuda.graph(graph, stream=s):
            _noalloc(T, M, brp, bci, words, B, C)
    torch.cuda.synchronize()
    for rep in range(3):
        C.fill_(0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, ref), rep
        _noalloc(T, Mo, obrp, obci, owords, Bo, Co)
        torch.cuda.synchronize()
        assert torch.equal(Co, o_ref), rep
