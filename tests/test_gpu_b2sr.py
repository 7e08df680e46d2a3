"""B2SR parity on the GPU: device csr2b2sr bit-exact against the oracle and
the hand-derived vectors; FBF SpMM against FP64 (<= 1e-5 relative) and,
with +-1 B, bit-exact against integer CSR SpMM (the reference's known
answer, test_spmm_full.cu:412-431); the sign-packed popcount path and the
dense packers bit-exact."""
import json
import os

import numpy as np
import pytest
import torch

import oracle
from spmm_eval_amd import api, graphs
from spmm_eval_amd._lib import SpmmError, call

pytestmark = pytest.mark.gpu
HAND = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "b2sr_hand.json")))


def dev_csr(g, dev):
    return torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)


def np_words(t):
    return t.cpu().numpy() if t.dtype != torch.uint16 else t.view(torch.int16).cpu().numpy().view(np.uint16)


def np_u32(t):
    return t.view(torch.int32).cpu().numpy().view(np.uint32)


def assert_b2sr_equal(dev_out, ref):
    brp, bci, words = dev_out
    np.testing.assert_array_equal(brp.cpu().numpy(), ref[0])
    np.testing.assert_array_equal(bci.cpu().numpy(), ref[1])
    signed = {torch.uint32: (torch.int32, np.uint32), torch.uint64: (torch.int64, np.uint64)}
    if words.dtype in signed:
        st, nt = signed[words.dtype]
        w = words.view(st).cpu().numpy().view(nt)
    else:
        w = np_words(words)
    np.testing.assert_array_equal(w, ref[2])


@pytest.mark.parametrize("T", [4, 8])
def test_hand_vectors(handle, dev, T):
    rp = torch.tensor(HAND["rowptr"], dtype=torch.int32, device=dev)
    ci = torch.tensor(HAND["colind"], dtype=torch.int32, device=dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, HAND["M"], HAND["K"], T)
    exp = HAND[f"T{T}"]
    assert brp.cpu().tolist() == exp["bsrRowPtr"]
    assert bci.cpu().tolist() == exp["bsrColInd"]
    assert words.cpu().tolist() == exp["bsrVal"]
    B = torch.tensor(HAND["B"], dtype=torch.float32, device=dev)
    Bp = torch.zeros(8, 4, device=dev)
    Bp[:, :2] = B  # N=4 keeps the vector kernel; N=2 the generic one
    C = api.spmm_b2sr(handle, T, HAND["M"], brp, bci, words, Bp)
    assert C[:, :2].cpu().tolist() == HAND["C"]
    C2 = api.spmm_b2sr(handle, T, HAND["M"], brp, bci, words, B.contiguous())
    assert C2.cpu().tolist() == HAND["C"]


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("shape", [(1, 1), (37, 101), (300, 300), (1000, 777)])
def test_csr2b2sr_random(handle, dev, T, shape):
    M, K = shape
    g = graphs.random_csr(M, K, 0.03, seed=M + T, long_rows=2, long_len=K)
    ref = oracle.csr2b2sr(g.rowptr, g.colind, T)
    rp, ci = dev_csr(g, dev)
    assert_b2sr_equal(api.csr2b2sr(handle, rp, ci, M, K, T), ref)


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_csr2b2sr_workload(handle, dev, T):
    g = graphs.workload_b2sr_graph()
    ref = oracle.csr2b2sr(g.rowptr, g.colind, T)
    rp, ci = dev_csr(g, dev)
    out = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    assert_b2sr_equal(out, ref)


def test_csr2b2sr_rejects_unsorted(handle, dev):
    rp = torch.tensor([0, 2], dtype=torch.int32, device=dev)
    ci = torch.tensor([5, 1], dtype=torch.int32, device=dev)
    with pytest.raises(SpmmError) as ei:
        api.csr2b2sr(handle, rp, ci, 1, 8, 4)
    assert ei.value.status == 1
    ci = torch.tensor([1, 9], dtype=torch.int32, device=dev)  # col >= K
    with pytest.raises(SpmmError):
        api.csr2b2sr(handle, rp, ci, 1, 8, 4)


@pytest.mark.parametrize("T", [2, 4, 16, 64])
def test_csr2b2sr_hub_rows_repeats_and_order(handle, dev, T):
    """Block rows of more than 64 entries take the bitmap path: rows with
    repeated columns (kept by the reference's reader) convert like the
    oracle; an unsorted long row is rejected like an unsorted short one."""
    r = np.random.default_rng(T)
    rows = []
    for i in range(3 * T):
        n = 400 if i % T == 1 else int(r.integers(0, 6))
        c = np.sort(r.integers(0, 3000, n)).astype(np.int32)  # repeats possible
        rows.append(c)
    rp = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    brp, bci, words = api.csr2b2sr(handle, torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev),
                                   len(rows), 3000, T)
    assert_b2sr_equal((brp, bci, words), oracle.csr2b2sr(rp, ci, T))
    bad = ci.copy()
    s0, s1 = rp[1], rp[2]
    bad[s0:s1] = bad[s0:s1][::-1]  # a long row in descending order
    with pytest.raises(SpmmError):
        api.csr2b2sr(handle, torch.from_numpy(rp).to(dev), torch.from_numpy(bad).to(dev),
                     len(rows), 3000, T)


@pytest.fixture(scope="module")
def b2sr_graph():
    return graphs.workload_b2sr_graph()


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("N", [32, 64, 128])
def test_fbf_workload(handle, dev, b2sr_graph, T, N):
    g = b2sr_graph
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    # known answer: +-1 B, exact
    S = np.random.default_rng(N + T).choice([-1, 1], size=(g.K, N)).astype(np.int32)
    C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(S.astype(np.float32)).to(dev))
    np.testing.assert_array_equal(C.cpu().numpy(), oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.float32))
    plan = api.B2srPlan(handle, T, g.M, brp)
    Cp = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(S.astype(np.float32)).to(dev),
                       plan=plan)
    assert torch.equal(Cp, C)  # planned schedule: same known answer
    # float B vs FP64, plain and planned
    B = graphs.uniform_pm1((g.K, N), seed=N * T)
    Cf = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(B).to(dev)).cpu().numpy()
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    assert np.linalg.norm(Cf - ref) / np.linalg.norm(ref) <= 1e-5
    Cfp = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(B).to(dev), plan=plan).cpu().numpy()
    assert np.linalg.norm(Cfp - ref) / np.linalg.norm(ref) <= 1e-5
    # BF16 B
    Bh = torch.from_numpy(B).to(dev).to(torch.bfloat16).view(torch.uint16)
    Ch = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bh).cpu().numpy()
    refq = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, np_words(Bh))
    assert np.linalg.norm(Ch - refq) / np.linalg.norm(refq) <= 1e-5
    Chp = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bh, plan=plan).cpu().numpy()
    assert np.linalg.norm(Chp - refq) / np.linalg.norm(refq) <= 1e-5


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("N", [1, 7, 100])
def test_fbf_odd_widths(handle, dev, T, N):
    g = graphs.random_csr(333, 444, 0.02, seed=N, long_rows=1, long_len=444)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    B = graphs.uniform_pm1((g.K, N), seed=N)
    C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(B).to(dev)).cpu().numpy()
    ref = oracle.b2sr_spmm_f64(T, g.M, *oracle.csr2b2sr(g.rowptr, g.colind, T), B)
    assert np.linalg.norm(C - ref) / max(np.linalg.norm(ref), 1e-30) <= 1e-5
    plan = api.B2srPlan(handle, T, g.M, brp)
    Cp = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(B).to(dev), plan=plan).cpu().numpy()
    assert np.linalg.norm(Cp - ref) / max(np.linalg.norm(ref), 1e-30) <= 1e-5


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("ldb,bf16", [(72, False), (70, False), (136, True), (132, True)])
def test_fbf_strided_operands(handle, dev, T, ldb, bf16):
    """B and C as column slices of wider arrays: an aligned leading dimension
    keeps the vector kernels (ldb 72 FP32, 136 BF16), an unaligned one runs
    the narrow kernel (70, 132); C's columns past N stay untouched; drop-in
    and planned forms; +-1 B exact."""
    N = 64
    g = graphs.random_csr(300, 500, 0.03, seed=T + ldb, long_rows=2, long_len=500)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    S = np.random.default_rng(T).choice([-1, 1], size=(g.K, ldb)).astype(np.float32)
    Bfull = torch.from_numpy(S).to(dev)
    if bf16:
        Bfull = api.to_bf16_bits(Bfull)
    B = Bfull[:, :N]
    exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S[:, :N].astype(np.int32)).astype(np.float32)
    plan = api.B2srPlan(handle, T, g.M, brp)
    for p in (None, plan):
        Cfull = torch.full((g.M, N + 12), 5.0, device=dev)
        api.spmm_b2sr(handle, T, g.M, brp, bci, words, B, C_out=Cfull[:, :N], plan=p)
        np.testing.assert_array_equal(Cfull[:, :N].cpu().numpy(), exact)
        assert (Cfull[:, N:] == 5.0).all()


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("N", [32, 64, 70, 128])
def test_sign_packed_popcount(handle, dev, b2sr_graph, T, N):
    g = b2sr_graph
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    X = graphs.uniform_pm1((g.K, N), seed=T + N)
    X[::97, 0] = -0.0
    Xd = torch.from_numpy(X).to(dev)
    Bw = api.pack_sign_rows(Xd)
    np.testing.assert_array_equal(np_u32(Bw), oracle.pack_sign_rows(X))
    C = api.spmm_b2sr_sign(handle, T, g.M, g.K, N, brp, bci, words, Bw).cpu().numpy()
    S = np.where(X >= 0, 1, -1).astype(np.int32)
    exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.int32)
    np.testing.assert_array_equal(C, exact)
    if N in (64, 70):  # without a handle: the per-row popcount kernel, same integers
        Cn = torch.empty((g.M, N), dtype=torch.int32, device=dev)
        call("spmm_b2sr_signpacked_i32", None, T, g.M, g.K, N, api._p(brp), api._p(bci),
             api._p(words), api._p(Bw), Bw.stride(0), api._p(Cn), Cn.stride(0), None)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(Cn.cpu().numpy(), exact)


@pytest.mark.parametrize("H,W", [(1, 1), (5, 31), (33, 32), (100, 100), (64, 129)])
def test_dense_packers(dev, H, W):
    X = graphs.uniform_pm1((H, W), seed=H * W)
    X[0, 0] = 0.0
    Xd = torch.from_numpy(X).to(dev)
    np.testing.assert_array_equal(np_u32(api.pack_sign_rows(Xd)), oracle.pack_sign_rows(X))
    np.testing.assert_array_equal(np_u32(api.pack_bin_rows(Xd)), oracle.pack_bin_rows(X))
    cols = api.pack_sign_cols(Xd)
    np.testing.assert_array_equal(np_u32(cols), oracle.pack_sign_cols(X))
    if H % 32 == 0:
        np.testing.assert_array_equal(api.unpack_cols(cols, H, W).cpu().numpy(),
                                      oracle.unpack_cols(np_u32(cols), H, W))
    sign = HAND["sign_pack"]
    Xh = torch.tensor(sign["X"], dtype=torch.float32, device=dev)
    assert np_u32(api.pack_sign_rows(Xh)).reshape(-1).tolist() == sign["sign"]
    assert np_u32(api.pack_bin_rows(Xh)).reshape(-1).tolist() == sign["bin"]


@pytest.mark.parametrize("T", [2, 4, 8, 32, 64])
@pytest.mark.parametrize("N", [64, 512])
def test_fbf_long_block_rows_repeated(handle, dev, T, N):
    """Block rows long enough to go through the long-row list (N=512 adds a
    second column tile), three calls in a row with another matrix between
    them: the list counters must come back to zero after every call."""
    g = graphs.random_csr(64, 8192, 0.002, seed=T + N, long_rows=3, long_len=3000)
    other = graphs.random_csr(300, 8192, 0.01, seed=5, long_rows=2, long_len=4000)
    B = graphs.uniform_pm1((8192, N), seed=N)
    Bd = torch.from_numpy(B).to(dev)
    outs = []
    for rep in range(3):
        for mat in ((g, other) if rep == 1 else (g,)):
            rp, ci = dev_csr(mat, dev)
            brp, bci, words = api.csr2b2sr(handle, rp, ci, mat.M, mat.K, T)
            C = api.spmm_b2sr(handle, T, mat.M, brp, bci, words, Bd).cpu().numpy()
            ref = oracle.csr_spmm_f64(mat.rowptr, mat.colind, np.ones(mat.nnz, np.float32), B)
            assert np.linalg.norm(C - ref) / np.linalg.norm(ref) <= 1e-5, (rep, mat.M)
            if mat is g:
                outs.append(C)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_fbf_dropin_unaligned_words(handle, dev, T):
    """bsrVal one word off the planned kernels' alignment (T bytes for T <= 8,
    16 above): the drop-in takes its row-per-subgroup kernels instead; +-1 B
    exact against integer CSR SpMM and equal to the aligned call."""
    N = 64
    g = graphs.random_csr(400, 3000, 0.01, seed=T, long_rows=3, long_len=2500)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    esz = words.element_size()
    raw = torch.zeros(words.numel() * esz + esz, dtype=torch.uint8, device=dev)
    raw[esz:] = words.view(torch.uint8)
    off = raw[esz:].view(words.dtype)  # one word further on
    assert off.data_ptr() % (T if T <= 8 else 16) != 0
    S = np.random.default_rng(T).choice([-1, 1], size=(g.K, N)).astype(np.float32)
    Bd = torch.from_numpy(S).to(dev)
    exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S.astype(np.int32)).astype(np.float32)
    C_off = api.spmm_b2sr(handle, T, g.M, brp, bci, off, Bd)
    np.testing.assert_array_equal(C_off.cpu().numpy(), exact)
    C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd)
    assert torch.equal(C, C_off)


@pytest.mark.parametrize("T", [2, 4, 8, 16])
def test_fbf_dropin_graph_replays(handle, dev, b2sr_graph, T):
    """The plan-free form captured in a hipGraph (two calls) and replayed
    three times, with an eager call on another matrix between replays: its
    device counters and class lists must be back to their idle state after
    every call, so every replay equals the eager result."""
    g = b2sr_graph
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    other = graphs.random_csr(300, 8192, 0.01, seed=7, long_rows=2, long_len=4000)
    orp, oci = dev_csr(other, dev)
    obrp, obci, owords = api.csr2b2sr(handle, orp, oci, other.M, other.K, T)
    Bo = torch.from_numpy(graphs.uniform_pm1((other.K, 64), seed=3)).to(dev)
    Bd = torch.from_numpy(graphs.uniform_pm1((g.K, 64), seed=T)).to(dev)
    C = torch.empty(g.M, 64, device=dev)
    api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, C_out=C)
    torch.cuda.synchronize()
    ref = C.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, C_out=C)
        s.synchronize()
        with torch.cuda.graph(graph, stream=s):
            api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, C_out=C)
            api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, C_out=C)
    torch.cuda.synchronize()
    o_ref = api.spmm_b2sr(handle, T, other.M, obrp, obci, owords, Bo)
    for rep in range(3):
        C.fill_(0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, ref), rep
        o = api.spmm_b2sr(handle, T, other.M, obrp, obci, owords, Bo)
        torch.cuda.synchronize()
        assert torch.equal(o, o_ref), rep


@pytest.mark.parametrize("T", [2, 4, 16, 64])
def test_fbf_planned_hub_rows_and_mismatch(handle, dev, T):
    """Rows of 3000-4000 entries reach the one-workgroup-per-row class; a
    plan for another T or M, or a CSR plan, is refused."""
    g = graphs.random_csr(300, 8192, 0.01, seed=T, long_rows=4, long_len=4000)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    B = graphs.uniform_pm1((g.K, 64), seed=3)
    plan = api.B2srPlan(handle, T, g.M, brp)
    C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, torch.from_numpy(B).to(dev), plan=plan).cpu().numpy()
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, np.ones(g.nnz, np.float32), B)
    assert np.linalg.norm(C - ref) / np.linalg.norm(ref) <= 1e-5
    Bd = torch.from_numpy(B).to(dev)
    other_T = 8 if T == 4 else 4
    with pytest.raises(SpmmError) as ei:
        api.spmm_b2sr(handle, other_T, g.M, brp, bci, words, Bd, plan=plan)
    assert ei.value.status == 5
    csr_plan = api.CsrPlan(handle, rp, g.nnz)
    with pytest.raises(SpmmError):
        api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, plan=csr_plan)
    with pytest.raises(SpmmError):
        api.spmm_csr(handle, rp, ci, torch.ones(g.nnz, device=dev), Bd, plan=plan)


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("N,bf16", [(64, False), (32, False), (256, True)])
def test_fbf_planned_dense_tiles(handle, dev, T, N, bf16):
    """Nearly full tiles: a chunk's set bits overflow one list window (the
    list is filled and drained several times per chunk); M not a multiple
    of T leaves a partial last block row. +-1 B: exact integer sums."""
    g = graphs.random_csr(130, 3000, 0.85, seed=T + N)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    S = np.random.default_rng(T).choice([-1, 1], size=(g.K, N)).astype(np.float32)
    Sd = torch.from_numpy(S).to(dev)
    if bf16:
        Sd = api.to_bf16_bits(Sd)
    plan = api.B2srPlan(handle, T, g.M, brp)
    C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Sd, plan=plan).cpu().numpy()
    exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S.astype(np.int32)).astype(np.float32)
    np.testing.assert_array_equal(C, exact)


def _mixed_rounds_graph(M, K, dense_cols, dense_p, sparse_per_row, seed):
    """Rows with a dense column block (tiles far denser than one bit each)
    followed by scattered columns: a block row's first round(s) hold more
    bits than the round-list slots, later rounds one or two a tile."""
    r = np.random.default_rng(seed)
    rows = []
    for i in range(M):
        d = np.flatnonzero(r.random(dense_cols) < dense_p) if i % 5 != 2 else np.zeros(0, np.int64)
        s = r.choice(np.arange(dense_cols, K), size=sparse_per_row if i % 9 else 0, replace=False)
        rows.append(np.unique(np.concatenate([d, s])).astype(np.int32))
    rowptr = np.zeros(M + 1, np.int32)
    rowptr[1:] = np.cumsum([len(x) for x in rows])
    colind = np.concatenate(rows).astype(np.int32)
    return graphs.Csr(rowptr, colind, np.ones(len(colind), np.float32), M, K)


@pytest.mark.parametrize("T", [16, 32, 64])
def test_fbf_planned_wide_rounds(handle, dev, T):
    """Planned wide-tile B2SR over block rows of many rounds (more than the
    16 whose row counts the gather stages at once), with rounds whose lists
    overflow their slots (walked from the words) between rounds merged into
    multi-round phases: +-1 B exact against integer CSR, float and BF16 B
    within 1e-5 of FP64, drop-in and planned forms."""
    g = _mixed_rounds_graph(150, 140000, 3 * T, 0.6, 150, seed=T)
    print(f"wide rounds T={T} nnz={g.nnz}", flush=True)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    plan = api.B2srPlan(handle, T, g.M, brp)
    for N in (32, 64):
        S = np.random.default_rng(N + T).choice([-1, 1], size=(g.K, N)).astype(np.int32)
        exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.float32)
        Sd = torch.from_numpy(S.astype(np.float32)).to(dev)
        for pl in (plan, None):
            C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Sd, plan=pl).cpu().numpy()
            np.testing.assert_array_equal(C, exact, err_msg=f"T={T} N={N} planned={pl is not None}")
        B = graphs.uniform_pm1((g.K, N), seed=N * T + 1)
        Bd = torch.from_numpy(B).to(dev)
        ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
        Cf = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, plan=plan).cpu().numpy()
        assert np.linalg.norm(Cf - ref) / np.linalg.norm(ref) <= 1e-5, (T, N)
        Bh = api.to_bf16_bits(Bd)
        refq = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, np_words(Bh))
        Ch = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bh, plan=plan).cpu().numpy()
        assert np.linalg.norm(Ch - refq) / np.linalg.norm(refq) <= 1e-5, (T, N)
        # a second call gives the same bits (the plan's round lists are rewritten)
        Cf2 = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, plan=plan).cpu().numpy()
        np.testing.assert_array_equal(Cf2, Cf)
    plan.close()


def test_harness_flow_from_mtx(handle, dev, tmp_path):
    """The reference harness end to end (test_spmm_full.cu:99-125, 344-369,
    412-431): a Matrix Market pattern file -> CSR -> B2SR-4 -> FBF SpMM,
    compared exactly with the CSR SpMM of the file's values on +-1 B."""
    g = graphs.chung_lu(5000, 60000, seed=11)
    rows = np.repeat(np.arange(g.M), np.diff(g.rowptr))
    f = tmp_path / "g.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n"
                 f"{g.M} {g.K} {g.nnz}\n" +
                 "\n".join(f"{a + 1} {b + 1}" for a, b in zip(rows, g.colind)) + "\n")
    rowptr, colind, val, K, _ = api.read_mtx(str(f))
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (rowptr, colind, val))
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, K, 4)
    S = torch.from_numpy(np.random.default_rng(5).choice([-1.0, 1.0], size=(K, 64)).astype(np.float32)).to(dev)
    plan = api.B2srPlan(handle, 4, g.M, brp)
    C_b2sr = api.spmm_b2sr(handle, 4, g.M, brp, bci, words, S, plan=plan)
    C_csr = api.spmm_csr(handle, rp, ci, v, S)
    torch.cuda.synchronize()
    assert torch.equal(C_b2sr, C_csr)  # "TEST PASSED: 1"


def b2sr_candidates(T, N, elem):
    """The candidate lane mappings spmm_b2sr_autotune times (variant, lanes,
    rows_scale): block-row kernels at every power-of-two lane count dividing
    N's 16-byte vectors (at most 8 column tiles) x rows_scale, word lanes,
    and B staged in LDS for N <= 8 and T <= 8."""
    e16 = 16 // elem
    out = []
    if N % e16 == 0:
        lanes = N // e16
        G = 64
        while G >= 1:
            if lanes % G == 0 and G * 8 >= min(lanes, 64):
                for rs in ((1, 2, 4) if T <= 8 else (1,)):
                    out.append((0, G, rs))
            G //= 2
    out.append((1, 0, 1))
    if N <= 8 and T <= 8:
        out.append((2, 0, 1))
    return out


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_b2sr_every_tuner_candidate(handle, dev, T):
    """Every candidate of the B2SR autotuner, pinned on the plan: +-1 B bit-exact
    against integer CSR, uniform B within 1e-5 of FP64 (FP32 and BF16 B), on a
    power-law graph whose hub block rows take every row class."""
    g = graphs.chung_lu(6000, 120000, seed=T, max_deg=2500)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    plan = api.B2srPlan(handle, T, g.M, brp)
    for N in (1, 4, 8, 32, 64, 128):
        S = np.random.default_rng(N).choice([-1, 1], size=(g.K, N)).astype(np.int32)
        exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.float32)
        Sd = torch.from_numpy(S.astype(np.float32)).to(dev)
        B = graphs.uniform_pm1((g.K, N), seed=N + T)
        Bd = torch.from_numpy(B).to(dev)
        ref = oracle.csr_spmm_f64(g.rowptr, g.colind, np.ones(g.nnz, np.float32), B)
        Bh = api.to_bf16_bits(Bd)
        refh = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, np.ones(g.nnz, np.float32), np_words(Bh))
        for elem, BB, rr in ((4, Bd, ref), (2, Bh, refh)):
            for var, G, rs in b2sr_candidates(T, N, elem):
                plan.set_geometry(N, elem == 2, var, G, rs)
                C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, BB, plan=plan).cpu().numpy()
                err = np.linalg.norm(C - rr) / max(np.linalg.norm(rr), 1e-30)
                assert err <= 1e-5, (T, N, elem, var, G, rs, err)
                if elem == 4:
                    Cs = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Sd, plan=plan).cpu().numpy()
                    assert np.array_equal(Cs, exact), (T, N, var, G, rs)
    plan.close()


@pytest.mark.parametrize("T", [2, 4, 8])
def test_b2sr_column_slices(handle, dev, T):
    """Planned B2SR at 2, 3 and 7 column slices (pinned on the plan): each
    launch takes the tiles of one range of block columns and the later ones
    add into C. +-1 B exact against integer CSR; FP32 and BF16 B within 1e-5
    of FP64; a power-law graph whose hub block rows take every row class,
    with K not a multiple of T times the slice count."""
    g = graphs.chung_lu(6001, 120000, seed=T + 40, max_deg=2500)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    plan = api.B2srPlan(handle, T, g.M, brp)
    ones = np.ones(g.nnz, np.float32)
    for N in (32, 64):
        S = np.random.default_rng(N + T).choice([-1, 1], size=(g.K, N)).astype(np.int32)
        exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.float32)
        Sd = torch.from_numpy(S.astype(np.float32)).to(dev)
        B = graphs.uniform_pm1((g.K, N), seed=N + T + 1)
        Bd = torch.from_numpy(B).to(dev)
        ref = oracle.csr_spmm_f64(g.rowptr, g.colind, ones, B)
        Bh = api.to_bf16_bits(Bd)
        refh = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, ones, np_words(Bh))
        for sl in (2, 3, 7):
            print(f"slices T={T} N={N} S={sl}", flush=True)
            plan.set_geometry(N, False, 0, 0, 1, sl)
            Cs = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Sd, plan=plan).cpu().numpy()
            np.testing.assert_array_equal(Cs, exact, err_msg=f"T={T} N={N} S={sl}")
            C = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, plan=plan).cpu().numpy()
            assert np.linalg.norm(C - ref) / np.linalg.norm(ref) <= 1e-5, (T, N, sl)
            plan.set_geometry(N, True, 0, 0, 1, sl)
            Ch = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bh, plan=plan).cpu().numpy()
            assert np.linalg.norm(Ch - refh) / np.linalg.norm(refh) <= 1e-5, (T, N, sl)
    with pytest.raises(SpmmError):  # slices need variant 0
        plan.set_geometry(32, False, 1, 0, 1, 4)
    plan.close()


@pytest.mark.parametrize("T,N", [(4, 64), (4, 8), (16, 32), (8, 4)])
def test_b2sr_autotune_keeps_a_candidate(handle, dev, T, N):
    g = graphs.chung_lu(20000, 400000, seed=5)
    rp, ci = dev_csr(g, dev)
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    plan = api.B2srPlan(handle, T, g.M, brp)
    Bd = torch.from_numpy(graphs.uniform_pm1((g.K, N), seed=3)).to(dev)
    C = torch.empty((g.M, N), device=dev)
    var, G, rs, sl = plan.autotune(brp, bci, words, Bd, C)
    assert (var, G, rs) in b2sr_candidates(T, N, 4)
    assert sl == 1 or (var == 0 and T <= 8 and sl in (2, 4, 6, 8, 12, 16)), sl
    C1 = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, plan=plan)  # the tuned entry
    plan.set_geometry(N, False, var, G, rs, sl)
    C2 = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Bd, plan=plan)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, np.ones(g.nnz, np.float32), Bd.cpu().numpy())
    assert np.linalg.norm(C1.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-5
    plan.close()


def test_b2sr_plan_set_rejects(handle, dev):
    g = graphs.chung_lu(500, 4000, seed=2)
    rp, ci = dev_csr(g, dev)
    brp, _, _ = api.csr2b2sr(handle, rp, ci, g.M, g.K, 16)
    plan = api.B2srPlan(handle, 16, g.M, brp)
    for args in ((8, False, 2), (64, False, 3), (6, False, 0), (64, False, 0, 3), (64, False, 0, 4, 3)):
        with pytest.raises(SpmmError) as ei:
            plan.set_geometry(*args)
        assert ei.value.status == 1, args
    plan.close()
