"""CSR SpMM parity on the GPU through the C-ABI, against the FP64 oracle.

Tolerances (stated per the path's contract): FP32 B -> relative Frobenius
error <= 1e-5 vs FP64 and every element within nnz_row * 2^-23 * sum|a||b|
(+1e-30); BF16 B -> the same 1e-5 bound against FP64 of the BF16-rounded B
(kernel exactness), and <= 1e-2 against FP64 of the FP32 source (precision
form). Values 1 with +-1 B -> bit-exact (the reference's own known-answer
check, test_spmm_full.cu:412-431).
"""
import numpy as np
import pytest
import torch

import oracle
from spmm_eval_amd import api, graphs
from spmm_eval_amd._lib import SpmmError

pytestmark = pytest.mark.gpu


def to_dev(g, dev):
    return (torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev),
            torch.from_numpy(g.val).to(dev))


def rel_fro(C, ref):
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(C.astype(np.float64) - ref) / (n if n > 0 else 1.0))


def elementwise_bound(g, B):
    absA = np.abs(g.val).astype(np.float64)
    rows = np.repeat(np.arange(g.M), g.degrees())
    # sum_j |a_ij| |b_j| per row and column, scaled by deg * eps
    import scipy.sparse as sp
    A = sp.csr_matrix((absA, g.colind, g.rowptr), shape=(g.M, g.K))
    mag = A @ np.abs(B.astype(np.float64))
    deg = np.maximum(g.degrees(), 1)[:, None]
    return deg * 2.0 ** -23 * mag + 1e-30


def check_f32(g, B, C):
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    assert rel_fro(C, ref) <= 1e-5
    assert (np.abs(C - ref) <= elementwise_bound(g, B)).all()


CASES = [
    # (M, K, density, long_rows, long_len, N)
    (1, 1, 1.0, 0, 0, 1),
    (50, 40, 0.2, 0, 0, 3),       # N not a multiple of 4 -> narrow kernel
    (300, 200, 0.05, 2, 200, 4),
    (513, 700, 0.02, 1, 700, 8),
    (1000, 1000, 0.01, 3, 1000, 32),
    (777, 3000, 0.003, 2, 3000, 64),
    (400, 5000, 0.002, 3, 5000, 64),  # rows of 5000 -> one workgroup per row
    (256, 256, 0.05, 1, 256, 96),     # 96 = 3 column tiles of 32
    (300, 300, 0.03, 1, 300, 128),
    (200, 200, 0.05, 1, 200, 256),
    (100, 100, 0.05, 1, 100, 512),
]


@pytest.mark.parametrize("M,K,dens,nl,ll,N", CASES)
def test_csr_f32_random(handle, dev, M, K, dens, nl, ll, N):
    g = graphs.random_csr(M, K, dens, seed=M + N, long_rows=nl, long_len=ll)
    B = graphs.uniform_pm1((K, N), seed=N)
    rp, ci, v = to_dev(g, dev)
    C = api.spmm_csr(handle, rp, ci, v, torch.from_numpy(B).to(dev)).cpu().numpy()
    check_f32(g, B, C)


def test_csr_f32_workload_small_and_known_answer(handle, dev):
    g, B = graphs.workload_csr_small()
    rp, ci, v = to_dev(g, dev)
    Bd = torch.from_numpy(B).to(dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    C1 = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan)
    C2 = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan)
    C3 = api.spmm_csr(handle, rp, ci, v, Bd)  # drop-in form
    torch.cuda.synchronize()
    assert torch.equal(C1, C2) and torch.equal(C1, C3)  # deterministic
    check_f32(g, B, C1.cpu().numpy())
    # +-1 B with values 1: integer sums, exact (test_spmm_full.cu:412-431)
    S = np.random.default_rng(7).choice([-1, 1], size=B.shape).astype(np.int32)
    Cs = api.spmm_csr(handle, rp, ci, v, torch.from_numpy(S.astype(np.float32)).to(dev), plan=plan)
    exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S)
    np.testing.assert_array_equal(Cs.cpu().numpy(), exact.astype(np.float32))
    assert sum(plan.bins()) == g.M


@pytest.mark.parametrize("N", [32, 64, 128])
def test_csr_bf16(handle, dev, N):
    g = graphs.random_csr(2000, 2500, 0.004, seed=N, long_rows=2, long_len=2500)
    B = graphs.uniform_pm1((g.K, N), seed=N + 1)
    Bh = oracle.f32_to_bf16(B)
    rp, ci, v = to_dev(g, dev)
    Bd = torch.from_numpy(Bh.view(np.int16)).to(dev).view(torch.uint16)
    C = api.spmm_csr(handle, rp, ci, v, Bd).cpu().numpy()
    ref_q = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bh)
    assert rel_fro(C, ref_q) <= 1e-5
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    assert rel_fro(C, ref) <= 1e-2


@pytest.mark.parametrize("tune", [(4, 1, 1), (8, 2, 1), (16, 1, 2), (32, 2, 4), (64, 1, 1)])
def test_csr_tune_variants(handle, dev, tune):
    g = graphs.chung_lu(5000, 80000, seed=3)
    g.val = graphs.row_normalised_uniform(g, 3)
    B = graphs.uniform_pm1((g.K, 128), seed=5)
    rp, ci, v = to_dev(g, dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    C = api.spmm_csr(handle, rp, ci, v, torch.from_numpy(B).to(dev), plan=plan, tune=tune)
    check_f32(g, B, C.cpu().numpy())


NARROW = [
    # (M, K, density, long_rows, long_len, N): every row class of the narrow
    # kernel (subgroup, wave, workgroup rows of 3000-5000 entries) and each
    # column-lane width, including N > 64 (column chunks)
    (500, 5000, 0.002, 3, 5000, 1),
    (600, 3000, 0.004, 4, 3000, 2),
    (400, 5000, 0.003, 2, 5000, 5),
    (300, 2000, 0.01, 2, 2000, 7),
    (300, 1000, 0.01, 2, 1000, 17),
    (200, 3000, 0.005, 2, 3000, 33),
    (150, 800, 0.02, 1, 800, 100),
]


@pytest.mark.parametrize("M,K,dens,nl,ll,N", NARROW)
def test_csr_narrow_widths(handle, dev, M, K, dens, nl, ll, N):
    g = graphs.random_csr(M, K, dens, seed=3 * M + N, long_rows=nl, long_len=ll)
    B = graphs.uniform_pm1((K, N), seed=N + 1)
    rp, ci, v = to_dev(g, dev)
    Bd = torch.from_numpy(B).to(dev)
    C = api.spmm_csr(handle, rp, ci, v, Bd).cpu().numpy()
    check_f32(g, B, C)
    plan = api.CsrPlan(handle, rp, g.nnz)
    Cp = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan).cpu().numpy()
    assert np.array_equal(Cp, C)  # same schedule, same sums
    Bh = api.to_bf16_bits(Bd)
    Ch = api.spmm_csr(handle, rp, ci, v, Bh, plan=plan).cpu().numpy()
    refq = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bh.view(torch.int16).cpu().numpy().view(np.uint16))
    assert rel_fro(Ch, refq) <= 1e-5


@pytest.mark.parametrize("N", [1, 3])
def test_csr_narrow_workload_graph(handle, dev, N):
    """SpMV-shaped products on the 89k-row power-law graph (hub rows of
    ~900 entries): f32 bound, and +-1 B with unit values bit-exact."""
    g = graphs.workload_b2sr_graph()
    rp, ci, v = to_dev(g, dev)
    B = graphs.uniform_pm1((g.K, N), seed=7 + N)
    C = api.spmm_csr(handle, rp, ci, v, torch.from_numpy(B).to(dev)).cpu().numpy()
    check_f32(g, B, C)
    S = np.random.default_rng(N).choice([-1, 1], size=(g.K, N)).astype(np.int32)
    ones = torch.ones(g.nnz, device=dev)
    Cs = api.spmm_csr(handle, rp, ci, ones, torch.from_numpy(S.astype(np.float32)).to(dev)).cpu().numpy()
    np.testing.assert_array_equal(Cs, oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.float32))


def test_csr_strided_and_empty(handle, dev):
    g = graphs.random_csr(300, 200, 0.05, seed=1)
    rp, ci, v = to_dev(g, dev)
    Bfull = torch.from_numpy(graphs.uniform_pm1((200, 70), seed=2)).to(dev)
    B = Bfull[:, :64]  # ldb = 70: not 16-byte aligned rows -> narrow kernel
    Cfull = torch.full((300, 80), 7.0, device=dev)
    api.spmm_csr(handle, rp, ci, v, B, C_out=Cfull[:, :64])
    check_f32(g, B.cpu().numpy(), Cfull[:, :64].cpu().numpy())
    assert (Cfull[:, 64:] == 7.0).all()  # columns past N untouched
    # nnz = 0: C is zeroed
    e = graphs.random_csr(10, 10, 0.0, seed=0)
    rp, ci, v = to_dev(e, dev)
    C = torch.full((10, 8), 3.0, device=dev)
    api.spmm_csr(handle, rp, ci, v, torch.ones(10, 8, device=dev), C_out=C)
    assert (C == 0).all()


def test_plan_mismatch_is_an_error(handle, dev):
    g = graphs.random_csr(100, 100, 0.05, seed=4)
    rp, ci, v = to_dev(g, dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    g2 = graphs.random_csr(101, 100, 0.05, seed=4)
    rp2, ci2, v2 = to_dev(g2, dev)
    with pytest.raises(SpmmError) as ei:
        api.spmm_csr(handle, rp2, ci2, v2, torch.ones(100, 8, device=dev), plan=plan)
    assert ei.value.status == 5


def test_csr_large_workload_properties(handle, dev):
    """Workload-3 graph at 1/8 of its nnz (same rows, same degree law): FP32
    and BF16 against FP64, and linearity C(aB1 + B2) = aC(B1) + C(B2)."""
    g = graphs.workload_large_graph(scale=0.125)
    N = 64
    B = graphs.uniform_pm1((g.K, N), seed=11)
    rp, ci, v = to_dev(g, dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    Bd = torch.from_numpy(B).to(dev)
    C = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan).cpu().numpy()
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    assert rel_fro(C, ref) <= 1e-5
    Bh = api.to_bf16_bits(Bd)
    Ch = api.spmm_csr(handle, rp, ci, v, Bh, plan=plan).cpu().numpy()
    assert rel_fro(Ch, ref) <= 1e-2
    B2 = torch.from_numpy(graphs.uniform_pm1((g.K, N), seed=12)).to(dev)
    lhs = api.spmm_csr(handle, rp, ci, v, 0.5 * Bd + B2, plan=plan)
    rhs = 0.5 * api.spmm_csr(handle, rp, ci, v, Bd, plan=plan) + api.spmm_csr(handle, rp, ci, v, B2, plan=plan)
    assert torch.linalg.norm(lhs - rhs) / torch.linalg.norm(rhs) < 1e-5


@pytest.mark.parametrize("bf16", [False, True])
def test_autotune_picks_a_valid_geometry(handle, dev, bf16):
    g = graphs.chung_lu(20000, 400000, seed=8)
    g.val = graphs.row_normalised_uniform(g, 8)
    N = 64
    B = graphs.uniform_pm1((g.K, N), seed=9)
    rp, ci, v = to_dev(g, dev)
    Bd = torch.from_numpy(B).to(dev)
    if bf16:
        Bd = api.to_bf16_bits(Bd)
    plan = api.CsrPlan(handle, rp, g.nnz)
    C = torch.empty((g.M, N), device=dev)
    G, vec, rs, sl = plan.autotune(rp, ci, v, Bd, C)
    assert G in (4, 8, 16, 32, 64) and vec in (1, 2) and rs in (1, 2, 4, 8)
    assert sl in (0, 2, 3, 4, 6, 8, 12, 16, 24, 32)
    C1 = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan)            # uses the tuned entry
    C2 = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=(G, vec, rs, sl))
    assert torch.equal(C1, C2)
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B) if not bf16 else \
        oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bd.view(torch.int16).cpu().numpy().view(np.uint16))
    assert rel_fro(C1.cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("S", [2, 3, 4, 6, 8, 16, 24, 32, 64])
@pytest.mark.parametrize("tune", [(8, 1, 1), (16, 1, 2), (4, 2, 1), (32, 1, 1)])
def test_csr_column_slices(handle, dev, S, tune):
    """Column-sliced launches (tune.slices = S): the entries of each row split
    by ranges of B's rows, one launch per (column tile, slice), partial sums
    added into C in slice order. FP32 within the oracle's bounds, +-1 B exact,
    two calls bitwise equal; BF16 within 1e-5 of FP64 of the rounded B."""
    g = graphs.chung_lu(12000, 600000, seed=S, max_deg=5000)
    g.val = graphs.row_normalised_uniform(g, S)
    K = g.K
    N = 128
    B = graphs.uniform_pm1((K, N), seed=S + 1)
    rp, ci, v = to_dev(g, dev)
    Bd = torch.from_numpy(B).to(dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    t = tune + (S,)
    C1 = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=t)
    C2 = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=t)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    check_f32(g, B, C1.cpu().numpy())
    Sg = np.random.default_rng(S).choice([-1, 1], size=(K, N)).astype(np.int32)
    ones = torch.ones(g.nnz, device=dev)
    Cs = api.spmm_csr(handle, rp, ci, ones, torch.from_numpy(Sg.astype(np.float32)).to(dev), plan=plan,
                      tune=t)
    np.testing.assert_array_equal(Cs.cpu().numpy(), oracle.csr_spmm_i64(g.rowptr, g.colind, Sg).astype(np.float32))
    Bh = api.to_bf16_bits(Bd)
    Ch = api.spmm_csr(handle, rp, ci, v, Bh, plan=plan, tune=t).cpu().numpy()
    refq = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bh.view(torch.int16).cpu().numpy().view(np.uint16))
    assert rel_fro(Ch, refq) <= 1e-5
    plan.close()


def test_csr_column_slices_edge_shapes(handle, dev):
    """Empty rows, rows wholly inside one slice, K not a multiple of S, a
    second K on the same plan (a new offset set), N with one column tile."""
    g = graphs.random_csr(700, 5003, 0.01, seed=4, long_rows=3, long_len=5003)
    rp, ci, v = to_dev(g, dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    for N, tune in ((32, (8, 1, 1, 16)), (64, (16, 1, 1, 4)), (128, (8, 1, 2, 8))):
        B = graphs.uniform_pm1((g.K, N), seed=N)
        C = api.spmm_csr(handle, rp, ci, v, torch.from_numpy(B).to(dev), plan=plan, tune=tune)
        check_f32(g, B, C.cpu().numpy())
    # the same structure as A[:, :4000] against a B of 4000 rows (columns >= 4000 dropped)
    keep = g.colind < 4000
    rows = np.repeat(np.arange(g.M), g.degrees())[keep]
    rp2 = np.zeros(g.M + 1, np.int32)
    rp2[1:] = np.cumsum(np.bincount(rows, minlength=g.M))
    g2 = graphs.Csr(rp2, g.colind[keep].copy(), g.val[keep].copy(), g.M, 4000)
    rpd, cid, vd = to_dev(g2, dev)
    plan2 = api.CsrPlan(handle, rpd, g2.nnz)
    B = graphs.uniform_pm1((4000, 64), seed=9)
    for S in (4, 8):
        C = api.spmm_csr(handle, rpd, cid, vd, torch.from_numpy(B).to(dev), plan=plan2, tune=(8, 1, 1, S))
        check_f32(g2, B, C.cpu().numpy())


def test_csr_column_slices_need_sorted_rows(handle, dev):
    g = graphs.random_csr(300, 3000, 0.02, seed=6)
    ci = g.colind.copy()
    r = int(np.argmax(g.degrees()))
    s, e = g.rowptr[r], g.rowptr[r + 1]
    ci[s:e] = ci[s:e][::-1]  # one row with descending columns
    gu = graphs.Csr(g.rowptr, ci, g.val, g.M, g.K)
    rp, cid, v = to_dev(gu, dev)
    B = torch.from_numpy(graphs.uniform_pm1((g.K, 64), seed=1)).to(dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    with pytest.raises(SpmmError) as ei:
        api.spmm_csr(handle, rp, cid, v, B, plan=plan, tune=(8, 1, 1, 4))
    assert ei.value.status == 2
    C = api.spmm_csr(handle, rp, cid, v, B, plan=plan, tune=(8, 1, 1))  # unsliced still runs
    check_f32(gu, B.cpu().numpy(), C.cpu().numpy())
    big = graphs.random_csr(2000, 40000, 0.002, seed=7)
    ci2 = big.colind.copy()
    r = int(np.argmax(big.degrees()))
    s, e = big.rowptr[r], big.rowptr[r + 1]
    ci2[s:e] = ci2[s:e][::-1]
    gb = graphs.Csr(big.rowptr, ci2, big.val, big.M, big.K)
    rp, cid, v = to_dev(gb, dev)
    B = torch.from_numpy(graphs.uniform_pm1((big.K, 128), seed=2)).to(dev)
    plan = api.CsrPlan(handle, rp, gb.nnz)
    C = torch.empty((gb.M, 128), device=dev)
    G, vec, rs, sl = plan.autotune(rp, cid, v, B, C)  # B is 20 MB: slices would be tried
    assert sl == 0
