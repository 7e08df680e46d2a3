"""Full-shape parity of the two largest workloads, every row checked.

1. The 232,965-row power-law graph (~114.5M nnz, heaviest row ~17k entries,
   values U(0,1] row-normalised), N = 64 and 128, FP32 and BF16 B, through
   the planned (autotuned) form and the plan-free drop-in (column-sliced at
   these sizes, bitwise equal to the planned form at its geometry): FP32 within 1e-5
   relative Frobenius of the FP64 oracle and every element within its
   rounding bound nnz_row * 2^-23 * sum|a||b|; BF16 within 1e-5 of FP64 of
   the BF16-rounded B (kernel exactness) and within 1e-2 of FP64 of the FP32
   source (the precision form's contract). Its pattern as B2SR-4: converted
   tile for tile like the oracle, +-1 B exact on every row. The reference
   checks its own workloads at full size in the same way
   (test_spmm_full.cu:412-431).
2. All 512 uneven graphs batched in one plan, FP32 and BF16: Y, dX and dval
   of every graph equal per-graph runs of the same kernels at the plan's
   geometry bit for bit, and the union is within 1e-5 of FP64 (of the rounded
   operands for BF16) and within 1e-2 of FP64 of the FP32 source.
"""
import numpy as np
import pytest
import torch

import oracle
from spmm_eval_amd import api, graphs
from spmm_eval_amd.batched import BatchPlan

pytestmark = pytest.mark.gpu


def rel_fro(C, ref):
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(C.astype(np.float64) - ref) / (n if n > 0 else 1.0))


@pytest.fixture(scope="module")
def large(dev):
    g = graphs.workload_large_graph()
    print(f"large graph: {g.M} rows, {g.nnz} nnz, max degree {int(g.degrees().max())}", flush=True)
    assert g.M == 232_965 and g.nnz > 110_000_000 and g.degrees().max() > 15_000
    d = tuple(torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    yield g, d
    torch.cuda.empty_cache()


@pytest.mark.parametrize("N", [64, 128])
def test_large_graph_f32_every_row(handle, dev, large, N):
    g, (rp, ci, v) = large
    B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 3 + N)
    Bd = torch.from_numpy(B).to(dev)
    plan = api.CsrPlan(handle, rp, g.nnz)
    C = torch.empty((g.M, N), device=dev)
    plan.autotune(rp, ci, v, Bd, C)
    Cp = api.spmm_csr(handle, rp, ci, v, Bd, C_out=C, plan=plan).cpu().numpy()
    Cd = api.spmm_csr(handle, rp, ci, v, Bd).cpu().numpy()  # plan-free drop-in
    # the drop-in runs column-sliced here; the planned form at its geometry gives the same bits
    geo = api.dropin_geometry(g.M, g.K, N, g.nnz)
    assert geo[3] > 1, geo
    Cg = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=geo).cpu().numpy()
    np.testing.assert_array_equal(Cd, Cg)
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    # sum |a||b| per element (values are positive), the rounding bound's scale
    mag = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, np.abs(B))
    bound = np.maximum(g.degrees(), 1)[:, None] * 2.0 ** -23 * mag + 1e-30
    for name, Cx in (("planned", Cp), ("drop-in", Cd)):
        assert rel_fro(Cx, ref) <= 1e-5, name
        bad = np.abs(Cx - ref) > bound
        assert not bad.any(), (name, int(bad.sum()), np.argwhere(bad)[:5].tolist())
    plan.close()


@pytest.mark.parametrize("N", [64, 128])
def test_large_graph_bf16_every_row(handle, dev, large, N):
    g, (rp, ci, v) = large
    B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 3 + N)
    Bh = oracle.f32_to_bf16(B)
    Bd = torch.from_numpy(Bh.view(np.int16)).to(dev).view(torch.uint16)
    plan = api.CsrPlan(handle, rp, g.nnz)
    C = torch.empty((g.M, N), device=dev)
    plan.autotune(rp, ci, v, Bd, C)
    Cp = api.spmm_csr(handle, rp, ci, v, Bd, C_out=C, plan=plan).cpu().numpy()
    Cd = api.spmm_csr(handle, rp, ci, v, Bd).cpu().numpy()
    geo = api.dropin_geometry(g.M, g.K, N, g.nnz, bf16=True)
    assert geo[3] > 1, geo
    Cg = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=geo).cpu().numpy()
    np.testing.assert_array_equal(Cd, Cg)
    ref_q = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bh)
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    for name, Cx in (("planned", Cp), ("drop-in", Cd)):
        assert rel_fro(Cx, ref_q) <= 1e-5, name
        err = rel_fro(Cx, ref)
        print(f"BF16 N={N} {name}: rel err vs FP64 of the FP32 source {err:.3e}", flush=True)
        assert err <= 1e-2, name
    plan.close()


def test_large_graph_b2sr4_every_row(handle, dev, large):
    """B2SR-4 on the same graph's pattern (the reddit-sized workload of the
    reference's FBF figures, backend/spmm.cu:869-870): the device csr2b2sr
    equals the oracle's tile for tile, and FBF with +-1 B equals integer CSR
    SpMM on every row, planned (autotuned) and plan-free, N = 64."""
    g, (rp, ci, _) = large
    T, N = 4, 64
    brp, bci, words = api.csr2b2sr(handle, rp, ci, g.M, g.K, T)
    ref = oracle.csr2b2sr(g.rowptr, g.colind, T)
    print(f"B2SR-4: {bci.numel()} tiles", flush=True)
    np.testing.assert_array_equal(brp.cpu().numpy(), ref[0])
    np.testing.assert_array_equal(bci.cpu().numpy(), ref[1])
    np.testing.assert_array_equal(words.cpu().numpy(), ref[2])
    del ref
    S = np.random.default_rng(5).choice([-1, 1], size=(g.K, N)).astype(np.int32)
    exact = oracle.csr_spmm_i64(g.rowptr, g.colind, S).astype(np.float32)
    Sd = torch.from_numpy(S.astype(np.float32)).to(dev)
    plan = api.B2srPlan(handle, T, g.M, brp)
    C = torch.empty((g.M, N), device=dev)
    print(f"B2SR-4 tuned: {plan.autotune(brp, bci, words, Sd, C)}", flush=True)
    for name, pl in (("planned", plan), ("drop-in", None)):
        Cx = api.spmm_b2sr(handle, T, g.M, brp, bci, words, Sd, plan=pl).cpu().numpy()
        np.testing.assert_array_equal(Cx, exact, err_msg=name)
    plan.close()


@pytest.fixture(scope="module")
def batch512():
    gs = graphs.workload_batched(512)
    assert len(gs) == 512
    return gs, graphs.concat_block_diagonal(gs)


@pytest.mark.parametrize("bf16", [False, True])
def test_batched_512_every_graph(handle, dev, batch512, bf16):
    gs, (U, row_off, nnz_off) = batch512
    N = 64
    X = graphs.normal((U.M, N), 9, 1)
    dY = graphs.normal((U.M, N), 9, 2)
    Xd, dYd = torch.from_numpy(X).to(dev), torch.from_numpy(dY).to(dev)
    if bf16:
        Xd, dYd = api.to_bf16_bits(Xd), api.to_bf16_bits(dYd)
    plan = BatchPlan(handle, gs, N=N, bf16=bf16, device=dev)
    Y, dX, dval = (t.clone() for t in plan.step(Xd, dYd, use_graph=True))
    torch.cuda.synchronize()
    info = plan.info()
    tune = (info["lanes_per_row"], info["vec"], info["rows_scale"])
    print(f"512 graphs, {U.M} rows, {U.nnz} nnz, bf16={bf16}: {info}", flush=True)
    for gi, g in enumerate(gs):
        r0, r1 = int(row_off[gi]), int(row_off[gi + 1])
        e0, e1 = int(nnz_off[gi]), int(nnz_off[gi + 1])
        rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
        p = api.CsrPlan(handle, rp, g.nnz)
        Yi = api.spmm_csr(handle, rp, ci, v, Xd[r0:r1], plan=p, tune=tune)
        tp = api.CsrTPlan(handle, rp, ci, g.K)
        if info["sddmm_form"] == 2:  # dval formed beside dX in the backward items
            dXi, dvi = api.spmm_csr_t_sddmm(tp, rp, ci, v, dYd[r0:r1], Xd[r0:r1], tune=tune)
        else:
            dXi = api.spmm_csr_t(handle, rp, ci, v, dYd[r0:r1], K=g.K, plan=tp, tune=tune)
            dvi = api.sddmm_csr(handle, rp, ci, dYd[r0:r1], Xd[r0:r1])
        torch.cuda.synchronize()
        assert torch.equal(Yi, Y[r0:r1]), ("Y", gi)
        assert torch.equal(dXi, dX[r0:r1]), ("dX", gi)
        assert torch.equal(dvi, dval[e0:e1]), ("dval", gi, info["sddmm_form"])
        p.close()
        tp.close()
    Xq = X if not bf16 else oracle.bf16_to_f32(Xd.view(torch.int16).cpu().numpy().view(np.uint16))
    dYq = dY if not bf16 else oracle.bf16_to_f32(dYd.view(torch.int16).cpu().numpy().view(np.uint16))
    cp, ri, cv, _ = oracle.csr2csc(U.rowptr, U.colind, U.val, U.K)
    Yc, dXc, dvc = Y.cpu().numpy(), dX.cpu().numpy(), dval.cpu().numpy()
    for Xr, dYr, tol in ((Xq, dYq, 1e-5), (X, dY, 1e-2)):
        assert rel_fro(Yc, oracle.csr_spmm_f64(U.rowptr, U.colind, U.val, Xr)) <= tol
        assert rel_fro(dXc, oracle.csr_spmm_f64(cp, ri, cv, dYr)) <= tol
        assert rel_fro(dvc, oracle.sddmm_f64(U.rowptr, U.colind, dYr, Xr)) <= tol
    plan.close()
