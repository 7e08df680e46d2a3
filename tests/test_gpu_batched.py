"""Batched uneven graphs: the one-plan block-diagonal fwd+bwd (one fused,
autotuned launch) equals per-graph runs of the same kernels at the same
geometry bit for bit, the hipGraph replay equals the eager step on every
replay (with eager work and synchronisations between replays), and both sit
within tolerance of the FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
from spmm_eval_amd import api, graphs
from spmm_eval_amd.batched import BatchPlan
from spmm_eval_amd._lib import SpmmError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    gs = graphs.workload_batched(ngraphs=48)
    return gs, graphs.concat_block_diagonal(gs)


@pytest.mark.parametrize("bf16", [False, True])
def test_batched_matches_per_graph_and_oracle(handle, dev, batch, bf16):
    gs, (U, row_off, nnz_off) = batch
    N = 64
    X = graphs.normal((U.M, N), 9, 1)
    dY = graphs.normal((U.M, N), 9, 2)
    Xd, dYd = torch.from_numpy(X).to(dev), torch.from_numpy(dY).to(dev)
    if bf16:
        Xd, dYd = api.to_bf16_bits(Xd), api.to_bf16_bits(dYd)
    plan = BatchPlan(handle, gs, N=N, bf16=bf16, device=dev)
    Y, dX, dval = plan.step(Xd, dYd, use_graph=False)
    torch.cuda.synchronize()
    Yg, dXg, dvalg = plan.step(Xd, dYd, use_graph=True)
    torch.cuda.synchronize()
    assert torch.equal(Y, Yg) and torch.equal(dX, dXg) and torch.equal(dval, dvalg)
    for rep in range(3):  # replays interleaved with eager work
        junk = torch.randn(1000, device=dev).sum()
        Yr, dXr, dvalr = plan.step(Xd, dYd, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(Y, Yr) and torch.equal(dX, dXr) and torch.equal(dval, dvalr), rep
        del junk
    # per-graph runs of the same kernels at the geometry the plan chose
    info = plan.info()
    tune = (info["lanes_per_row"], info["vec"], info["rows_scale"])
    for gi in (0, 7, len(gs) - 1):
        g = gs[gi]
        r0, r1 = int(row_off[gi]), int(row_off[gi + 1])
        e0, e1 = int(nnz_off[gi]), int(nnz_off[gi + 1])
        rp, ci = torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)
        v = torch.from_numpy(g.val).to(dev)
        p = api.CsrPlan(handle, rp, g.nnz)
        Yi = api.spmm_csr(handle, rp, ci, v, Xd[r0:r1], plan=p, tune=tune)
        assert torch.equal(Yi, Y[r0:r1])
        if info["sddmm_form"] == 2:  # dval formed in the backward items (the tuner's pick)
            tp = api.CsrTPlan(handle, rp, ci, g.K)
            _, dvi = api.spmm_csr_t_sddmm(tp, rp, ci, v, dYd[r0:r1], Xd[r0:r1], tune=tune)
        else:
            dvi = api.sddmm_csr(handle, rp, ci, dYd[r0:r1], Xd[r0:r1])
        assert torch.equal(dvi, dval[e0:e1]), (gi, info["sddmm_form"])
    # oracle on the union
    Xf = X if not bf16 else oracle.bf16_to_f32(Xd.view(torch.int16).cpu().numpy().view(np.uint16))
    dYf = dY if not bf16 else oracle.bf16_to_f32(dYd.view(torch.int16).cpu().numpy().view(np.uint16))
    ref = oracle.csr_spmm_f64(U.rowptr, U.colind, U.val, Xf)
    assert np.linalg.norm(Y.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-5
    cp, ri, cv, _ = oracle.csr2csc(U.rowptr, U.colind, U.val, U.K)
    ref_dx = oracle.csr_spmm_f64(cp, ri, cv, dYf)
    assert np.linalg.norm(dX.cpu().numpy() - ref_dx) / np.linalg.norm(ref_dx) <= 1e-5
    ref_dv = oracle.sddmm_f64(U.rowptr, U.colind, dYf, Xf)
    assert np.linalg.norm(dval.cpu().numpy() - ref_dv) / np.linalg.norm(ref_dv) <= 1e-5
    plan.close()


@pytest.mark.parametrize("bf16", [False, True])
def test_batched_sddmm_forms(handle, dev, batch, bf16):
    """Every SDDMM form of the fused launch pinned on one plan: entry chunks
    (0) and forward rows (1) give the same bits; the backward items forming
    dval beside dX (2) equal the per-graph fused backward bit for bit and the
    other forms to 1e-6; each form's graph replay equals its eager step; a
    form the geometry cannot run is refused."""
    gs, (U, row_off, nnz_off) = batch
    N = 64
    Xd = torch.from_numpy(graphs.normal((U.M, N), 11, 1)).to(dev)
    dYd = torch.from_numpy(graphs.normal((U.M, N), 11, 2)).to(dev)
    if bf16:
        Xd, dYd = api.to_bf16_bits(Xd), api.to_bf16_bits(dYd)
    plan = BatchPlan(handle, gs, N=N, bf16=bf16, device=dev)
    lanes = 8 if bf16 else 16  # one column tile at one vector per lane
    out = {}
    for form in (0, 1, 2):
        for pers in (False, True):
            plan.set_geometry(lanes, 1, 1, pers, form)
            Y, dX, dval = (t.clone() for t in plan.step(Xd, dYd, use_graph=False))
            Yg, dXg, dvalg = plan.step(Xd, dYd, use_graph=True)
            torch.cuda.synchronize()
            assert torch.equal(Y, Yg) and torch.equal(dX, dXg) and torch.equal(dval, dvalg)
            out[(form, pers)] = (Y, dX, dval)
    for key, (Y, dX, dval) in out.items():
        assert torch.equal(Y, out[(0, False)][0]), key
        assert torch.equal(dX, out[(0, False)][1]), key
        # the persistent launch runs the same items: the same bits as the one-shot launch
        assert torch.equal(dval, out[(key[0], False)][2]), key
    assert torch.equal(out[(1, False)][2], out[(0, False)][2])
    Y2, dX2, dv2 = out[(2, False)]
    assert torch.equal(dX2, out[(0, False)][1])  # same transposed rows, same sums
    dv0 = out[(0, False)][2]
    assert float((dv2 - dv0).abs().max()) <= 1e-6 * float(dv0.abs().max())
    for gi in (0, 5, len(gs) - 1):
        g = gs[gi]
        r0, r1 = int(row_off[gi]), int(row_off[gi + 1])
        e0, e1 = int(nnz_off[gi]), int(nnz_off[gi + 1])
        rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
        tp = api.CsrTPlan(handle, rp, ci, g.K)
        dXi, dvi = api.spmm_csr_t_sddmm(tp, rp, ci, v, dYd[r0:r1], Xd[r0:r1], tune=(lanes, 1, 1))
        torch.cuda.synchronize()
        assert torch.equal(dXi, dX2[r0:r1]) and torch.equal(dvi, dv2[e0:e1]), gi
    with pytest.raises(SpmmError):  # two column tiles cannot form dval in the backward items
        plan.set_geometry(lanes // 2, 1, 1, False, 2)
    with pytest.raises(SpmmError):
        plan.set_geometry(lanes, 1, 3, False, 0)
    plan.close()
