"""The plan-free CSR drop-ins (spmm_csr_f32 / _bf16, spmm_csr_t_f32 / _bf16):
no host analysis, no allocation after the first call at a size, no
synchronisation. Their rows run the planned kernel's per-class code, so the
output equals the planned form at the drop-in's lane geometry bit for bit
(spmm_csr_dropin_geometry: the default geometry, or column slices when a
column tile of B is large and rows are long, with per-row slice offsets built
on the device each call; rows whose columns descend are summed whole, found
on the device, with no synchronisation); they replay from a captured graph
with eager work between replays; the handle state that grows for a larger
graph keeps every result right; and growing during a capture is refused
(the earlier state would be left to a graph that outgrew it)."""
import numpy as np
import pytest
import torch

import oracle
from spmm_eval_amd import api, graphs
from spmm_eval_amd._lib import SpmmError

pytestmark = pytest.mark.gpu


def rel(C, ref):
    return float(np.linalg.norm(C.astype(np.float64) - ref) / max(np.linalg.norm(ref), 1e-300))


@pytest.fixture(scope="module")
def hubgraph(dev):
    # hub rows of several thousand entries (workgroup class), medium and short rows
    g = graphs.chung_lu(30000, 1_500_000, seed=17, max_deg=6000)
    g.val = graphs.row_normalised_uniform(g, 17)
    d = tuple(torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    return g, d


@pytest.mark.parametrize("N,bf16,sliced", [(32, False, False), (64, False, False), (128, False, True),
                                           (64, True, False), (128, True, False), (256, True, True)])
def test_dropin_equals_planned_at_its_geometry(handle, dev, hubgraph, N, bf16, sliced):
    g, (rp, ci, v) = hubgraph
    assert int(g.degrees().max()) > 2000
    B = graphs.uniform_pm1((g.K, N), seed=N)
    Bd = torch.from_numpy(B).to(dev)
    if bf16:
        Bd = api.to_bf16_bits(Bd)
    geo = api.dropin_geometry(g.M, g.K, N, g.nnz, bf16)
    assert (geo[3] > 1) == sliced, geo
    C_drop = api.spmm_csr(handle, rp, ci, v, Bd)
    C_drop2 = api.spmm_csr(handle, rp, ci, v, Bd)
    plan = api.CsrPlan(handle, rp, g.nnz)
    C_plan = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=geo)
    torch.cuda.synchronize()
    assert torch.equal(C_drop, C_plan) and torch.equal(C_drop, C_drop2)
    if not sliced:  # nothing tuned: the planned default is the same geometry
        assert torch.equal(C_drop, api.spmm_csr(handle, rp, ci, v, Bd, plan=plan))
    ref = oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bd.view(torch.int16).cpu().numpy().view(np.uint16)) \
        if bf16 else oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    assert rel(C_drop.cpu().numpy(), ref) <= 1e-5
    plan.close()


@pytest.mark.parametrize("bf16", [False, True])
def test_transposed_dropin_equals_planned(handle, dev, hubgraph, bf16):
    g, (rp, ci, v) = hubgraph
    N = 64
    B = graphs.uniform_pm1((g.M, N), seed=5)
    Bd = torch.from_numpy(B).to(dev)
    if bf16:
        Bd = api.to_bf16_bits(Bd)
    Ct = api.spmm_csr_t(handle, rp, ci, v, Bd, K=g.K)
    tp = api.CsrTPlan(handle, rp, ci, g.K)
    Cp = api.spmm_csr_t(handle, rp, ci, v, Bd, K=g.K, plan=tp)
    torch.cuda.synchronize()
    assert torch.equal(Ct, Cp)
    cp, ri, cv, _ = oracle.csr2csc(g.rowptr, g.colind, g.val, g.K)
    Bf = oracle.bf16_to_f32(Bd.view(torch.int16).cpu().numpy().view(np.uint16)) if bf16 else B
    assert rel(Ct.cpu().numpy(), oracle.csr_spmm_f64(cp, ri, cv, Bf)) <= 1e-5
    tp.close()


def test_dropins_replay_from_a_graph(dev, hubgraph):
    """Four drop-in calls (CSR, transposed CSR, BF16 CSR, column-sliced CSR)
    captured once and replayed three times, with eager drop-in calls on
    another graph through the same handle between replays; every replay
    equals the eager result."""
    g, (rp, ci, v) = hubgraph
    h = api.Handle()
    N = 64
    Bd = torch.from_numpy(graphs.uniform_pm1((g.K, N), seed=8)).to(dev)
    Bh = api.to_bf16_bits(Bd)
    Bw = torch.from_numpy(graphs.uniform_pm1((g.K, 128), seed=10)).to(dev)
    assert api.dropin_geometry(g.M, g.K, 128, g.nnz)[3] > 1  # sliced
    C1 = torch.empty((g.M, N), device=dev)
    C2 = torch.empty((g.K, N), device=dev)
    C3 = torch.empty((g.M, N), device=dev)
    C4 = torch.empty((g.M, 128), device=dev)
    outs = (C1, C2, C3, C4)

    def calls():
        api.spmm_csr(h, rp, ci, v, Bd, C_out=C1)
        api.spmm_csr_t(h, rp, ci, v, Bd, K=g.K, C_out=C2)
        api.spmm_csr(h, rp, ci, v, Bh, C_out=C3)
        api.spmm_csr(h, rp, ci, v, Bw, C_out=C4)

    calls()  # grows the handle's state and workspace
    torch.cuda.synchronize()
    ref = [C.clone() for C in outs]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        calls()
        s.synchronize()
        with torch.cuda.graph(gr, stream=s):
            calls()
    torch.cuda.synchronize()
    small = graphs.chung_lu(3000, 40000, seed=3)
    srp, sci, sv = (torch.from_numpy(x).to(dev) for x in (small.rowptr, small.colind, small.val))
    Bs = torch.from_numpy(graphs.uniform_pm1((small.K, N), seed=9)).to(dev)
    for rep in range(3):
        for C in outs:
            C.fill_(0)
        gr.replay()
        torch.cuda.synchronize()
        for C, R in zip(outs, ref):
            assert torch.equal(C, R), rep
        Cs = api.spmm_csr(h, srp, sci, sv, Bs)  # eager, same handle state
        torch.cuda.synchronize()
        assert rel(Cs.cpu().numpy(), oracle.csr_spmm_f64(small.rowptr, small.colind, small.val,
                                                          Bs.cpu().numpy())) <= 1e-5
    h.close()


@pytest.mark.parametrize("bf16", [False, True])
def test_sliced_dropin_strided_operands(handle, dev, hubgraph, bf16):
    """B and C as column slices of wider row-major arrays (ldb, ldc > N): the
    sliced drop-in equals the planned form at its geometry on the same views,
    and C's columns outside the view are untouched."""
    g, (rp, ci, v) = hubgraph
    N = 128 if not bf16 else 256
    geo = api.dropin_geometry(g.M, g.K, N, g.nnz, bf16)
    assert geo[3] > 1
    Bw = torch.from_numpy(graphs.uniform_pm1((g.K, N + 64), seed=33)).to(dev)
    if bf16:
        Bw = api.to_bf16_bits(Bw)
    B = Bw[:, 32:32 + N]  # 16-byte aligned view, ldb = N + 64
    Cw = torch.full((g.M, N + 32), 7.0, device=dev)
    Cp = torch.full((g.M, N + 32), 7.0, device=dev)
    api.spmm_csr(handle, rp, ci, v, B, C_out=Cw[:, 16:16 + N])
    plan = api.CsrPlan(handle, rp, g.nnz)
    api.spmm_csr(handle, rp, ci, v, B, C_out=Cp[:, 16:16 + N], plan=plan, tune=geo)
    torch.cuda.synchronize()
    assert torch.equal(Cw, Cp)
    assert bool((Cw[:, :16] == 7.0).all()) and bool((Cw[:, 16 + N:] == 7.0).all())
    Bf = oracle.bf16_to_f32(B.contiguous().view(torch.int16).cpu().numpy().view(np.uint16)) if bf16 \
        else B.cpu().numpy()
    assert rel(Cw[:, 16:16 + N].cpu().numpy(), oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, Bf)) <= 1e-5
    plan.close()


def test_sliced_dropin_edge_rows(handle, dev, hubgraph):
    """Empty rows (every seventh), rows whose entries all lie in one slice (the
    columns of every fifth row clipped into the first 1,000), and K not a
    multiple of the slice count: the sliced drop-in equals the planned form at
    its geometry, and empty rows come out as zeros."""
    g0, _ = hubgraph
    keep = []
    for r in range(g0.M):
        seg = g0.colind[g0.rowptr[r]:g0.rowptr[r + 1]]
        if r % 7 == 3:
            seg = seg[:0]
        elif r % 5 == 1:
            seg = np.unique(seg % 1000)
        keep.append(seg)
    rowptr = np.zeros(g0.M + 1, np.int32)
    rowptr[1:] = np.cumsum([len(x) for x in keep])
    colind = np.concatenate(keep).astype(np.int32)
    val = graphs.uniform_pm1((len(colind),), seed=41).astype(np.float32)
    K = g0.K + 5  # B has rows no entry uses; ceil(K / S) splits unevenly
    g = graphs.Csr(rowptr, colind, val, g0.M, K)
    N = 128
    geo = api.dropin_geometry(g.M, K, N, g.nnz)
    assert geo[3] > 1 and K % geo[3] != 0, geo
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    B = graphs.uniform_pm1((K, N), seed=42)
    Bd = torch.from_numpy(B).to(dev)
    Cd = api.spmm_csr(handle, rp, ci, v, Bd)
    plan = api.CsrPlan(handle, rp, g.nnz)
    Cp = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=geo)
    torch.cuda.synchronize()
    assert torch.equal(Cd, Cp)
    assert bool((Cd[3::7] == 0).all())
    assert rel(Cd.cpu().numpy(), oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)) <= 1e-5
    plan.close()


@pytest.mark.parametrize("bf16", [False, True])
def test_sliced_dropin_rows_not_ascending(handle, dev, hubgraph, bf16):
    """A row whose columns descend cannot be split by column ranges: the
    offsets kernel flags it on the device and the first slice launch sums
    every row whole (the later ones do nothing). The result stays within the
    oracle's bound, and the next call on ascending rows, through the same
    handle state, is again bitwise the planned sliced form."""
    g, (rp, ci, v) = hubgraph
    N = 128 if not bf16 else 256
    geo = api.dropin_geometry(g.M, g.K, N, g.nnz, bf16)
    assert geo[3] > 1
    r = int(np.argmax(g.degrees()))
    cu = g.colind.copy()
    cu[g.rowptr[r]:g.rowptr[r + 1]] = cu[g.rowptr[r]:g.rowptr[r + 1]][::-1]
    ciu = torch.from_numpy(cu).to(dev)
    B = graphs.uniform_pm1((g.K, N), seed=21)
    Bd = torch.from_numpy(B).to(dev)
    if bf16:
        Bd = api.to_bf16_bits(Bd)
        Bf = oracle.bf16_to_f32(Bd.view(torch.int16).cpu().numpy().view(np.uint16))
    else:
        Bf = B
    for _ in range(2):
        Cu = api.spmm_csr(handle, rp, ciu, v, Bd)
        torch.cuda.synchronize()
        assert rel(Cu.cpu().numpy(), oracle.csr_spmm_f64(g.rowptr, cu, g.val, Bf)) <= 1e-5
    Cs = api.spmm_csr(handle, rp, ci, v, Bd)
    plan = api.CsrPlan(handle, rp, g.nnz)
    Cp = api.spmm_csr(handle, rp, ci, v, Bd, plan=plan, tune=geo)
    torch.cuda.synchronize()
    assert torch.equal(Cs, Cp)
    plan.close()


def test_state_regrowth_keeps_results(dev):
    h = api.Handle()
    for n, nnz, seed in ((2000, 20000, 1), (40000, 800000, 2), (500, 4000, 3), (40000, 900000, 4)):
        g = graphs.chung_lu(n, nnz, seed=seed)
        rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
        B = graphs.uniform_pm1((g.K, 32), seed=seed)
        C = api.spmm_csr(h, rp, ci, v, torch.from_numpy(B).to(dev))
        Ct = api.spmm_csr_t(h, rp, ci, v, torch.from_numpy(B).to(dev), K=g.K)
        torch.cuda.synchronize()
        assert rel(C.cpu().numpy(), oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)) <= 1e-5
        cp, ri, cv, _ = oracle.csr2csc(g.rowptr, g.colind, g.val, g.K)
        assert rel(Ct.cpu().numpy(), oracle.csr_spmm_f64(cp, ri, cv, B)) <= 1e-5
    h.close()


def test_growth_during_capture_is_refused(dev):
    g = graphs.chung_lu(2000, 20000, seed=6)
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    Bd = torch.from_numpy(graphs.uniform_pm1((g.K, 32), seed=6)).to(dev)
    h = api.Handle()  # fresh: its state has never grown
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    gr = torch.cuda.CUDAGraph()
    x = torch.zeros(16, device=dev)
    status = None
    with torch.cuda.stream(s):
        with torch.cuda.graph(gr, stream=s):
            try:
                api.spmm_csr(h, rp, ci, v, Bd)
            except SpmmError as e:
                status = e.status
            x.add_(1)
    torch.cuda.synchronize()
    assert status == 2  # SPMM_ERR_UNSUPPORTED, nothing queued
    gr.replay()
    torch.cuda.synchronize()
    assert float(x[0]) == 1.0
    C = api.spmm_csr(h, rp, ci, v, Bd)  # eager: grows, then runs
    torch.cuda.synchronize()
    assert rel(C.cpu().numpy(), oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, Bd.cpu().numpy())) <= 1e-5
    h.close()
