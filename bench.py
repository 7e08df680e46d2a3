"""SpMM benchmark on one MI355X.

Headline step (`value`): the FBF SpMM the reference harness launches --
B2SR-4 bit tiles x FP32 B -> FP32 C (spmm4_full_full_new2,
backend/spmm.cu:877-956, test_spmm_full.cu:344-369) -- on the seeded
89,250-row power-law 0/1 graph (~900k nnz) with hidden N = 64, inputs
resident in HBM. Metric: algorithmic GB/s (gather model: every stored bit
reads one full B row), with GFLOP/s = 2 nnz N / t.

`extra` holds the other workloads (skipped with --quick): CSR FP32 on the
16,384-row graph (N=32) and on the 232,965-row ~114.6M-nnz graph (N=64 and
128, FP32 and BF16 B); GCN-style fwd+bwd (Y = A X, dX = A^T dY,
dval = SDDMM) in FP32 and BF16 features; 512 uneven graphs batched in one
captured hipGraph vs 512 separate FP32 launches; aligned B2SR-8/16 binary
SpMV (binary and FP32 x) on the headline graph beside CSR at N = 1; the B2SR
grid (tiles 2..64 x N = 1/32/64/128) beside CSR. Errors there are measured
against an FP64 numpy recomputation on sampled rows.

`cpu_baseline`: the OpenMP FP32 CSR SpMM of the oracle library on the same
graph and B (3 warm-ups, median of 15), same byte count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ bytes
def csr_bytes(M, nnz, N, s_val=4, s_b=4):
    return 4 * (M + 1) + nnz * (4 + s_val) + nnz * N * s_b + M * N * 4


def b2sr_bytes(M, T, nb, nnz, N, s_b=4):
    MB = (M + T - 1) // T
    tile = {2: 2, 4: 4, 8: 8, 16: 32, 32: 128, 64: 512}[T]
    return 4 * (MB + 1) + nb * (4 + tile) + nnz * N * s_b + M * N * 4


def sddmm_bytes(M, nnz, N, s=4):
    return 4 * (M + 1) + 8 * nnz + nnz * N * s + M * N * s


def b2sr_lanes(N, elem_bytes=4):
    """Lanes per row the FBF kernels use for an aligned N (widest power of
    two dividing N / (16 / elem_bytes), at most 64)."""
    lanes = N // (16 // elem_bytes)
    t = 64
    while t > 1 and (t > lanes or lanes % t):
        t //= 2
    return t


def profiled_traffic(profile: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from a committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by what it reads for a
    device copy of known size). None when the summary has no row for it."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    key = f"`void spmm::{kernel}`"
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == key:
            try:
                return int(cells[4])
            except ValueError:
                return None
    return None


# ----------------------------------------------------------------- timing
def time_launches(fn, steps, warmup):
    """Runs fn() warmup times, then `steps` times back to back between two
    events on the current stream (the timed region: ms per step); then a
    second pass of `steps` launches, each between its own event pair on the
    same stream (mean kernel duration, kept out of the timed region because
    the extra event records between launches cost a few microseconds each).
    Returns (ms per step, mean ms per launch)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(s)
    for _ in range(steps):
        fn()
    t1.record(s)
    torch.cuda.synchronize()
    total = t0.elapsed_time(t1) / steps
    pairs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in pairs:
        a.record(s)
        fn()
        b.record(s)
    torch.cuda.synchronize()
    per = float(np.mean([a.elapsed_time(b) for a, b in pairs]))
    return total, per


def time_loop(fn, steps, warmup):
    """Back-to-back launches between two events (no per-launch events)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(s)
    for _ in range(steps):
        fn()
    t1.record(s)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def time_graph(fn, out, steps=100, replays=5):
    """Kernels too short for host launch pacing: `steps` back-to-back calls
    captured in one hipGraph, replayed `replays` times between two events
    (ms per call). Every replay's `out` must equal the eager result."""
    import torch

    fn()
    torch.cuda.synchronize()
    ref = out.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        fn()  # warm the capture stream
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            for _ in range(steps):
                fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    total = 0.0
    for _ in range(replays):
        out.fill_(0)
        torch.cuda.synchronize()
        t0.record()
        g.replay()
        t1.record()
        torch.cuda.synchronize()
        total += t0.elapsed_time(t1)
        if not torch.equal(out, ref):
            raise RuntimeError("graph replay differs from the eager result")
    return total / (replays * steps)


def sampled_rel_err(g, B64, C, rows):
    """Relative Frobenius error of C on `rows` against an FP64 recompute."""
    num = den = 0.0
    for i in rows:
        s, e = g.rowptr[i], g.rowptr[i + 1]
        ref = g.val[s:e].astype(np.float64) @ B64[g.colind[s:e]]
        num += float(((C[i] - ref) ** 2).sum())
        den += float((ref ** 2).sum())
    return (num / den) ** 0.5 if den > 0 else 0.0


# ------------------------------------------------------------------ parts
def headline(args, h, dev):
    import torch

    from spmm_eval_amd import api, graphs

    T, N = args.tile, args.hidden
    g = graphs.workload_b2sr_graph()
    rp, ci = torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)
    t = time.time()
    brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, T)
    torch.cuda.synchronize()
    conv_ms = (time.time() - t) * 1e3
    nb = bci.numel()
    B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 2)
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((g.M, N), device=dev)

    # row schedule built once per structure, outside the timed region (the
    # reference converts and uploads its B2SR arrays before timing too)
    bplan = api.B2srPlan(h, T, g.M, brp)
    torch.cuda.synchronize()

    def step():
        api.spmm_b2sr(h, T, g.M, brp, bci, words, Bd, C_out=C, plan=bplan)

    ms_step, ms_launch = time_launches(step, args.steps, args.warmup)
    # the plan-free drop-in form (bulk kernel + hub-row list kernel)
    Cd = torch.empty((g.M, N), device=dev)
    dropin_ms = time_loop(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bd, C_out=Cd),
                          args.steps, args.warmup)
    dropin_diff = float((C - Cd).abs().max())
    nbytes = b2sr_bytes(g.M, T, nb, g.nnz, N)
    # the same product through the CSR kernel (the reference prints its
    # cuSPARSE CSR time beside the B2SR one, test_spmm_full.cu:505)
    v = torch.ones(g.nnz, device=dev)
    plan = api.CsrPlan(h, rp, g.nnz)
    Ccsr = torch.empty((g.M, N), device=dev)
    csr_ms = time_loop(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=Ccsr, plan=plan), args.steps, args.warmup)
    del plan
    flops = 2.0 * g.nnz * N
    # exactness spot check: +-1 B against an integer recompute of 2000 rows
    S = np.random.default_rng(3).choice([-1, 1], size=(g.K, N)).astype(np.float32)
    Cs = api.spmm_b2sr(h, T, g.M, brp, bci, words, torch.from_numpy(S).to(dev), plan=bplan).cpu().numpy()
    rows = np.random.default_rng(4).choice(g.M, 2000, replace=False)
    exact = all(np.array_equal(Cs[i], S[g.colind[g.rowptr[i]:g.rowptr[i + 1]]].sum(axis=0)) for i in rows)
    err = sampled_rel_err(g, B.astype(np.float64), C.cpu().numpy(), rows)
    same = float((C - Ccsr).abs().max())
    return dict(g=g, B=B, T=T, N=N, nb=nb, nbytes=nbytes, flops=flops, ms_step=ms_step,
                ms_launch=ms_launch, conv_ms=conv_ms, exact=exact, err=err, csr_ms=csr_ms,
                csr_bytes=csr_bytes(g.M, g.nnz, N), csr_maxdiff=same, dropin_ms=dropin_ms,
                dropin_diff=dropin_diff)


def cpu_baseline(hd):
    import oracle

    g, B = hd["g"], hd["B"]
    C = np.empty((g.M, B.shape[1]), np.float32)
    for _ in range(3):
        oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B, out=C)
    ts = []
    for _ in range(15):
        t = time.perf_counter()
        oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B, out=C)
        ts.append(time.perf_counter() - t)
    ms = float(np.median(ts)) * 1e3
    return {
        "value": round(hd["nbytes"] / (ms * 1e-3) / 1e9, 2),
        "unit": "GB/s (algorithmic, same byte count)",
        "ms_per_step": round(ms, 3),
        "cores": oracle.omp_threads(),
        "kind": "port",
        "sample": f"OpenMP FP32 CSR SpMM (oracle/spmm_oracle.c orc_csr_spmm_f32_omp) on the full "
                  f"89,250-row graph, N={B.shape[1]}, values 1; 3 warm-ups, median of 15 runs",
    }


def extra_csr(args, h, dev):
    import torch

    from spmm_eval_amd import api, graphs

    out = {}
    # workload 1: 16,384 rows, N=32
    g, B = graphs.workload_csr_small()
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    plan = api.CsrPlan(h, rp, g.nnz)
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((g.M, 32), device=dev)
    geo = plan.autotune(rp, ci, v, Bd, C)
    # ~8 us a call: timed as back-to-back launches in a captured graph
    ms = time_graph(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=C, plan=plan), C, steps=200)
    nbytes = csr_bytes(g.M, g.nnz, 32)
    rows = np.random.default_rng(5).choice(g.M, 2000, replace=False)
    out["csr_f32_16k_n32"] = {"geometry": geo,
        "nnz": g.nnz, "ms": round(ms, 4), "GBps": round(nbytes / ms / 1e6, 1),
        "pct_hbm_peak": round(100 * nbytes / ms / 1e6 / HBM_PEAK_GBS, 1),
        "GFLOPs": round(2 * g.nnz * 32 / ms / 1e6, 1),
        "rel_err_fp64_sampled": sampled_rel_err(g, B.astype(np.float64), C.cpu().numpy(), rows)}
    del plan
    # workload 3: 232,965 rows, ~114.6M nnz
    t = time.time()
    g = graphs.workload_large_graph()
    log(f"large graph: {g.M} rows, {g.nnz} nnz, max degree {g.degrees().max()} ({time.time() - t:.1f}s)")
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    t = time.time()
    plan = api.CsrPlan(h, rp, g.nnz)
    analyze_ms = (time.time() - t) * 1e3
    rows = np.random.default_rng(6).choice(g.M, 1500, replace=False)
    for N in (64, 128):
        B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 3 + N)
        B64 = B.astype(np.float64)
        Bd = torch.from_numpy(B).to(dev)
        Bh = api.to_bf16_bits(Bd)
        C = torch.empty((g.M, N), device=dev)
        for name, BB, sb in (("f32", Bd, 4), ("bf16", Bh, 2)):
            geo = plan.autotune(rp, ci, v, BB, C)
            ms = time_loop(lambda: api.spmm_csr(h, rp, ci, v, BB, C_out=C, plan=plan), args.big_steps, 3)
            nbytes = csr_bytes(g.M, g.nnz, N, s_b=sb)
            out[f"csr_{name}_233k_n{N}"] = {"geometry": geo,
                "nnz": g.nnz, "ms": round(ms, 3), "GBps": round(nbytes / ms / 1e6, 1),
                "pct_hbm_peak": round(100 * nbytes / ms / 1e6 / HBM_PEAK_GBS, 1),
                "GFLOPs": round(2 * g.nnz * N / ms / 1e6, 1),
                "rel_err_fp64_sampled": sampled_rel_err(g, B64, C.cpu().numpy(), rows)}
        del Bd, Bh, C
    out["csr_233k_analyze_ms"] = round(analyze_ms, 2)
    del plan, rp, ci, v
    torch.cuda.empty_cache()
    return out


def extra_spmv_aligned(args, h, dev):
    """Aligned B2SR-8/16 repack and the binary SpMV over it on the headline
    graph, beside CSR SpMM at N = 1 (the csrmv the reference checks against)."""
    import torch

    from spmm_eval_amd import api, graphs

    out = {}
    g = graphs.workload_b2sr_graph()
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    rng = np.random.default_rng(11)
    xf = torch.from_numpy(rng.uniform(-1, 1, g.K).astype(np.float32)).to(dev)
    plan = api.CsrPlan(h, rp, g.nnz)
    y = torch.empty((g.M, 1), device=dev)
    ms_csr = time_graph(lambda: api.spmm_csr(h, rp, ci, v, xf[:, None], C_out=y, plan=plan), y)
    out["spmv_csr_f32_89k"] = {"ms": round(ms_csr, 5),
                               "GBps": round(csr_bytes(g.M, g.nnz, 1) / ms_csr / 1e6, 1)}
    del plan
    for T in (8, 16):
        t0 = time.time()
        brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, T)
        nrp, nci, nval = api.b2sr_aligned(h, T, brp, bci, words)
        torch.cuda.synchronize()
        conv_ms = (time.time() - t0) * 1e3
        MB, KB, nt = (g.M + T - 1) // T, (g.K + T - 1) // T, nci.numel()
        xb = torch.from_numpy(rng.integers(0, 1 << T, KB).astype(np.int32)).to(dev)
        xb = xb.to(torch.uint8) if T == 8 else xb.to(torch.int16).view(torch.uint16)
        cnt = torch.empty(g.M, dtype=torch.int32, device=dev)
        yb = torch.empty(MB, dtype=xb.dtype, device=dev)
        yf = torch.empty(g.M, device=dev)
        ms_bin = time_graph(lambda: api.spmv_b2sr_aligned(h, T, g.M, g.K, nrp, nci, nval, xb,
                                                          y_count=cnt, y_bin=yb), cnt)
        ms_f32 = time_graph(lambda: api.spmv_b2sr_aligned(h, T, g.M, g.K, nrp, nci, nval, xf,
                                                          y=yf), yf)
        struct = (MB + 1) * 4 + nt * 4 + nval.numel() * 4
        b_bin = struct + KB * T // 8 + g.M * 4 + MB * T // 8
        b_f32 = struct + g.K * 4 + g.M * 4
        out[f"spmv_b2sr{T}_aligned_89k"] = {
            "tiles_padded": nt, "tiles": int(bci.numel()), "convert_ms_host_clock": round(conv_ms, 2),
            "bin_ms": round(ms_bin, 5), "bin_GBps": round(b_bin / ms_bin / 1e6, 1),
            "f32_ms": round(ms_f32, 5), "f32_GBps": round(b_f32 / ms_f32 / 1e6, 1)}
        del brp, bci, words, nrp, nci, nval
    torch.cuda.empty_cache()
    return out


def extra_b2sr_grid(args, h, dev):
    """The B2SR workload's grid on the headline graph: planned FBF SpMM for
    tiles 4/8/16/32 (and 2/64) x hidden 32/64/128, plus N = 1 (the
    narrow kernel), beside CSR at the same N; hipGraph-replay timing."""
    import torch

    from spmm_eval_amd import api, graphs

    g = graphs.workload_b2sr_graph()
    rp, ci = torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)
    v = torch.ones(g.nnz, device=dev)
    cp = api.CsrPlan(h, rp, g.nnz)
    widths = (1, 32, 64, 128)
    Bs = {N: torch.from_numpy(graphs.uniform_pm1((g.K, N), seed=N + 17)).to(dev) for N in widths}
    Cs = {N: torch.empty((g.M, N), device=dev) for N in widths}
    out = {"csr_us": {}, "b2sr_us": {}}
    for N in widths:
        ms = time_graph(lambda: api.spmm_csr(h, rp, ci, v, Bs[N], C_out=Cs[N], plan=cp), Cs[N],
                        steps=50, replays=3)
        out["csr_us"][f"N{N}"] = round(ms * 1e3, 2)
    for T in (2, 4, 8, 16, 32, 64):
        brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, T)
        plan = api.B2srPlan(h, T, g.M, brp)
        for N in widths:
            ms = time_graph(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bs[N], C_out=Cs[N],
                                                  plan=plan), Cs[N], steps=50, replays=3)
            out["b2sr_us"][f"T{T}_N{N}"] = round(ms * 1e3, 2)
        plan.close()
        del brp, bci, words
    torch.cuda.empty_cache()
    return {"b2sr_grid_89k": out}


def extra_fwd_bwd(args, h, dev):
    import torch

    from spmm_eval_amd import api, graphs

    g, X, dY = graphs.workload_gcn(N=128)
    N = X.shape[1]
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    trp, tci, tval, _ = api.csr_transpose(h, rp, ci, v, g.K)
    pf, pb = api.CsrPlan(h, rp, g.nnz), api.CsrPlan(h, trp, g.nnz)
    Xd, dYd = torch.from_numpy(X).to(dev), torch.from_numpy(dY).to(dev)
    Xh, dYh = api.to_bf16_bits(Xd), api.to_bf16_bits(dYd)
    Y = torch.empty((g.M, N), device=dev)
    dX = torch.empty((g.K, N), device=dev)
    dval = torch.empty(g.nnz, device=dev)

    def fwd(XX):
        api.spmm_csr(h, rp, ci, v, XX, C_out=Y, plan=pf)

    def bwd_x(GG):
        api.spmm_csr(h, trp, tci, tval, GG, K=g.M, C_out=dX, plan=pb)

    def bwd_v(XX, GG):
        api.sddmm_csr(h, rp, ci, GG, XX, out=dval, plan=pf)

    def step(XX, GG):
        fwd(XX)
        bwd_x(GG)
        bwd_v(XX, GG)

    # the backward as one pass over A's transposed plan (dX and dval together)
    pt = api.CsrTPlan(h, rp, ci, g.K)
    dX2 = torch.empty((g.K, N), device=dev)
    dval2 = torch.empty(g.nnz, device=dev)

    def bwd_fused(XX, GG, tune):
        api.spmm_csr_t_sddmm(pt, rp, ci, v, GG, XX, dX_out=dX2, dval_out=dval2, tune=tune)

    def tune_fused(XX, GG):
        lanes = N // (16 // XX.element_size())
        cands = [(lanes // vec, vec, rs) for vec in (1, 2) for rs in (1, 2, 4)
                 if lanes % vec == 0 and lanes // vec <= 64]
        return min(cands, key=lambda t: time_loop(lambda: bwd_fused(XX, GG, t), 30, 3))

    res = {}
    for name, XX, GG, s in (("f32", Xd, dYd, 4), ("bf16", Xh, dYh, 2)):
        geo = (pf.autotune(rp, ci, v, XX, Y), pb.autotune(trp, tci, tval, GG, dX, K=g.M))
        ms = time_loop(lambda: step(XX, GG), 100, 10)
        parts = {"fwd_ms": round(time_loop(lambda: fwd(XX), 100, 10), 4),
                 "dX_ms": round(time_loop(lambda: bwd_x(GG), 100, 10), 4),
                 "sddmm_ms": round(time_loop(lambda: bwd_v(XX, GG), 100, 10), 4)}
        tf = tune_fused(XX, GG)
        fused_ms = time_loop(lambda: (fwd(XX), bwd_fused(XX, GG, tf)), 100, 10)
        torch.cuda.synchronize()
        parts["fused"] = {"step_ms": round(fused_ms, 4),
                          "bwd_ms": round(time_loop(lambda: bwd_fused(XX, GG, tf), 100, 10), 4),
                          "geometry": list(tf),
                          "max_abs_diff_dX": float((dX2 - dX).abs().max()),
                          "max_rel_diff_dval": float(((dval2 - dval).abs().max()
                                                      / dval.abs().max().clamp_min(1e-30)))}
        nbytes = 2 * csr_bytes(g.M, g.nnz, N, s_b=s) + sddmm_bytes(g.M, g.nnz, N, s)
        rows = np.random.default_rng(7).choice(g.M, 1500, replace=False)
        Yc, dvc = Y.cpu().numpy(), dval.cpu().numpy()
        e_y = sampled_rel_err(g, X.astype(np.float64), Yc, rows)
        num = den = 0.0
        for i in rows:
            s0, e0 = g.rowptr[i], g.rowptr[i + 1]
            ref = X[g.colind[s0:e0]].astype(np.float64) @ dY[i].astype(np.float64)
            num += float(((dvc[s0:e0] - ref) ** 2).sum())
            den += float((ref ** 2).sum())
        res[name] = {"ms": round(ms, 4), "GBps": round(nbytes / ms / 1e6, 1), "geometry": geo, **parts,
                     "rel_err_Y_fp64": e_y, "rel_err_dval_fp64": (num / den) ** 0.5}
    res["bf16_speedup_vs_f32"] = round(res["f32"]["ms"] / res["bf16"]["ms"], 3)
    # BF16 step with the fused backward against the plain FP32 step (three
    # separate FP32 launches) and against the FP32 step with the fused backward
    res["bf16_fused_vs_f32_plain"] = round(res["f32"]["ms"] / res["bf16"]["fused"]["step_ms"], 3)
    res["bf16_fused_vs_f32_fused"] = round(res["f32"]["fused"]["step_ms"]
                                           / res["bf16"]["fused"]["step_ms"], 3)
    return {"gcn_fwd_bwd_89k_n128": res}


def extra_batched(args, h, dev):
    import torch

    from spmm_eval_amd import api, graphs
    from spmm_eval_amd.batched import BatchPlan

    gs = graphs.workload_batched(512)
    N = 64
    plan = BatchPlan(h, gs, N=N, bf16=True, device=dev)
    M = plan.M
    X = torch.from_numpy(graphs.normal((M, N), 21, 1)).to(dev)
    dY = torch.from_numpy(graphs.normal((M, N), 21, 2)).to(dev)
    Xh, dYh = api.to_bf16_bits(X), api.to_bf16_bits(dY)
    ms_graph = time_loop(lambda: plan.step(Xh, dYh, use_graph=True), 50, 5)
    # 512 separate plain FP32 launches per operation
    per = []
    ro = plan.row_off
    for i, g in enumerate(gs):
        rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
        trp, tci, tval, _ = api.csr_transpose(h, rp, ci, v, g.K)
        per.append((g, rp, ci, v, trp, tci, tval, api.CsrPlan(h, rp, g.nnz), api.CsrPlan(h, trp, g.nnz),
                    int(ro[i]), int(ro[i + 1]), int(plan.nnz_off[i]), int(plan.nnz_off[i + 1])))
    Y = torch.empty((M, N), device=dev)
    dX = torch.empty((M, N), device=dev)
    dval = torch.empty(plan.nnz, device=dev)

    def separate():
        for g, rp, ci, v, trp, tci, tval, pf, pb, r0, r1, e0, e1 in per:
            api.spmm_csr(h, rp, ci, v, X[r0:r1], C_out=Y[r0:r1], plan=pf)
            api.spmm_csr(h, trp, tci, tval, dY[r0:r1], K=g.M, C_out=dX[r0:r1], plan=pb)
            api.sddmm_csr(h, rp, ci, dY[r0:r1], X[r0:r1], out=dval[e0:e1])

    ms_sep = time_loop(separate, 5, 1)
    info = plan.info()
    # each SDDMM form at the tuned geometry (form 2 only on one column tile)
    forms = {}
    for form in (0, 1, 2):
        if form == 2 and info["lanes_per_row"] * info["vec"] * 8 != N:
            continue
        plan.set_geometry(info["lanes_per_row"], info["vec"], info["rows_scale"],
                          info["persistent"], form)
        forms[str(form)] = round(time_loop(lambda: plan.step(Xh, dYh, use_graph=True), 20, 3), 4)
    plan.close()
    return {"batched_512_fwd_bwd_n64": {"fused_launch": info, "sddmm_form_ms": forms,
        "graphs": len(gs), "rows": M, "nnz": plan.nnz,
        "bf16_one_graph_ms": round(ms_graph, 4), "f32_separate_launches_ms": round(ms_sep, 3),
        "speedup": round(ms_sep / ms_graph, 2)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--tile", type=int, default=4)
    ap.add_argument("--hidden", type=int, default=64)
    ap.add_argument("--big-steps", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="headline only (no extra workloads)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", choices=["csr", "fwd_bwd", "batched", "spmv_aligned", "b2sr_grid"], default=None,
                    help="run one extra workload alone and print its result (for profiling)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="run only a device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    if args.copy_calib:
        n = args.copy_calib * (1 << 20) // 4
        a = torch.ones(n, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_MiB": args.copy_calib}))
        return

    from spmm_eval_amd import api

    h = api.Handle()
    if args.only:
        fn = {"csr": extra_csr, "fwd_bwd": extra_fwd_bwd, "batched": extra_batched,
              "spmv_aligned": extra_spmv_aligned, "b2sr_grid": extra_b2sr_grid}[args.only]
        print(json.dumps(fn(args, h, dev)))
        h.close()
        return
    log(f"headline: B2SR-{args.tile} FBF SpMM, N={args.hidden}")
    hd = headline(args, h, dev)
    gbs = hd["nbytes"] / (hd["ms_step"] * 1e-3) / 1e9
    ach = hd["nbytes"] / (hd["ms_launch"] * 1e-3) / 1e9
    kbuf = "true" if hd["g"].K * hd["N"] * 4 < 2 ** 31 else "false"  # B through a buffer resource
    if hd["T"] <= 8:  # block-row kernels: narrow tiles, wide tiles
        kname = f"b2sr_brow_kernel<{hd['T']}, float, {b2sr_lanes(hd['N'])}, {kbuf}>"
    else:
        kname = f"b2sr_wide_kernel<{hd['T']}, float, {min(32, b2sr_lanes(hd['N']))}, {kbuf}>"
    traffic = profiled_traffic(f"b2sr_fbf_t{hd['T']}_n{hd['N']}_89k.md", kname)
    result = {
        "metric": "SpMM HBM GB/s (algorithmic) -- B2SR-4 FBF SpMM, FP32 B",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hd["ms_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Chung-Lu power-law graph, uniform[-1,1] B)",
        "config": {"workload": "b2sr_fbf_spmm", "rows": hd["g"].M, "nnz": hd["g"].nnz,
                   "tile": hd["T"], "nnzb": hd["nb"], "hidden": hd["N"], "parallelism": "single GPU"},
        "gflops": round(hd["flops"] / (hd["ms_step"] * 1e-3) / 1e9, 1),
        "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/ counter summary (FETCH_SIZE x2 + WRITE_SIZE x1, "
                                       "factors from a 256 MiB copy)" if traffic else None,
                     "kernel": kname,
                     "bytes_per_launch": hd["nbytes"], "mean_launch_ms": round(hd["ms_launch"], 5)},
        "checks": {"pm1_exact_2000_rows": bool(hd["exact"]), "rel_err_fp64_sampled": hd["err"]},
        "csr2b2sr_ms": round(hd["conv_ms"], 2),
        "dropin_unplanned": {"ms": round(hd["dropin_ms"], 5),
                             "GBps": round(hd["nbytes"] / hd["dropin_ms"] / 1e6, 1),
                             "max_abs_diff_vs_planned": hd["dropin_diff"]},
        "csr_same_graph": {"ms": round(hd["csr_ms"], 5),
                           "GBps": round(hd["csr_bytes"] / hd["csr_ms"] / 1e6, 1),
                           "max_abs_diff_vs_b2sr": hd["csr_maxdiff"]},
    }
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(hd)
    if not args.quick:
        extra = {}
        for fn in (extra_csr, extra_fwd_bwd, extra_batched, extra_spmv_aligned, extra_b2sr_grid):
            log(fn.__name__)
            try:
                extra.update(fn(args, h, dev))
            except Exception as e:  # report and keep the headline line
                extra[fn.__name__] = {"error": repr(e)}
        result["extra"] = extra
    h.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
