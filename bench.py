"""SpMM benchmark on one MI355X.

Headline step (`value`): CSR FP32 SpMM C = A B -- the cuSPARSE call of
evalCSRSpmmFloatCuSPARSE (test_spmm_full.cu:212-342) that the library
replaces -- on the largest workload that fits one GPU: the seeded
232,965-row power-law graph (~114.5M nnz, heaviest row ~17k entries,
values U(0,1] row-normalised), B 232,965 x 128 FP32 U[-1,1], inputs
resident in HBM, row schedule built and autotuned once outside the timed
region (as the reference uploads its operands before timing).

Byte counts. `value` and `roofline` use the COMPULSORY bytes of a launch:
every array read or written once (rowptr, colind, val, B, C). That is a
lower bound on HBM traffic, so `roofline.frac` <= 1 holds by construction.
The gather-model count (every stored entry reads one full B row; the
per-entry figure of the measurement plan) is reported as
`effective_GBps_gather_model`: B rows served by L2 or the MALL count there
each time, so it exceeds the HBM peak and is not an HBM figure.

Timing. The timed region is `--steps` launches after `--warmup` between two
events on the launch stream (ms_per_step). Beside it every GPU figure is also
the median of 5 batches of >= 100 back-to-back launches after 10 warm-ups
(`median_of_5`), and kernels too short for host launch pacing are timed as
captured hipGraph replays (median of 5 replays of 100-200 launches).

`cpu_baseline`: the oracle library's OpenMP FP32 CSR SpMM on the same graph
and B on all host threads (3 warm-ups, median of 10). `extra.cpu` holds the
same kind of baseline for every other workload (forward, CPU dX and CPU
SDDMM for the fwd+bwd comparisons).

`extra` (skipped with --quick): the same graph at N = 64 and in BF16, and its
A^T dY (the backward's dX) through the transposed plan; the
16,384-row N=32 workload; B2SR-4 FBF SpMM on the 89,250-row 0/1 graph
(planned and plan-free) with the tile x N grid, the sign-packed popcount
SpMM and the dense packers; BMM32 (XNOR/popcount matmul); GCN-style
fwd+bwd (Y = A X, dX = A^T dY, dval = SDDMM) in FP32 and BF16; 512 uneven
graphs batched in one captured launch vs 512 separate FP32 launches; aligned
B2SR-8/16 binary SpMV. GPU errors there are against an FP64 numpy
recomputation on sampled rows (the heaviest rows always included).
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
def csr_gather_bytes(M, nnz, N, s_val=4, s_b=4):
    """Gather model: every stored entry reads one full B row."""
    return 4 * (M + 1) + nnz * (4 + s_val) + nnz * N * s_b + M * N * 4


def csr_compulsory_bytes(M, K, nnz, N, s_b=4):
    """Each array once: rowptr, colind, val, B, C."""
    return 4 * (M + 1) + 8 * nnz + K * N * s_b + M * N * 4


TILE_BYTES = {2: 2, 4: 4, 8: 8, 16: 32, 32: 128, 64: 512}


def b2sr_gather_bytes(M, T, nb, nnz, N, s_b=4):
    MB = (M + T - 1) // T
    return 4 * (MB + 1) + nb * (4 + TILE_BYTES[T]) + nnz * N * s_b + M * N * 4


def b2sr_compulsory_bytes(M, K, T, nb, N, s_b=4):
    MB = (M + T - 1) // T
    return 4 * (MB + 1) + nb * (4 + TILE_BYTES[T]) + K * N * s_b + M * N * 4


def sddmm_gather_bytes(M, nnz, N, s=4):
    return 4 * (M + 1) + 8 * nnz + nnz * N * s + M * N * s


def rate(nbytes, ms):
    return round(nbytes / (ms * 1e-3) / 1e9, 1)


def profiled_traffic(profile: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from a committed counter summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by its calibration factor);
    None when the summary has no row for that kernel."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    key = f"`{kernel}`"
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 3 and cells[0] == key:
            try:
                return int(cells[-1])
            except ValueError:
                return None
    return None


# ----------------------------------------------------------------- timing
def _events():
    import torch

    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def time_region(fn, steps, warmup):
    """The contract's timed region: `warmup` untimed launches, a sync, then
    exactly `steps` launches between two events on the launch stream (the
    library launches on torch's current stream) and a sync. ms per step."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    t0, t1 = _events()
    t0.record(s)
    for _ in range(steps):
        fn()
    t1.record(s)
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / steps


def median_batches(fn, launches=100, batches=5, warmup=10):
    """Median over `batches` batches of `launches` back-to-back launches
    (events on the launch stream around each batch), after `warmup`
    launches. Returns (median ms per launch, [ms per launch of each batch])."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    per = []
    for _ in range(batches):
        t0, t1 = _events()
        t0.record(s)
        for _ in range(launches):
            fn()
        t1.record(s)
        torch.cuda.synchronize()
        per.append(t0.elapsed_time(t1) / launches)
    return float(np.median(per)), [round(x, 5) for x in per]


def time_graph(fn, out, steps=100, replays=5):
    """Kernels too short for host launch pacing: `steps` back-to-back calls
    captured in one hipGraph, replayed `replays` times, each replay between
    two events; median ms per call. Every replay's `out` must equal the
    eager result."""
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
    per = []
    for _ in range(replays):
        out.fill_(0)
        torch.cuda.synchronize()
        t0, t1 = _events()
        t0.record()
        g.replay()
        t1.record()
        torch.cuda.synchronize()
        per.append(t0.elapsed_time(t1) / steps)
        if not torch.equal(out, ref):
            raise RuntimeError("graph replay differs from the eager result")
    return float(np.median(per))


def cpu_time(fn, warmup=3, reps=10):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)) * 1e3


def check_rows(g, rng_seed, n=1500, heavy=16):
    """Rows for the sampled FP64 check: the `heavy` heaviest rows plus `n`
    random ones."""
    deg = g.degrees()
    top = np.argsort(deg)[-heavy:]
    rnd = np.random.default_rng(rng_seed).choice(g.M, n, replace=False)
    return np.unique(np.concatenate([top, rnd]))


def sampled_rel_err(g, B64, C, rows):
    """Relative Frobenius error of C on `rows` against an FP64 recompute."""
    num = den = 0.0
    for i in rows:
        s, e = g.rowptr[i], g.rowptr[i + 1]
        ref = g.val[s:e].astype(np.float64) @ B64[g.colind[s:e]]
        num += float(((C[i] - ref) ** 2).sum())
        den += float((ref ** 2).sum())
    return (num / den) ** 0.5 if den > 0 else 0.0


def bf16_round(x):
    """FP32 -> BF16 (round to nearest even) -> FP32, for the error reference
    of BF16 runs."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


# ------------------------------------------------------------------ parts
def headline(args, h, dev):
    import torch

    from spmm_eval_amd import api, graphs

    N = args.hidden
    t = time.time()
    g = graphs.workload_large_graph()
    log(f"large graph: {g.M} rows, {g.nnz} nnz, max degree {int(g.degrees().max())} "
        f"({time.time() - t:.1f}s)")
    B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 3 + N)
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((g.M, N), device=dev)
    t = time.time()
    plan = api.CsrPlan(h, rp, g.nnz)
    torch.cuda.synchronize()
    analyze_ms = (time.time() - t) * 1e3
    if args.geometry:  # a fixed geometry (profiling runs: no autotuner dispatches in the trace)
        geo = tuple(int(x) for x in args.geometry.split(","))
        tune = geo
    else:
        geo = plan.autotune(rp, ci, v, Bd, C)
        tune = None

    def step():
        api.spmm_csr(h, rp, ci, v, Bd, C_out=C, plan=plan, tune=tune)

    ms_step = time_region(step, args.steps, args.warmup)
    ms_med, batches = median_batches(step, launches=args.batch_launches)
    rows = check_rows(g, 6)
    err = sampled_rel_err(g, B.astype(np.float64), C.cpu().numpy(), rows)
    # the plan-free drop-in (spmm_csr_f32: schedule built inside the call)
    dropin_ms = dropin_err = None
    if not args.no_dropin:
        Cd = torch.empty((g.M, N), device=dev)
        dropin_ms, _ = median_batches(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=Cd),
                                      launches=20, warmup=3)
        dropin_err = sampled_rel_err(g, B.astype(np.float64), Cd.cpu().numpy(), rows)
    return dict(g=g, B=B, N=N, geo=geo, plan=plan, dev_arrays=(rp, ci, v), ms_step=ms_step,
                ms_med=ms_med, batches=batches, err=err, rows=rows, analyze_ms=analyze_ms,
                dropin_ms=dropin_ms, dropin_err=dropin_err)


def cpu_baseline(hd):
    import oracle

    g, B = hd["g"], hd["B"]
    C = np.empty((g.M, B.shape[1]), np.float32)
    ms = cpu_time(lambda: oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B, out=C))
    nbytes = csr_compulsory_bytes(g.M, g.K, g.nnz, B.shape[1])
    return {
        "value": rate(nbytes, ms),
        "unit": "GB/s (compulsory bytes, same count as value)",
        "ms_per_step": round(ms, 3),
        "cores": oracle.omp_threads(),
        "kind": "port",
        "sample": f"the whole workload: OpenMP FP32 CSR SpMM (oracle/spmm_oracle.c "
                  f"orc_csr_spmm_f32_omp, dynamic chunks of 64 rows) on the {g.M}-row "
                  f"{g.nnz}-nnz graph, N={B.shape[1]}, same B; 3 warm-ups, median of 10",
        "gpu_speedup": round(ms / hd["ms_step"], 1),
    }


def extra_large_variants(args, h, dev, hd):
    """The headline graph at the other widths and precision: FP32 N=64, BF16
    N=64 and N=128 (FP32 accumulate), each autotuned, with CPU baselines; its
    A^T dY through the transposed plan; and B2SR-4 FBF on its pattern."""
    import torch

    import oracle
    from spmm_eval_amd import api, graphs

    g = hd["g"]
    rp, ci, v = hd["dev_arrays"]
    plan = hd["plan"]
    rows = hd["rows"]
    out, cpu = {}, {}
    for N in (64, 128):
        B = hd["B"] if N == hd["N"] else graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 3 + N)
        Bd = torch.from_numpy(B).to(dev)
        Bh = api.to_bf16_bits(Bd)
        C = torch.empty((g.M, N), device=dev)
        for name, BB, sb in (("f32", Bd, 4), ("bf16", Bh, 2)):
            if name == "f32" and N == hd["N"]:
                continue  # the headline itself
            geo = plan.autotune(rp, ci, v, BB, C)
            ms, _ = median_batches(lambda: api.spmm_csr(h, rp, ci, v, BB, C_out=C, plan=plan),
                                   launches=100)
            Bref = (bf16_round(B) if name == "bf16" else B).astype(np.float64)
            out[f"csr_{name}_233k_n{N}"] = {
                "geometry": geo, "ms_median_of_5": round(ms, 4),
                "compulsory_GBps": rate(csr_compulsory_bytes(g.M, g.K, g.nnz, N, sb), ms),
                "effective_GBps_gather_model": rate(csr_gather_bytes(g.M, g.nnz, N, s_b=sb), ms),
                "GFLOPs": round(2 * g.nnz * N / ms / 1e6, 1),
                "rel_err_fp64_sampled": sampled_rel_err(g, Bref, C.cpu().numpy(), rows),
                "rel_err_vs_fp32_source_sampled": sampled_rel_err(g, B.astype(np.float64),
                                                                  C.cpu().numpy(), rows)}
            if name == "bf16":
                Bhn = Bh.cpu().numpy().view(np.uint16)
                Cc = np.empty((g.M, N), np.float32)
                cms = cpu_time(lambda: oracle.csr_spmm_bf16_omp(g.rowptr, g.colind, g.val, Bhn, out=Cc),
                               reps=5)
            else:
                Cc = np.empty((g.M, N), np.float32)
                cms = cpu_time(lambda: oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B, out=Cc),
                               reps=5)
            cpu[f"csr_{name}_233k_n{N}"] = {"ms": round(cms, 2), "gpu_speedup": round(cms / ms, 1)}
        del Bd, Bh, C
    # the backward's dX = A^T dY on the same graph: the transposed plan (values read through
    # its entry permutation), autotuned, column slices included; checked on sampled rows of
    # an FP64 recompute over scipy's transpose
    import scipy.sparse as sps

    tp = api.CsrTPlan(h, rp, ci, g.K)
    At = sps.csr_matrix((g.val, g.colind, g.rowptr), shape=(g.M, g.K)).T.tocsr()
    gt = graphs.Csr(At.indptr.astype(np.int32), At.indices.astype(np.int32),
                    At.data.astype(np.float32), g.K, g.M)
    del At
    rows_t = check_rows(gt, 9)
    N = 128
    dY = graphs.uniform_pm1((g.M, N), seed=graphs.SEED_BASE + 7)
    dYd = torch.from_numpy(dY).to(dev)
    dX = torch.empty((g.K, N), device=dev)
    for name, dYx, sb in (("f32", dYd, 4), ("bf16", api.to_bf16_bits(dYd), 2)):
        geo = tp.autotune(v, dYx, dX)
        ms, _ = median_batches(lambda: api.spmm_csr_t(h, rp, ci, v, dYx, g.K, C_out=dX, plan=tp),
                               launches=20)
        ref = (bf16_round(dY) if name == "bf16" else dY).astype(np.float64)
        out[f"csr_t_{name}_233k_n{N}"] = {
            "geometry": geo, "ms_median_of_5": round(ms, 4),
            # each array once, the permutation included
            "compulsory_GBps": rate(csr_compulsory_bytes(g.K, g.M, g.nnz, N, sb) + 4 * g.nnz, ms),
            "rel_err_fp64_sampled": sampled_rel_err(gt, ref, dX.cpu().numpy(), rows_t)}
    tp.close()
    del dYd, dX, gt
    torch.cuda.empty_cache()
    # B2SR-4 FBF on the same graph's pattern (binary A): the reddit-sized workload of the
    # reference's only published figures for this path (backend/spmm.cu:869-870,
    # spmm4_full_full_new2, one un-warmed launch on an unstated sm_70 GPU)
    T = 4
    gb = graphs.Csr(g.rowptr, g.colind, np.ones(g.nnz, np.float32), g.M, g.K)
    torch.cuda.synchronize()
    t0, t1 = _events()
    t0.record()
    brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, T)
    t1.record()
    torch.cuda.synchronize()
    conv_ms = t0.elapsed_time(t1)
    nb = bci.numel()
    bplan = api.B2srPlan(h, T, g.M, brp)
    for N, pub in ((64, 97.027), (128, 166.101)):
        B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 11 + N)
        Bd = torch.from_numpy(B).to(dev)
        C = torch.empty((g.M, N), device=dev)
        choice = bplan.autotune(brp, bci, words, Bd, C)
        ms, _ = median_batches(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bd, C_out=C,
                                                     plan=bplan), launches=20)
        err = sampled_rel_err(gb, B.astype(np.float64), C.cpu().numpy(), rows)
        ms_d, _ = median_batches(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bd, C_out=C),
                                 launches=20)
        err_d = sampled_rel_err(gb, B.astype(np.float64), C.cpu().numpy(), rows)
        S = np.random.default_rng(N).choice([-1, 1], size=(g.K, N)).astype(np.float32)
        Cs = api.spmm_b2sr(h, T, g.M, brp, bci, words, torch.from_numpy(S).to(dev),
                           plan=bplan).cpu().numpy()
        exact = all(np.array_equal(Cs[i], S[g.colind[g.rowptr[i]:g.rowptr[i + 1]]].sum(axis=0))
                    for i in rows)
        out[f"b2sr4_fbf_233k_n{N}"] = {
            "nnzb": nb, "csr2b2sr_ms_device": round(conv_ms, 2), "tuned": list(choice),
            "planned_ms_median_of_5": round(ms, 4), "dropin_ms_median_of_5": round(ms_d, 4),
            "compulsory_GBps": rate(b2sr_compulsory_bytes(g.M, g.K, T, nb, N), ms),
            "effective_GBps_gather_model": rate(b2sr_gather_bytes(g.M, T, nb, g.nnz, N), ms),
            "pm1_exact_sampled_rows": bool(exact), "rel_err_fp64_sampled": err,
            "dropin_rel_err_fp64_sampled": err_d,
            "reference_published_ms": pub,
            "reference_published_on": "reddit, unstated sm_70 GPU, one un-warmed launch"}
        del Bd, C
    bplan.close()
    del brp, bci, words
    torch.cuda.empty_cache()
    out["cpu"] = cpu
    return out


def extra_csr_small(args, h, dev):
    """16,384 rows, ~260k nnz, values 1, B 16,384 x 32 (config 1)."""
    import torch

    import oracle
    from spmm_eval_amd import api, graphs

    g, B = graphs.workload_csr_small()
    rp, ci, v = (torch.from_numpy(x).to(dev) for x in (g.rowptr, g.colind, g.val))
    plan = api.CsrPlan(h, rp, g.nnz)
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((g.M, 32), device=dev)
    geo = plan.autotune(rp, ci, v, Bd, C)
    # ~7 us a call: back-to-back launches in a captured graph
    ms = time_graph(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=C, plan=plan), C, steps=200)
    Cd = torch.empty((g.M, 32), device=dev)
    ms_d, _ = median_batches(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=Cd), launches=100)
    # the plan-free drop-in is asynchronous and capturable: also timed like the planned form
    ms_dg = time_graph(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=Cd), Cd, steps=200)
    dropin = {"ms_eager_median_of_5": round(ms_d, 5), "ms_graph_replay_median_of_5": round(ms_dg, 5),
              "rel_err_fp64_sampled": None}
    rows = check_rows(g, 5, n=2000)
    dropin["rel_err_fp64_sampled"] = sampled_rel_err(g, B.astype(np.float64), Cd.cpu().numpy(), rows)
    nb = csr_gather_bytes(g.M, g.nnz, 32)
    Cc = np.empty((g.M, 32), np.float32)
    cms = cpu_time(lambda: oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B, out=Cc))
    return {"csr_f32_16k_n32": {
        "geometry": geo, "nnz": g.nnz, "ms_graph_replay_median_of_5": round(ms, 5),
        "effective_GBps_gather_model": rate(nb, ms),
        "pct_of_8TBps_gather_model": round(100 * nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 1),
        "compulsory_GBps": rate(csr_compulsory_bytes(g.M, g.K, g.nnz, 32), ms),
        "GFLOPs": round(2 * g.nnz * 32 / ms / 1e6, 1),
        "rel_err_fp64_sampled": sampled_rel_err(g, B.astype(np.float64), C.cpu().numpy(), rows),
        "dropin": dropin,
        "cpu": {"ms": round(cms, 4), "gpu_speedup": round(cms / ms, 1)}}}


def extra_b2sr(args, h, dev):
    """The reference harness's own step -- B2SR-4 FBF SpMM
    (spmm4_full_full_new2, test_spmm_full.cu:344-369) -- on the 89,250-row
    0/1 graph, planned and plan-free, beside CSR on the same graph; the tile x
    N grid; the sign-packed popcount SpMM (int32 counts); the packers."""
    import torch

    import oracle
    from spmm_eval_amd import api, graphs

    g = graphs.workload_b2sr_graph()
    rp, ci = torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)
    v = torch.ones(g.nnz, device=dev)
    out = {}
    # csr2b2sr conversion (both passes + the count read-back), after a first
    # call that sizes the handle's workspace; median of 5 calls
    brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, 4)
    conv = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0, t1 = _events()
        t0.record()
        brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, 4)
        t1.record()
        torch.cuda.synchronize()
        conv.append(t0.elapsed_time(t1))
    conv_ms = float(np.median(conv))
    N, T = 64, 4
    nb = bci.numel()
    B = graphs.uniform_pm1((g.K, N), seed=graphs.SEED_BASE + 2)
    Bd = torch.from_numpy(B).to(dev)
    C = torch.empty((g.M, N), device=dev)
    bplan = api.B2srPlan(h, T, g.M, brp)
    ms = time_graph(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bd, C_out=C, plan=bplan), C)
    Cd = torch.empty((g.M, N), device=dev)
    ms_d = time_graph(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bd, C_out=Cd), Cd)
    cplan = api.CsrPlan(h, rp, g.nnz)
    Cc = torch.empty((g.M, N), device=dev)
    ms_c = time_graph(lambda: api.spmm_csr(h, rp, ci, v, Bd, C_out=Cc, plan=cplan), Cc)
    S = np.random.default_rng(3).choice([-1, 1], size=(g.K, N)).astype(np.float32)
    Cs = api.spmm_b2sr(h, T, g.M, brp, bci, words, torch.from_numpy(S).to(dev), plan=bplan).cpu().numpy()
    rows = check_rows(g, 4, n=2000)
    exact = all(np.array_equal(Cs[i], S[g.colind[g.rowptr[i]:g.rowptr[i + 1]]].sum(axis=0)) for i in rows)
    cb = b2sr_compulsory_bytes(g.M, g.K, T, nb, N)
    out["b2sr4_fbf_89k_n64"] = {
        "nnzb": nb, "planned_ms_graph_replay": round(ms, 5),
        "compulsory_GBps": rate(cb, ms), "effective_GBps_gather_model":
            rate(b2sr_gather_bytes(g.M, T, nb, g.nnz, N), ms),
        "dropin_ms_graph_replay": round(ms_d, 5),
        "dropin_max_abs_diff_vs_planned": float((C - Cd).abs().max()),
        "csr_same_graph_ms": round(ms_c, 5),
        "csr2b2sr_ms_device": round(conv_ms, 3),
        "pm1_exact_sampled_rows": bool(exact),
        "rel_err_fp64_sampled": sampled_rel_err(g, B.astype(np.float64), C.cpu().numpy(), rows)}
    Cn = np.empty((g.M, N), np.float32)
    out["b2sr4_fbf_89k_n64"]["cpu_csr_f32_same_graph_ms"] = round(
        cpu_time(lambda: oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B, out=Cn)), 3)
    del bplan, cplan
    # sign-packed popcount SpMM (B signs packed 32 to a word, int32 counts)
    sign = {}
    for NN in (32, 64, 128):
        Bs = torch.from_numpy(graphs.uniform_pm1((g.K, NN), seed=NN + 31)).to(dev)
        Bw = api.pack_sign_rows(Bs)
        for TT in (4, 8, 16, 32):
            brp2, bci2, w2 = (brp, bci, words) if TT == 4 else api.csr2b2sr(h, rp, ci, g.M, g.K, TT)
            Ci = torch.empty((g.M, NN), dtype=torch.int32, device=dev)
            ms_s = time_graph(lambda: api.spmm_b2sr_sign(h, TT, g.M, g.K, NN, brp2, bci2, w2, Bw,
                                                         C_out=Ci), Ci)
            nbb = bci2.numel()
            MB = (g.M + TT - 1) // TT
            comp = 4 * (MB + 1) + nbb * (4 + TILE_BYTES[TT]) + g.K * NN // 8 + g.M * NN * 4
            sign[f"T{TT}_N{NN}"] = {"ms": round(ms_s, 5), "compulsory_GBps": rate(comp, ms_s)}
    out["b2sr_signpacked_89k"] = sign
    # dense packers on the 89,250 x 128 B
    Bs = torch.from_numpy(graphs.uniform_pm1((g.K, 128), seed=77)).to(dev)
    W = torch.empty((g.K, 4), dtype=torch.uint32, device=dev)
    ms_p = time_graph(lambda: W.copy_(api.pack_sign_rows(Bs)), W)
    out["pack_sign_rows_89k_x128"] = {"ms_incl_copy": round(ms_p, 5)}
    return out


def extra_b2sr_grid(args, h, dev):
    """Planned FBF SpMM for tiles 2..64 x N = 1/32/64/128 on the 89k graph
    (N = 1 runs the narrow kernel), beside CSR at the same N; then the B2SR
    autotuner's pick for each (T, N) (variant, lanes per row, rows_scale)
    and its time."""
    import torch

    from spmm_eval_amd import api, graphs

    g = graphs.workload_b2sr_graph()
    rp, ci = torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev)
    v = torch.ones(g.nnz, device=dev)
    cp = api.CsrPlan(h, rp, g.nnz)
    widths = tuple(int(x) for x in args.grid_widths.split(","))
    Bs = {N: torch.from_numpy(graphs.uniform_pm1((g.K, N), seed=N + 17)).to(dev) for N in widths}
    Cs = {N: torch.empty((g.M, N), device=dev) for N in widths}
    out = {"csr_us": {}, "b2sr_us": {}, "b2sr_tuned_us": {}, "b2sr_tuned_choice": {}}
    for N in widths:
        ms = time_graph(lambda: api.spmm_csr(h, rp, ci, v, Bs[N], C_out=Cs[N], plan=cp), Cs[N],
                        steps=50, replays=5)
        out["csr_us"][f"N{N}"] = round(ms * 1e3, 2)
    for T in (int(x) for x in args.grid_tiles.split(",")):
        brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, T)
        plan = api.B2srPlan(h, T, g.M, brp)
        for N in widths:
            ms = time_graph(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bs[N], C_out=Cs[N],
                                                  plan=plan), Cs[N], steps=50, replays=5)
            out["b2sr_us"][f"T{T}_N{N}"] = round(ms * 1e3, 2)
            if args.grid_no_tune:
                continue
            # the autotuner's pick among the lane mappings (variant, lanes, rows_scale)
            choice = plan.autotune(brp, bci, words, Bs[N], Cs[N])
            ms = time_graph(lambda: api.spmm_b2sr(h, T, g.M, brp, bci, words, Bs[N], C_out=Cs[N],
                                                  plan=plan), Cs[N], steps=50, replays=5)
            out["b2sr_tuned_us"][f"T{T}_N{N}"] = round(ms * 1e3, 2)
            out["b2sr_tuned_choice"][f"T{T}_N{N}"] = list(choice)
        plan.close()
        del brp, bci, words
    torch.cuda.empty_cache()
    return {"b2sr_grid_89k": out}


def extra_bmm(args, h, dev):
    """XNOR/popcount BMM32 (backend/mm.cu:17-125) on square sizes, float and
    sign-binarised outputs."""
    import torch

    from spmm_eval_amd import api

    out = {}
    rng = np.random.default_rng(12)
    for n in (1024, 4096):
        Kw = n // 32
        A = torch.from_numpy(rng.integers(0, 2 ** 32, (Kw, n), dtype=np.uint64).astype(np.uint32)).to(dev)
        Bw = torch.from_numpy(rng.integers(0, 2 ** 32, (Kw, n), dtype=np.uint64).astype(np.uint32)).to(dev)
        Cf = api.bmm32_xnor_f32(A, Bw, n, n, Kw)
        ms_f = time_graph(lambda: api.bmm32_xnor_f32(A, Bw, n, n, Kw, out=Cf), Cf, steps=50)
        Cb = api.bmm32_xnor_bin(A, Bw, n, n, n)
        ms_b = time_graph(lambda: api.bmm32_xnor_bin(A, Bw, n, n, n, out=Cb), Cb, steps=50)
        ops = 2.0 * n * n * n  # one XNOR + one popcount-add per bit pair
        out[f"bmm32_{n}"] = {"f32_out_ms": round(ms_f, 5), "bin_out_ms": round(ms_b, 5),
                             "f32_out_TOPs_binary": round(ops / ms_f / 1e9, 1),
                             "bin_out_TOPs_binary": round(ops / ms_b / 1e9, 1),
                             "f32_out_C_GBps": rate(n * n * 4, ms_f)}
    return out


def extra_fwd_bwd(args, h, dev):
    import torch

    import oracle
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
        return min(cands, key=lambda t: median_batches(lambda: bwd_fused(XX, GG, t), 30, 3, 3)[0])

    def med(fn):
        return round(median_batches(fn, launches=100)[0], 4)

    res = {}
    for name, XX, GG, s in (("f32", Xd, dYd, 4), ("bf16", Xh, dYh, 2)):
        geo = (pf.autotune(rp, ci, v, XX, Y), pb.autotune(trp, tci, tval, GG, dX, K=g.M))
        ms = med(lambda: step(XX, GG))
        parts = {"fwd_ms": med(lambda: fwd(XX)), "dX_ms": med(lambda: bwd_x(GG)),
                 "sddmm_ms": med(lambda: bwd_v(XX, GG))}
        tf = tune_fused(XX, GG)
        fused_ms = med(lambda: (fwd(XX), bwd_fused(XX, GG, tf)))
        torch.cuda.synchronize()
        parts["fused"] = {"step_ms": fused_ms, "bwd_ms": med(lambda: bwd_fused(XX, GG, tf)),
                          "geometry": list(tf),
                          "max_abs_diff_dX": float((dX2 - dX).abs().max()),
                          "max_rel_diff_dval": float(((dval2 - dval).abs().max()
                                                      / dval.abs().max().clamp_min(1e-30)))}
        nbytes = 2 * csr_gather_bytes(g.M, g.nnz, N, s_b=s) + sddmm_gather_bytes(g.M, g.nnz, N, s)
        rows = check_rows(g, 7)
        Yc, dvc = Y.cpu().numpy(), dval.cpu().numpy()
        Xr = (bf16_round(X) if name == "bf16" else X).astype(np.float64)
        dYr = (bf16_round(dY) if name == "bf16" else dY).astype(np.float64)
        e_y = sampled_rel_err(g, X.astype(np.float64), Yc, rows)
        num = den = 0.0
        for i in rows:
            s0, e0 = g.rowptr[i], g.rowptr[i + 1]
            ref = X[g.colind[s0:e0]].astype(np.float64) @ dY[i].astype(np.float64)
            num += float(((dvc[s0:e0] - ref) ** 2).sum())
            den += float((ref ** 2).sum())
        e_yr = sampled_rel_err(g, Xr, Yc, rows)
        res[name] = {"ms": ms, "effective_GBps_gather_model": rate(nbytes, ms), "geometry": geo,
                     **parts, "rel_err_Y_fp64_vs_fp32_source": e_y,
                     "rel_err_Y_fp64_of_rounded_operands": e_yr,
                     "rel_err_dval_fp64_vs_fp32_source": (num / den) ** 0.5}
    res["bf16_vs_f32_three_launches_each"] = round(res["f32"]["ms"] / res["bf16"]["ms"], 3)
    res["bf16_vs_f32_fused_backward_each"] = round(res["f32"]["fused"]["step_ms"]
                                                   / res["bf16"]["fused"]["step_ms"], 3)
    res["bf16_fused_vs_f32_three_launches"] = round(res["f32"]["ms"] / res["bf16"]["fused"]["step_ms"], 3)
    # CPU fwd + dX + SDDMM (OpenMP, all host threads), FP32 and BF16 operands
    tcp, tri, tcv, _ = oracle.csr2csc(g.rowptr, g.colind, g.val, g.K)
    Yn = np.empty((g.M, N), np.float32)
    dXn = np.empty((g.K, N), np.float32)
    dvn = np.empty(g.nnz, np.float32)
    Xhn, dYhn = Xh.cpu().numpy().view(np.uint16), dYh.cpu().numpy().view(np.uint16)
    cpu = {}
    for name, XX, GG in (("f32", X, dY), ("bf16", Xhn, dYhn)):
        spmm = oracle.csr_spmm_bf16_omp if name == "bf16" else oracle.csr_spmm_f32_omp
        f = cpu_time(lambda: spmm(g.rowptr, g.colind, g.val, XX, out=Yn))
        b = cpu_time(lambda: spmm(tcp, tri, tcv, GG, out=dXn))
        sd = cpu_time(lambda: oracle.sddmm_omp(g.rowptr, g.colind, GG, XX, out=dvn))
        cpu[name] = {"fwd_ms": round(f, 3), "dX_ms": round(b, 3), "sddmm_ms": round(sd, 3),
                     "step_ms": round(f + b + sd, 3),
                     "gpu_speedup_three_launches": round((f + b + sd) / res[name]["ms"], 1)}
    res["cpu"] = cpu
    res["cpu_threads"] = oracle.omp_threads()
    return {"gcn_fwd_bwd_89k_n128": res}


def extra_batched(args, h, dev):
    import torch

    import oracle
    from spmm_eval_amd import api, graphs
    from spmm_eval_amd.batched import BatchPlan

    gs = graphs.workload_batched(512)
    N = 64
    plan = BatchPlan(h, gs, N=N, bf16=True, device=dev)
    M = plan.M
    X = torch.from_numpy(graphs.normal((M, N), 21, 1)).to(dev)
    dY = torch.from_numpy(graphs.normal((M, N), 21, 2)).to(dev)
    Xh, dYh = api.to_bf16_bits(X), api.to_bf16_bits(dY)
    ms_graph, _ = median_batches(lambda: plan.step(Xh, dYh, use_graph=True), launches=100)
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

    ms_sep, _ = median_batches(separate, launches=3, batches=5, warmup=1)
    # the same 1,536 launches captured in one hipGraph: device time without
    # the host's pacing of the calls
    try:
        ms_sep_graph = round(time_graph(separate, Y, steps=1, replays=5), 3)
    except Exception as e:  # noqa: BLE001 -- reported, not fatal
        log(f"separate launches not capturable: {e}")
        ms_sep_graph = None
    info = plan.info()
    forms = {}
    for form in (0, 1, 2):
        if form == 2 and info["lanes_per_row"] * info["vec"] * 8 != N:
            continue
        plan.set_geometry(info["lanes_per_row"], info["vec"], info["rows_scale"],
                          info["persistent"], form)
        forms[str(form)] = round(median_batches(lambda: plan.step(Xh, dYh, use_graph=True), 30, 5, 3)[0], 4)
    # CPU: the same fwd + dX + SDDMM over the block-diagonal union, FP32
    u, _, _ = graphs.concat_block_diagonal(gs)
    tcp, tri, tcv, _ = oracle.csr2csc(u.rowptr, u.colind, u.val, u.K)
    Xn, dYn = X.cpu().numpy(), dY.cpu().numpy()
    Yn, dXn, dvn = np.empty((M, N), np.float32), np.empty((M, N), np.float32), np.empty(u.nnz, np.float32)

    def cpu_step():
        oracle.csr_spmm_f32_omp(u.rowptr, u.colind, u.val, Xn, out=Yn)
        oracle.csr_spmm_f32_omp(tcp, tri, tcv, dYn, out=dXn)
        oracle.sddmm_omp(u.rowptr, u.colind, dYn, Xn, out=dvn)

    cms = cpu_time(cpu_step, warmup=2, reps=5)
    plan.close()
    return {"batched_512_fwd_bwd_n64": {"fused_launch": info, "sddmm_form_ms": forms,
        "graphs": len(gs), "rows": M, "nnz": plan.nnz,
        "bf16_one_graph_ms_median_of_5": round(ms_graph, 4),
        "f32_separate_launches_ms_median_of_5": round(ms_sep, 3),
        "speedup": round(ms_sep / ms_graph, 2),
        "f32_separate_launches_one_graph_ms_median_of_5": ms_sep_graph,
        "speedup_vs_captured_separate": round(ms_sep_graph / ms_graph, 2) if ms_sep_graph else None,
        "cpu_f32_union_step_ms": round(cms, 2), "cpu_threads": oracle.omp_threads()}}


def extra_spmv_aligned(args, h, dev):
    """Aligned B2SR-8/16 repack and the binary SpMV over it on the 89k graph,
    beside CSR SpMM at N = 1 (the csrmv the reference checks against)."""
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
                               "effective_GBps": rate(csr_gather_bytes(g.M, g.nnz, 1), ms_csr)}
    del plan
    for T in (8, 16):
        brp, bci, words = api.csr2b2sr(h, rp, ci, g.M, g.K, T)
        nrp, nci, nval = api.b2sr_aligned(h, T, brp, bci, words)
        torch.cuda.synchronize()
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
            "tiles_padded": nt, "tiles": int(bci.numel()),
            "bin_ms": round(ms_bin, 5), "bin_GBps": rate(b_bin, ms_bin),
            "f32_ms": round(ms_f32, 5), "f32_GBps": rate(b_f32, ms_f32)}
        del brp, bci, words, nrp, nci, nval
    torch.cuda.empty_cache()
    return out


def gather_calib(args, h, dev):
    """Gather-shaped calibration of the L2-miss counters and measured gather
    ceilings, with the product CSR kernel on synthetic CSR matrices:
    (a) every row of a 1 GiB B read exactly once in random order (a random
        permutation of the columns, 8 entries per row): the bytes the kernel
        must fetch past L2 are known (B + colind + val + rowptr + the plan's
        row list and spans), so FETCH_SIZE of this dispatch calibrates the
        counter for random 256-byte (N=64) and 512-byte (N=128) rows;
    (b) uniform random columns, 232,965 rows x 128 entries, from B footprints
        of 2 MiB (L2-resident), 64 MiB (MALL-resident) and 1 GiB (HBM): the
        gather rate each level of the memory system sustains."""
    import torch

    from spmm_eval_amd import api

    rng = np.random.default_rng(2024)
    res = {"permutation": {}, "ceilings": {}}
    for N in (64, 128):
        K = (1 << 30) // (N * 4)
        d = 8
        M = K // d
        perm = rng.permutation(K).astype(np.int32).reshape(M, d)
        ci = np.sort(perm, axis=1).reshape(-1)
        rp = np.arange(0, K + 1, d, dtype=np.int32)
        rpd, cid = torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev)
        v = torch.ones(K, device=dev)
        B = torch.rand((K, N), device=dev)
        C = torch.empty((M, N), device=dev)
        plan = api.CsrPlan(h, rpd, K)
        tune = (16 if N == 64 else 32, 1, 1)
        api.spmm_csr(h, rpd, cid, v, B, C_out=C, plan=plan, tune=tune)
        ms, _ = median_batches(lambda: api.spmm_csr(h, rpd, cid, v, B, C_out=C, plan=plan, tune=tune),
                               launches=10, batches=3, warmup=2)
        fetch = K * N * 4 + 8 * K + 4 * (M + 1) + 12 * M  # B, colind+val, rowptr, rows+spans
        res["permutation"][f"N{N}"] = {"rows_of_B": K, "B_bytes": K * N * 4, "fetch_bytes": fetch,
                                       "write_bytes": M * N * 4, "grid_rows": M, "lanes": tune[0],
                                       "ms": round(ms, 4), "GBps": rate(fetch + M * N * 4, ms)}
        log(f"permutation gather N={N}: {ms:.3f} ms")
        del rpd, cid, v, B, C, plan
        torch.cuda.empty_cache()
    M, d = 232_965, 128
    for N in (64, 128):
        for label, nbytes in (("L2_2MiB", 2 << 20), ("MALL_64MiB", 64 << 20), ("HBM_1GiB", 1 << 30)):
            K = nbytes // (N * 4)
            ci = np.sort(rng.integers(0, K, size=(M, d), dtype=np.int32), axis=1).reshape(-1)
            rp = np.arange(0, M * d + 1, d, dtype=np.int32)
            rpd, cid = torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev)
            v = torch.ones(M * d, device=dev)
            B = torch.rand((K, N), device=dev)
            C = torch.empty((M, N), device=dev)
            plan = api.CsrPlan(h, rpd, M * d)
            geo = plan.autotune(rpd, cid, v, B, C)
            ms, _ = median_batches(lambda: api.spmm_csr(h, rpd, cid, v, B, C_out=C, plan=plan),
                                   launches=20, batches=5, warmup=3)
            res["ceilings"][f"N{N}_{label}"] = {
                "geometry": geo, "ms": round(ms, 4), "gathers": M * d,
                "G_gathers_per_s": round(M * d / ms / 1e6, 2),
                "gather_GBps": rate(M * d * N * 4, ms)}
            log(f"ceiling N={N} {label}: {ms:.3f} ms")
            del rpd, cid, v, B, C, plan
            torch.cuda.empty_cache()
    return {"gather_calibration": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--hidden", type=int, default=128)
    ap.add_argument("--batch-launches", type=int, default=100,
                    help="launches per batch of the median-of-5 timing")
    ap.add_argument("--quick", action="store_true", help="headline only (no extra workloads)")
    ap.add_argument("--geometry", default=None,
                    help="headline geometry lanes,vec,rows_scale,slices instead of autotuning it")
    ap.add_argument("--no-dropin", action="store_true", help="skip the headline's drop-in timing")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--only", choices=["csr_small", "fwd_bwd", "batched", "spmv_aligned", "b2sr",
                                       "b2sr_grid", "bmm", "gather_calib"], default=None,
                    help="run one extra workload alone and print its result (for profiling)")
    ap.add_argument("--grid-tiles", default="2,4,8,16,32,64",
                    help="tile sizes of the B2SR grid (comma list)")
    ap.add_argument("--grid-widths", default="1,32,64,128",
                    help="widths N of the B2SR grid (comma list)")
    ap.add_argument("--grid-no-tune", action="store_true",
                    help="B2SR grid: the plan's default lane mapping only (profiling one kernel)")
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
        fn = {"csr_small": extra_csr_small, "fwd_bwd": extra_fwd_bwd, "batched": extra_batched,
              "spmv_aligned": extra_spmv_aligned, "b2sr": extra_b2sr, "b2sr_grid": extra_b2sr_grid,
              "bmm": extra_bmm, "gather_calib": gather_calib}[args.only]
        print(json.dumps(fn(args, h, dev)))
        h.close()
        return
    log(f"headline: CSR FP32 SpMM, 233k rows, N={args.hidden}")
    hd = headline(args, h, dev)
    g, N = hd["g"], hd["N"]
    comp = csr_compulsory_bytes(g.M, g.K, g.nnz, N)
    gath = csr_gather_bytes(g.M, g.nnz, N)
    G, VEC, _, SL = hd["geo"]
    tiles = N // (G * VEC * 4)
    if SL > 1:  # tiles x SL launches per step: slice 0 stores C, the others add into it
        kname = f"void spmm::csr_sliced_kernel<float, {G}, {VEC}, false/true> x {SL} slices"
        launches = tiles * SL
    else:
        kname = f"void spmm::csr_spmm_kernel<float, {G}, {VEC}>"
        launches = 1
    traffic = profiled_traffic(f"csr_f32_233k_n{N}.md", kname)
    ach = comp / (hd["ms_med"] * 1e-3) / 1e9
    result = {
        "metric": f"SpMM HBM GB/s (compulsory bytes) -- CSR FP32 SpMM, 232,965-row power-law graph, N={N}",
        "value": rate(comp, hd["ms_step"]),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hd["ms_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Chung-Lu power-law graph, values U(0,1] row-normalised, B U[-1,1])",
        "config": {"workload": "csr_spmm_f32_powerlaw_233k", "rows": g.M, "nnz": g.nnz,
                   "max_row_degree": int(g.degrees().max()), "hidden": N,
                   "plan": "row schedule analysed once, geometry autotuned",
                   "geometry_lanes_vec_rows_slices": list(hd["geo"]), "parallelism": "single GPU"},
        "gflops": round(2.0 * g.nnz * N / (hd["ms_step"] * 1e-3) / 1e9, 1),
        "pct_hbm_peak": round(100 * comp / (hd["ms_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
        "effective_GBps_gather_model": rate(gath, hd["ms_step"]),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": f"profiles/csr_f32_233k_n{N}.md (FETCH_SIZE and WRITE_SIZE passes "
                                       "scaled by the factors measured there)" if traffic else None,
                     "kernel": kname, "kernel_launches_per_step": launches,
                     "bytes_per_launch": comp,
                     "bytes_model": "compulsory: rowptr + colind + val + B + C, each once",
                     "mean_launch_ms": round(hd["ms_med"], 5),
                     "launch_ms_source": "median of 5 batches of back-to-back launches, events on the "
                                         "launch stream"},
        "timing": {"timed_region_ms_per_step": round(hd["ms_step"], 5),
                   "median_of_5_batches_ms": round(hd["ms_med"], 5), "batches_ms": hd["batches"],
                   "launches_per_batch": args.batch_launches},
        "checks": {"rel_err_fp64_sampled_rows_incl_heaviest": hd["err"], "bar": 1e-5},
        "analyze_ms_host_clock": round(hd["analyze_ms"], 2),
        "dropin_plan_free": {"ms_median_of_5": round(hd["dropin_ms"], 4) if hd["dropin_ms"] else None,
                             "rel_err_fp64_sampled": hd["dropin_err"],
                             "geometry_lanes_vec_rows_slices": list(api.dropin_geometry(g.M, g.K, N, g.nnz)),
                             "form": "rows sorted and slice offsets built on the device per call"},
    }
    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(hd)
    if not args.quick:
        extra = {}
        try:
            log("extra_large_variants")
            extra.update(extra_large_variants(args, h, dev, hd))
        except Exception as e:
            extra["extra_large_variants"] = {"error": repr(e)}
        hd["plan"].close()
        del hd
        import torch as _t

        _t.cuda.empty_cache()
        for fn in (extra_csr_small, extra_b2sr, extra_b2sr_grid, extra_bmm, extra_fwd_bwd,
                   extra_batched, extra_spmv_aligned):
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
