"""Seeded synthetic inputs of the SpMM workloads.

The graph structure comes from the native Chung-Lu generator
(``csrc/graphgen.cpp``); values and dense operands come from numpy's PCG64
with per-purpose seeds, so every consumer (GPU run, oracle, CPU baseline)
sees identical arrays.

Workloads (sizes of the reference's Flickr / Reddit runs,
``backend/spmm.cu:869-875``, ``Makefile:12-22``):
  1  csr_small : 16,384 rows, ~260k nnz, values 1.0, B 16,384 x 32 U[-1,1]
  2  b2sr      : 89,250 rows, ~900k nnz, 0/1, B 89,250 x {32,64,128}
  3  large     : 232,965 rows, ~114.6M nnz (max degree ~20k), values U(0,1]
                 row-normalised, B 232,965 x {64,128} FP32 and BF16
  4  gcn       : the workload-2 graph with D^-1/2 A D^-1/2 values,
                 X 89,250 x 128 N(0,1), dY N(0,1)
  5  batched   : 512 graphs, n log-uniform in [64, 20000], mean degree
                 U(8, 64), hidden 64
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib

SEED_BASE = 0x5EED0000
GAMMA = 2.1


@dataclass
class Csr:
    rowptr: np.ndarray  # int32 [M+1]
    colind: np.ndarray  # int32 [nnz]
    val: np.ndarray     # float32 [nnz]
    M: int
    K: int

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr)


def chung_lu(n: int, target_nnz: int, seed: int, max_deg: float | None = None,
             gamma: float = GAMMA) -> Csr:
    """Symmetric power-law adjacency, no self-loops, no duplicates, sorted
    column indices, values 1.0. ``max_deg`` defaults to sqrt(target_nnz)
    (the structural cut-off of Chung-Lu graphs)."""
    if max_deg is None:
        max_deg = math.sqrt(max(target_nnz, 1))
    G = _lib.gen()
    h = G.gen_chung_lu(int(n), int(target_nnz), float(gamma), float(max_deg), C.c_uint64(seed))
    if not h:
        raise ValueError("gen_chung_lu rejected its arguments")
    try:
        nnz = G.gen_graph_nnz(h)
        rowptr = np.empty(n + 1, np.int32)
        colind = np.empty(max(nnz, 1), np.int32)
        if G.gen_graph_copy(h, rowptr.ctypes.data, colind.ctypes.data) != 0:
            raise ValueError("graph too large for int32 indices")
    finally:
        G.gen_graph_free(h)
    colind = colind[:nnz]
    return Csr(rowptr, colind, np.ones(nnz, np.float32), n, n)


def rng(seed: int, stream: int) -> np.random.Generator:
    return np.random.default_rng([seed, stream])


def uniform_pm1(shape, seed: int, stream: int = 1) -> np.ndarray:
    return rng(seed, stream).uniform(-1.0, 1.0, size=shape).astype(np.float32)


def normal(shape, seed: int, stream: int) -> np.ndarray:
    return rng(seed, stream).standard_normal(size=shape, dtype=np.float32)


def sym_norm_values(g: Csr) -> np.ndarray:
    """D^-1/2 A D^-1/2 for binary symmetric A (no self-loops added)."""
    deg = g.degrees().astype(np.float64)
    inv = np.where(deg > 0, 1.0 / np.sqrt(np.maximum(deg, 1)), 0.0)
    rows = np.repeat(np.arange(g.M), g.degrees())
    return (inv[rows] * inv[g.colind]).astype(np.float32)


def row_normalised_uniform(g: Csr, seed: int, stream: int = 2) -> np.ndarray:
    """U(0,1] values divided by their row sum."""
    u = 1.0 - rng(seed, stream).random(g.nnz)  # (0, 1]
    rows = np.repeat(np.arange(g.M), g.degrees())
    s = np.bincount(rows, weights=u, minlength=g.M)
    return (u / s[rows]).astype(np.float32)


# --------------------------------------------------------------- workloads
def workload_csr_small(seed_off: int = 1) -> tuple[Csr, np.ndarray]:
    seed = SEED_BASE + seed_off
    g = chung_lu(16_384, 260_000, seed)
    B = uniform_pm1((g.K, 32), seed)
    return g, B


def workload_b2sr_graph(seed_off: int = 2) -> Csr:
    return chung_lu(89_250, 900_000, SEED_BASE + seed_off)


def workload_large_graph(seed_off: int = 3, scale: float = 1.0) -> Csr:
    """232,965 rows, ~114.6M nnz, max degree ~20k; `scale` shrinks the
    target nnz (tests use a fraction)."""
    seed = SEED_BASE + seed_off
    g = chung_lu(232_965, int(114_600_000 * scale), seed, max_deg=20_000 * max(scale, 0.05) ** 0.5)
    g.val = row_normalised_uniform(g, seed)
    return g


def workload_gcn(seed_off: int = 4, N: int = 128) -> tuple[Csr, np.ndarray, np.ndarray]:
    seed = SEED_BASE + seed_off
    g = chung_lu(89_250, 900_000, SEED_BASE + 2)  # the workload-2 graph
    g.val = sym_norm_values(g)
    X = normal((g.K, N), seed, 1)
    dY = normal((g.M, N), seed, 2)
    return g, X, dY


def batched_sizes(ngraphs: int = 512, seed_off: int = 5) -> list[tuple[int, float]]:
    r = rng(SEED_BASE + seed_off, 0)
    n = np.rint(np.exp(r.uniform(math.log(64), math.log(20_000), ngraphs))).astype(int)
    d = r.uniform(8, 64, ngraphs)
    return [(int(a), float(b)) for a, b in zip(n, d)]


def workload_batched(ngraphs: int = 512, seed_off: int = 5) -> list[Csr]:
    out = []
    for gi, (n, d) in enumerate(batched_sizes(ngraphs, seed_off)):
        g = chung_lu(n, int(n * d), SEED_BASE + seed_off + 1000 * (gi + 1))
        g.val = sym_norm_values(g)
        out.append(g)
    return out


def concat_block_diagonal(graphs: list[Csr]) -> tuple[Csr, np.ndarray, np.ndarray]:
    """Block-diagonal union with global column indices. Returns
    (csr, row_off[G+1], nnz_off[G+1])."""
    row_off = np.zeros(len(graphs) + 1, np.int64)
    nnz_off = np.zeros(len(graphs) + 1, np.int64)
    for i, g in enumerate(graphs):
        row_off[i + 1] = row_off[i] + g.M
        nnz_off[i + 1] = nnz_off[i] + g.nnz
    rowptr = np.zeros(row_off[-1] + 1, np.int64)
    colind = np.empty(nnz_off[-1], np.int32)
    val = np.empty(nnz_off[-1], np.float32)
    for i, g in enumerate(graphs):
        rowptr[row_off[i] + 1: row_off[i + 1] + 1] = g.rowptr[1:] + nnz_off[i]
        colind[nnz_off[i]: nnz_off[i + 1]] = g.colind + row_off[i]
        val[nnz_off[i]: nnz_off[i + 1]] = g.val
    M = int(row_off[-1])
    return Csr(rowptr.astype(np.int32), colind, val, M, M), row_off.astype(np.int32), nnz_off.astype(np.int32)


def random_csr(M: int, K: int, density: float, seed: int, empty_rows: bool = True,
               long_rows: int = 0, long_len: int = 0) -> Csr:
    """Small unstructured CSR for edge cases: sorted, duplicate-free rows,
    optional empty rows and a few very long rows."""
    r = rng(seed, 7)
    rows = []
    for i in range(M):
        if i < long_rows:
            k = min(K, long_len)
        else:
            k = r.binomial(K, density)
            if empty_rows and i % 7 == 3:
                k = 0
        rows.append(np.sort(r.choice(K, size=k, replace=False)) if k else np.zeros(0, np.int64))
    rowptr = np.zeros(M + 1, np.int32)
    rowptr[1:] = np.cumsum([len(x) for x in rows])
    colind = np.concatenate(rows).astype(np.int32) if M else np.zeros(0, np.int32)
    val = r.uniform(-1, 1, len(colind)).astype(np.float32)
    return Csr(rowptr, colind, val, M, K)
