"""Batched uneven graphs through one plan (spmm_batch_plan_create /
spmm_batch_fwd_bwd of the C-ABI): the graphs' CSR arrays are concatenated
block-diagonally, analysed once, and one fwd+bwd step (Y = A X,
dX = A^T dY, dval = SDDMM(dY, X)) runs as one fused (autotuned, optionally
persistent) kernel launch, replayed as a captured hipGraph."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .api import Handle, _p, _stream
from .graphs import Csr
from ._lib import Tune, call


class BatchPlan:
    def __init__(self, h: Handle, graphs_: list[Csr], N: int, bf16: bool, device, stream=None):
        self.h = h
        self.N = N
        self.bf16 = bf16
        row_off = np.zeros(len(graphs_) + 1, np.int32)
        nnz_off = np.zeros(len(graphs_) + 1, np.int32)
        for i, g in enumerate(graphs_):
            row_off[i + 1] = row_off[i] + g.M
            nnz_off[i + 1] = nnz_off[i] + g.nnz
        self.M, self.nnz = int(row_off[-1]), int(nnz_off[-1])
        rowptr = np.zeros(self.M + 1, np.int32)
        for i, g in enumerate(graphs_):
            rowptr[row_off[i] + 1: row_off[i + 1] + 1] = g.rowptr[1:] + nnz_off[i]
        self.row_off, self.nnz_off = row_off, nnz_off
        self.rowptr = torch.from_numpy(rowptr).to(device)
        self.colind = torch.from_numpy(np.concatenate([g.colind for g in graphs_])).to(device)
        self.val = torch.from_numpy(np.concatenate([g.val for g in graphs_])).to(device)
        p = C.c_void_p()
        call("spmm_batch_plan_create", h.h, len(graphs_), row_off.ctypes.data, nnz_off.ctypes.data,
             _p(self.rowptr), _p(self.colind), _p(self.val), N, int(bf16), C.byref(p), _stream(stream))
        self.p = p
        dev = self.rowptr.device
        self.Y = torch.empty((self.M, N), dtype=torch.float32, device=dev)
        self.dX = torch.empty((self.M, N), dtype=torch.float32, device=dev)
        self.dval = torch.empty(max(self.nnz, 1), dtype=torch.float32, device=dev)[: self.nnz]

    def info(self) -> dict:
        """Autotuned geometry of the fused launch."""
        t, pers, srows, ms = Tune(), C.c_int(), C.c_int(), C.c_float()
        call("spmm_batch_plan_info", self.p, C.byref(t), C.byref(pers), C.byref(srows), C.byref(ms))
        return {"lanes_per_row": t.lanes_per_row, "vec": t.vec, "rows_scale": t.rows_scale,
                "persistent": bool(pers.value), "sddmm_form": int(srows.value),
                "tuned_ms": round(ms.value, 4)}

    def set_geometry(self, lanes_per_row: int, vec: int, rows_scale: int, persistent: bool,
                     sddmm_form: int):
        """Pin the fused launch's geometry and SDDMM form (0 entry chunks,
        1 forward rows, 2 fused into the backward items)."""
        call("spmm_batch_plan_set", self.p, C.byref(Tune(lanes_per_row, vec, rows_scale)),
             int(persistent), int(sddmm_form))

    def step(self, X: torch.Tensor, dY: torch.Tensor, use_graph: bool = True, stream=None):
        call("spmm_batch_fwd_bwd", self.p, _p(X), _p(dY), _p(self.Y), _p(self.dX), _p(self.dval),
             int(use_graph), _stream(stream))
        return self.Y, self.dX, self.dval

    def close(self):
        if self.p:
            _lib.lib().spmm_batch_plan_destroy(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
