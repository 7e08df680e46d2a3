"""The compiled C++ caller of the C-ABI (spmm_eval_amd/spmm_harness, built
from csrc/harness.cpp by build()): the reference harness's flow
(test_spmm_full.cu:474-514) through include/spmm_mi355x.h and
libspmm_mi355x.so, linked as a C++ program would link them.

Without a GPU: its host checks (argument validation of every entry point it
calls, the Matrix-Market reader on a file it writes). With a GPU: the whole
flow on a Matrix-Market file and on the 89,250-row power-law graph, B2SR-4
equal to CSR and to a host integer product on every element."""
import os
import subprocess

import numpy as np
import pytest

from spmm_eval_amd import graphs

HARNESS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "spmm_eval_amd", "spmm_harness")


def run(*args, timeout=120):
    assert os.path.exists(HARNESS), "spmm_harness is not built (run __graft_entry__.build())"
    return subprocess.run([HARNESS, *args], capture_output=True, text=True, timeout=timeout)


def test_harness_host_checks():
    r = run("--host-checks")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


def test_harness_usage_error():
    r = run("--chung-lu", "10", "20")
    assert r.returncode == 2 and "usage" in r.stderr


def write_mtx(path, g, symmetric=False):
    rows = np.repeat(np.arange(g.M), g.degrees())
    keep = rows >= g.colind if symmetric else np.ones(len(rows), bool)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern " + ("symmetric" if symmetric else "general") + "\n")
        f.write(f"{g.M} {g.K} {int(keep.sum())}\n")
        for r, c in zip(rows[keep], g.colind[keep]):
            f.write(f"{r + 1} {c + 1}\n")


@pytest.mark.gpu
@pytest.mark.parametrize("N", [32, 64])
def test_harness_mtx_end_to_end(tmp_path, N):
    g = graphs.chung_lu(3000, 40000, seed=21)
    p = tmp_path / "g.mtx"
    write_mtx(str(p), g)
    r = run(str(p), str(N))
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "TEST PASSED: 1" in r.stdout
    assert "mismatches B2SR vs CSR: 0, CSR vs host integer product: 0" in r.stdout


@pytest.mark.gpu
def test_harness_powerlaw_89k():
    r = run("--chung-lu", "89250", "900000", str(graphs.SEED_BASE + 2), "64")
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "TEST PASSED: 1" in r.stdout
    assert "graph: 89250 x 89250" in r.stdout
