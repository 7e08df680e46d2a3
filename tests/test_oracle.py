"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: hand-derived vectors (tests/golden/b2sr_hand.json), a textbook or
library routine (numpy dense matmul, scipy.sparse transpose, torch's BF16
rounding, numpy.packbits), brute force on tiny inputs, and invariants. Each
check is chosen so that a plausible slip (wrong bit order, transposed
operand, dropped padding rule, off-by-one in block rows) fails; the
`planted` tests show that for a few of them.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from spmm_eval_amd.graphs import Csr, chung_lu, random_csr

HERE = os.path.dirname(os.path.abspath(__file__))
HAND = json.load(open(os.path.join(HERE, "golden", "b2sr_hand.json")))


def dense(g: Csr, binary=False):
    A = np.zeros((g.M, g.K), np.float64)
    rows = np.repeat(np.arange(g.M), g.degrees())
    A[rows, g.colind] = 1.0 if binary else g.val
    return A


def brute_b2sr(A01: np.ndarray, T: int):
    """Independent formulation: pad the dense 0/1 matrix to whole tiles,
    cut it into T x T tiles, keep the tiles with any 1 (row-major tile
    order), and read each tile row as a big-endian bit string."""
    M, K = A01.shape
    MB, KB = -(-M // T), -(-K // T)
    P = np.zeros((MB * T, KB * T), np.uint8)
    P[:M, :K] = A01
    tiles = P.reshape(MB, T, KB, T).transpose(0, 2, 1, 3)  # [MB][KB][T][T]
    nz = tiles.reshape(MB, KB, -1).any(axis=2)
    rowptr = np.concatenate([[0], np.cumsum(nz.sum(axis=1))]).astype(np.int32)
    br, bc = np.nonzero(nz)
    bits = tiles[br, bc]  # [nb][T][T]
    weights = (1 << np.arange(T - 1, -1, -1, dtype=np.uint64))
    words = (bits.astype(np.uint64) * weights).sum(axis=2).reshape(-1)
    return rowptr, bc.astype(np.int32), words


# ------------------------------------------------------------ hand vectors
def hand_csr():
    rp = np.array(HAND["rowptr"], np.int32)
    ci = np.array(HAND["colind"], np.int32)
    return Csr(rp, ci, np.ones(len(ci), np.float32), HAND["M"], HAND["K"])


@pytest.mark.parametrize("T", [4, 8])
def test_csr2b2sr_hand_vectors(T):
    g = hand_csr()
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    exp = HAND[f"T{T}"]
    assert brp.tolist() == exp["bsrRowPtr"]
    assert bci.tolist() == exp["bsrColInd"]
    assert words.tolist() == exp["bsrVal"]


def test_b2sr_spmm_hand_vectors():
    g = hand_csr()
    B = np.array(HAND["B"], np.float32)
    for T in (4, 8):
        brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
        C = oracle.b2sr_spmm_f64(T, g.M, brp, bci, words, B)
        assert C.tolist() == HAND["C"]


def test_pack_hand_vectors():
    X = np.array(HAND["sign_pack"]["X"], np.float32)
    assert oracle.pack_sign_rows(X).reshape(-1).tolist() == HAND["sign_pack"]["sign"]
    assert oracle.pack_bin_rows(X).reshape(-1).tolist() == HAND["sign_pack"]["bin"]


def test_bmm_hand_vectors():
    for a, b, c in HAND["bmm"]["cases"]:
        A = np.array([a], np.uint32)
        B = np.array([b], np.uint32)
        assert oracle.bmm32_xnor_f32(A, B, 1, 1, 1)[0, 0] == c
        bit = oracle.bmm32_xnor_bin(A, B, 1, 1, 32)[0, 0] >> 31
        assert bit == (1 if c >= 0 else 0)


# ------------------------------------------------------------------- CSR
@pytest.mark.parametrize("M,K,N", [(1, 1, 1), (37, 53, 7), (64, 20, 33), (200, 150, 64)])
def test_csr_spmm_f64_matches_dense(M, K, N):
    g = random_csr(M, K, 0.1, seed=M * 7 + N, long_rows=2, long_len=K)
    B = np.random.default_rng(1).uniform(-1, 1, (K, N)).astype(np.float32)
    C = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    ref = dense(g) @ B.astype(np.float64)
    np.testing.assert_allclose(C, ref, rtol=1e-12, atol=1e-12)
    # also against scipy's sparse product
    A = sp.csr_matrix((g.val.astype(np.float64), g.colind, g.rowptr), shape=(M, K))
    np.testing.assert_allclose(C, A @ B.astype(np.float64), rtol=1e-12, atol=1e-12)


def test_planted_transpose_is_caught():
    g = random_csr(30, 30, 0.2, seed=3, empty_rows=False)
    B = np.random.default_rng(2).uniform(-1, 1, (30, 5)).astype(np.float32)
    C = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    wrong = dense(g).T @ B  # a transposed operand must not pass the dense check
    assert not np.allclose(C, wrong)


def test_csr_spmm_bf16_and_f32_forms():
    g = random_csr(120, 90, 0.15, seed=11)
    B = np.random.default_rng(3).uniform(-1, 1, (90, 24)).astype(np.float32)
    Bh = oracle.f32_to_bf16(B)
    C64 = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, oracle.bf16_to_f32(Bh))
    np.testing.assert_array_equal(oracle.csr_spmm_f64_bf16(g.rowptr, g.colind, g.val, Bh), C64)
    C32 = oracle.csr_spmm_f32_omp(g.rowptr, g.colind, g.val, B)
    ref = oracle.csr_spmm_f64(g.rowptr, g.colind, g.val, B)
    assert np.linalg.norm(C32 - ref) / np.linalg.norm(ref) < 1e-6
    C16 = oracle.csr_spmm_bf16_omp(g.rowptr, g.colind, g.val, Bh)
    assert np.linalg.norm(C16 - C64) / np.linalg.norm(C64) < 1e-6


def test_csr_spmm_i64_integer():
    g = random_csr(80, 60, 0.2, seed=5)
    B = np.random.default_rng(4).choice([-1, 1], size=(60, 9)).astype(np.int32)
    C = oracle.csr_spmm_i64(g.rowptr, g.colind, B)
    np.testing.assert_array_equal(C, (dense(g, binary=True) @ B).astype(np.int64))


def test_bf16_rounding_matches_torch():
    import torch

    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(10000).astype(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 1e-40, -1e-40, 3.4e38, 1.0 + 2 ** -8,
                  1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9], np.float32),
        # exact ties: low 16 bits = 0x8000
        (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
    ])
    ours = oracle.f32_to_bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(ours, ref)
    nan = oracle.f32_to_bf16(np.array([np.nan], np.float32))
    assert np.isnan(oracle.bf16_to_f32(nan))[0]


# ---------------------------------------------------------------- B2SR
@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("shape", [(1, 1), (13, 29), (70, 70), (257, 100)])
def test_csr2b2sr_matches_brute_force(T, shape):
    M, K = shape
    g = random_csr(M, K, 0.05, seed=M + K + T, long_rows=1, long_len=K)
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    rb, cb, wb = brute_b2sr(dense(g, binary=True).astype(np.uint8), T)
    np.testing.assert_array_equal(brp, rb)
    np.testing.assert_array_equal(bci, cb)
    np.testing.assert_array_equal(words.astype(np.uint64), wb)
    # invariants: one bit per stored entry, blocks ascending per block row
    assert int(np.unpackbits(words.view(np.uint8)).sum()) == g.nnz
    for br in range(len(brp) - 1):
        seg = bci[brp[br]:brp[br + 1]]
        assert (np.diff(seg) > 0).all()


def test_planted_bit_order_is_caught():
    g = hand_csr()
    _, _, words = oracle.csr2b2sr(g.rowptr, g.colind, 4)
    lsb_first = np.array([int(f"{w:04b}"[::-1], 2) for w in words])
    assert lsb_first.tolist() != HAND["T4"]["bsrVal"]


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_b2sr_spmm_matches_dense(T):
    g = chung_lu(300, 3000, seed=T)
    B = np.random.default_rng(T).uniform(-1, 1, (300, 40)).astype(np.float32)
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    C = oracle.b2sr_spmm_f64(T, g.M, brp, bci, words, B)
    np.testing.assert_allclose(C, dense(g, binary=True) @ B.astype(np.float64), rtol=0, atol=1e-12)


@pytest.mark.parametrize("T", [4, 32])
@pytest.mark.parametrize("N", [1, 31, 64, 70])
def test_b2sr_sign_matches_integer_csr(T, N):
    g = chung_lu(200, 2000, seed=N)
    X = np.random.default_rng(N).uniform(-1, 1, (200, N)).astype(np.float32)
    X[0, 0] = -0.0  # counts as +1 (x >= 0)
    Bw = oracle.pack_sign_rows(X)
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    C = oracle.b2sr_spmm_sign_i32(T, g.M, N, brp, bci, words, Bw)
    S = np.where(X >= 0, 1, -1).astype(np.int32)
    np.testing.assert_array_equal(C, oracle.csr_spmm_i64(g.rowptr, g.colind, S))


# ---------------------------------------------------------------- packers
@pytest.mark.parametrize("H,W", [(1, 1), (5, 31), (33, 32), (40, 100)])
def test_packers_match_packbits(H, W):
    X = np.random.default_rng(H * W).uniform(-1, 1, (H, W)).astype(np.float32)
    X[X.shape[0] // 2, 0] = 0.0
    Wc = -(-W // 32)
    for fn, pred in ((oracle.pack_sign_rows, X >= 0), (oracle.pack_bin_rows, X > 0)):
        bits = np.zeros((H, Wc * 32), np.uint8)
        bits[:, :W] = pred
        ref = np.packbits(bits, axis=1, bitorder="big").view(">u4").astype(np.uint32)
        np.testing.assert_array_equal(fn(X), ref)
    cols = oracle.pack_sign_cols(X)
    Hp = 32 * -(-H // 32)
    assert cols.shape == (Wc, Hp)
    np.testing.assert_array_equal(cols[:, :H].T, oracle.pack_sign_rows(X))
    assert (cols[:, H:] == 0).all()
    if H % 32 == 0:  # Bit32ColTo's layout assumes whole 32-row groups
        np.testing.assert_array_equal(oracle.unpack_cols(cols, H, W), (X >= 0).astype(np.float32))


# ---------------------------------------------------------------- BMM
def _unpack_pm1(words):  # [Kw][R] -> [R][32*Kw] of +-1
    bits = np.unpackbits(np.ascontiguousarray(words.T).astype(">u4").view(np.uint8), axis=1, bitorder="big")
    return bits.astype(np.int64) * 2 - 1


@pytest.mark.parametrize("M,N,Kw", [(1, 1, 1), (32, 64, 3), (70, 33, 5)])
def test_bmm_matches_pm1_matmul(M, N, Kw):
    r = np.random.default_rng(M + N)
    A = r.integers(0, 2 ** 32, (Kw, M), dtype=np.uint64).astype(np.uint32)
    B = r.integers(0, 2 ** 32, (Kw, N), dtype=np.uint64).astype(np.uint32)
    ref = _unpack_pm1(A) @ _unpack_pm1(B).T
    np.testing.assert_array_equal(oracle.bmm32_xnor_f32(A, B, M, N, Kw), ref.astype(np.float32))
    Kbits = 32 * Kw
    bits = (ref >= 0).astype(np.uint8)
    Nw = -(-N // 32)
    padded = np.zeros((M, Nw * 32), np.uint8)
    padded[:, :N] = bits
    ref_bin = np.packbits(padded, axis=1, bitorder="big").view(">u4").astype(np.uint32).T
    np.testing.assert_array_equal(oracle.bmm32_xnor_bin(A, B, M, N, Kbits), ref_bin)


# ------------------------------------------------------------- adjoint
@pytest.mark.parametrize("M,K", [(1, 1), (40, 25), (150, 300)])
def test_csr2csc_matches_scipy(M, K):
    g = random_csr(M, K, 0.1, seed=M + K, long_rows=1, long_len=K)
    cp, ri, cv, perm = oracle.csr2csc(g.rowptr, g.colind, g.val, K)
    A = sp.csr_matrix((g.val, g.colind, g.rowptr), shape=(M, K)).tocsc()
    A.sort_indices()
    np.testing.assert_array_equal(cp, A.indptr)
    np.testing.assert_array_equal(ri, A.indices)
    np.testing.assert_array_equal(cv, A.data)
    np.testing.assert_array_equal(g.val[perm], cv)


def test_sddmm_matches_dense():
    g = random_csr(60, 45, 0.2, seed=9)
    r = np.random.default_rng(9)
    dY = r.standard_normal((60, 17)).astype(np.float32)
    X = r.standard_normal((45, 17)).astype(np.float32)
    out = oracle.sddmm_f64(g.rowptr, g.colind, dY, X)
    rows = np.repeat(np.arange(60), g.degrees())
    full = dY.astype(np.float64) @ X.astype(np.float64).T
    np.testing.assert_allclose(out, full[rows, g.colind], rtol=1e-12, atol=1e-12)
    # the timed CPU baselines: FP32 within FP32 rounding of the dense
    # product, BF16 equal to the dense product of the rounded operands
    o32 = oracle.sddmm_omp(g.rowptr, g.colind, dY, X)
    np.testing.assert_allclose(o32, full[rows, g.colind], rtol=1e-5, atol=1e-5)
    dYh, Xh = oracle.f32_to_bf16(dY), oracle.f32_to_bf16(X)
    o16 = oracle.sddmm_omp(g.rowptr, g.colind, dYh, Xh)
    fullh = oracle.bf16_to_f32(dYh).astype(np.float64) @ oracle.bf16_to_f32(Xh).astype(np.float64).T
    np.testing.assert_allclose(o16, fullh[rows, g.colind], rtol=1e-5, atol=1e-5)


def test_sddmm_is_value_gradient():
    """d/dval <A X, G> = SDDMM(G, X): central differences on the FP64 dense
    form (exact up to rounding since the loss is linear in val)."""
    g = random_csr(20, 15, 0.3, seed=4, empty_rows=False)
    r = np.random.default_rng(4)
    X = r.standard_normal((15, 6))
    G = r.standard_normal((20, 6))
    grad = oracle.sddmm_f64(g.rowptr, g.colind, G.astype(np.float32), X.astype(np.float32))
    rows = np.repeat(np.arange(20), g.degrees())
    X32, G32 = X.astype(np.float32).astype(np.float64), G.astype(np.float32).astype(np.float64)
    for e in range(0, g.nnz, 7):
        h = 1e-3
        vp, vm = g.val.astype(np.float64).copy(), g.val.astype(np.float64).copy()
        vp[e] += h
        vm[e] -= h
        Ap = sp.csr_matrix((vp, g.colind, g.rowptr), shape=(20, 15))
        Am = sp.csr_matrix((vm, g.colind, g.rowptr), shape=(20, 15))
        fd = ((Ap @ X32) * G32).sum() - ((Am @ X32) * G32).sum()
        assert abs(fd / (2 * h) - grad[e]) < 1e-9


# ------------------------------------------ against the reference's code
def _ref_or_skip():
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return R


@pytest.mark.parametrize("M,K", [(1, 1), (70, 70), (400, 400)])
def test_csr2csc_matches_reference_code(M, K):
    """oracle.csr2csc against the reference's csr2csc (readMtx.hpp:319-345)
    compiled from its own source. Square shapes only: the reference passes
    (ncols, nrows) to coo2csc (:341), which swaps them again (:311-316), so
    its transpose sizes the output by the row count and is only correct when
    M == K (every workload here is square). The oracle transposes any shape;
    test_csr2csc_matches_scipy pins the rectangular case."""
    R = _ref_or_skip()
    g = random_csr(M, K, 0.05, seed=M, long_rows=1, long_len=K)
    nnz = g.nnz
    cp, ri, cv = np.empty(K + 1, np.int32), np.empty(max(nnz, 1), np.int32), np.empty(max(nnz, 1), np.float32)
    R.ref_csr2csc(M, K, g.rowptr, g.colind if nnz else np.zeros(1, np.int32),
                  g.val if nnz else np.zeros(1, np.float32), cp, ri, cv)
    ocp, ori, ocv, _ = oracle.csr2csc(g.rowptr, g.colind, g.val, K)
    np.testing.assert_array_equal(cp, ocp)
    np.testing.assert_array_equal(ri[:nnz], ori)
    np.testing.assert_array_equal(cv[:nnz], ocv)


def test_generator_csr_matches_reference_coo2csr():
    """The synthetic graphs are in the CSR form the reference's coo2csr
    (readMtx.hpp:246-304) builds from the same entries in shuffled order."""
    R = _ref_or_skip()
    g = chung_lu(500, 6000, seed=9)
    rows = np.repeat(np.arange(g.M), g.degrees()).astype(np.int32)
    perm = np.random.default_rng(0).permutation(g.nnz)
    rp, ci, cv = np.empty(g.M + 1, np.int32), np.empty(g.nnz, np.int32), np.empty(g.nnz, np.float32)
    R.ref_coo2csr(g.M, g.K, g.nnz, np.ascontiguousarray(rows[perm]), np.ascontiguousarray(g.colind[perm]),
                  np.ascontiguousarray(g.val[perm]), rp, ci, cv)
    np.testing.assert_array_equal(rp, g.rowptr)
    np.testing.assert_array_equal(ci, g.colind)


# ---- aligned B2SR-8/16 repack and the binary SpMV over it ----

def test_aligned_pack_hand_vector():
    """One T = 8 block row of 5 tiles -> 2 groups of 4 (3 empty column-0
    tiles of padding); word j of a group is its tiles' row-j bytes, first
    tile highest (hand-computed)."""
    T = 8
    brp = np.array([0, 5, 5], np.int32)  # second block row empty
    bci = np.array([1, 3, 4, 6, 9], np.int32)
    tiles = np.zeros((5, T), np.uint8)
    tiles[0, 0], tiles[1, 0], tiles[2, 0], tiles[3, 0] = 0x80, 0x01, 0xFF, 0x10
    tiles[4, 7] = 0xA5
    nrp, nci, nval = oracle.b2sr_aligned_pack(T, brp, bci, tiles.reshape(-1))
    np.testing.assert_array_equal(nrp, [0, 8, 8])
    np.testing.assert_array_equal(nci, [1, 3, 4, 6, 9, 0, 0, 0])
    assert nval.shape == (16,)
    assert nval[0] == 0x8001FF10 and (nval[1:8] == 0).all()
    assert (nval[8:15] == 0).all() and nval[15] == 0xA5000000


def _unaligned(T, nrp, nval):
    """Independent reading of the layout: a uint32 word in big-endian byte
    order is its GS tiles' row pieces in order."""
    GS = 32 // T
    be = nval.astype(">u4").view(np.uint8 if T == 8 else ">u2")  # [groups*T*GS]
    g = be.reshape(-1, T, GS).transpose(0, 2, 1)                   # [groups][GS][T]
    return g.reshape(-1, T).astype(oracle.WORD_DTYPE[T])           # padded tiles x rows


@pytest.mark.parametrize("T", [8, 16])
@pytest.mark.parametrize("shape", [(1, 1), (70, 70), (300, 517)])
def test_aligned_pack_round_trip(T, shape):
    g = random_csr(*shape, density=0.05, seed=shape[0] + T)
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    nrp, nci, nval = oracle.b2sr_aligned_pack(T, brp, bci, words)
    GS = 32 // T
    n_tiles = np.diff(brp)
    np.testing.assert_array_equal(np.diff(nrp), -(-n_tiles // GS) * GS)
    back = _unaligned(T, nrp, nval) if len(nval) else np.zeros((0, T))
    old = words.reshape(-1, T)
    for br in range(len(brp) - 1):
        a, n, p = brp[br], n_tiles[br], nrp[br]
        np.testing.assert_array_equal(back[p:p + n], old[a:a + n])
        np.testing.assert_array_equal(nci[p:p + n], bci[a:a + n])
        assert (back[p + n:nrp[br + 1]] == 0).all() and (nci[p + n:nrp[br + 1]] == 0).all()


def test_aligned_pack_rejects_other_T():
    with pytest.raises(ValueError):
        oracle.b2sr_aligned_pack(4, np.array([0, 1], np.int32), np.array([0], np.int32),
                                 np.zeros(4, np.uint8))


def _pack_x(x01, T):
    """x bit-packed per block column, bit T-1-c = x[T*bc + c] (numpy.packbits)."""
    KB = -(-len(x01) // T)
    xp = np.zeros(KB * T, np.uint8)
    xp[:len(x01)] = x01
    b = np.packbits(xp.reshape(KB, T), axis=1)  # big-endian bytes per block column
    return b[:, 0] if T == 8 else b.view(">u2")[:, 0].astype(np.uint16)


@pytest.mark.parametrize("T", [8, 16])
@pytest.mark.parametrize("shape", [(1, 1), (33, 70), (250, 250), (400, 1030)])
def test_aligned_spmv_matches_dense(T, shape):
    g = random_csr(*shape, density=0.04, seed=7 + T)
    x = (np.random.default_rng(T).random(g.K) < 0.5).astype(np.uint8)
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    nrp, nci, nval = oracle.b2sr_aligned_pack(T, brp, bci, words)
    cnt, yb = oracle.spmv_b2sr_aligned(T, nrp, nci, nval, _pack_x(x, T))
    want = dense(g, binary=True).astype(np.int64) @ x
    MB = len(brp) - 1
    full = np.zeros(MB * T, np.int64)
    full[:g.M] = want
    np.testing.assert_array_equal(cnt, full)
    np.testing.assert_array_equal(yb, _pack_x((full > 0).astype(np.uint8), T))


def test_planted_group_order_is_caught():
    """Reading the group's tiles lowest-first (a plausible slip) breaks the
    dense check above."""
    T = 8
    g = random_csr(64, 256, density=0.2, seed=1)
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    nrp, nci, nval = oracle.b2sr_aligned_pack(T, brp, bci, words)
    bad = nval.astype(">u4").view("<u4").astype(np.uint32)  # byte-reversed words
    x = (np.random.default_rng(0).random(g.K) < 0.5).astype(np.uint8)
    cnt, _ = oracle.spmv_b2sr_aligned(T, nrp, nci, bad, _pack_x(x, T))
    assert not np.array_equal(cnt[:g.M], dense(g, binary=True).astype(np.int64) @ x)


@pytest.mark.parametrize("T", [2, 4, 8, 16, 32, 64])
def test_csr2b2sr_matches_torch_bsr(T):
    """A library routine as a second pin: torch's CPU CSR -> BSR conversion
    of the same 0/1 matrix (padded to whole blocks) has the same block row
    pointers and block columns, and its blocks pack to the same words."""
    import torch

    g = random_csr(300, 500, 0.02, seed=T, long_rows=1, long_len=500)
    MB, KB = -(-g.M // T), -(-g.K // T)
    rp = np.concatenate([g.rowptr, np.full(MB * T - g.M, g.rowptr[-1], np.int32)])
    A = torch.sparse_csr_tensor(torch.from_numpy(rp).long(), torch.from_numpy(g.colind).long(),
                                torch.ones(g.nnz), size=(MB * T, KB * T))
    bsr = A.to_sparse_bsr((T, T))
    brp, bci, words = oracle.csr2b2sr(g.rowptr, g.colind, T)
    np.testing.assert_array_equal(bsr.crow_indices().numpy(), brp)
    np.testing.assert_array_equal(bsr.col_indices().numpy(), bci)
    nz = bsr.values().numpy() != 0  # [nb][T][T]
    weights = np.array([1 << (T - 1 - c) for c in range(T)], dtype=object)
    ref = [int((nz[b, r].astype(object) * weights).sum()) for b in range(nz.shape[0]) for r in range(T)]
    assert [int(x) for x in words] == ref
