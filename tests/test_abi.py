"""C-ABI checks that need no GPU: the library loads, exports every function
include/spmm_mi355x.h declares, and rejects bad arguments before touching
the device (the status codes of the header)."""
import ctypes as C

import pytest

from spmm_eval_amd import _lib


def test_library_exports_every_declared_symbol():
    names = _lib.header_functions()
    assert len(names) >= 25
    L = _lib.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    assert set(_lib._SIGS) | {"spmm_status_string", "spmm_last_error"} >= set(names)


def test_status_strings():
    L = _lib.lib()
    assert L.spmm_status_string(0) == b"SPMM_OK"
    assert L.spmm_status_string(2) == b"SPMM_ERR_UNSUPPORTED"
    assert L.spmm_version() >= 100


NULL = C.c_void_p(0)


@pytest.mark.parametrize("call,args", [
    ("bmm32_xnor_f32", (-1, 1, 1, NULL, NULL, NULL, 1, NULL)),
    ("bmm32_xnor_f32", (4, 4, 1, NULL, NULL, C.c_void_p(16), 4, NULL)),   # null A
    ("bmm32_xnor_bin", (4, 4, -3, NULL, NULL, NULL, NULL)),
    ("spmm_b2sr_f32", (NULL, 5, 4, 4, 4, NULL, NULL, NULL, NULL, 4, NULL, 4, NULL)),  # T=5
    ("spmm_b2sr_signpacked_i32", (NULL, 4, 4, 4, 64, C.c_void_p(16), NULL, NULL, C.c_void_p(16), 1,
                                  C.c_void_p(16), 64, NULL)),  # ldbw < 2
    ("csr2b2sr_nnzb", (NULL, 4, 4, 4, NULL, NULL, 4, NULL, None, NULL)),  # null handle
    ("spmm_csr_f32_planned", (NULL, 4, 4, 4, 4, NULL, NULL, NULL, NULL, 4, NULL, 4, None, NULL)),
    ("pack_sign_rows_u32", (4, 40, C.c_void_p(16), 39, C.c_void_p(16), NULL)),  # ldx < W
    ("sddmm_csr_f32", (NULL, 2, 2, 4, 2, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), 2,
                       C.c_void_p(16), 4, C.c_void_p(16), NULL)),  # lddy < N
    ("spmm_batch_plan_create", (NULL, 1, NULL, NULL, NULL, NULL, NULL, 8, 0, None, NULL)),
    ("b2sr_aligned_nnzb", (C.c_void_p(16), 4, 3, C.c_void_p(16), C.c_void_p(16), None, NULL)),
    ("b2sr_aligned_pack", (C.c_void_p(16), 16, 3, 3, NULL, NULL, NULL, NULL, NULL, NULL, NULL)),  # 3 % 2
    ("spmv_b2sr_aligned_bin", (C.c_void_p(16), 8, 16, 16, 4, C.c_void_p(16), C.c_void_p(16),
                               C.c_void_p(16), C.c_void_p(16), NULL, NULL, NULL)),  # no output
    ("spmv_b2sr_aligned_f32", (C.c_void_p(16), 32, 16, 16, 4, NULL, NULL, NULL, NULL,
                               C.c_void_p(16), NULL)),  # T = 32 has no aligned form
])
def test_invalid_arguments_rejected_on_host(call, args):
    L = _lib.lib()
    st = getattr(L, call)(*args)
    assert st == 1, (call, st, L.spmm_last_error())
    assert L.spmm_last_error()  # a message is recorded


def test_zero_sized_calls_are_noops():
    L = _lib.lib()
    assert L.bmm32_xnor_f32(0, 5, 1, NULL, NULL, NULL, 5, NULL) == 0
    assert L.spmm_b2sr_f32(NULL, 4, 0, 4, 4, NULL, NULL, NULL, NULL, 4, NULL, 4, NULL) == 0
    assert L.pack_sign_rows_u32(0, 0, NULL, 0, NULL, NULL) == 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="missing"):
        _lib.lib()


def test_dropin_geometry_on_the_host():
    """spmm_csr_dropin_geometry (host only): the plan-free CSR SpMM slices a
    column tile of B of about 15 MB or more into slices of about 10 MB when
    rows average at least 16 entries a slice -- the autotuner's slice counts
    on the 232,965-row graph -- and otherwise runs the default geometry."""
    from spmm_eval_amd import api

    M, nnz = 232_965, 114_506_280
    assert api.dropin_geometry(M, M, 128, nnz) == (32, 1, 1, 12)
    assert api.dropin_geometry(M, M, 64, nnz) == (16, 1, 1, 6)
    assert api.dropin_geometry(M, M, 128, nnz, bf16=True) == (16, 1, 1, 6)
    assert api.dropin_geometry(M, M, 64, nnz, bf16=True) == (8, 1, 1, 3)
    # short rows (mean degree 10) or a small B: one launch at the default geometry
    assert api.dropin_geometry(89_250, 89_250, 128, 900_036) == (8, 1, 1, 0)
    assert api.dropin_geometry(16_384, 16_384, 32, 259_972) == (8, 1, 1, 0)
    # 15 MB tile: two slices; just below: none
    assert api.dropin_geometry(30_000, 30_000, 128, 1_500_000)[3] == 2
    assert api.dropin_geometry(29_000, 29_000, 128, 1_500_000)[3] == 0
    # a tile wider than 32 lanes, or N not a whole number of 16-byte vectors
    assert api.dropin_geometry(M, M, 256, nnz)[3] == 0
    assert api.dropin_geometry(100, 100, 7, 500)[:2] == (0, 0)
    with pytest.raises(_lib.SpmmError):  # elem_bytes neither 4 nor 2
        _lib.call("spmm_csr_dropin_geometry", 10, 10, 64, 100, 3, C.byref(_lib.Tune()))
    with pytest.raises(_lib.SpmmError):
        _lib.call("spmm_csr_dropin_geometry", -1, 10, 64, 100, 4, C.byref(_lib.Tune()))
