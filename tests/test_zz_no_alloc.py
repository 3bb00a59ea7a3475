"""muxi_spmv_merge_based_noalloc_f32: the merge-path SpMV with its scratch in a static device array.

On the GPU its y equals muxi_spmv_merge_based_f32's bit for bit (the same kernels on the same partitions), on
two calls in a row, the second with a smaller matrix, and again on a third call back at the first size (the
array keeps the previous call's contents, which no kernel may read). On the CPU: argument errors, nnz == 0 and
the size limit are answered before any HIP call.
"""
import ctypes as C

import pytest
import torch

from muxi_spmv_amd import _lib, api, tune

LIMIT = 357_890_048  # numRows + nnz the 2 MiB scratch holds (174,751 partitions of 2,048 path items)


def _call(fn, A, x, y):
    s = torch.cuda.current_stream().cuda_stream or None
    return fn(C.byref(A.c), C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), C.c_void_p(s) if s else None)


def test_noalloc_rejects_before_any_device_work():
    lib = _lib.load()
    assert lib.muxi_spmv_merge_based_noalloc_f32(None, None, None, None) == -2
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 4, 4, 0
    assert lib.muxi_spmv_merge_based_noalloc_f32(C.byref(a), None, None, None) == 0  # nnz == 0: y untouched
    a.nnz = 5  # device arrays missing
    assert lib.muxi_spmv_merge_based_noalloc_f32(C.byref(a), None, None, None) == -2
    a.d_rowPtr = a.d_colIdx = a.d_values = 16  # never dereferenced: every case below is rejected first
    assert lib.muxi_spmv_merge_based_noalloc_f32(C.byref(a), None, None, None) == -3
    a.numRows, a.nnz = 1_000, LIMIT - 1_000 + 1
    assert lib.muxi_spmv_merge_based_noalloc_f32(C.byref(a), C.c_void_p(16), C.c_void_p(16), None) == -2
    msg = lib.muxi_spmv_last_error().decode()
    assert "static scratch" in msg and f"up to {LIMIT}" in msg, msg
    a.numRows, a.nnz = 2_000_000_000, 500_000_000
    assert lib.muxi_spmv_merge_based_noalloc_f32(C.byref(a), C.c_void_p(16), C.c_void_p(16), None) == -2
    assert "INT32_MAX" in lib.muxi_spmv_last_error().decode()


@pytest.mark.gpu
def test_noalloc_matches_merge_based_bitwise_on_successive_calls():
    lib = _lib.load()
    cases = [("powerlaw", 400_000, (2.0, 1, 50_000), 11), ("uniform", 3_000, (2, 9, 0), 12),
             ("powerlaw", 400_000, (2.0, 1, 50_000), 13)]
    for fam, R, p, seed in cases:
        A = tune.gen_matrix(fam, R, R, *p, seed=seed)
        try:
            L = A.numRows + A.nnz
            assert L <= LIMIT
            print(f"noalloc case {fam} R={R} nnz={A.nnz} seed={seed}", flush=True)
            x = tune.gen_vector(A.numCols, seed=seed)
            y_ref = torch.full((A.numRows,), float("nan"), device="cuda")
            y = torch.full((A.numRows,), float("nan"), device="cuda")
            assert _call(lib.muxi_spmv_merge_based_f32, A, x, y_ref) == 0
            assert _call(lib.muxi_spmv_merge_based_noalloc_f32, A, x, y) == 0, lib.muxi_spmv_last_error()
            torch.cuda.synchronize()
            assert torch.equal(y, y_ref), (fam, R, float((y - y_ref).abs().max()))
            y64 = api.oracle_f64(A, x)
            err = float(torch.linalg.vector_norm(y.double() - y64) / torch.linalg.vector_norm(y64))
            assert err <= 1e-6, err
        finally:
            A.free()


@pytest.mark.gpu
def test_noalloc_repeated_on_a_side_stream():
    """On a non-default stream, several calls back to back on the same stream each give the same y."""
    lib = _lib.load()
    A = tune.gen_matrix("uniform", 50_000, 50_000, 3, 12, 0, seed=21)
    try:
        x = tune.gen_vector(A.numCols, seed=21)
        y_ref = torch.empty(A.numRows, device="cuda")
        assert _call(lib.muxi_spmv_merge_based_f32, A, x, y_ref) == 0
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        ys = [torch.full((A.numRows,), float("nan"), device="cuda") for _ in range(4)]
        torch.cuda.synchronize()  # the fills above ran on the default stream
        with torch.cuda.stream(s):
            for y in ys:
                assert _call(lib.muxi_spmv_merge_based_noalloc_f32, A, x, y) == 0
        s.synchronize()
        for y in ys:
            assert torch.equal(y, y_ref)
    finally:
        A.free()
