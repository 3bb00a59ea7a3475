"""The C-ABI library loads on a CPU host, exports every declared symbol, and
its host-side error paths return the reference's status codes without
touching a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from muxi_spmv_amd import _lib
import oracle

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").rglob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(muxi_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = declared_functions()
    assert len(decl) > 30
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers the header exactly
    assert set(decl) == set(_lib.SIGNATURES), set(decl) ^ set(_lib.SIGNATURES)


def test_enum_values_match_reference():
    # src/utils/common.h:32-71
    assert _lib.STATUS_NAMES[-1] == "SPMV_ERROR_INVALID_HANDLE"
    assert _lib.STATUS_NAMES[-9] == "SPMV_ERROR_INTERNAL"
    assert (_lib.SPMV_FORMAT_CSR, _lib.SPMV_FORMAT_CSR5) == (0, 4)
    assert (_lib.SPMV_TYPE_FLOAT, _lib.SPMV_TYPE_DOUBLE, _lib.SPMV_TYPE_BF16) == (0, 1, 3)


def test_struct_layouts():
    # field order of CSRMatrix<T> (sparse_formats.h:26-40): 3 ints, 6 pointers, 2 bools
    assert [f[0] for f in _lib.muxi_csr_t._fields_][:6] == ["numRows", "numCols", "nnz", "rowPtr", "colIdx",
                                                             "values"]
    assert C.sizeof(_lib.muxi_csr_t) == 4 * 3 + 4 + 8 * 6 + 8
    assert C.sizeof(_lib.spmv_opts_t) == 6 * 4 + 8


def test_default_opts():
    o = _lib.load().muxi_spmv_default_opts()
    assert (o.operation, o.sync, o.check_inputs, o.use_tensor_core, o.use_streaming, o.stream_id) == (0,) * 6
    assert not o.stream


def test_null_handle_errors():
    lib = _lib.load()
    assert lib.muxi_spmv_create_handle(None) == -1
    assert lib.muxi_spmv_destroy_handle(None) == -1
    assert lib.muxi_spmv_set_stream(None, None) == -1
    assert b"NULL" in lib.muxi_spmv_last_error()
    assert lib.muxi_spmv_exec_f32(None, 0, None, None, None, 1.0, 0.0) == -1


def test_null_matrix_and_vector_errors():
    lib = _lib.load()
    assert lib.muxi_spmv_csr_f32(None, None, None, 1.0, 0.0, None) == -2
    assert lib.muxi_spmv_merge_based_f64(None, None, None, None) == -2
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 10, 10, 5  # device arrays missing
    assert lib.muxi_spmv_csr_f32(C.byref(a), None, None, 1.0, 0.0, None) == -2
    p = C.c_void_p()
    assert lib.muxi_plan_create(C.byref(a), 0, 0, 5, 8, 0, None, C.byref(p)) == -2
    assert lib.muxi_plan_create(None, 0, 0, 5, 8, 0, None, C.byref(p)) == -2
    assert lib.muxi_plan_exec(None, None, None, 1.0, 0.0, 1, None) == -1
    assert lib.muxi_csr5_preprocess(None, None, 0, None) == -2
    assert lib.muxi_spmv_ellpack_f32(None, None, None, None) == -2


def test_nnz_zero_is_success_without_device_work():
    # spmv_csr.cu:61-63, spmv_csr5.cu:229-231 / :285-287: nnz == 0 returns success, y untouched
    lib = _lib.load()
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 4, 4, 0
    assert lib.muxi_spmv_csr_f32(C.byref(a), None, None, 1.0, 0.0, None) == 0
    assert lib.muxi_spmv_merge_based_f32(C.byref(a), None, None, None) == 0
    c5 = _lib.muxi_csr5_t()
    assert lib.muxi_csr5_preprocess(C.byref(a), C.byref(c5), 0, None) == 0
    assert c5.numTiles == 0
    assert lib.muxi_spmv_csr5_f32(C.byref(c5), None, None, 1.0, 0.0, None) == 0


def test_int32_overflow_guard():
    lib = _lib.load()
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 2_000_000_000, 10, 500_000_000
    a.d_rowPtr = a.d_colIdx = a.d_values = 16  # never dereferenced: rejected first
    assert lib.muxi_spmv_csr_f32(C.byref(a), C.c_void_p(16), C.c_void_p(16), 1.0, 0.0, None) == -2
    assert b"INT32_MAX" in lib.muxi_spmv_last_error()


def test_unsupported_type_combination():
    lib = _lib.load()
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 4, 4, 0
    a.d_rowPtr = 16
    p = C.c_void_p()
    # bf16 values with FP64 x/accumulation is not built
    assert lib.muxi_plan_create(C.byref(a), _lib.SPMV_TYPE_BF16, _lib.SPMV_TYPE_DOUBLE, 5, 8, 0, None,
                                C.byref(p)) == -8


@pytest.mark.parametrize("R,N,ePP,mult", [(2_000_000, 14_700_000, 16, 8), (100_000, 1_000_000, 16, 8),
                                          (10, 5, 16, 8), (2_000_000, 14_700_000, 64, 4), (0, 0, 16, 8)])
def test_partition_count_matches_oracle(R, N, ePP, mult):
    assert _lib.load().muxi_merge_partition_count(R, N, ePP, mult, 104) == \
        oracle.merge_partition_count(R, N, ePP, mult, 104, 4)


def test_format_conversion_host_side_errors():
    """Argument errors of the conversions are reported before any device work."""
    lib = _lib.load()
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 10, 10, 5  # device arrays missing
    coo, bsr, out = _lib.muxi_coo_t(), _lib.muxi_bsr_t(), _lib.muxi_csr_t()
    assert lib.muxi_csr_to_coo(None, 0, C.byref(coo), None) == -2
    assert lib.muxi_csr_to_coo(C.byref(a), 2, C.byref(coo), None) == -8  # HALF values: no
    assert lib.muxi_csr_to_coo(C.byref(a), 0, C.byref(coo), None) == -2
    assert lib.muxi_csr_transpose(C.byref(a), 0, C.byref(out), None) == -2
    assert lib.muxi_csr_to_bsr(C.byref(a), 0, 0, C.byref(bsr), None) == -2  # block size 0
    bsr.numRows, bsr.numCols, bsr.blockSize, bsr.numBlockRows, bsr.numBlockCols = 10, 10, 4, 2, 3  # wrong count
    assert lib.muxi_bsr_to_csr(C.byref(bsr), 0, C.byref(out), None) == -2
    coo.numRows, coo.numCols, coo.nnz = 3, 3, 2  # no arrays
    assert lib.muxi_coo_to_csr(C.byref(coo), 0, C.byref(out), None) == -2
    for f in (lib.muxi_csr_free_device, lib.muxi_coo_free_device, lib.muxi_bsr_free_device):
        assert f(None) == -2
    assert lib.muxi_csr_free_device(C.byref(a)) == 0  # not owned: untouched
    assert C.sizeof(_lib.muxi_bsr_t) == 6 * 4 + 6 * 8 + 8 and C.sizeof(_lib.muxi_coo_t) == 4 * 3 + 4 + 6 * 8 + 8


def test_host_plan_host_side_errors():
    lib = _lib.load()
    p = C.c_void_p()
    assert lib.muxi_host_plan_create(None, 0, 0, 4, 0, 0, C.byref(p)) == -2
    a = _lib.muxi_csr_t()
    a.numRows, a.numCols, a.nnz = 10, 10, 5  # device arrays missing
    assert lib.muxi_host_plan_create(C.byref(a), 0, 0, 4, 0, 0, C.byref(p)) == -2
    assert lib.muxi_host_plan_create(C.byref(a), 0, 0, 4, 0, 0, None) == -1
    assert lib.muxi_host_plan_exec(None, None, None, 1.0, 0.0) == -1
    assert lib.muxi_host_plan_timing(None, None, None, None, None) == -1
    assert lib.muxi_host_plan_destroy(None) == -1
    assert lib.muxi_host_alloc(None, 16) == -3
    assert lib.muxi_host_free(None) == 0


def test_python_wrappers_reject_bad_vectors_before_any_launch():
    # Every wrapper checks x and y (dtype, contiguity, length, device) before it hands raw addresses to the
    # library: a short y would be written out of bounds, a host tensor read as a device address.
    import torch

    from muxi_spmv_amd import api
    from muxi_spmv_amd._lib import SpmvError

    rp = np.array([0, 2, 3, 5], np.int32)
    ci = np.array([0, 2, 1, 0, 3], np.int32)
    A = api.CSRMatrix(rp, ci, np.ones(5, np.float32), numCols=4, device="cpu")
    good_x, good_y = torch.zeros(4), torch.zeros(3)
    cases = [
        (torch.zeros(3), good_y, "x needs >= 4"),
        (good_x, torch.zeros(2), "y needs >= 3"),
        (torch.zeros(4, dtype=torch.float64), good_y, "x must be torch.float32"),
        (torch.zeros(8)[::2], good_y, "x must be contiguous"),
        (good_x, good_y, "x must be on the GPU"),
        ([0.0] * 4, good_y, "x must be a torch tensor"),
    ]
    for fn in (api.spmv_csr, api.spmv_merge_based, api.spmv_sparse_optimized):
        for x, y, msg in cases:
            with pytest.raises(SpmvError, match=msg) as ei:
                fn(A, x, y)
            assert ei.value.status == -3
    # transpose: x has numRows entries, y numCols
    with pytest.raises(SpmvError, match="y needs >= 4"):
        api.spmv_csr(A, torch.zeros(3), torch.zeros(3), opts=_opts_transpose())
    # a plan checks against its x type and nrhs without touching the device
    p = api.Plan.__new__(api.Plan)
    p.A, p.xtype, p.nrhs, p._p = A, torch.float32, 2, None
    with pytest.raises(SpmvError, match="x needs >= 8"):
        p.exec(torch.zeros(4), torch.zeros(6))
    with pytest.raises(SpmvError, match="y needs >= 6"):
        p.exec(torch.zeros(8), torch.zeros(3))


def _opts_transpose():
    o = _lib.spmv_opts_t()
    o.operation = _lib.SPMV_OP_TRANSPOSE
    return o


def test_checked_plan_argument_errors_before_any_device_work():
    lib = _lib.load()
    p = C.c_void_p(1)
    a = _lib.muxi_csr_t()
    for bad in (0.0, -1e-3, 1.0, float("nan")):
        assert lib.muxi_plan_create_checked(C.byref(a), 0, 0, bad, None, C.byref(p)) == -9
        assert b"rtol" in lib.muxi_spmv_last_error()
        assert p.value is None  # the output is cleared
    n = C.c_int(7)
    assert lib.muxi_plan_checked_rows(None, C.byref(n), None, 0, None) == -1
