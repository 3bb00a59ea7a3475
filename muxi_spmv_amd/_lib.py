"""ctypes binding of libmuxi_spmv_amd.so (the C-ABI in include/muxi_spmv/muxi_spmv.h).

Loading never falls back to anything: a missing or stale library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MUXI_SPMV_LIB", _PKG / "libmuxi_spmv_amd.so"))

# ---- enums (src/utils/common.h:32-71)
SPMV_SUCCESS = 0
STATUS_NAMES = {
    0: "SPMV_SUCCESS", -1: "SPMV_ERROR_INVALID_HANDLE", -2: "SPMV_ERROR_INVALID_MATRIX",
    -3: "SPMV_ERROR_INVALID_VECTOR", -4: "SPMV_ERROR_MEMORY_ALLOC", -5: "SPMV_ERROR_MEMORY_COPY",
    -6: "SPMV_ERROR_DEVICE", -7: "SPMV_ERROR_UNSUPPORTED_FORMAT", -8: "SPMV_ERROR_UNSUPPORTED_TYPE",
    -9: "SPMV_ERROR_INTERNAL",
}
SPMV_FORMAT_CSR, SPMV_FORMAT_COO, SPMV_FORMAT_CSR2, SPMV_FORMAT_BSR, SPMV_FORMAT_CSR5 = 0, 1, 2, 3, 4
SPMV_TYPE_FLOAT, SPMV_TYPE_DOUBLE, SPMV_TYPE_HALF, SPMV_TYPE_BF16 = 0, 1, 2, 3
SPMV_OP_NORMAL, SPMV_OP_TRANSPOSE = 0, 1

KERNELS = {"auto": 0, "scalar": 1, "vector": 2, "vwave": 3, "adaptive": 4, "merge": 5, "csr5": 6, "ell": 7,
           "xslab": 8, "lslab": 9}
KERNEL_NAMES = {v: k for k, v in KERNELS.items()}


class spmv_opts_t(C.Structure):
    _fields_ = [("operation", C.c_int), ("sync", C.c_int), ("check_inputs", C.c_int),
                ("use_tensor_core", C.c_int), ("use_streaming", C.c_int), ("stream_id", C.c_int),
                ("stream", C.c_void_p)]


class muxi_csr_t(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int),
                ("rowPtr", C.c_void_p), ("colIdx", C.c_void_p), ("values", C.c_void_p),
                ("d_rowPtr", C.c_void_p), ("d_colIdx", C.c_void_p), ("d_values", C.c_void_p),
                ("ownsHostMemory", C.c_bool), ("ownsDeviceMemory", C.c_bool)]


class muxi_csr5_t(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int),
                ("rowPtr", C.c_void_p), ("colIdx", C.c_void_p), ("values", C.c_void_p),
                ("sigma", C.c_int), ("numTiles", C.c_int),
                ("tile_row_ptr", C.POINTER(C.c_int)), ("tile_nnz_offset", C.POINTER(C.c_int)),
                ("d_rowPtr", C.c_void_p), ("d_colIdx", C.c_void_p), ("d_values", C.c_void_p),
                ("d_tile_row_ptr", C.c_void_p), ("d_tile_nnz_offset", C.c_void_p),
                ("ownsCsrData", C.c_bool), ("ownsTileData", C.c_bool), ("ownsDeviceMemory", C.c_bool),
                ("conversionTimeMs", C.c_double)]


class muxi_ell_t(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("numElements", C.c_int), ("nnz", C.c_int),
                ("d_colIdx", C.c_void_p), ("d_values", C.c_void_p)]


class muxi_coo_t(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int),
                ("rowIdx", C.c_void_p), ("colIdx", C.c_void_p), ("values", C.c_void_p),
                ("d_rowIdx", C.c_void_p), ("d_colIdx", C.c_void_p), ("d_values", C.c_void_p),
                ("ownsHostMemory", C.c_bool), ("ownsDeviceMemory", C.c_bool)]


class muxi_csr2_t(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int),
                ("rowPtr", C.c_void_p), ("values", C.c_void_p), ("metadata", C.c_void_p),
                ("d_rowPtr", C.c_void_p), ("d_values", C.c_void_p), ("d_metadata", C.c_void_p),
                ("ownsHostMemory", C.c_bool), ("ownsDeviceMemory", C.c_bool)]


class muxi_bsr_t(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("numBlockRows", C.c_int), ("numBlockCols", C.c_int),
                ("numBlocks", C.c_int), ("blockSize", C.c_int),
                ("rowPtr", C.c_void_p), ("colIdx", C.c_void_p), ("values", C.c_void_p),
                ("d_rowPtr", C.c_void_p), ("d_colIdx", C.c_void_p), ("d_values", C.c_void_p),
                ("ownsHostMemory", C.c_bool), ("ownsDeviceMemory", C.c_bool)]


class muxi_spmv_handle_t(C.Structure):
    _fields_ = [("deviceId", C.c_int), ("stream", C.c_void_p), ("opts", spmv_opts_t), ("warpSize", C.c_int),
                ("smCount", C.c_int), ("maxThreadsPerBlock", C.c_int), ("sharedMemPerSM", C.c_size_t)]


class muxi_perf_result_t(C.Structure):
    """PerformanceResult (performance_benchmark.h:21-53), as muxi_spmv_bench.h lays it out."""
    _fields_ = [("timeMs", C.c_double), ("gflops", C.c_double), ("bandwidthGBs", C.c_double),
                ("bandwidthUtilization", C.c_double), ("iterations", C.c_int), ("correctnessPassed", C.c_bool),
                ("minTimeMs", C.c_double), ("maxTimeMs", C.c_double), ("avgTimeMs", C.c_double),
                ("stdDevMs", C.c_double), ("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int),
                ("sparsity", C.c_double), ("deviceId", C.c_int), ("warpSize", C.c_int),
                ("peakBandwidthGBs", C.c_double), ("useCSR5", C.c_bool), ("conversionTimeMs", C.c_double),
                ("totalTimeMs", C.c_double), ("numTiles", C.c_int), ("tileSize", C.c_int),
                ("conversionOverheadPercent", C.c_double)]


class muxi_bench_config_t(C.Structure):
    """BenchmarkConfig (performance_benchmark.h:58-71), as muxi_spmv_bench.h lays it out."""
    _fields_ = [("warmupIterations", C.c_int), ("measureIterations", C.c_int), ("checkCorrectness", C.c_bool),
                ("syncAfterEach", C.c_bool), ("stream", C.c_void_p), ("floatType", C.c_int),
                ("compareCusparse", C.c_bool)]


_VP, _I, _D, _F, _Z = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_size_t
_P_CSR, _P_CSR5, _P_ELL = C.POINTER(muxi_csr_t), C.POINTER(muxi_csr5_t), C.POINTER(muxi_ell_t)
_P_OPTS = C.POINTER(spmv_opts_t)
_P_COO, _P_BSR = C.POINTER(muxi_coo_t), C.POINTER(muxi_bsr_t)
_P_H = C.POINTER(muxi_spmv_handle_t)

# name -> (restype, argtypes). Every symbol declared in include/muxi_spmv/*.h appears here.
SIGNATURES: dict[str, tuple] = {
    "muxi_spmv_default_opts": (spmv_opts_t, []),
    "muxi_spmv_create_handle": (_I, [C.POINTER(_P_H)]),
    "muxi_spmv_destroy_handle": (_I, [_P_H]),
    "muxi_spmv_set_stream": (_I, [_P_H, _VP]),
    "muxi_spmv_last_error": (C.c_char_p, []),
    "muxi_spmv_version": (C.c_char_p, []),
    "muxi_spmv_csr_f32": (_I, [_P_CSR, _VP, _VP, _F, _F, _P_OPTS]),
    "muxi_spmv_csr_f64": (_I, [_P_CSR, _VP, _VP, _D, _D, _P_OPTS]),
    "muxi_spmv_exec_f32": (_I, [_P_H, _I, _VP, _VP, _VP, _F, _F]),
    "muxi_spmv_exec_f64": (_I, [_P_H, _I, _VP, _VP, _VP, _D, _D]),
    "muxi_spmv_exec_transpose_f32": (_I, [_P_H, _I, _VP, _VP, _VP, _F, _F]),
    "muxi_spmv_exec_transpose_f64": (_I, [_P_H, _I, _VP, _VP, _VP, _D, _D]),
    "muxi_spmv_merge_based_f32": (_I, [_P_CSR, _VP, _VP, _VP]),
    "muxi_spmv_merge_based_f64": (_I, [_P_CSR, _VP, _VP, _VP]),
    "muxi_spmv_merge_based_v2_f32": (_I, [_P_CSR, _VP, _VP, _VP, C.c_bool]),
    "muxi_spmv_merge_based_v2_f64": (_I, [_P_CSR, _VP, _VP, _VP, C.c_bool]),
    "muxi_spmv_merge_based_noalloc_f32": (_I, [_P_CSR, _VP, _VP, _VP]),
    "muxi_spmv_sparse_optimized_f32": (_I, [_P_CSR, _VP, _VP, _VP]),
    "muxi_spmv_sparse_optimized_f64": (_I, [_P_CSR, _VP, _VP, _VP]),
    "muxi_merge_partition_count": (_I, [_I, _I, _I, _I, _I]),
    "muxi_merge_partitions": (_I, [_I, _I, _I, _VP, _VP]),
    "muxi_merge_path_coords": (_I, [_VP, _I, _I, _VP, _I, _VP, _VP, _VP]),
    "muxi_csr5_preprocess": (_I, [_P_CSR, _P_CSR5, _I, _VP]),
    "muxi_csr5_destroy": (_I, [_P_CSR5]),
    "muxi_spmv_csr5_f32": (_I, [_P_CSR5, _VP, _VP, _F, _F, _P_OPTS]),
    "muxi_spmv_csr5_f64": (_I, [_P_CSR5, _VP, _VP, _D, _D, _P_OPTS]),
    "muxi_csr_to_ellpack_f32": (_I, [_P_CSR, _P_ELL, _I]),
    "muxi_csr_to_ellpack_f64": (_I, [_P_CSR, _P_ELL, _I]),
    "muxi_spmv_ellpack_f32": (_I, [_P_ELL, _VP, _VP, _VP]),
    "muxi_spmv_ellpack_f64": (_I, [_P_ELL, _VP, _VP, _VP]),
    "muxi_ellpack_destroy": (_I, [_P_ELL]),
    "muxi_plan_create": (_I, [_P_CSR, _I, _I, _I, _I, _I, _VP, C.POINTER(_VP)]),
    "muxi_plan_create_rows": (_I, [_P_CSR, _I, _I, _I, _I, _I, _VP, C.c_longlong, C.c_longlong, C.POINTER(_VP)]),
    "muxi_plan_destroy": (_I, [_VP]),
    "muxi_plan_create_spmm": (_I, [_P_CSR, _I, _I, _I, _I, _VP, C.POINTER(_VP)]),
    "muxi_plan_create_checked": (_I, [_P_CSR, _I, _I, C.c_double, _VP, C.POINTER(_VP)]),
    "muxi_plan_checked_rows": (_I, [_VP, C.POINTER(_I), C.POINTER(_I), _I, _VP]),
    "muxi_plan_exec": (_I, [_VP, _VP, _VP, _D, _D, _I, _VP]),
    "muxi_plan_info": (_I, [_VP, C.POINTER(_I), C.POINTER(_I), C.POINTER(_Z)]),
    "muxi_plan_exec_ev": (_I, [_VP, _VP, _VP, _D, _D, _I, _VP, _VP, _VP]),
    "muxi_event_create": (_I, [C.POINTER(_VP)]),
    "muxi_event_destroy": (_I, [_VP]),
    "muxi_event_record": (_I, [_VP, _VP]),
    "muxi_event_elapsed_ms": (_I, [_VP, _VP, C.POINTER(C.c_float)]),
    "muxi_spmv_oracle_f64": (_I, [_P_CSR, _I, _I, _VP, _VP, _VP]),
    "muxi_convert_values": (_I, [_VP, _I, _VP, _I, _Z, _VP]),
    "muxi_device_copy": (_I, [_VP, _VP, _Z, _VP]),
    # muxi_spmv_prep.h
    "muxi_reorder_columns": (_I, [_I, _I, _VP, _VP, _VP, _VP, _VP]),
    "muxi_reorder_x_f32": (_I, [_VP, _VP, _VP, _I]),
    "muxi_sort_columns_within_rows_f32": (_I, [_I, _VP, _VP, _VP]),
    "muxi_reorder_columns_minrow": (_I, [_I, _I, _VP, _VP, _VP, _VP]),
    "muxi_reorder_columns_by_frequency": (_I, [_I, _I, _VP, _VP, _VP, _VP]),
    "muxi_reorder_rows_by_length_f32": (_I, [_I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "muxi_rcm": (_I, [_I, _VP, _VP, _VP]),
    "muxi_permute_symmetric": (_I, [_I, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP]),
    "muxi_bandwidth": (_I, [_I, _VP, _VP]),
    "muxi_mtx_read_header": (_I, [C.c_char_p, C.c_void_p]),
    "muxi_mtx_read_csr_f32": (_I, [C.c_char_p, _I, _P_CSR]),
    "muxi_mtx_read_csr_f64": (_I, [C.c_char_p, _I, _P_CSR]),
    "muxi_mtx_write_csr_f32": (_I, [C.c_char_p, _P_CSR, _I]),
    "muxi_mtx_write_csr_f64": (_I, [C.c_char_p, _P_CSR, _I]),
    "muxi_csr_free_host": (_I, [_P_CSR]),
    "muxi_mtx_read_coo_f32": (_I, [C.c_char_p, _P_COO]),
    "muxi_mtx_read_coo_f64": (_I, [C.c_char_p, _P_COO]),
    "muxi_mtx_write_coo_f32": (_I, [C.c_char_p, _P_COO]),
    "muxi_mtx_write_coo_f64": (_I, [C.c_char_p, _P_COO]),
    "muxi_coo_free_host": (_I, [_P_COO]),
    "muxi_host_plan_create": (_I, [_P_CSR, _I, _I, _I, _I, _I, C.POINTER(_VP)]),
    "muxi_host_plan_destroy": (_I, [_VP]),
    "muxi_host_plan_exec": (_I, [_VP, _VP, _VP, _D, _D]),
    "muxi_host_plan_timing": (_I, [_VP, C.POINTER(_F), C.POINTER(_F), C.POINTER(_F), C.POINTER(_F)]),
    "muxi_host_plan_chunk": (_I, [_VP, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "muxi_host_alloc": (_I, [C.POINTER(_VP), _Z]),
    "muxi_host_free": (_I, [_VP]),
    "muxi_csr_to_coo": (_I, [_P_CSR, _I, _P_COO, _VP]),
    "muxi_coo_to_csr": (_I, [_P_COO, _I, _P_CSR, _VP]),
    "muxi_csr_to_bsr": (_I, [_P_CSR, _I, _I, _P_BSR, _VP]),
    "muxi_bsr_to_csr": (_I, [_P_BSR, _I, _P_CSR, _VP]),
    "muxi_csr_transpose": (_I, [_P_CSR, _I, _P_CSR, _VP]),
    "muxi_csr_free_device": (_I, [_P_CSR]),
    "muxi_coo_free_device": (_I, [_P_COO]),
    "muxi_bsr_free_device": (_I, [_P_BSR]),
    "muxi_shard_rows": (_I, [_VP, _I, _I, _VP]),
    "muxi_remap_indices_dev": (_I, [_VP, _VP, _VP, _I, _VP]),
    "muxi_gather_dev": (_I, [_VP, _VP, _VP, _I, _I, _VP]),
    "muxi_scatter_dev": (_I, [_VP, _VP, _VP, _I, _I, _VP]),
    "muxi_permute_rows_dev": (_I, [_I, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _VP]),
    # muxi_spmv_batch.h
    "muxi_batch_plan_create": (_I, [_P_CSR, _I, _I, _I, C.POINTER(_VP), C.POINTER(_VP), _I, _VP,
                                    C.POINTER(_VP)]),
    "muxi_batch_plan_destroy": (_I, [_VP]),
    "muxi_batch_exec": (_I, [_VP, _D, _D, _VP]),
    "muxi_batch_exec_ev": (_I, [_VP, _D, _D, _VP, _VP, _VP]),
    "muxi_batch_exec_per_matrix": (_I, [_VP, _I, _I, _VP]),
    "muxi_batch_info": (_I, [_VP, C.POINTER(_I), C.POINTER(_I), C.POINTER(_Z)]),
    "muxi_batch_schedule": (_I, [_VP, C.POINTER(_I), C.POINTER(_I)]),
    "muxi_batch_schedule_for": (_I, [_D, _D, C.POINTER(_I), C.POINTER(_I)]),
    "muxi_spmm_schedule_for": (_I, [_I, _D]),
    "muxi_batch_graph_create": (_I, [_VP, _D, _D, _I, _I, _VP, C.POINTER(_VP)]),
    "muxi_graph_launch": (_I, [_VP, _VP]),
    "muxi_graph_destroy": (_I, [_VP]),
    # muxi_spmv_tune.h
    "muxi_gen_matrix": (_I, [_I, _I, _I, _D, _D, _D, _I, _I, C.c_uint64, _VP, _P_CSR]),
    "muxi_gen_permute": (_I, [_P_CSR, _I, C.c_uint64, _VP]),
    "muxi_gen_free": (_I, [_P_CSR]),
    "muxi_gen_vector": (_I, [_VP, _I, C.c_long, C.c_uint64, _VP]),
    "muxi_matrix_features": (_I, [_P_CSR, _VP, C.POINTER(C.c_double)]),
    "muxi_flush_caches": (_I, [_VP, _Z, _VP]),
    "muxi_mlp_param_count": (C.c_long, [_I, _I, _I]),
    "muxi_mlp_workspace_bytes": (_Z, [_I, _I, _I, _I]),
    "muxi_mlp_forward": (_I, [_VP, _I, _I, _I, _VP, _I, _VP, _VP, _VP]),
    "muxi_mlp_backward": (_I, [_VP, _I, _I, _I, _VP, _I, _VP, _VP, _VP, _VP, _VP, _VP]),
    "muxi_mlp_adam": (_I, [_VP, _VP, _VP, _VP, C.c_long, _F, _F, _F, _F, _I, _VP]),
    "muxi_selector_set": (_I, [_VP, _I, _I, _I, _VP, _VP, _VP, _VP]),
    "muxi_selector_clear": (_I, []),
    "muxi_selector_reset": (_I, []),
    "muxi_selector_state": (_I, []),
    "muxi_selector_pick": (_I, [_P_CSR, _I, _VP, C.POINTER(_I), C.POINTER(_I)]),
    # muxi_spmv_bench.h
    "muxi_bench_default_config": (muxi_bench_config_t, []),
    "muxi_calculate_gflops": (_D, [_I, _D]),
    "muxi_calculate_bandwidth": (_D, [_I, _I, _I, _D, _I]),
    "muxi_measure_csr_performance": (_I, [_P_CSR, _VP, _VP, C.POINTER(muxi_bench_config_t), _P_H,
                                          C.POINTER(muxi_perf_result_t)]),
    "muxi_measure_csr5_performance": (_I, [_P_CSR, _P_CSR5, _VP, _VP, C.POINTER(muxi_bench_config_t), _P_H,
                                           C.POINTER(muxi_perf_result_t), _I]),
    "muxi_write_results_json": (_I, [C.POINTER(muxi_perf_result_t), _I, C.c_char_p]),
    "muxi_compare_with_rocsparse": (_I, [_P_CSR, _VP, _VP, C.POINTER(muxi_bench_config_t), _I,
                                         C.POINTER(_D), C.POINTER(_D), C.POINTER(_D)]),
    "muxi_print_performance_result": (None, [C.POINTER(muxi_perf_result_t)]),
}


class muxi_column_reorder_stats_t(C.Structure):
    _fields_ = [("numReorders", C.c_int), ("avgDistanceBefore", C.c_float), ("avgDistanceAfter", C.c_float),
                ("improvementRatio", C.c_float)]

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads the library once; raises if it is missing or lacks a declared symbol."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} not built: run `python -m muxi_spmv_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError on a missing export
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class SpmvError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)} {detail}".strip())


def check(status: int, where: str) -> None:
    if status != SPMV_SUCCESS:
        detail = load().muxi_spmv_last_error().decode(errors="replace")
        raise SpmvError(status, where, detail)
