"""ctypes bindings of libsaruman_hip.so (declarations in include/saruman_hip.h).

The library is built in-tree (saruman_amd/lib/libsaruman_hip.so).  Loading fails
loudly when it is missing: there is no Python or CPU fallback for the alignment.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libsaruman_hip.so"

SRM_OK = 0
SRM_KERNEL = {"auto": 0, "plain": 1, "banded": 2}
SRM_CAND_STOP_ROW0 = 1
SRM_CAND_BYTES = 2
SRM_CODE_PAD = 256
SRM_MAX_L = 1000
SRM_MAX_W = 900

CAND_DTYPE = np.dtype([("read_off", "<u4"), ("win_off", "<u4"), ("ops_off", "<u4"),
                       ("L", "<u2"), ("w", "<u2"), ("flag", "u1"), ("k", "u1"),
                       ("bits", "u1"), ("pad", "u1")])
RES_DTYPE = np.dtype([("score", "<i2"), ("n_ops", "<u2"), ("end_row", "<u2"),
                      ("end_col", "<u2")])
assert CAND_DTYPE.itemsize == 20 and RES_DTYPE.itemsize == 8

_vp = C.c_void_p
_lib = None

# (name, restype, argtypes)
_PROTOS = [
    ("srm_version", C.c_char_p, []),
    ("srm_last_error", C.c_char_p, []),
    ("srm_device_count", C.c_int, []),
    ("srm_device_sync", C.c_int, [C.c_int]),
    ("srm_ctx_create", C.c_int, [C.c_int, C.c_size_t, C.POINTER(_vp)]),
    ("srm_ctx_destroy", None, [_vp]),
    ("srm_ctx_pool_used", C.c_size_t, [_vp]),
    ("srm_ctx_set_pool_cap", C.c_int, [_vp, C.c_size_t]),
    ("srm_debug_fail_allocs", None, [C.c_int]),
    ("srm_ctx_set_genome", C.c_int, [_vp, _vp, C.c_size_t]),
    ("srm_job_create", C.c_int, [_vp, C.POINTER(_vp)]),
    ("srm_job_destroy", None, [_vp]),
    ("srm_job_release", C.c_int, [_vp]),
    ("srm_job_bytes_bound", C.c_size_t, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_size_t,
                                         C.c_size_t, C.c_size_t, C.c_size_t, C.c_int, C.c_int,
                                         C.c_size_t, C.c_int]),
    ("srm_job_upload", C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_size_t, _vp,
                                 C.c_size_t, _vp, C.c_size_t, _vp, C.c_size_t]),
    ("srm_job_run", C.c_int, [_vp, C.c_int]),
    ("srm_job_download", C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    ("srm_job_last_ms", C.c_float, [_vp]),
    ("srm_job_kernel_ms", C.c_float, [_vp, C.c_int]),
    ("srm_job_use_graph", C.c_int, [_vp, C.c_int]),
    ("srm_job_path_count", C.c_int, [_vp, C.c_int]),
    ("srm_job_records", C.c_longlong, [_vp, _vp, _vp, _vp, C.c_int, C.c_size_t, C.c_char_p, _vp,
                                       C.c_size_t, _vp]),
    ("srm_align_static", C.c_int, [C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_int, _vp, C.c_size_t,
                                   _vp, C.c_size_t, _vp, _vp, C.c_size_t]),
    ("srm_job_set_records", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_size_t, C.c_char_p]),
    ("srm_job_records_text", C.c_longlong, [_vp, _vp, C.c_size_t, _vp]),
    ("srm_format_record", C.c_int, [C.c_char_p, C.c_uint, C.c_int, C.c_int, C.c_int, C.c_int,
                                    _vp, _vp, _vp, _vp, C.c_char_p, C.c_int]),
    ("srm_slot_rows", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
    ("srm_pack_codes", C.c_longlong, [_vp, C.c_size_t, _vp, C.c_size_t]),
    ("srm_ctx_borrow_genome", C.c_int, [_vp, _vp]),
    ("srm_ctx_set_tuning", C.c_int, [_vp, _vp]),
    ("srm_job_band_info", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_float)]),
    ("srm_band_occupancy", C.c_int, [C.c_int, C.c_int]),
    ("srm_calibrate_copy", C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_float)]),
    ("srm_format_records", C.c_longlong, [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          C.c_char_p, _vp, C.c_size_t, C.POINTER(C.c_longlong),
                                          C.c_int]),
    # host candidate generation and synthetic workloads
    ("srm_filter_params", C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    ("srm_index_build", C.c_int, [_vp, C.c_size_t, _vp, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("srm_index_build_at", C.c_int, [_vp, C.c_size_t, _vp, C.c_int, C.c_int, C.c_uint32,
                                     C.POINTER(_vp)]),
    ("srm_index_destroy", None, [_vp]),
    ("srm_filter_run_chunk", C.c_int, [_vp, _vp, C.c_size_t, C.c_uint64, _vp, _vp, _vp, _vp,
                                       C.c_size_t, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("srm_filter_run", C.c_int, [_vp, _vp, C.c_size_t, _vp, _vp, _vp, _vp, _vp, C.c_size_t,
                                 C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
    ("srm_filter_sizes", C.c_int, [_vp, _vp, _vp, _vp]),
    ("srm_filter_copy", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("srm_filter_destroy", None, [_vp]),
    ("srm_dfilter_create", C.c_int, [_vp, _vp, _vp, C.c_size_t, C.POINTER(_vp)]),
    ("srm_dfilter_run", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, C.c_int, C.c_int,
                                  C.POINTER(_vp)]),
    ("srm_dfilter_run2", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, C.c_int,
                                   C.POINTER(_vp), C.POINTER(_vp)]),
    ("srm_dfilter_stats", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("srm_host_alloc", _vp, [C.c_size_t]),
    ("srm_host_free", None, [_vp]),
    ("srm_dfilter_destroy", None, [_vp]),
    ("srm_synth_genome", C.c_int, [C.c_uint64, C.c_size_t, C.c_double, _vp]),
    ("srm_synth_repeats", C.c_longlong, [C.c_uint64, _vp, C.c_size_t, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_double, C.c_double]),
    ("srm_synth_reads", C.c_int, [C.c_uint64, _vp, _vp, _vp, C.c_int, C.c_size_t, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_double, _vp, _vp, _vp, _vp, _vp]),
    ("srm_fastx_load", C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.POINTER(_vp)]),
    ("srm_fastx_sizes", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("srm_fastx_copy", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("srm_fastx_destroy", None, [_vp]),
    ("srm_genome_chunks_load", C.c_int, [C.c_char_p, C.c_uint32, C.c_int, C.POINTER(_vp)]),
    ("srm_genome_chunks_sizes", C.c_int, [_vp, _vp, _vp]),
    ("srm_genome_chunks_copy", C.c_int, [_vp, _vp, _vp, _vp]),
    ("srm_genome_chunks_destroy", None, [_vp]),
    # serving
    ("srm_server_create", C.c_int, [_vp, _vp, C.c_size_t, _vp, C.c_int, C.POINTER(_vp)]),
    ("srm_server_submit", C.c_longlong, [_vp, _vp, _vp, _vp, _vp, C.c_size_t]),
    ("srm_server_wait", C.c_int, [_vp, C.c_longlong, C.c_int]),
    ("srm_server_result", C.c_longlong, [_vp, C.c_longlong, _vp, C.c_size_t, _vp, C.c_int]),
    ("srm_server_hold", None, [_vp, C.c_int]),
    ("srm_server_queue", None, [_vp, _vp, _vp, _vp]),
    ("srm_server_memory", None, [_vp, _vp, _vp]),
    ("srm_server_destroy", None, [_vp]),
    # drop-in entry points (src/cuda-functions.h:55-78)
    ("cuda_needleman_wunsch", None, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp,
                                     _vp, _vp, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t,
                                     C.c_int]),
    ("cuda_allocate_memory", None, [C.c_int]),
    ("cuda_free_memory", None, [C.c_int]),
    ("get_cuda_infos", C.c_ulong, []),
]

# every function declared in include/saruman_hip.h (checked by tests/test_abi.py)
EXPORTED = [p[0] for p in _PROTOS]
GLOBALS = ["read_names", "read_length_value", "be_verbose_value", "use_gpu_value",
           "cuda_memory_per_thread", "forward_output_control", "reverse_output_control",
           "forward_alignments_done", "reverse_alignments_done", "forward_alignment_candidates",
           "reverse_alignment_candidates", "main_alloc_ptr_for", "main_alloc_ptr_rev",
           "alignment_time"]


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python __graft_entry__.py)")
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in _PROTOS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != SRM_OK:
        msg = load().srm_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def ptr(a) -> int | None:
    if a is None:
        return None
    return a.ctypes.data
