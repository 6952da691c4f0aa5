"""ctypes mirror of include/pmdd_engine.h and include/pmdd_host.h.

Loads the in-tree shared libraries built by ``make -C pmdd_amd/csrc``.  A
missing library is an error: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libpmdd_host.so")
ENGINE_LIB = os.path.join(LIB_DIR, "libpmdd_engine.so")
CLI_BIN = os.path.join(_PKG, "bin", "pmdd")

PMDD_OK = 0
PMDD_E_INVALID = -1
PMDD_E_NODEVICE = -2
PMDD_E_HIP = -3
PMDD_E_NOMEM = -4
PMDD_E_OVERFLOW = -5
PMDD_E_UNSUPPORTED = -6
PMDD_E_REJECTED = -7

PMDD_DD_BDD = 0
PMDD_DD_MDD = 1

DEFAULT_MAX_CAND = 20_000_000
NOALLOC_SCRATCH_BYTES = 128 << 20
DEFAULT_MAX_PROD = 625_000
MAX_PROD_ADAPTIVE = 0

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class DDView(C.Structure):
    _fields_ = [
        ("d", C.c_int32),
        ("num_layers", C.c_int32),
        ("kind", C.c_int32),
        ("root", C.c_int32),
        ("terminal", C.c_int32),
        ("reserved", C.c_int32),
        ("layer_off", _i64p),
        ("in_off", _i64p),
        ("arc_src", _i32p),
        ("arc_val", _i32p),
        ("arc_w", _i32p),
        ("min_weight", _i32p),
        ("sd_partner", _i32p),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("problem_type", C.c_int32),
        ("state_dominance", C.c_int32),
        ("count_pair_tests", C.c_int32),
        ("max_cand", C.c_int64),
        ("max_prod", C.c_int64),
        ("phase_timing", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("filter_threads", C.c_int32 * 3),
        ("cross_workgroups", C.c_int32),
        ("cross_split_max", C.c_int32),
        ("refill_waves", C.c_int32),
        ("source", C.c_int32),
    ]


class Frontier(C.Structure):
    _fields_ = [
        ("d", C.c_int32),
        ("reserved", C.c_int32),
        ("count", C.c_int64),
        ("points", _i32p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("wall_pack_transfer_s", C.c_double),
        ("wall_join_s", C.c_double),
        ("kernel_expand_td_s", C.c_double),
        ("kernel_dominance_s", C.c_double),
        ("kernel_total_s", C.c_double),
        ("gpu_mem_peak_used_bytes", C.c_int64),
        ("gpu_mem_peak_reserved_bytes", C.c_int64),
        ("work_candidates_total", C.c_int64),
        ("work_candidates_peak", C.c_int64),
        ("work_frontier_survivors_total", C.c_int64),
        ("work_frontier_peak_points", C.c_int64),
        ("work_join_products_total", C.c_int64),
        ("dominance_filtered_total", C.c_int64),
        ("dominance_pair_tests", C.c_int64),
        ("kernel_filter_s", C.c_double),
        ("filter_launches", C.c_int64),
        ("filter_candidates", C.c_int64),
        ("algorithmic_bytes", C.c_int64),
        ("kernel_frontier_check_s", C.c_double),
        ("frontier_check_launches", C.c_int64),
        ("frontier_check_bytes", C.c_int64),
        ("kernel_cross_s", C.c_double),
        ("cross_launches", C.c_int64),
        ("cross_bytes", C.c_int64),
        ("pair_tests_filter", C.c_int64),
        ("pair_tests_frontier_check", C.c_int64),
        ("pair_tests_cross", C.c_int64),
        ("pair_tests_state_dominance", C.c_int64),
        ("kernel_state_dominance_s", C.c_double),
        ("state_dominance_launches", C.c_int64),
        ("join_bytes", C.c_int64),
        ("graph_launches", C.c_int64),
        ("join_batches", C.c_int64),
        ("join_batch_peak_products", C.c_int64),
        ("ctx_mem_peak_bytes", C.c_int64),
        ("mem_need_peak_bytes", C.c_int64),
        ("mem_estimate_bytes", C.c_int64),
        ("nomem_retries", C.c_int32),
        ("reserved2", C.c_int32),
        ("queue_wait_s", C.c_double),
        ("layer_coupling", C.c_int32),
        ("std_capacity", C.c_int32),
        ("std_candidates_per_layer", C.POINTER(C.c_double)),
        ("std_frontier_survivors_per_layer", C.POINTER(C.c_double)),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("slots", C.c_int32),
        ("memory_budget", C.c_int64),
        ("pack", C.c_int32),
        ("pack_inflight", C.c_int32),
    ]


class ServerInfo(C.Structure):
    _fields_ = [
        ("budget", C.c_int64),
        ("reserved_peak", C.c_int64),
        ("admitted", C.c_int64),
        ("rejected", C.c_int64),
        ("memory_waits", C.c_int64),
        ("trims", C.c_int64),
        ("nomem_retries", C.c_int64),
        ("packed", C.c_int64),
        ("pack_steps", C.c_int64),
        ("pack_requests_per_step_x100", C.c_int64),
        ("running_peak", C.c_int32),
        ("slots", C.c_int32),
    ]


PMDD_PENDING = 1
PMDD_BUILD_STATE_DOMINANCE = 1


class DDStructure(C.Structure):
    _fields_ = [
        ("original_width", C.c_int64),
        ("reduced_width", C.c_int64),
        ("original_num_nodes", C.c_int64),
        ("reduced_num_nodes", C.c_int64),
        ("max_num_in_arcs", C.c_int64),
        ("total_num_in_arcs", C.c_int64),
        ("num_layers", C.c_int32),
        ("d", C.c_int32),
    ]


# Every symbol each header declares, checked by the CPU test suite.
ENGINE_SYMBOLS = [
    "pmdd_ctx_create",
    "pmdd_ctx_destroy",
    "pmdd_device_count",
    "pmdd_ctx_get_tuning",
    "pmdd_ctx_set_tuning",
    "pmdd_enumerate",
    "pmdd_topdown_enumerate",
    "pmdd_topdown_mdd_enumerate",
    "pmdd_coupled_mdd_enumerate",
    "pmdd_coupled_bdd_enumerate",
    "pmdd_upload",
    "pmdd_enumerate_resident",
    "pmdd_resident_free",
    "pmdd_enumerate_resident_noalloc",
    "pmdd_frontier_free",
    "pmdd_engine_version",
    "pmdd_server_create",
    "pmdd_server_submit",
    "pmdd_server_poll",
    "pmdd_server_wait",
    "pmdd_server_wait_any",
    "pmdd_server_destroy",
    "pmdd_server_get_info",
    "pmdd_debug_fail_alloc",
]
HOST_SYMBOLS = [
    "pmdd_dd_build_from_file",
    "pmdd_dd_get_view",
    "pmdd_dd_get_structure",
    "pmdd_dd_layer_size",
    "pmdd_dd_free",
    "pmdd_frontier_sort",
]

_host = None
_engine = None


def _require(path: str) -> None:
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `make -C pmdd_amd/csrc` (or __graft_entry__.build())"
        )


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        _require(HOST_LIB)
        lib = C.CDLL(HOST_LIB)
        lib.pmdd_dd_build_from_file.argtypes = [
            C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
        lib.pmdd_dd_build_from_file.restype = C.c_int
        lib.pmdd_dd_get_view.argtypes = [C.c_void_p]
        lib.pmdd_dd_get_view.restype = C.POINTER(DDView)
        lib.pmdd_dd_get_structure.argtypes = [C.c_void_p, C.POINTER(DDStructure)]
        lib.pmdd_dd_get_structure.restype = None
        lib.pmdd_dd_layer_size.argtypes = [C.c_void_p, C.c_int]
        lib.pmdd_dd_layer_size.restype = C.c_int64
        lib.pmdd_dd_free.argtypes = [C.c_void_p]
        lib.pmdd_dd_free.restype = None
        lib.pmdd_frontier_sort.argtypes = [C.POINTER(Frontier)]
        lib.pmdd_frontier_sort.restype = C.c_int
        _host = lib
    return _host


def engine_lib() -> C.CDLL:
    global _engine
    if _engine is None:
        _require(ENGINE_LIB)
        lib = C.CDLL(ENGINE_LIB)
        err = [C.c_char_p, C.c_size_t]
        lib.pmdd_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)] + err
        lib.pmdd_ctx_create.restype = C.c_int
        lib.pmdd_ctx_destroy.argtypes = [C.c_void_p]
        lib.pmdd_ctx_destroy.restype = None
        lib.pmdd_device_count.argtypes = []
        lib.pmdd_device_count.restype = C.c_int
        lib.pmdd_ctx_get_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(Tuning)]
        lib.pmdd_ctx_get_tuning.restype = C.c_int
        lib.pmdd_ctx_set_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(Tuning)]
        lib.pmdd_ctx_set_tuning.restype = C.c_int
        lib.pmdd_enumerate.argtypes = [
            C.c_void_p, C.POINTER(DDView), C.POINTER(Opts), C.POINTER(Frontier),
            C.POINTER(Stats)] + err
        lib.pmdd_enumerate.restype = C.c_int
        lib.pmdd_topdown_enumerate.argtypes = [
            C.c_void_p, C.POINTER(DDView), C.c_int, C.c_int, C.c_int64, C.POINTER(Frontier),
            C.POINTER(Stats)] + err
        lib.pmdd_topdown_enumerate.restype = C.c_int
        lib.pmdd_topdown_mdd_enumerate.argtypes = [
            C.c_void_p, C.POINTER(DDView), C.c_int64, C.POINTER(Frontier),
            C.POINTER(Stats)] + err
        lib.pmdd_topdown_mdd_enumerate.restype = C.c_int
        lib.pmdd_coupled_mdd_enumerate.argtypes = [
            C.c_void_p, C.POINTER(DDView), C.c_int64, C.c_int64, C.POINTER(Frontier),
            C.POINTER(Stats)] + err
        lib.pmdd_coupled_mdd_enumerate.restype = C.c_int
        lib.pmdd_coupled_bdd_enumerate.argtypes = [
            C.c_void_p, C.POINTER(DDView), C.c_int, C.c_int, C.c_int64, C.c_int64,
            C.POINTER(Frontier), C.POINTER(Stats)] + err
        lib.pmdd_coupled_bdd_enumerate.restype = C.c_int
        lib.pmdd_upload.argtypes = [C.c_void_p, C.POINTER(DDView), C.POINTER(C.c_void_p)] + err
        lib.pmdd_upload.restype = C.c_int
        lib.pmdd_enumerate_resident.argtypes = [
            C.c_void_p, C.c_void_p, C.POINTER(Opts), C.POINTER(Frontier),
            C.POINTER(Stats)] + err
        lib.pmdd_enumerate_resident.restype = C.c_int
        lib.pmdd_resident_free.argtypes = [C.c_void_p]
        lib.pmdd_resident_free.restype = None
        lib.pmdd_enumerate_resident_noalloc.argtypes = [
            C.c_void_p, C.c_void_p, C.POINTER(Opts), _i32p, C.c_int64, C.POINTER(C.c_int64),
            C.POINTER(Stats)] + err
        lib.pmdd_enumerate_resident_noalloc.restype = C.c_int
        lib.pmdd_frontier_free.argtypes = [C.POINTER(Frontier)]
        lib.pmdd_frontier_free.restype = None
        lib.pmdd_engine_version.argtypes = []
        lib.pmdd_engine_version.restype = C.c_char_p
        lib.pmdd_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)] + err
        lib.pmdd_server_create.restype = C.c_int
        lib.pmdd_server_submit.argtypes = [
            C.c_void_p, C.POINTER(DDView), C.POINTER(Opts), C.POINTER(C.c_int64)] + err
        lib.pmdd_server_submit.restype = C.c_int
        for fn in (lib.pmdd_server_poll, lib.pmdd_server_wait):
            fn.argtypes = [C.c_void_p, C.c_int64, C.POINTER(Frontier), C.POINTER(Stats)] + err
            fn.restype = C.c_int
        lib.pmdd_server_wait_any.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(Frontier),
                                             C.POINTER(Stats)] + err
        lib.pmdd_server_wait_any.restype = C.c_int
        lib.pmdd_server_destroy.argtypes = [C.c_void_p]
        lib.pmdd_server_destroy.restype = None
        lib.pmdd_server_get_info.argtypes = [C.c_void_p, C.POINTER(ServerInfo)]
        lib.pmdd_server_get_info.restype = C.c_int
        lib.pmdd_debug_fail_alloc.argtypes = [C.c_int64]
        lib.pmdd_debug_fail_alloc.restype = None
        _engine = lib
    return _engine


class PmddError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message
