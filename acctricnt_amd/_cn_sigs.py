"""ctypes signatures of every function include/acc_cn.h declares."""
from __future__ import annotations

import ctypes as C

i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p



class ServerConfig(C.Structure):
    """acc_server_config"""
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_int32), ("streams", C.c_int32),
                ("slots", C.c_int32), ("no_index", C.c_int32)]


class ServerStats(C.Structure):
    """acc_server_stats"""
    _fields_ = [(name, C.c_uint64) for name in (
        "submitted", "completed", "failed", "cancelled", "batches", "batched_queries", "max_batch_seen",
        "admissions", "evictions", "graph_bytes", "slot_bytes", "resident_graphs", "pending", "index_queries")]


# name -> (restype, argtypes); kept in step with include/acc_cn.h
SIGNATURES = {
    "acc_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "acc_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "acc_ctx_destroy": (None, [vp]),
    "acc_ctx_synchronize": (C.c_int, [vp]),
    "acc_ctx_stream": (vp, [vp]),
    "acc_last_error": (C.c_char_p, []),
    "acc_graph_upload": (C.c_int, [vp, C.c_uint32, C.c_uint64, u32p, i32p, C.c_uint32, C.POINTER(vp)]),
    "acc_graph_release": (None, [vp]),
    "acc_graph_info": (C.c_int, [vp, u32p, u64p, u64p]),
    "acc_graph_download": (C.c_int, [vp, u32p, i32p, i32p]),
    "acc_graph_set_shard": (C.c_int, [vp, C.c_uint32, C.c_uint32, u32p]),
    "acc_cn_compute": (C.c_int, [vp, C.c_int, C.c_uint32]),
    "acc_cn_download": (C.c_int, [vp, i32p, u64p]),
    "acc_cn_all_edges": (C.c_int, [vp, C.c_int, C.c_uint32, i32p, u64p]),
    "acc_cn_device_counts": (vp, [vp]),
    "acc_cn_timings": (C.c_int, [vp, f32p, C.c_int, C.c_char_p, C.c_int]),
    "acc_triangle_total": (C.c_int, [vp, C.c_int, u64p]),
    "acc_triangle_launch": (C.c_int, [vp]),
    "acc_triangle_total_device": (C.c_int, [vp, C.c_uint32, C.c_uint64, vp, vp, vp]),
    "acc_graph_set_param": (C.c_int, [vp, C.c_int, C.c_int64]),
    "acc_graph_get_param": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int64)]),
    "acc_graph_set_partition": (C.c_int, [vp, C.c_uint32, u32p]),
    "acc_cn_graph_totals": (C.c_int, [vp, u64p]),
    "acc_cn_capture": (C.c_int, [vp, C.c_int, C.c_uint32]),
    "acc_cn_replay": (C.c_int, [vp]),
    "acc_tune": (C.c_int, [vp, C.c_uint32, C.c_char_p, C.c_int]),
    "acc_graph_set_vertex_ids": (C.c_int, [vp, i32p]),
    "acc_scan_ws_create": (C.c_int, [vp, C.POINTER(vp)]),
    "acc_scan_ws_destroy": (None, [vp]),
    "acc_scan_ws_info": (C.c_int, [vp, u64p, C.POINTER(vp)]),
    "acc_scan_run": (C.c_int, [vp, C.c_int64, C.c_int64, C.c_int32, C.c_uint32]),
    "acc_scan_synchronize": (C.c_int, [vp]),
    "acc_scan_results": (C.c_int, [vp, i32p, i32p, C.POINTER(C.c_uint8), i32p, C.POINTER(C.c_int64)]),
    "acc_scan_pairs": (C.c_int, [vp, i32p, C.c_uint64, u64p]),
    "acc_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "acc_server_create": (C.c_int, [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
    "acc_server_destroy": (None, [vp]),
    "acc_server_add_graph": (C.c_int, [vp, C.c_uint32, C.c_uint64, u32p, i32p, C.c_uint32, i32p]),
    "acc_server_remove_graph": (C.c_int, [vp, C.c_int32]),
    "acc_server_graph_info": (C.c_int, [vp, C.c_int32, u32p, u64p, C.POINTER(C.c_int), u64p]),
    "acc_server_submit": (C.c_int, [vp, C.c_int32, C.c_int64, C.c_int64, C.c_int32, u64p]),
    "acc_server_wait": (C.c_int, [vp, C.c_uint64, C.c_int32]),
    "acc_server_poll": (C.c_int, [vp, C.c_uint64, C.POINTER(C.c_int)]),
    "acc_server_result": (C.c_int, [vp, C.c_uint64, C.POINTER(C.POINTER(C.c_int8)), C.POINTER(i32p), u32p,
                                    C.POINTER(C.c_int64),
                                    C.POINTER(C.c_double)]),
    "acc_server_release": (C.c_int, [vp, C.c_uint64]),
    "acc_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
}


def declare(lib) -> None:
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
