"""ctypes bindings of the package's two shared libraries.

``libacc_graph.so`` (host CSR utilities, include/acc_graph.h) and
``libacc_cn.so`` (the HIP all-edge common-neighbor library, include/acc_cn.h)
are built in-tree by ``acctricnt_amd/csrc/Makefile``.  Loading fails loudly
when a library is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
GRAPH_LIB_PATH = os.path.join(PKG_DIR, "libacc_graph.so")
CN_LIB_PATH = os.path.join(PKG_DIR, "libacc_cn.so")

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p


class AccError(RuntimeError):
    """Raised when a library call returns a negative ACC_E* code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_graph_lib = None
_cn_lib = None


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def graph_lib():
    """Load libacc_graph.so (host-only utilities)."""
    global _graph_lib
    if _graph_lib is not None:
        return _graph_lib
    if not os.path.exists(GRAPH_LIB_PATH):
        raise ImportError(f"{GRAPH_LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(GRAPH_LIB_PATH)
    _sig(lib, "acc_graph_last_error", C.c_char_p)
    _sig(lib, "acc_hgraph_n", C.c_uint32, vp)
    _sig(lib, "acc_hgraph_m", C.c_uint64, vp)
    _sig(lib, "acc_hgraph_offsets", u32p, vp)
    _sig(lib, "acc_hgraph_adj", i32p, vp)
    _sig(lib, "acc_hgraph_free", None, vp)
    _sig(lib, "acc_hgraph_from_edges", C.c_int, C.c_uint32, C.c_uint64, i32p, i32p, C.POINTER(vp))
    _sig(lib, "acc_hgraph_from_csr", C.c_int, C.c_uint32, C.c_uint64, u32p, i32p, C.POINTER(vp))
    _sig(lib, "acc_csr_validate", C.c_int, C.c_uint32, C.c_uint64, u32p, i32p)
    _sig(lib, "acc_hgraph_relabel_degree", C.c_int, vp, i32p)
    _sig(lib, "acc_gen_er", C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, C.POINTER(vp))
    _sig(lib, "acc_gen_rmat", C.c_int, C.c_uint32, C.c_int, C.c_uint64, C.c_double, C.c_double,
         C.c_double, C.c_uint64, C.POINTER(vp))
    _sig(lib, "acc_gen_chung_lu", C.c_int, C.c_uint32, C.c_double, C.c_double, C.c_double,
         C.c_uint64, C.POINTER(vp))
    _sig(lib, "acc_hgraph_write_bins", C.c_int, vp, C.c_char_p)
    _sig(lib, "acc_hgraph_read_bins", C.c_int, C.c_char_p, C.POINTER(vp))
    _sig(lib, "acc_convert_edge_list", C.c_int, C.c_char_p, C.c_char_p)
    _sig(lib, "acc_hgraph_reverse_slots", C.c_int, vp, u32p)
    _sig(lib, "acc_write_scan_result", C.c_int, C.c_char_p, C.c_uint32, i32p, i32p, i32p, i32p, C.c_uint64)
    _sig(lib, "acc_scan_result_path", C.c_int, C.c_char_p, C.c_double, C.c_int, C.c_char_p, C.c_int)
    _sig(lib, "acc_csr_verify_counts", C.c_int, C.c_uint32, C.c_uint64, u32p, i32p, i32p, C.c_int32, C.c_int,
         C.POINTER(C.c_uint64), C.POINTER(C.c_uint64))
    _graph_lib = lib
    return lib


def check_graph(rc: int) -> None:
    if rc != 0:
        raise AccError(rc, graph_lib().acc_graph_last_error().decode())


def cn_lib():
    """Load libacc_cn.so, the HIP library.  Raises ImportError if it is not built."""
    global _cn_lib
    if _cn_lib is not None:
        return _cn_lib
    if not os.path.exists(CN_LIB_PATH):
        raise ImportError(f"{CN_LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(CN_LIB_PATH)
    from . import _cn_sigs

    _cn_sigs.declare(lib)
    _cn_lib = lib
    return lib


def check_cn(rc: int) -> None:
    if rc != 0:
        raise AccError(rc, cn_lib().acc_last_error().decode())


def ptr(a, ctype):
    """Pointer to the data of a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(ctype))
