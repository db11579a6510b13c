"""acctricnt_amd: all-edge common-neighbor counting on MI355X (gfx950).

For every slot (u, v) of a sorted, symmetric int32 CSR the library computes
|N(u) ∩ N(v)| with hand-written HIP kernels, the quantity beneath triangle
counting and the SCAN structural-clustering core check of
RapidsAtHKUST/AccTriCnt.  Native code:

* ``libacc_cn.so``    the HIP library behind the C ABI in include/acc_cn.h;
* ``libacc_graph.so`` host CSR utilities (include/acc_graph.h).

Python modules: :mod:`.graph` (CSR, generators, on-disk format) and
:mod:`.cn` (contexts, resident graphs, counting).
"""
from . import graph  # noqa: F401

__all__ = ["graph", "cn"]


def __getattr__(name):
    if name == "cn":
        from . import cn as _cn

        return _cn
    raise AttributeError(name)
