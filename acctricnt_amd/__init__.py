"""acctricnt_amd: all-edge common-neighbor counting on MI355X (gfx950).

For every slot (u, v) of a sorted, symmetric int32 CSR the library computes
|N(u) ∩ N(v)| with hand-written HIP kernels, the quantity beneath triangle
counting and the SCAN structural-clustering core check of
RapidsAtHKUST/AccTriCnt.  Native code:

* ``libacc_cn.so``    the HIP library behind the C ABI in include/acc_cn.h;
* ``libacc_graph.so`` host CSR utilities (include/acc_graph.h).

Python modules: :mod:`.graph` (CSR, generators, on-disk format),
:mod:`.cn` (contexts, resident graphs, counting, batches, SCAN) and
:mod:`.serve` (SCAN serving over resident graphs).  Importing the package
loads no native library; the first call that needs one loads it and fails
loudly if it is not built.
"""
from . import cn, graph, serve  # noqa: F401

__all__ = ["graph", "cn", "serve"]
