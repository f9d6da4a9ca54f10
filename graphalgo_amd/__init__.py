"""graphalgo_amd -- masked SpGEMM C<M> = A*B over CSR on MI355X (gfx950), with
triangle counting, k-truss / truss decomposition and batched betweenness
centrality built on it.  The reference is UAlex322/GraphAlgo (C++/OpenMP);
its C++ API is mirrored in include/graphalgo/, this package mirrors it in
Python over the same C ABI (include/graphalgo.h)."""
from . import graphs
from .graphs import CSR

__all__ = ["graphs", "CSR", "api"]


def __getattr__(name):
    # the GPU API loads libgraphalgo.so lazily so graph utilities work on
    # machines without the HIP runtime
    if name == "api":
        from . import api
        return api
    raise AttributeError(name)
