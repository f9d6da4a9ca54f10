"""graphalgo_amd -- masked SpGEMM C<M> = A*B over CSR on MI355X (gfx950), with
triangle counting, k-truss / truss decomposition and batched betweenness
centrality built on it.  The reference is UAlex322/GraphAlgo (C++/OpenMP);
its C++ API is mirrored in include/graphalgo/, this package mirrors it in
Python over the same C ABI (include/graphalgo.h).

`graphalgo_amd.api` is the GPU interface (it loads libgraphalgo.so on first
use); `graphalgo_amd.graphs` is the host graph substrate."""
from . import graphs
from .graphs import CSR

__all__ = ["graphs", "CSR"]
