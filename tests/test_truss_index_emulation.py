"""CPU emulation of the truss decomposition's index maintenance: the
compaction kernels (k_keep_tile_sums / k_keep_scatter / k_keep_finish) and
the mirrored support pass (lower-half supports, then k_sup_mirror copying
each to its transposed entry), replayed in numpy over full decompositions
and compared with the oracle's truss numbers (src/main.cpp:89-146 over all
k).

The emulation keeps the kernels' conventions: newpos[j] = new position of a
survivor, -1 - (survivors before j) for a removed entry; new row pointers =
the survivor count before each old row start; mirror indices carried through
the old -> new map, negative ones skipped by the mirror copy.  A planted
fault -- supports left unmirrored in the upper half -- makes keep decisions
asymmetric, and the invariant check catches it.
"""
import numpy as np
import pytest
import scipy.sparse as sp


def _lower_supports(n, rp, col):
    """S(i,j) = (C*C)(i,j) for the stored entries with j < i, zero above."""
    C = sp.csr_matrix((np.ones(len(col), dtype=np.int64), col, rp), shape=(n, n))
    P = (C @ C).tocsr()
    P.sort_indices()
    rows = np.repeat(np.arange(n), np.diff(rp))
    prow = np.repeat(np.arange(n), np.diff(P.indptr))
    pkey = prow.astype(np.int64) * n + P.indices
    key = rows.astype(np.int64) * n + col
    at = np.minimum(np.searchsorted(pkey, key), max(len(pkey) - 1, 0))
    S = np.where((len(pkey) > 0) & (pkey[at] == key), P.data[at], 0).astype(np.int32)
    S[col > rows] = 0
    return S


def _mirror_copy(S, mirror):
    """k_sup_mirror: an upper entry p (its mirror m > p) takes S[m]."""
    p = np.arange(len(S))
    up = mirror > p
    S = S.copy()
    S[up] = S[mirror[up]]
    return S


def _compact(n, rp, col, eid, mirror, sup, thr):
    keep = sup >= thr
    rank = np.cumsum(keep) - keep  # survivors before j
    newpos = np.where(keep, rank, -1 - rank).astype(np.int32)
    total = int(keep.sum())
    nold = np.nonzero(keep)[0]
    before = lambda j: np.where(j >= len(col), total, np.where(newpos[np.minimum(j, len(col) - 1)] >= 0,  # noqa: E731
                                                               newpos[np.minimum(j, len(col) - 1)],
                                                               -1 - newpos[np.minimum(j, len(col) - 1)]))
    nrp = np.empty(n + 1, dtype=np.int32)
    nrp[:n] = before(rp[:n]) if len(col) else 0
    nrp[n] = total
    mo = mirror[nold]
    mirror_new = np.where(mo >= 0, newpos[np.maximum(mo, 0)], -1).astype(np.int32)
    return nrp, col[keep], eid[keep], mirror_new, total


def _check_mirrors(n, rp, col, mirror):
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert (mirror >= 0).all(), "a survivor's transposed entry was removed"
    assert np.array_equal(mirror[mirror], np.arange(len(col)))
    assert np.array_equal(rows[mirror], col) and np.array_equal(col[mirror], rows)


def emulate_truss(g, mirror_writes=True):
    n = g.n
    rp = np.asarray(g.rowptr, dtype=np.int32)
    col = np.asarray(g.col, dtype=np.int32)
    rows = np.repeat(np.arange(n), np.diff(rp))
    key = rows.astype(np.int64) * n + col
    mirror = np.searchsorted(key, col.astype(np.int64) * n + rows).astype(np.int32)
    eid = np.arange(len(col), dtype=np.int32)
    truss = np.full(len(col), 2, dtype=np.int32)
    k = 3
    while len(col):
        while True:
            S = _lower_supports(n, rp, col)
            if mirror_writes:
                S = _mirror_copy(S, mirror)
            nrp, ncol, neid, nmirror, total = _compact(n, rp, col, eid, mirror, S, k - 2)
            if total == len(col):
                break
            rp, col, eid, mirror = nrp, ncol, neid, nmirror
            _check_mirrors(n, rp, col, mirror)
        truss[eid] = k
        k += 1
    return truss, k - 1 if (truss > 2).any() else 2


@pytest.mark.parametrize("which", ["rmat9", "rmat10", "G43"])
def test_index_maintenance_matches_oracle(G, oracle, small_graphs, which):
    g = {"rmat9": lambda: G.rmat(9, 8, seed=3), "rmat10": lambda: G.rmat(10, 16, seed=5),
         "G43": lambda: small_graphs["G43.mtx"][1]}[which]()
    truss, _ = emulate_truss(g)
    want, kmax, *_ = oracle.truss_decomposition(g)
    assert np.array_equal(truss, want)
    assert int(truss.max()) == kmax


def test_missing_mirror_writes_are_caught(G):
    """The failure mode of the short-row kernel without mirror writes: upper
    supports stay zero, an entry and its transpose get different keep
    decisions, and a survivor's mirror index turns negative."""
    with pytest.raises(AssertionError):
        emulate_truss(G.rmat(9, 8, seed=3), mirror_writes=False)
