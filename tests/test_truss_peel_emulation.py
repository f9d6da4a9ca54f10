"""CPU emulation of peeling by support updates (k_peel_scan / k_peel_mark /
k_peel_destroy in algorithms.hip), replayed in Python over full truss
decompositions and single-k runs, against the oracle's truss numbers and
the reference's iteration counts (src/main.cpp:89-146).

It keeps the kernels' conventions: supports are computed once (the count
lives on the lower, canonical entry); every entry carries the round that
removes its edge (ALIVE while none does): round r removes the edges stamped
r, a smaller stamp is an edge removed earlier; a destroyed triangle is
handled by its edge of least canonical position among those stamped r; a
decrement that takes a support from thr to thr - 1 stamps the edge r + 1 and
appends it to the next round's list; a new k scans for alive canonical
entries below the threshold (canonical: the transposed entry lies before
it); dead entries are dropped when they are at least half of the stored
ones (only at a k boundary).  Planted mistakes (no ownership rule, or
decrements that ignore the earlier-removed state) must change the result.
"""
import numpy as np
import pytest


ALIVE = 1 << 40


class PeelEmu:
    def __init__(self, g, own_rule=True, skip_dead=True):
        self.n = g.n
        self.rp = np.asarray(g.rowptr, dtype=np.int64).copy()
        self.col = np.asarray(g.col, dtype=np.int64).copy()
        self.own_rule, self.skip_dead = own_rule, skip_dead
        rows = np.repeat(np.arange(self.n), np.diff(self.rp))
        key = rows * self.n + self.col
        self.mirror = np.searchsorted(key, self.col * self.n + rows)
        self.eid = np.arange(len(self.col))
        self.sup = self._supports()
        self.stamp = np.full(len(self.col), ALIVE, dtype=np.int64)

    def _supports(self):
        sets = [set(self.col[self.rp[i]:self.rp[i + 1]].tolist()) for i in range(self.n)]
        rows = np.repeat(np.arange(self.n), np.diff(self.rp))
        return np.array([len(sets[i] & sets[j]) for i, j in zip(rows, self.col)], dtype=np.int64)

    def row(self, i):
        return range(self.rp[i], self.rp[i + 1])

    def find(self, i, w):
        b, e = self.rp[i], self.rp[i + 1]
        p = b + np.searchsorted(self.col[b:e], w)
        return p if p < e and self.col[p] == w else -1

    def cid(self, p, r):  # canonical position of the entry at p in row r
        return p if self.col[p] < r else self.mirror[p]

    def scan(self, thr, r):
        out = []
        for p in range(len(self.col)):
            m = self.mirror[p]
            if m < p and self.sup[p] < thr and self.stamp[p] == ALIVE:
                self.stamp[p] = self.stamp[m] = r
                out.append(p)
        return out

    def destroy(self, cur, thr, r, truss, tval):
        nxt = []
        for c in cur:
            mc = self.mirror[c]
            j, i = self.col[c], self.col[mc]
            if truss is not None:
                truss[self.eid[c]] = truss[self.eid[mc]] = tval
            for q in self.row(i):
                w = self.col[q]
                sa = self.stamp[q]
                if (self.skip_dead and sa < r) or w == j:
                    continue
                lo = self.find(j, w)
                if lo < 0:
                    continue
                sb = self.stamp[lo]
                if self.skip_dead and sb < r:
                    continue
                ca, cb = self.cid(q, i), self.cid(lo, j)
                if self.own_rule and not ((sa != r or c < ca) and (sb != r or c < cb)):
                    continue
                for s_, cx, p in ((sa, ca, q), (sb, cb, lo)):
                    if s_ > r:
                        old = self.sup[cx]
                        self.sup[cx] -= 1
                        if old == thr:
                            self.stamp[p] = self.stamp[self.mirror[p]] = r + 1
                            nxt.append(cx)
        return nxt

    def compact(self):
        keep = self.stamp == ALIVE
        newpos = np.cumsum(keep) - 1
        rows = np.repeat(np.arange(self.n), np.diff(self.rp))
        self.rp = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=self.n))])
        self.mirror = newpos[self.mirror[keep]]
        self.col, self.eid, self.sup = self.col[keep], self.eid[keep], self.sup[keep]
        self.stamp = np.full(len(self.col), ALIVE, dtype=np.int64)

    def run(self, ks=None):
        """ks None: truss decomposition -> (truss, kmax, passes); else a k list
        -> [(entries, passes) per k]."""
        all_k = ks is None
        truss = np.full(len(self.col), 2, dtype=np.int64) if all_k else None
        alive = stored = len(self.col)
        passes, kidx, kmax, out = (0 if all_k and alive == 0 else 1), 0, 2, []
        k = 3 if all_k else ks[0]
        r = 0
        cur = self.scan(k - 2, r + 1) if alive else []
        while True:
            if not cur:
                if all_k and alive == 0:
                    break
                if all_k:
                    kmax, k = k, k + 1
                else:
                    out.append((alive, passes))
                    kidx += 1
                    if kidx == len(ks):
                        break
                    k = ks[kidx]
                if alive == 0:
                    continue
                if 2 * alive <= stored:
                    self.compact()
                    stored = alive
                cur = self.scan(k - 2, r + 1)
                continue
            r += 1
            nxt = self.destroy(cur, k - 2, r, truss, k - 1)
            alive -= 2 * len(cur)
            if not all_k or alive:  # a decomposition stops without a product once the graph is empty
                passes += 1
            cur = nxt
        return (truss, kmax, passes) if all_k else out


def _graphs(G, small_graphs):
    return {"rmat9": G.rmat(9, 8, seed=3), "rmat10": G.rmat(10, 16, seed=5), "G43": small_graphs["G43.mtx"][1],
            "iris": small_graphs["iris_dataset_30NN.mtx"][1]}


@pytest.mark.parametrize("which", ["rmat9", "rmat10", "G43", "iris"])
def test_peel_decomposition_matches_oracle(G, oracle, small_graphs, which):
    g = _graphs(G, small_graphs)[which]
    truss, kmax, passes = PeelEmu(g).run()
    want, okmax, opasses, *_ = oracle.truss_decomposition(g)
    assert np.array_equal(truss, want) and kmax == okmax
    assert passes == opasses  # 1 + the rounds that removed something = the support products


@pytest.mark.parametrize("which", ["rmat10", "G43"])
def test_peel_single_k_matches_reference_iterations(G, oracle, small_graphs, golden, which):
    g = _graphs(G, small_graphs)[which]
    for k in (3, 4, 5, 7):
        (entries, passes), = PeelEmu(g).run([k])
        rp, col, it = oracle.ktruss(g, k)
        assert (entries, passes) == (len(col), it), k
    if which == "G43":
        per_k = {k: (nnz, it) for k, nnz, it in golden["graphs"]["G43.mtx"]["truss"]["per_k"]}
        for k in (3, 4):
            assert tuple(PeelEmu(g).run([k])[0]) == per_k[k]


def test_peel_k_list_matches_single_runs(G, oracle):
    g = G.rmat(10, 16, seed=7)
    got = PeelEmu(g).run([3, 5, 6, 9])
    for (entries, _), k in zip(got, (3, 5, 6, 9)):
        assert entries == len(oracle.ktruss(g, k)[1])


def test_planted_mistakes_are_caught(G, oracle):
    g = G.rmat(10, 16, seed=5)
    want = oracle.truss_decomposition(g)[0]
    assert not np.array_equal(PeelEmu(g, own_rule=False).run()[0], want)
    assert not np.array_equal(PeelEmu(g, skip_dead=False).run()[0], want)
