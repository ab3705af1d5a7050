"""The leaf count the approx_n_clusters search reads per probe, pinned on the oracle.

For min_cluster_size m >= 2 the condensed tree's leaves (extract_leaves,
src/clustering/condensed_tree.rs:264-294) number one more than the linkage rows whose
smaller child holds at least m points: those rows are exactly the splits, every split
opens two clusters, and its parent cluster stops being a leaf. The library's search
(src/lib.rs:377-447) uses this count instead of condensing at every probe; here it is
checked against the oracle's condense + extract_leaves on trees with ties and on every
probe size."""
import numpy as np

import oracle as O


def test_leaf_count_is_one_plus_splits():
    rng = np.random.default_rng(0)
    for trial in range(40):
        n = int(rng.integers(2, 600))
        e = rng.normal(size=(n, int(rng.integers(1, 4)))).astype(np.float32)
        if trial % 2:
            e = (np.round(e * 2) / 2).astype(np.float32)
        u, v, w = O.mst(e, int(rng.integers(1, 6)))
        L, R, D, S = O.linkage(u, v, w, n)
        sz = np.concatenate([np.ones(n, np.int64), S])
        small = np.minimum(sz[L], sz[R])
        for m in range(2, max(n // 2, 2) + 2):
            P, Ch, lam, CS = O.condense(L, R, D, S, n, m)
            assert len(O.extract_leaves(P, Ch, CS, n)) == 1 + int((small >= m).sum()), (n, m)
