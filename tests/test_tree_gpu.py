"""The condensed tree, layers and n-cluster search on the device against the oracle.

condense_tree (src/clustering/condensed_tree.rs:127-248) runs as pointer-jumping kernels
over the linkage; its rows (parent, child, lambda, child size, in the reference's walk
order) must equal the oracle's restatement bit for bit, on trees with tied and zero
distances, on a chain (a single path of depth n - 1, the deepest a linkage can be), and for
every kind of min_cluster_size: 0 and 1 (every point its own cluster), 2, small, around n/2,
n and above n (one cluster, all points fall out at the root). Layers and the search are
compared end to end, including min_cluster_size 1 and the tiny n where the search's upper
bracket n/2 falls below 2."""
import numpy as np
import pytest

import evoc_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def _linkage(n, dim, seed, ties=False):
    rng = np.random.default_rng(seed)
    ms = 1 if ties else 5  # with min_samples 1 a duplicate point is at distance 0
    e = rng.normal(size=(n, dim)).astype(np.float32)
    if ties:
        e = (np.round(e * 2) / 2).astype(np.float32)  # duplicates: zero and equal distances
    u, v, w = O.mst(e, ms)
    return O.linkage(u, v, w, n)


def _same(a, b):
    P, Ch, lam, CS = a
    oP, oCh, olam, oCS = b
    assert len(P) == len(oP)
    assert np.array_equal(P, oP)
    assert np.array_equal(Ch, oCh)
    assert np.array_equal(CS, oCS)
    assert np.array_equal(lam.view(np.uint32), olam.view(np.uint32))


@pytest.mark.parametrize("n,dim,seed,ties", [(2, 2, 1, False), (3, 2, 2, True), (7, 3, 3, False),
                                             (100, 2, 4, True), (1000, 3, 5, False),
                                             (1000, 2, 6, True), (20011, 4, 7, False)])
def test_condense_tree_bit_exact(n, dim, seed, ties):
    L, R, D, S = _linkage(n, dim, seed, ties)
    if ties and n >= 100:
        assert (D == 0).any()  # lambda = inf rows are exercised
    for mcs in sorted({0, 1, 2, 3, 5, 17, max(n // 2, 1), n, n + 5}):
        got = H.condense_tree(L, R, D, S, n, mcs)
        want = O.condense(L, R, D, S, n, mcs)
        _same(got, want)


def test_condense_tree_single_point():
    P, Ch, lam, CS = H.condense_tree(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), 1, 5)
    assert len(P) == 0


def test_condense_tree_chain():
    # points on a line with growing gaps: each merge adds one point to the growing cluster,
    # so the linkage is a path of depth n - 1
    n = 100_000
    u = np.arange(n - 1, dtype=np.int32)
    v = u + 1
    w = ((np.arange(n - 1) + 1) * 1e-3).astype(np.float32)
    L, R, D, S = O.linkage(u, v, w, n)
    depth = 0
    node = 2 * n - 2
    while node >= n:
        node = L[node - n]
        depth += 1
    assert depth >= n - 2
    for mcs in (2, 5, 1000, n // 2):
        _same(H.condense_tree(L, R, D, S, n, mcs), O.condense(L, R, D, S, n, mcs))


def test_condense_tree_large_random():
    L, R, D, S = _linkage(200_000, 2, 11)
    for mcs in (5, 60):
        _same(H.condense_tree(L, R, D, S, 200_000, mcs), O.condense(L, R, D, S, 200_000, mcs))


@pytest.mark.parametrize("base_mcs", [1, 2, 9])
def test_cluster_layers_every_min_cluster_size(base_mcs):
    rng = np.random.default_rng(20 + base_mcs)
    c = rng.normal(scale=6.0, size=(15, 4))
    e = (c[rng.integers(0, 15, 2500)] + rng.normal(size=(2500, 4))).astype(np.float32)
    gl, gs, gp = H.cluster_layers(e, 5, base_mcs, 0.2, 10)
    ol, os_, op = O.cluster_layers(e, 5, base_mcs, 0.2, 10, fp64=True)
    assert gl.shape == ol.shape
    assert np.array_equal(gl, ol)
    assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))
    assert np.array_equal(gp.view(np.uint64), op.view(np.uint64))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9])
def test_search_tiny_inputs(n):
    e = np.random.default_rng(n).normal(size=(n, 3)).astype(np.float32)
    for target in (1, 2, 4):
        gl, gs = H.search_for_n_clusters(e, 2, target)
        ol, os_ = O.search_n_clusters(e, 2, target)
        assert np.array_equal(gl, ol)
        assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))


def test_search_many_targets_bit_exact():
    rng = np.random.default_rng(31)
    c = rng.normal(scale=8.0, size=(120, 3))
    e = (c[rng.integers(0, 120, 30000)] + rng.normal(size=(30000, 3))).astype(np.float32)
    for target in (3, 50, 120, 400):
        gl, gs = H.search_for_n_clusters(e, 5, target)
        ol, os_ = O.search_n_clusters(e, 5, target)
        assert np.array_equal(gl, ol)
        assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))
