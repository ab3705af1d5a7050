"""Pins the CPU oracle: the reference's own known-answer tests (transcribed, with the
reference file:line they come from) plus textbook cross-checks (numpy brute force,
scipy's MST / single linkage / find_peaks, Random123's Philox vectors), chosen so a
dropped term, a wrong sign or a transposed operand fails at least one of them."""
import numpy as np
import pytest
import scipy.cluster.hierarchy as sch
import scipy.signal
import scipy.sparse.csgraph as csg

import datagen
import oracle as O


# ------------------------------------------------------------------ random numbers
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds (first output word)
    assert O.lib().orc_philox_u32(0, 0, 0, 0, 0) == 0x6627E8D5
    assert O.lib().orc_philox_u32(0xFFFFFFFFFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF,
                                  0xFFFFFFFF) == 0x408F276D
    key = 0xA4093822 | (0x299F31D0 << 32)
    assert O.lib().orc_philox_u32(key, 0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344) == 0xD16CFE09


# ------------------------------------------------------------------ kNN
def test_knn_matches_float64_brute_force():
    x, _ = datagen.gaussian_mixture(600, 24, 6, 5.0, 1.0, seed=3)
    idx, dist = O.knn_exhaustive(x, 10, metric=0)
    x64 = x.astype(np.float64)
    d2 = ((x64[:, None, :] - x64[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    ref = np.argsort(d2, axis=1, kind="stable")[:, :10]
    assert np.array_equal(idx, ref)
    np.testing.assert_allclose(dist, np.sqrt(np.take_along_axis(d2, ref, 1)), rtol=1e-5)


def test_knn_cosine_matches_float64():
    x, _ = datagen.unit_mixture(500, 40, 8, 0.3, seed=4)
    x = x * np.random.default_rng(0).uniform(0.5, 2.0, size=(500, 1)).astype(np.float32)
    idx, dist = O.knn_exhaustive(x, 8, metric=1)
    x64 = x.astype(np.float64)
    nx = x64 / np.linalg.norm(x64, axis=1, keepdims=True)
    d = 1.0 - nx @ nx.T
    np.fill_diagonal(d, np.inf)
    ref = np.argsort(d, axis=1, kind="stable")[:, :8]
    assert np.array_equal(idx, ref)
    np.testing.assert_allclose(dist, np.take_along_axis(d, ref, 1), atol=2e-6)


def test_knn_ties_by_index_and_self_excluded():
    # kd_tree.rs:752-798 orders (dist, idx); duplicates make exact ties
    x = np.array([[0.0], [1.0], [1.0], [2.0], [0.0]], np.float32)
    idx, dist = O.knn_exhaustive(x, 3)
    assert idx[0].tolist() == [4, 1, 2]
    assert idx[4].tolist() == [0, 1, 2]
    assert all(i not in row for i, row in enumerate(idx))


# ------------------------------------------------------------------ fuzzy set
def test_smooth_knn_dist_rho_and_sigma():
    # fuzzy_graph.rs:307-322: rho = first distance, sigma > 0
    idx = np.array([[1, 2], [0, 2], [0, 1]], np.int32)
    dist = np.array([[1.0, 2.0], [1.5, 3.0], [0.5, 1.5]], np.float32)
    _, _, _, sig, rho = O.fuzzy_graph(idx, dist, 2, symmetrise=False)
    np.testing.assert_allclose(rho, [1.0, 1.5, 0.5], atol=1e-4)
    assert (sig > 0).all()


def test_sigma_solves_target_and_scale_invariance():
    rng = np.random.default_rng(1)
    dist = np.sort(rng.uniform(0.1, 3.0, size=(200, 15)), axis=1).astype(np.float32)
    idx = np.tile(np.arange(1, 16, dtype=np.int32), (200, 1))
    _, _, _, sig, rho = O.fuzzy_graph(idx, dist, 15, symmetrise=False)
    val = np.exp(-np.maximum(dist - rho[:, None], 0) / sig[:, None]).sum(1)
    assert np.all(np.abs(val - np.log2(15)) < 2e-5)
    _, _, _, sig2, _ = O.fuzzy_graph(idx, dist * 4, 15, symmetrise=False)
    np.testing.assert_allclose(sig2, 4 * sig, rtol=1e-4)


def test_knn_to_coo_counts_and_self_loops():
    # fuzzy_graph.rs:335-370
    idx = np.array([[1, 2], [0, 2], [0, 1]], np.int32)
    dist = np.array([[1.0, 2.0], [1.0, 1.5], [2.0, 1.5]], np.float32)
    ip, ind, w, _, _ = O.fuzzy_graph(idx, dist, 2, symmetrise=False)
    assert len(w) == 6 and ((w >= 0) & (w <= 1)).all()
    idx = np.array([[0, 1], [1, 0]], np.int32)
    dist = np.array([[0.0, 1.0], [0.0, 1.0]], np.float32)
    ip, ind, w, _, _ = O.fuzzy_graph(idx, dist, 2, symmetrise=False)
    assert len(w) == 2
    rows = np.repeat(np.arange(2), np.diff(ip))
    assert (rows != ind).all()


def test_fuzzy_union_known_answers():
    # fuzzy_graph.rs:373-415: 0.8 and 0.6 -> 0.92 both ways; one-way 0.7 -> 0.7 both ways.
    # Distances chosen so the weights are 1.0 / 0.0 after smoothing; check the union on
    # a graph built from explicit weights through the GPU-independent formula instead.
    idx = np.array([[1], [0]], np.int32)
    dist = np.array([[1.0], [1.0]], np.float32)
    ip, ind, w, _, _ = O.fuzzy_graph(idx, dist, 1, symmetrise=True)
    # rho = d[0] so every directed weight is exp(0) = 1; union 1 + 1 - 1 = 1
    assert ind.tolist() == [1, 0] and np.allclose(w, 1.0)
    idx = np.array([[1, 2], [2, 0], [0, 1]], np.int32)
    dist = np.array([[1.0, 3.0], [1.0, 2.0], [1.0, 5.0]], np.float32)
    ip, ind, w, sig, rho = O.fuzzy_graph(idx, dist, 2, symmetrise=True)
    wd = {}
    for i in range(3):
        for t in range(2):
            wd[(i, idx[i, t])] = np.exp(-max(dist[i, t] - rho[i], 0) / sig[i])
    for i in range(3):
        for p in range(ip[i], ip[i + 1]):
            j = ind[p]
            a, b = wd.get((i, j), 0.0), wd.get((j, i), 0.0)
            assert abs(w[p] - (a + b - a * b)) < 1e-6
        assert list(ind[ip[i]:ip[i + 1]]) == sorted(ind[ip[i]:ip[i + 1]])


# ------------------------------------------------------------------ embedding
def _triangle():
    return np.array([0, 2, 4, 6]), np.array([1, 2, 0, 2, 0, 1], np.int32), np.ones(6, np.float32)


def test_embedding_shapes_determinism_and_seeds():
    ip, ind, w = _triangle()
    for dim in (2, 4, 8, 16):
        assert O.embedding(ip, ind, w, dim, seed=42).shape == (3, dim)
    a = O.embedding(ip, ind, w, 4, seed=123)
    b = O.embedding(ip, ind, w, 4, seed=123)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(O.embedding(ip, ind, w, 4, seed=1), O.embedding(ip, ind, w, 4, seed=2))
    assert np.isfinite(O.embedding(ip, ind, w, 8, seed=42)).all()


def test_embedding_zero_epochs_keeps_init_and_no_edges():
    # embedding.rs:379-390, :326-332
    init = np.array([[1.0, 2.0], [3.0, 4.0]], np.float32)
    out = O.embedding(np.array([0, 0, 0]), np.zeros(0, np.int32), np.zeros(0, np.float32), 2,
                      n_epochs=0, init=init)
    assert np.array_equal(out, init)
    one = O.embedding(np.array([0, 0]), np.zeros(0, np.int32), np.zeros(0, np.float32), 2)
    assert one.shape == (1, 2)


def test_embedding_attraction_contracts():
    # embedding.rs:393-419: repulsion disabled, two linked nodes move closer
    init = np.array([[0.0, 0.0], [10.0, 0.0]], np.float32)
    out = O.embedding(np.array([0, 1, 2]), np.array([1, 0], np.int32), np.ones(2, np.float32), 2,
                      n_epochs=100, negative_sample_rate=0.0, init=init)
    assert ((out[0] - out[1]) ** 2).sum() < 100.0


def test_embedding_disconnected_components_finite():
    ip = np.array([0, 1, 2, 3, 4])
    ind = np.array([1, 0, 3, 2], np.int32)
    out = O.embedding(ip, ind, np.ones(4, np.float32), 4)
    assert out.shape == (4, 4) and np.isfinite(out).all()


# ------------------------------------------------------------------ MST
def test_mst_known_answers():
    # mst.rs:176-283
    assert len(O.mst(np.zeros((1, 2)), 1)[0]) == 0
    u, v, w = O.mst(np.array([[0.0, 0.0], [3.0, 4.0]]), 1)
    assert len(w) == 1 and abs(w[0] - 5.0) < 1e-10
    u, v, w = O.mst(np.array([[0.0, 0.0], [1.0, 0.0], [0.5, 0.866]]), 1)
    assert len(w) == 2 and 1.9 < w.sum() < 2.1
    pts = np.array([[0, 0], [0.1, 0], [0, 0.1], [10, 10], [10.1, 10], [10, 10.1]], np.float64)
    u, v, w = O.mst(pts, 1)
    assert len(w) == 5 and (w > 5.0).sum() == 1
    line = np.array([[0.0], [1.0], [2.0], [3.0]])
    assert O.mst(line, 2)[2].sum() >= O.mst(line, 1)[2].sum()
    u, v, w = O.mst(np.arange(5, dtype=np.float64)[:, None], 1)
    assert abs(w.sum() - 4.0) < 1e-10
    u, v, w = O.mst(np.array([[0.0], [1.0], [3.0], [6.0], [10.0]]), 1)
    assert set(u) | set(v) == {0, 1, 2, 3, 4}
    u, v, w = O.mst(np.array([[0.0, 0.0], [0.01, 0.0], [10.0, 0.0]]), 2)
    assert (w > 1.0).all()
    u, v, w = O.mst(np.array([[0.0] * 8, [1.0] * 8, [2.0] * 8]), 1)
    assert (w > 0).all() and np.isfinite(w).all()


def _mr_dense(x, ms):
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    dd = d2.copy()
    np.fill_diagonal(dd, np.inf)
    core = np.sort(dd, axis=1)[:, min(ms, len(x) - 1) - 1] if ms > 0 else np.zeros(len(x))
    return np.maximum(np.maximum(d2, core[:, None]), core[None, :]), core


def test_core_distances_and_mst_total_vs_scipy():
    x, _ = datagen.make_blobs(60, 3, 4, 30.0, 1.0, seed=5)
    for ms in (1, 5):
        mr, core = _mr_dense(x, ms)
        np.testing.assert_allclose(O.core_sq(x, ms), core, rtol=1e-12)
        u, v, w = O.mst(x, ms)
        ref = csg.minimum_spanning_tree(np.sqrt(mr) + 0.0).toarray()
        # scipy drops zero entries; our mutual-reachability weights are all > 0 here
        np.testing.assert_allclose(np.sort(w), np.sort(ref[ref > 0]), rtol=1e-12)
        assert abs(w.sum() - ref.sum()) < 1e-9


def test_mst_edge_set_matches_canonical_kruskal():
    # unique MST under (w^2, min, max): Kruskal on the dense graph must agree exactly
    rng = np.random.default_rng(9)
    x = np.round(rng.normal(size=(80, 3)) * 2) / 2  # many exact ties
    mr, _ = _mr_dense(x, 3)
    n = len(x)
    cand = sorted((mr[i, j], i, j) for i in range(n) for j in range(i + 1, n))
    par = list(range(n))

    def f(a):
        while par[a] != a:
            par[a] = par[par[a]]
            a = par[a]
        return a
    ref = []
    for w2, i, j in cand:
        a, b = f(i), f(j)
        if a != b:
            par[a] = b
            ref.append((i, j))
    u, v, w = O.mst(x, 3)
    assert sorted(zip(u.tolist(), v.tolist())) == sorted(ref)


# ------------------------------------------------------------------ linkage
def test_linkage_known_answers():
    # linkage.rs:378-507
    L, R, D, S = O.linkage([0, 1], [1, 2], np.array([1.0, 2.0]), 3)
    assert D.tolist() == [1.0, 2.0] and S.tolist() == [2, 3]
    L, R, D, S = O.linkage([1, 0], [2, 1], np.array([5.0, 1.0]), 3)
    assert D.tolist() == [1.0, 5.0]
    L, R, D, S = O.linkage([0, 1, 2], [1, 2, 3], np.array([1.0, 2.0, 3.0]), 4)
    assert S.tolist() == [2, 3, 4]
    L, R, D, S = O.linkage([0, 2, 1], [1, 3, 2], np.array([3.0, 1.0, 5.0]), 4)
    assert (np.diff(D) >= 0).all()
    L, R, D, S = O.linkage([0, 2, 0], [1, 3, 2], np.array([1.0, 1.0, 10.0]), 4)
    assert S.tolist() == [2, 2, 4]
    assert (L >= R).all()


def test_linkage_partitions_match_scipy_single():
    x, _ = datagen.make_blobs(25, 3, 4, 30.0, 0.5, seed=42)
    n = len(x)
    u, v, w = O.mst(x, 1)
    L, R, D, S = O.linkage(u, v, w, n)
    Z = np.column_stack([R, L, D, S]).astype(np.float64)
    ref = sch.linkage(x, method="single")
    np.testing.assert_allclose(np.sort(Z[:, 2]), np.sort(ref[:, 2]), rtol=1e-12)
    for t in np.quantile(D, [0.2, 0.5, 0.9, 0.99]):
        a = sch.fcluster(Z, t, criterion="distance")
        b = sch.fcluster(ref, t, criterion="distance")
        # same partition up to relabelling
        assert len(set(zip(a, b))) == len(set(a)) == len(set(b))


# ------------------------------------------------------------------ condensed tree / labels
def _tree(x, ms, mcs):
    n = len(x)
    u, v, w = O.mst(x, ms)
    L, R, D, S = O.linkage(u, v, w, n)
    return O.condense(L, R, D, S, n, mcs)


def test_condensed_two_clusters_labels_strengths():
    # condensed_tree.rs:512-573
    x = np.array([[0, 0], [0.1, 0], [0, 0.1], [10, 10], [10.1, 10], [10, 10.1]], np.float64)
    P, C, lam, CS = _tree(x, 1, 2)
    lv = O.extract_leaves(P, C, CS, 6)
    assert len(lv) == 2
    lab = O.label_vector(P, C, lv, 6)
    assert lab[0] == lab[1] == lab[2] and lab[3] == lab[4] == lab[5] and lab[0] != lab[3]
    assert (lab >= 0).all()
    st = O.strengths(P, C, lam.astype(np.float32), CS, lv, lab)
    assert ((st >= 0) & (st <= 1)).all()
    P2, C2, _, CS2 = _tree(x, 1, 6)
    assert len(O.extract_leaves(P2, C2, CS2, 6)) <= 1


def test_condense_empty_and_noise_pair():
    P, C, lam, CS = O.condense(np.zeros(0, np.int64), np.zeros(0, np.int64),
                               np.zeros(0, np.float32), np.zeros(0, np.int64), 0, 2)
    assert len(P) == 0
    x = np.array([[0, 0], [0.05, 0], [5, 5], [10, 0], [0, 10]], np.float64)
    P, C, lam, CS = _tree(x, 1, 2)
    lab = O.label_vector(P, C, O.extract_leaves(P, C, CS, 5), 5)
    if lab[0] >= 0 and lab[1] >= 0:
        assert lab[0] == lab[1]


def test_labels_vs_bruteforce_nearest_selected_ancestor():
    # the union-find representative equals the nearest selected ancestor
    x, _ = datagen.make_blobs(40, 4, 3, 15.0, 1.0, seed=11)
    n = len(x)
    P, C, lam, CS = _tree(x, 5, 5)
    par = {c: p for p, c in zip(P, C)}
    clusters = sorted(set(P[CS > 1].tolist()) | set(C[CS > 1].tolist()))
    rng = np.random.default_rng(0)
    for _ in range(5):
        sel = sorted(rng.choice(clusters, size=min(4, len(clusters)), replace=False).tolist())
        lab = O.label_vector(P, C, sel, n)
        root = P.min()
        for i in range(n):
            a = par.get(i)
            while a is not None and a not in sel:
                a = par.get(a)
            exp = sel.index(a) if (a is not None and a > root) else -1
            assert lab[i] == exp


# ------------------------------------------------------------------ persistence / layers
def test_find_peaks_known_answers_and_scipy():
    # persistence.rs:504-536
    assert O.find_peaks([0, 1, 0, 2, 0]).tolist() == [1, 3]
    assert O.find_peaks([0, 1, 1, 0]).tolist() in ([1], [2])
    assert len(O.find_peaks([1, 2, 3, 4, 5])) == 0
    assert len(O.find_peaks([])) == 0 and len(O.find_peaks([1, 2])) == 0
    rng = np.random.default_rng(2)
    for _ in range(50):
        x = rng.integers(0, 4, size=30).astype(np.float64)
        assert O.find_peaks(x).tolist() == scipy.signal.find_peaks(x)[0].tolist()


def _two():
    return np.array([[0, 0], [0.1, 0], [0, 0.1], [0.1, 0.1], [0.05, 0.05],
                     [10, 10], [10.1, 10], [10, 10.1], [10.1, 10.1], [10.05, 10.05]], np.float64)


def _three():
    return np.vstack([_two(), [[20, 0], [20.1, 0], [20, 0.1], [20.1, 0.1], [20.05, 0.05]]])


def test_barcode_and_persistence_shape():
    # persistence.rs:538-566
    x = _two()
    P, C, lam, CS = _tree(x, 1, 2)
    for fp64 in (False, True):
        sizes, pers, b, d, ld = O.persistence(P, C, lam.astype(np.float32), CS, 10, 2, fp64)
        assert len(b) > 0 and d[0] == 10
        assert len(sizes) == len(pers) and len(sizes) > 0 and (pers >= 0).all()


def test_persistence_curve_brute_force():
    # each barcode adds (death - birth) * lambda_death to bins [birth, death)
    x, _ = datagen.make_blobs(50, 5, 3, 12.0, 1.0, seed=21)
    n = len(x)
    P, C, lam, CS = _tree(x, 5, 5)
    sizes, pers, b, d, ld = O.persistence(P, C, lam.astype(np.float32), CS, n, 5, True)
    ref = np.zeros(len(sizes))
    for i in range(1, len(b)):
        if d[i] <= b[i]:
            continue
        for k, s in enumerate(sizes):
            if b[i] <= s < d[i]:
                ref[k] += (d[i] - b[i]) * ld[i]
    np.testing.assert_allclose(pers, ref, rtol=1e-12)
    assert np.array_equal(sizes, np.unique(b))


def test_cluster_layers_known_answers():
    # persistence.rs:568-671
    labels, st, pers = O.cluster_layers(_two(), 1, 2, 0.2, 10)
    base = labels[0]
    assert base.max() + 1 >= 2
    assert base[0] == base[1] == base[4] and base[5] == base[6] == base[9] and base[0] != base[5]
    labels, _, _ = O.cluster_layers(_three(), 1, 2, 0.2, 10)
    assert labels[0].max() + 1 >= 3
    counts = [l.max() + 1 for l in labels]
    assert counts == sorted(counts, reverse=True)
    assert len(O.cluster_layers(_three(), 1, 2, 0.2, 2)[0]) <= 2
    for lab, s in zip(*O.cluster_layers(_two(), 1, 2, 0.2, 10)[:2]):
        assert ((s[lab >= 0] >= 0) & (s[lab >= 0] <= 1)).all()


def test_select_diverse_peaks_known_answers():
    # persistence.rs:673-700
    assert len(O.select_diverse_peaks([], [], [], [], [], 0.2, 10)) == 0
    sel = O.select_diverse_peaks([1], [0.0, 5.0, 0.0], [2.0, 3.0, 4.0], [2.0, 2.0], [10.0, 5.0],
                                 0.2, 10)
    assert sel.tolist() == [1]


def _acc(pred, gt):
    c = 0
    for p in set(pred.tolist()) - {-1}:
        c += np.bincount(gt[pred == p]).max()
    return c / len(pred)


def test_integration_blobs_mst_and_labels():
    # tests/integration_tests.rs:17-189
    x, gt = datagen.make_blobs(20, 2, 4, 50.0, 0.5, seed=42)
    u, v, w = O.mst(x, 5)
    assert len(u) == len(x) - 1
    cross = [(a, b) for a, b in zip(u, v) if gt[a] != gt[b]]
    assert len(cross) == 1
    x, gt = datagen.make_blobs(40, 2, 4, 100.0, 0.3, seed=42)
    P, C, lam, CS = _tree(x, 5, 5)
    lv = O.extract_leaves(P, C, CS, len(x))
    lab = O.label_vector(P, C, lv, len(x))
    assert (lab < 0).sum() <= 2 and _acc(lab, gt) > 0.95
    x, gt = datagen.make_blobs(30, 3, 4, 50.0, 0.5, seed=99)
    P, C, lam, CS = _tree(x, 5, 5)
    assert len(O.extract_leaves(P, C, CS, len(x))) >= 3


def test_integration_search_for_n_clusters():
    # tests/integration_tests.rs:252-333
    x, gt = datagen.make_blobs(50, 2, 4, 50.0, 0.5, seed=42)
    lab, st = O.search_n_clusters(x, 5, 2)
    assert lab.max() + 1 == 2 and _acc(lab, gt) > 0.95
    x, gt = datagen.make_blobs(40, 5, 4, 40.0, 0.5, seed=7)
    lab, _ = O.search_n_clusters(x, 5, 5)
    assert abs((lab.max() + 1) - 5) <= 1


def test_integration_degenerate_inputs():
    # tests/integration_tests.rs:504-545
    u, v, w = O.mst(np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0]]), 5)
    assert len(u) == 2
    x = np.ones((10, 3))
    u, v, w = O.mst(x, 5)
    assert np.isfinite(w).all() and (w >= 0).all()


def test_end_to_end_blobs_recovered():
    # fixture and bar of tests/integration_tests_gpu.rs:22-67 (two blobs, accuracy > 0.95)
    x, gt = datagen.make_blobs(100, 2, 8, 20.0, 5.0, seed=42)
    r = O.evoc(x.astype(np.float32))
    pers = r["persistence"]
    best = max(i for i in range(len(pers)) if pers[i] == pers.max()) if len(pers) > 1 else 0
    assert _acc(r["labels"][best], gt) > 0.95


def test_fp64_persistence_selects_the_reference_f32_layers():
    """The library sums barcode persistence in fp64; the reference sums in T = f32
    (src/clustering/persistence.rs:135-173) and find_peaks / select_diverse_peaks compare
    those sums exactly. Over 40 seeded embeddings of varied size and cluster count the two
    modes of the oracle select identical layers (labels, order and best layer); the
    library is bit-exact against the fp64 mode (tests/test_stages_gpu.py)."""
    agree = 0
    total = 40
    for seed in range(total):
        rng = np.random.default_rng(seed)
        n = int(rng.integers(500, 2500))
        c = int(rng.integers(3, 40))
        e, _ = datagen.gaussian_mixture(n, 8, c, float(rng.uniform(1.5, 5.0)), 1.0, seed=seed)
        l64, s64, p64 = O.cluster_layers(e, 5, 5, 0.2, 10, fp64=True)
        l32, s32, p32 = O.cluster_layers(e, 5, 5, 0.2, 10, fp64=False)
        same = l64.shape == l32.shape and np.array_equal(l64, l32) and \
            np.array_equal(s64.view(np.uint32), s32.view(np.uint32))
        if same:
            b64 = max(i for i in range(len(p64)) if p64[i] == p64.max()) if len(p64) > 1 else 0
            b32 = max(i for i in range(len(p32)) if p32[i] == p32.max()) if len(p32) > 1 else 0
            same = b64 == b32
        agree += same
    print(f"fp64 vs f32 persistence: identical layer selection on {agree}/{total} embeddings")
    assert agree == total


# ------------------------------------------------------------------ embedding trajectories
# Two nodes joined by one edge of weight 1 (src/graph/embedding.rs:48-298, defaults
# noise 0.5, alpha 0.1, beta1 0.5, beta2 0.9, eps 1e-7). Expected positions are literals
# derived by hand in float64 from the closed forms written next to them; f32 arithmetic
# agrees to a few ulps. Schedule: eps = max_w / w = 1, eps_neg = 1.5 / rate; both "next"
# times start at their period, so epoch 0 samples nothing.
_IP2 = np.array([0, 1, 2])
_IND2 = np.array([1, 0], np.int32)
_W2 = np.ones(2, np.float32)


def test_embedding_epoch_zero_is_a_no_op():
    init = np.array([[0.0, 0.0], [1.0, 0.0]], np.float32)
    out = O.embedding(_IP2, _IND2, _W2, 2, n_epochs=1, init=init, seed=3)
    assert np.array_equal(out, init)


def test_embedding_first_step_bias_terms_and_lr():
    """Epoch 1 (t = 2): lr = 0.1 (1 - 1/2) = 0.05. Attraction at d = 1:
    coeff = (-2*0.5*1 - 2) / (2 - 0.5 + 1) = -1.2, gradient of node 0 = 1.2 along x.
    Adam from zero moments: m = 0.5 * 1.2, v = 0.1 * 1.44, step =
    lr * sqrt(1 - 0.81) / (1 - 0.25) * m / (sqrt(v) + eps') = 0.0459468239.
    A t = 1 bias correction would give 0.05, lr without the decay 0.0918936,
    a sampled epoch 0 two steps."""
    init = np.array([[0.0, 0.0], [1.0, 0.0]], np.float32)
    out = O.embedding(_IP2, _IND2, _W2, 2, n_epochs=2, init=init, seed=3)
    np.testing.assert_allclose(out, [[0.0459468239, 0.0], [0.9540531761, 0.0]], rtol=0, atol=2e-7)


def test_embedding_second_step_attractive_coefficient():
    """Three epochs. Epoch 1 (lr = 0.0666667) moves each node 0.0612624319 inward; epoch 2
    (t = 3, lr = 0.0333333) sees d2 = 1 - 2 * 0.0612624319 and the gradient
    g2 = -coeff(d2) d2 = 1.2016598677 with coeff(d) = (-d - 2) / (2 d^2 - 0.5 d + 1);
    m2 = 0.6 + 0.5 (g2 - 0.6), v2 = 0.144 + 0.1 (g2^2 - 0.144), and the step
    lr * sqrt(1 - 0.729) / 0.875 * m2 / sqrt(v2) = 0.0341290262. The second step depends on
    g2 / g1, so it pins the attractive coefficient (a clipped or differently scaled one
    changes it)."""
    init = np.array([[0.0, 0.0], [1.0, 0.0]], np.float32)
    out = O.embedding(_IP2, _IND2, _W2, 2, n_epochs=3, init=init, seed=3)
    np.testing.assert_allclose(out, [[0.0953914581, 0.0], [0.9046085419, 0.0]], rtol=0, atol=2e-7)


def test_embedding_negative_counts_and_gamma_schedule():
    """d = 2, negative_sample_rate 3 (eps_neg = 0.5, next_neg = 0.5), three epochs, seed 12:
    the pinned Philox draws of node 0 (the smaller endpoint) all land on node 1.
    Epoch 1: n_neg = floor((1 - 0.5) / 0.5) = 1 and gamma = 0.5 + 1/2 = 1.0. Attraction
    -coeff(2) * 2 = +1.0 and repulsion gamma*4 / ((1 + 0.25*4) * 4) * (-2) = -1.0 cancel:
    node 0 has gradient 0 and does not move (any other gamma moves it). Node 1 (no
    negatives) steps 0.0666667 * 0.9189365 inward. next_neg becomes 0.5 + 1 * 0.5 = 1.0.
    Epoch 2: n_neg = floor((2 - 1) / 0.5) = 2 and gamma = 1.5: node 0's gradient is
    -coeff(d) d + 2 * clip(-gamma*4 d / ((1 + d^2/4) d^2)) = -2.1793687 at d = 1.9387376,
    a step of -0.0313563360; node 1 ends at 1.9045601471."""
    for dr, ep in ((0, 1), (0, 2), (1, 2)):
        r = O.lib().orc_philox_u32(12, dr, ep, 0, 0x454E4547)
        assert (r * 2) >> 32 == 1   # draw lands on node 1, not on node 0 itself
    init = np.array([[0.0, 0.0], [2.0, 0.0]], np.float32)
    out = O.embedding(_IP2, _IND2, _W2, 2, n_epochs=3, negative_sample_rate=3.0, init=init,
                      seed=12)
    np.testing.assert_allclose(out, [[-0.0313563360, 0.0], [1.9045601471, 0.0]], rtol=0,
                               atol=4e-7)
