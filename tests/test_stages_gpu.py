"""Stage-by-stage parity of the HIP library against the CPU oracle on identical inputs.

Bars (from the task's north star): fuzzy weights within 1e-5 relative; embedding, label
propagation layout, core distances, MST, condensed tree, labels, strengths and fp64
persistence bit-exact when both sides get the same input."""
import numpy as np
import pytest
from sklearn.metrics import adjusted_rand_score as ari

import datagen
import evoc_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph():
    x, gt = datagen.gaussian_mixture(3000, 32, 12, 4.0, 1.0, seed=13)
    idx, dist = O.knn_exhaustive(x, 15)
    return x, gt, idx, dist


def test_fuzzy_graph_parity(graph):
    x, gt, idx, dist = graph
    for sym in (True, False):
        gp, gi, gw, gs, gr = H.fuzzy_graph(idx, dist, 15.0, symmetrise=sym)
        op, oi, ow, os_, orh = O.fuzzy_graph(idx, dist, 15.0, symmetrise=sym)
        assert np.array_equal(gp, op) and np.array_equal(gi, oi)
        np.testing.assert_allclose(gw, ow, rtol=1e-5, atol=0)
        np.testing.assert_array_equal(gr, orh)
        np.testing.assert_allclose(gs, os_, rtol=1e-5)
        print("weights bit-identical: %.4f" % np.mean(gw.view(np.uint32) == ow.view(np.uint32)))


def test_fuzzy_graph_hub_rows():
    # one node in every kNN list: its symmetrised row is longer than a wavefront
    n, k = 400, 5
    rng = np.random.default_rng(0)
    idx = np.stack([rng.permutation(np.delete(np.arange(n), i))[:k] for i in range(n)]).astype(np.int32)
    idx[1:, 0] = 0
    for i in range(1, n):
        if len(set(idx[i])) < k:
            idx[i] = [0] + [j for j in rng.permutation(n) if j not in (0, i)][: k - 1]
    dist = np.sort(rng.uniform(0.1, 2.0, size=(n, k)), axis=1).astype(np.float32)
    gp, gi, gw, _, _ = H.fuzzy_graph(idx, dist, float(k))
    op, oi, ow, _, _ = O.fuzzy_graph(idx, dist, float(k))
    assert gp[1] - gp[0] > 64
    assert np.array_equal(gp, op) and np.array_equal(gi, oi)
    np.testing.assert_allclose(gw, ow, rtol=1e-5)


def test_embedding_parity_bit_exact(graph):
    x, gt, idx, dist = graph
    op, oi, ow, _, _ = O.fuzzy_graph(idx, dist, 15.0)
    init = O.label_prop_init(op, oi, ow, 8, data=x, seed=7)
    for dim, ini in ((8, init), (4, None)):
        g = H.embedding(op, oi, ow, dim, n_epochs=30, init=ini, seed=11)
        o = O.embedding(op, oi, ow, dim, n_epochs=30, init=ini, seed=11)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


def test_embedding_asymmetric_graph_parity():
    rng = np.random.default_rng(5)
    n = 300
    rows = [sorted(rng.choice(n, size=rng.integers(1, 8), replace=False).tolist()) for _ in range(n)]
    ip = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    ind = np.concatenate(rows).astype(np.int32)
    w = rng.uniform(0.05, 1.0, size=len(ind)).astype(np.float32)
    g = H.embedding(ip, ind, w, 4, n_epochs=20, seed=3)
    o = O.embedding(ip, ind, w, 4, n_epochs=20, seed=3)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


def _hub_graph(seed=9, n=400, hub_size=200):
    """Symmetric column-sorted CSR: 4 random edges per node plus three hubs with about
    hub_size edges each (several 64-slot chunks), symmetric weights in (0.05, 1)."""
    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, size=4, replace=False):
            if int(j) != i:
                adj[i].add(int(j))
                adj[int(j)].add(i)
    for h in (0, 7, 123):
        for j in rng.choice(n, size=hub_size, replace=False):
            if int(j) != h:
                adj[h].add(int(j))
                adj[int(j)].add(h)
    wt = {}
    rows = [sorted(a) for a in adj]
    ip = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    ind = np.concatenate(rows).astype(np.int32)
    w = np.empty(len(ind), np.float32)
    for i, r in enumerate(rows):
        for q, j in enumerate(r):
            key = (min(i, j), max(i, j))
            if key not in wt:
                wt[key] = np.float32(rng.uniform(0.05, 1.0))
            w[ip[i] + q] = wt[key]
    assert max(len(r) for r in rows) >= 128
    return ip, ind, w


def test_embedding_hub_nodes_parity():
    """Nodes with 64 or more edges take the wavefront-per-node epoch kernel. The hubs
    span several 64-slot chunks, and a high negative-sample rate stages more than 256
    gradient terms per chunk; the layout must equal the oracle's bit for bit, for the
    specialised dims and a generic one."""
    ip, ind, w = _hub_graph()
    for dim, rate in ((8, 1.0), (16, 5.0), (5, 5.0)):
        g = H.embedding(ip, ind, w, dim, n_epochs=25, negative_sample_rate=rate, seed=13)
        o = O.embedding(ip, ind, w, dim, n_epochs=25, negative_sample_rate=rate, seed=13)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), (dim, rate)


def test_embedding_workgroup_hub_nodes_parity():
    """Nodes with 1024 or more edges take one workgroup each: 16 wavefronts stage the
    terms of 16 chunks at once and one adds them in slot order. Hubs of about 1100 and
    2900 edges end inside the first and the third group of 1024 slots; a negative-sample
    rate of 5 stages more than 1024 terms per group. Bit for bit equal to the oracle."""
    for n, hub, dims in ((1200, 1100, ((8, 1.0), (5, 5.0))), (3000, 2900, ((8, 5.0), (16, 1.0)))):
        ip, ind, w = _hub_graph(seed=21, n=n, hub_size=hub)
        assert np.diff(ip).max() >= 1024
        for dim, rate in dims:
            g = H.embedding(ip, ind, w, dim, n_epochs=12, negative_sample_rate=rate, seed=17)
            o = O.embedding(ip, ind, w, dim, n_epochs=12, negative_sample_rate=rate, seed=17)
            assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), (n, dim, rate)


def test_label_prop_init_hub_graph_parity():
    """The input level's label votes and expansion run one wavefront per node on the
    GPU; hubs with ~200 neighbours exercise their multi-chunk paths."""
    ip, ind, w = _hub_graph(seed=21, n=600)
    for seed in (3, 8):
        g = H.label_prop_init(ip, ind, w, 8, data=None, seed=seed)
        o = O.label_prop_init(ip, ind, w, 8, data=None, seed=seed)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), seed


def _outlier_graph(seed=5, n_a=1500, n_b=900, n_c=400):
    """Symmetric column-sorted CSR built so that label propagation leaves outliers: a
    cluster of weight-1.0 edges, chains of weight-0.3 edges (no vote can pass 1.0)
    hooked to the cluster by a few weak edges, and isolated pairs; node ids shuffled so
    that outliers reach earlier and later outliers in their searches."""
    rng = np.random.default_rng(seed)
    n = n_a + n_b + n_c
    wt = {}

    def edge(i, j, w):
        if i != j:
            wt[(min(i, j), max(i, j))] = np.float32(w)

    for i in range(n_a):
        for j in rng.choice(n_a, size=8, replace=False):
            edge(i, int(j), 1.0)
    for q in range(n_b - 1):
        edge(n_a + q, n_a + q + 1, 0.3)
    for q in range(0, n_b, 50):
        edge(n_a + q, int(rng.integers(n_a)), 0.2)
    for q in range(0, n_c - 1, 2):
        edge(n_a + n_b + q, n_a + n_b + q + 1, 0.4)
    perm = rng.permutation(n)
    adj = [dict() for _ in range(n)]
    for (i, j), w in wt.items():
        a, b = int(perm[i]), int(perm[j])
        adj[a][b] = w
        adj[b][a] = w
    rows = [sorted(a.items()) for a in adj]
    ip = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    ind = np.array([j for r in rows for j, _ in r], np.int32)
    w = np.array([x for r in rows for _, x in r], np.float32)
    return ip, ind, w


def test_label_prop_init_outliers_parity():
    ip, ind, w = _outlier_graph()
    x = np.random.default_rng(2).normal(size=(ip.shape[0] - 1, 12)).astype(np.float32)
    for seed, data in ((1, None), (7, x)):
        g = H.label_prop_init(ip, ind, w, 8, data=data, seed=seed)
        o = O.label_prop_init(ip, ind, w, 8, data=data, seed=seed)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), seed


def test_label_prop_init_parity(graph):
    x, gt, idx, dist = graph
    op, oi, ow, _, _ = O.fuzzy_graph(idx, dist, 15.0)
    g = H.label_prop_init(op, oi, ow, 8, data=x, seed=42)
    o = O.label_prop_init(op, oi, ow, 8, data=x, seed=42)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    g2 = H.label_prop_init(op, oi, ow, 8, data=None, seed=3)
    o2 = O.label_prop_init(op, oi, ow, 8, data=None, seed=3)
    assert np.array_equal(g2.view(np.uint32), o2.view(np.uint32))


def _embedding(n=4000, seed=3):
    x, gt = datagen.gaussian_mixture(n, 8, 15, 3.0, 1.0, seed=seed)
    return x, gt


def test_core_distances_bit_exact():
    e, _ = _embedding()
    for ms in (1, 5, 16):
        assert np.array_equal(H.core_sq(e, ms).view(np.uint32), O.core_sq(e, ms).view(np.uint32))


def test_mst_bit_exact_with_ties():
    e, _ = _embedding(3000)
    for ms in (1, 5):
        gu, gv, gw = H.mst(e, ms)
        ou, ov, ow = O.mst(e, ms)
        order = np.lexsort((ov, ou, ow))
        assert np.array_equal(gu, ou[order]) and np.array_equal(gv, ov[order])
        assert np.array_equal(gw.view(np.uint32), ow[order].view(np.uint32))
    tied = np.round(np.random.default_rng(4).normal(size=(1500, 4)) * 2).astype(np.float32) / 2
    gu, gv, gw = H.mst(tied, 3)
    ou, ov, ow = O.mst(tied, 3)
    assert sorted(zip(gu.tolist(), gv.tolist())) == sorted(zip(ou.tolist(), ov.tolist()))


def test_mst_and_core_deep_tree_ragged():
    # 20011 points: 313 tiles (the last one partial), nine levels of kd splits, and a
    # block of exact duplicates so that equal coordinates straddle split boundaries
    e, _ = _embedding(20011, seed=12)
    e[5000:5400] = e[5000]
    assert np.array_equal(H.core_sq(e, 5).view(np.uint32), O.core_sq(e, 5).view(np.uint32))
    gu, gv, gw = H.mst(e, 5)
    ou, ov, ow = O.mst(e, 5)
    order = np.lexsort((ov, ou, ow))
    assert np.array_equal(gu, ou[order]) and np.array_equal(gv, ov[order])
    assert np.array_equal(gw.view(np.uint32), ow[order].view(np.uint32))
    e3 = np.ascontiguousarray(e[:, :3])
    assert np.array_equal(H.core_sq(e3, 12).view(np.uint32), O.core_sq(e3, 12).view(np.uint32))


def test_cluster_layers_bit_exact():
    e, gt = _embedding(3000, seed=8)
    gl, gs, gp = H.cluster_layers(e, 5, 5, 0.2, 10)
    ol, os_, op = O.cluster_layers(e, 5, 5, 0.2, 10, fp64=True)
    assert gl.shape == ol.shape
    assert np.array_equal(gl, ol)
    assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))
    assert np.array_equal(gp.view(np.uint64), op.view(np.uint64))


def test_search_for_n_clusters_bit_exact():
    e, gt = _embedding(3000, seed=9)
    for target in (5, 15, 40):
        gl, gs = H.search_for_n_clusters(e, 5, target)
        ol, os_ = O.search_n_clusters(e, 5, target)
        assert np.array_equal(gl, ol)
        assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))


def _acc(pred, gt):
    c = 0
    for p in set(pred.tolist()) - {-1}:
        c += np.bincount(gt[pred == p]).max()
    return c / len(pred)


def test_end_to_end_same_knn_identical_labels(graph):
    x, gt, idx, dist = graph
    r = H.evoc(x, precomputed_knn=(idx, dist), seed=5)
    o = O.evoc(x, O.params(), seed=5, pre_idx=idx, pre_dist=dist)
    assert np.array_equal(r.embedding.view(np.uint32), o["embedding"].view(np.uint32))
    assert np.array_equal(r.cluster_layers, o["labels"])
    assert np.array_equal(r.persistence_scores, o["persistence"])


def test_end_to_end_vs_oracle_ari():
    x, gt = datagen.gaussian_mixture(4000, 64, 20, 5.0, 1.0, seed=42)
    r = H.evoc(x, seed=42)
    o = O.evoc(x, O.params(), seed=42)
    assert np.array_equal(r.nn_indices, o["knn_idx"])
    pers = o["persistence"]
    ob = max(i for i in range(len(pers)) if pers[i] == pers.max()) if len(pers) > 1 else 0
    score = ari(r.best_labels(), o["labels"][ob])
    print("ARI GPU vs CPU labels:", score, "layers", len(r.cluster_layers), len(o["labels"]))
    assert score >= 0.98


def test_reference_gpu_integration_blobs():
    # tests/integration_tests_gpu.rs:22-67 and :122-185
    x, gt = datagen.make_blobs(100, 2, 8, 20.0, 5.0, seed=42)
    r = H.evoc(x, seed=42)
    assert _acc(r.best_labels(), gt) > 0.95
    o = O.evoc(x.astype(np.float32), O.params(), seed=42)
    pers = o["persistence"]
    ob = max(i for i in range(len(pers)) if pers[i] == pers.max()) if len(pers) > 1 else 0
    assert abs(r.n_clusters() - (o["labels"][ob].max() + 1)) <= 2
    self_loops = sum(i in row for i, row in enumerate(r.nn_indices))
    assert self_loops <= 0.01 * len(x)


def test_reference_gpu_integration_approx_n_clusters_and_precomputed():
    # tests/integration_tests_gpu.rs:187-268
    x, gt = datagen.make_blobs(100, 3, 8, 20.0, 5.0, seed=42)
    idx, dist = O.knn_exhaustive(x.astype(np.float32), 15)
    r = H.evoc(x, precomputed_knn=(idx, dist), seed=42)
    assert np.array_equal(r.nn_indices, idx)
    r = H.evoc(x, params=H.EvocParams(approx_n_clusters=3), seed=42)
    assert len(r.cluster_layers) == 1 and r.persistence_scores.tolist() == [0.0]
    assert abs(r.n_clusters() - 3) <= 1


def test_strided_input_matches_contiguous():
    x, _ = datagen.make_blobs(80, 2, 6, 20.0, 3.0, seed=1)
    xf = np.asfortranarray(x.astype(np.float32))
    a = H.evoc(x.astype(np.float32), seed=3)
    b = H.evoc(xf, seed=3)
    assert np.array_equal(a.cluster_layers, b.cluster_layers)


def test_result_arrays_outlive_the_result():
    """The arrays of an EvocResult view the library's result buffers: one array kept alone
    keeps them, and later calls (which allocate and free results of their own) leave it
    intact."""
    import gc

    x, _ = datagen.gaussian_mixture(3000, 32, 8, 5.0, 1.0, seed=11)
    r = H.evoc(x, seed=3)
    emb, lab, nn = r.embedding, r.cluster_layers, r.nn_indices
    keep = (emb.copy(), lab.copy(), nn.copy())
    del r
    gc.collect()
    for seed in (4, 5):
        r2 = H.evoc(x, seed=seed)
        assert r2.embedding.shape == emb.shape
        del r2
        gc.collect()
    assert np.array_equal(emb.view(np.uint32), keep[0].view(np.uint32))
    assert np.array_equal(lab, keep[1])
    assert np.array_equal(nn, keep[2])
    emb[0, 0] = 1.0  # writable views


def test_cluster_layers_agree_with_reference_f32_persistence():
    """The GPU's fp64 persistence sums against the oracle's reference-order f32 mode
    (src/clustering/persistence.rs:135-173): identical layers and labels."""
    agree = 0
    for seed in range(6):
        e, _ = _embedding(3000 + 500 * seed, seed=40 + seed)
        gl, gs, gp = H.cluster_layers(e, 5, 5, 0.2, 10)
        ol, os_, op = O.cluster_layers(e, 5, 5, 0.2, 10, fp64=False)
        agree += gl.shape == ol.shape and np.array_equal(gl, ol)
    print(f"GPU fp64 vs oracle f32 persistence: identical layers on {agree}/6 embeddings")
    assert agree == 6


def test_non_finite_input_rejected_on_device_paths():
    """evoc_hip_run and evoc_hip_run_device check the matrix on the device before any stage
    reads it; a NaN or inf anywhere is EVOC_ERR_INVALID_ARGUMENT."""
    from devbuf import DeviceArray
    x, _ = datagen.gaussian_mixture(600, 16, 4, 4.0, 1.0, seed=2)
    for bad in (np.nan, np.inf):
        y = x.copy()
        y[417, 9] = bad
        with pytest.raises(H.EvocError, match="non-finite"):
            H.evoc(y, seed=1)
        dev = DeviceArray(y)
        with pytest.raises(H.EvocError, match="non-finite"):
            H.evoc_device(dev.ptr.value, 600, 16, seed=1)
    r = H.evoc(x, seed=1)   # the library is still usable afterwards
    assert r.cluster_layers.shape[1] == 600


def test_more_than_32_neighbours_and_min_samples_end_to_end():
    """n_neighbours and min_samples between 33 and 64 (the wider register lists of the
    fuzzy set and the core distances): fuzzy set against the oracle, core distances bit for
    bit, and the whole pipeline on a shared kNN graph identical to the oracle's."""
    x, gt = datagen.gaussian_mixture(2500, 32, 10, 4.0, 1.0, seed=77)
    idx, dist = O.knn_exhaustive(x, 48)
    gp, gi, gw, gs, gr = H.fuzzy_graph(idx, dist, 48.0)
    op, oi, ow, os_, orh = O.fuzzy_graph(idx, dist, 48.0)
    assert np.array_equal(gp, op) and np.array_equal(gi, oi)
    np.testing.assert_allclose(gw, ow, rtol=1e-5, atol=0)
    e, _ = _embedding()
    for ms in (33, 40, 64):
        assert np.array_equal(H.core_sq(e, ms).view(np.uint32), O.core_sq(e, ms).view(np.uint32))
    p = H.EvocParams(n_neighbours=48, min_samples=40)
    r = H.evoc(x, params=p, precomputed_knn=(idx, dist), seed=3)
    o = O.evoc(x, O.params(n_neighbours=48, min_samples=40), seed=3, pre_idx=idx, pre_dist=dist)
    assert np.array_equal(r.embedding.view(np.uint32), o["embedding"].view(np.uint32))
    assert np.array_equal(r.cluster_layers, o["labels"])
    assert np.array_equal(r.persistence_scores, o["persistence"])
