"""Pins the oracle's label-propagation initialisation and CSR algebra: the reference's own
unit tests transcribed (file:line of each), plus textbook cross-checks that a plausible
mistake fails (numpy SVD for the PCA, a dense float64 restatement of the expansion, brute
force for the votes and the outlier walk).

Reference tests: src/utils/sparse.rs:656-951 (CSR algebra),
src/graph/label_prop.rs:574-738 (label propagation), src/graph/fuzzy_graph.rs:335-454
(directed weights and the fuzzy union). The reference runs them in f64; the oracle stores
f32, so exact small-integer results are compared exactly and the rest to f32 rounding."""
import numpy as np
import pytest

import oracle as O
from oracle import Csr


def make_3x3():
    # [[1, 0, 2], [0, 3, 0], [4, 0, 5]]  (src/utils/sparse.rs:641-649)
    return Csr([0, 2, 3, 5], [0, 2, 1, 0, 2], [1.0, 2.0, 3.0, 4.0, 5.0], 3, 3)


def close(a, b, tol=1e-6):
    return abs(float(a) - float(b)) <= tol * max(1.0, abs(float(b)))


# ------------------------------------------------------------------ sparse.rs:656-951
def test_from_coo_basic():
    c = O.csr_from_coo([0, 0, 1, 2, 2], [0, 2, 1, 0, 2], [1, 2, 3, 4, 5], 3)
    assert (c.nrows, c.ncols, c.nnz()) == (3, 3, 5)
    assert c.indptr.tolist() == [0, 2, 3, 5]
    assert c.indices[:5].tolist() == [0, 2, 1, 0, 2]
    assert c.data[:5].tolist() == [1.0, 2.0, 3.0, 4.0, 5.0]


def test_from_coo_duplicates_summed():
    c = O.csr_from_coo([0, 0, 0], [1, 1, 2], [1.0, 3.0, 5.0], 2)
    assert c.indptr.tolist() == [0, 2, 2]
    assert c.indices[:2].tolist() == [1, 2]
    assert c.data[:2].tolist() == [4.0, 5.0]


def test_from_coo_unsorted_input_sorted_by_row_then_column():
    c = O.csr_from_coo([2, 0, 1, 0, 2], [2, 2, 1, 0, 0], [5, 2, 3, 1, 4], 3)
    assert c.indptr.tolist() == [0, 2, 3, 5]
    assert c.indices[:5].tolist() == [0, 2, 1, 0, 2]
    assert c.data[:5].tolist() == [1.0, 2.0, 3.0, 4.0, 5.0]


def test_from_coo_empty():
    c = O.csr_from_coo([], [], [], 5)
    assert c.nrows == 5 and c.nnz() == 0
    assert c.indptr.tolist() == [0, 0, 0, 0, 0, 0]


def test_from_partition_basic():
    c = O.csr_from_partition([2, 0, 1, 2], 3)
    assert (c.nrows, c.ncols, c.nnz()) == (4, 3, 4)
    assert c.indices.tolist() == [2, 0, 1, 2]
    assert (c.data == 1.0).all()


def test_transpose_roundtrip():
    a = make_3x3()
    at = O.csr_op("transpose", a)
    assert (at.nrows, at.ncols, at.nnz()) == (3, 3, 5)
    assert at.rows()[0] == [(0, 1.0), (2, 4.0)]
    att = O.csr_op("transpose", at)
    assert att.indptr.tolist() == a.indptr.tolist()
    assert att.indices.tolist() == a.indices.tolist()
    assert att.data.tolist() == a.data.tolist()


def test_transpose_non_square():
    m = Csr([0, 2, 3], [0, 1, 2], [1.0, 2.0, 3.0], 2, 3)
    mt = O.csr_op("transpose", m)
    assert (mt.nrows, mt.ncols) == (3, 2)
    assert mt.indptr.tolist() == [0, 1, 2, 3]
    assert mt.indices.tolist() == [0, 0, 1]
    assert mt.data.tolist() == [1.0, 2.0, 3.0]


def test_matmul_identity():
    a = make_3x3()
    eye = Csr([0, 1, 2, 3], [0, 1, 2], [1.0, 1.0, 1.0], 3, 3)
    r = O.csr_op("matmul", a, eye)
    assert r.data.tolist() == a.data.tolist()
    assert r.indices.tolist() == a.indices.tolist()


def test_matmul_a_times_at():
    a = make_3x3()
    d = O.csr_op("matmul", a, O.csr_op("transpose", a)).to_dense()
    assert d.tolist() == [[5.0, 0.0, 14.0], [0.0, 9.0, 0.0], [14.0, 0.0, 41.0]]


def test_matmul_non_square():
    a = Csr([0, 2, 3], [0, 1, 2], [1.0, 2.0, 3.0], 2, 3)
    b = Csr([0, 1, 2, 3], [0, 1, 0], [4.0, 5.0, 6.0], 3, 2)
    c = O.csr_op("matmul", a, b)
    assert (c.nrows, c.ncols) == (2, 2)
    assert c.to_dense().tolist() == [[4.0, 10.0], [18.0, 0.0]]


def test_matmul_dense_basic():
    rhs = np.array([[1, 0], [0, 1], [1, 1]], np.float32)
    r = O.csr_matmul_dense(make_3x3(), rhs)
    assert r.tolist() == [[3.0, 2.0], [0.0, 3.0], [9.0, 5.0]]


def test_elementwise_mul_with_transpose():
    a = make_3x3()
    h = O.csr_op("elementwise_mul", a, O.csr_op("transpose", a))
    d = h.to_dense()
    assert (d[0, 0], d[0, 2], d[1, 1], d[2, 0], d[2, 2]) == (1.0, 8.0, 9.0, 8.0, 25.0)
    assert h.nnz() == 5


def test_elementwise_mul_disjoint():
    a = Csr([0, 1, 1], [0], [5.0], 2, 2)
    b = Csr([0, 0, 1], [1], [7.0], 2, 2)
    assert O.csr_op("elementwise_mul", a, b).nnz() == 0


def test_normalise_cols_l2_unit_norms():
    m = Csr([0, 1, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 3.0, 4.0], 3, 2)
    r = O.csr_op("normalise_cols_l2", m)
    sq = np.zeros(2)
    for p in range(4):
        sq[r.indices[p]] += float(r.data[p]) ** 2
    assert close(np.sqrt(sq[0]), 1.0) and close(np.sqrt(sq[1]), 1.0)
    c0, c1 = np.sqrt(10.0), np.sqrt(20.0)
    for got, exp in zip(r.data[:4], [1 / c0, 2 / c1, 3 / c0, 4 / c1]):
        assert close(got, exp)


def test_normalise_cols_l2_empty_column():
    r = O.csr_op("normalise_cols_l2", Csr([0, 1, 1], [0], [3.0], 2, 2))
    assert r.data[0] == 1.0


def test_normalise_rows_l1_unit_sums():
    m = Csr([0, 1, 2, 4], [0, 1, 0, 1], [1.0, 2.0, 3.0, 4.0], 3, 2)
    r = O.csr_op("normalise_rows_l1", m)
    for i in range(3):
        assert close(np.abs(r.data[r.indptr[i]:r.indptr[i + 1]]).sum(), 1.0)
    assert r.data[0] == 1.0 and r.data[1] == 1.0
    assert close(r.data[2], 3 / 7) and close(r.data[3], 4 / 7)


def test_normalise_rows_l1_empty_row():
    r = O.csr_op("normalise_rows_l1", Csr([0, 0, 1], [0], [5.0], 2, 2))
    assert r.data[0] == 1.0


def test_clip_values_basic():
    r = O.csr_op("clip_values", Csr([0, 3], [0, 1, 2], [-1.0, 0.5, 2.0], 1, 3))
    assert r.data.tolist() == [0.0, 0.5, 1.0]


def test_to_adjacency_list_and_dense_roundtrip():
    a = make_3x3()
    assert a.rows() == [[(0, 1.0), (2, 2.0)], [(1, 3.0)], [(0, 4.0), (2, 5.0)]]
    assert a.to_dense().tolist() == [[1.0, 0.0, 2.0], [0.0, 3.0, 0.0], [4.0, 0.0, 5.0]]


def _rand_csr(rng, nr, nc, density):
    m = (rng.random((nr, nc)) < density) * rng.uniform(0.1, 2.0, (nr, nc))
    ip, ix, dv = [0], [], []
    for i in range(nr):
        for j in np.nonzero(m[i])[0]:
            ix.append(j)
            dv.append(m[i, j])
        ip.append(len(ix))
    return Csr(ip, ix, dv, nr, nc), m.astype(np.float32)


def test_csr_algebra_matches_dense_float64():
    """Every operation against its dense float64 definition on random matrices."""
    rng = np.random.default_rng(4)
    for _ in range(10):
        a, da = _rand_csr(rng, 13, 9, 0.3)
        b, db = _rand_csr(rng, 9, 11, 0.3)
        c, dc = _rand_csr(rng, 13, 9, 0.4)
        da64, db64, dc64 = da.astype(np.float64), db.astype(np.float64), dc.astype(np.float64)
        np.testing.assert_allclose(O.csr_op("transpose", a).to_dense(), da.T)
        np.testing.assert_allclose(O.csr_op("matmul", a, b).to_dense(), da64 @ db64, rtol=1e-5)
        np.testing.assert_allclose(O.csr_op("elementwise_mul", a, c).to_dense(), da64 * dc64,
                                   rtol=1e-6)
        cn = np.sqrt((da64 ** 2).sum(0))
        np.testing.assert_allclose(O.csr_op("normalise_cols_l2", a).to_dense(),
                                   da64 / np.where(cn > 0, cn, 1.0), rtol=1e-5)
        rn = np.abs(da64).sum(1, keepdims=True)
        np.testing.assert_allclose(O.csr_op("normalise_rows_l1", a).to_dense(),
                                   da64 / np.where(rn > 0, rn, 1.0), rtol=1e-5)
        rhs = rng.normal(size=(9, 4)).astype(np.float32)
        np.testing.assert_allclose(O.csr_matmul_dense(a, rhs), da64 @ rhs, rtol=1e-5, atol=1e-5)


# ------------------------------------------------------------------ label_prop.rs:574-738
def test_iteration_skips_labelled_nodes():
    g = Csr([0, 1, 2, 3], [1, 0, 1], [1.0, 1.0, 1.0], 3, 3)
    assert O.lp_iteration(g, [0, 1, 2]).tolist() == [0, 1, 2]


def test_iteration_propagates_strong_label():
    g = Csr([0, 2, 2, 2], [1, 2], [2.0, 0.3], 3, 3)
    assert O.lp_iteration(g, [-1, 5, 7])[0] == 5


def test_iteration_stays_unlabelled_below_threshold():
    g = Csr([0, 1, 1], [1], [0.5], 2, 2)
    assert O.lp_iteration(g, [-1, 3])[0] == -1


def test_iteration_ignores_unlabelled_neighbours():
    g = Csr([0, 1, 1], [1], [5.0], 2, 2)
    assert O.lp_iteration(g, [-1, -1])[0] == -1


def test_iteration_threshold_is_strict_and_last_tie_wins():
    # a vote of exactly 1.0 is adopted through the tie branch (`v == max_vote` with max_vote
    # starting at the threshold 1.0), one just below is not; of two labels with equal totals
    # above the threshold the later one wins (label_prop.rs:67-82)
    g = Csr([0, 1, 1], [1], [1.0], 2, 2)
    assert O.lp_iteration(g, [-1, 4])[0] == 4
    g = Csr([0, 1, 1], [1], [np.nextafter(np.float32(1.0), np.float32(0))], 2, 2)
    assert O.lp_iteration(g, [-1, 4])[0] == -1
    # votes in first-appearance order: 8: 1.5, 9: 0.25 + 1.25 = 1.5, 6: 1.0
    g = Csr([0, 4, 4, 4, 4, 4], [1, 2, 3, 4], [1.5, 0.25, 1.0, 1.25], 5, 5)
    assert O.lp_iteration(g, [-1, 8, 9, 6, 9])[0] == 9
    assert O.lp_iteration(g, [-1, 9, 8, 6, 8])[0] == 8
    # a later, larger total replaces an earlier tie
    g = Csr([0, 3, 3, 3, 3], [1, 2, 3], [1.5, 1.5, 2.0], 4, 4)
    assert O.lp_iteration(g, [-1, 1, 2, 3])[0] == 3


def test_iteration_matches_brute_force_votes():
    rng = np.random.default_rng(7)
    for _ in range(20):
        n = 40
        a, dense = _rand_csr(rng, n, n, 0.15)
        labels = np.where(rng.random(n) < 0.5, rng.integers(0, 6, n), -1)
        got = O.lp_iteration(a, labels)
        for i in range(n):
            if labels[i] >= 0:
                assert got[i] == labels[i]
                continue
            tot = {}
            order = []
            for j in a.indices[a.indptr[i]:a.indptr[i + 1]]:
                l = int(labels[j])
                if l not in tot:
                    order.append(l)
                    tot[l] = np.float32(0)
                tot[l] = np.float32(tot[l] + dense[i, j])
            best, mv = -1, np.float32(1.0)
            for l in order:
                if l < 0:
                    continue
                if tot[l] > mv:
                    best, mv = l, tot[l]
                elif tot[l] == mv:
                    best = l
            assert got[i] == best


def test_outliers_finds_labelled_neighbour():
    g = Csr([0, 1, 3, 4], [1, 0, 2, 1], [1.0] * 4, 3, 3)
    lab = O.lp_outliers(g, [-1, -1, 42], 99)
    assert lab[0] >= 0 and lab[1] == 42


def test_outliers_random_fallback_for_isolated():
    g = Csr([0, 0, 0], [], [], 2, 2)
    lab = O.lp_outliers(g, [-1, 5], 99)
    assert 0 <= lab[0] <= 5


def test_outliers_nothing_labelled_is_a_no_op():
    g = Csr([0, 1, 2], [1, 0], [1.0, 1.0], 2, 2)
    assert O.lp_outliers(g, [-1, -1], 3).tolist() == [-1, -1]


def test_outliers_walk_matches_stack_restatement():
    """label_outliers (label_prop.rs:105-140): nodes in index order, each walks a stack
    (last pushed first) for at most 64 pops, taking the first labelled neighbour met;
    labels assigned to earlier outliers are visible to later ones."""
    rng = np.random.default_rng(12)
    for _ in range(20):
        n = 60
        a, _ = _rand_csr(rng, n, n, 0.04)
        labels = np.where(rng.random(n) < 0.2, rng.integers(0, 9, n), -1)
        got = O.lp_outliers(a, labels, 5)
        exp = labels.copy()
        walked = np.zeros(n, bool)
        for i in range(n):
            if exp[i] >= 0:
                continue
            queue = [i]
            for _ in range(64):
                if not queue:
                    break
                node = queue.pop()
                hit = None
                for j in a.indices[a.indptr[node]:a.indptr[node + 1]]:
                    if exp[j] >= 0:
                        hit = exp[j]
                        break
                    queue.append(int(j))
                if hit is not None:
                    exp[i] = hit
                    walked[i] = True
                    break
            if not walked[i]:
                exp[i] = got[i]   # random fallback: checked for range only
                assert 0 <= got[i] <= labels.max()
        assert np.array_equal(got, exp)


def test_remap_contiguous():
    m, lab = O.lp_remap([10, 10, 20, 20, 10])
    assert m == 2 and lab.tolist() == [0, 0, 1, 1, 0]


def test_remap_handles_negatives():
    m, lab = O.lp_remap([5, -1, 5, -1])
    assert m == 3 and lab[0] == 0 and lab[2] == 0
    assert 1 <= lab[1] <= 2 and 1 <= lab[3] <= 2 and lab[1] != lab[3]


def test_loop_returns_valid_partition():
    n = 5
    rows, cols = [], []
    for i in range(n):
        j = (i + 1) % n
        rows += [i, j]
        cols += [j, i]
    g = O.csr_from_coo(rows, cols, [1.0] * len(rows), n)
    part, np_ = O.lp_loop(g, 3, 10, 42)
    assert len(part) == n and np_ > 0 and (part < np_).all() and (part >= 0).all()


def test_loop_seeds_and_partition_count():
    """A graph of isolated nodes: the seeds keep their labels, every other node falls back
    to a random label in 0..=max_label (label_prop.rs:136-138), so at most approx_n_parts
    parts remain and the remap makes them contiguous."""
    n = 50
    g = Csr(np.zeros(n + 1, np.int64), [], [], n, n)
    part, m = O.lp_loop(g, 20, 5, 3)
    assert 1 <= m <= 20 and set(part.tolist()) == set(range(m))


def test_unit_range_normalisation():
    d = O.unit_range([[0.0, 10.0], [4.0, 20.0], [8.0, 30.0]])
    assert ((d >= -1.0) & (d <= 1.0)).all()
    assert d[0, 0] == -1.0 and d[2, 0] == 1.0 and d[0, 1] == -1.0 and d[2, 1] == 1.0
    assert d[1, 0] == 0.0 and d[1, 1] == 0.0
    c = O.unit_range([[3.0, 1.0], [3.0, 2.0]])   # constant column left unchanged
    assert c[:, 0].tolist() == [3.0, 3.0]


def test_pca_init_matches_numpy_svd_up_to_sign():
    """pca_init (label_prop.rs:288-321): U S of the centred data, first k components, each
    rescaled to [-1, 1]. numpy's SVD gives the same columns up to sign."""
    rng = np.random.default_rng(3)
    for n, d, dim in ((40, 12, 4), (60, 5, 8), (30, 9, 9)):
        x = (rng.normal(size=(n, d)) * np.linspace(3, 0.5, d)).astype(np.float32)
        got = O.pca_init(x, dim)
        k = min(dim, n, d)
        xc = x.astype(np.float64) - x.astype(np.float64).mean(0)
        u, s, _ = np.linalg.svd(xc, full_matrices=False)
        us = u[:, :k] * s[:k]
        lo, hi = us.min(0), us.max(0)
        ref = (us - (hi + lo) / 2) / ((hi - lo) / 2)
        for c in range(k):
            sgn = 1.0 if np.dot(ref[:, c], got[:, c]) >= 0 else -1.0
            np.testing.assert_allclose(got[:, c], sgn * ref[:, c], atol=2e-5)
        assert (got[:, k:] == 0).all()


def test_random_init_unit_rows():
    a = O.random_init(200, 6, 9)
    np.testing.assert_allclose(np.linalg.norm(a.astype(np.float64), axis=1), 1.0, atol=1e-6)
    assert np.array_equal(a, O.random_init(200, 6, 9))
    assert not np.array_equal(a, O.random_init(200, 6, 10))
    assert (np.abs(a) <= 1.0).all() and a.std() > 0.2


def test_expansion_matches_dense_float64():
    """The expansion of label_prop_init_inner (label_prop.rs:487-516) against its dense
    float64 definition: 0.5 (rows_l1((G .* G^T) R_n) L + rows_l1(R_n) L), columns centred,
    times 0.1, with R_n the column-L2-normalised partition indicator."""
    rng = np.random.default_rng(8)
    for _ in range(5):
        n, np_, dim = 50, 7, 4
        g, dg = _rand_csr(rng, n, n, 0.12)
        part = rng.integers(0, np_, n)
        part[:np_] = np.arange(np_)
        lay = rng.normal(size=(np_, dim)).astype(np.float32)
        got = O.lp_expand(g, part, np_, lay, 0.1)
        G = dg.astype(np.float64)
        R = np.zeros((n, np_))
        R[np.arange(n), part] = 1.0
        Rn = R / np.sqrt(R.sum(0))
        E = (G * G.T) @ Rn
        E = E / np.where(E.sum(1, keepdims=True) > 0, E.sum(1, keepdims=True), 1.0)
        D = Rn / Rn.sum(1, keepdims=True)
        res = 0.5 * (E @ lay + D @ lay)
        res = 0.1 * (res - res.mean(0))
        np.testing.assert_allclose(got, res, atol=2e-6)


def test_init_small_graph_random_base():
    # label_prop.rs:705-737: a 10-node complete graph is below the base threshold
    n = 10
    rows, cols = [], []
    for i in range(n):
        for j in range(i + 1, n):
            rows += [i, j]
            cols += [j, i]
    g = O.csr_from_coo(rows, cols, [0.5] * len(rows), n)
    out = O.label_prop_init(g.indptr, g.indices[:g.nnz()].astype(np.int32), g.data[:g.nnz()], 4,
                            seed=42)
    assert out.shape == (n, 4) and np.isfinite(out).all()
    np.testing.assert_allclose(out, O.random_init(n, 4, 42))


# ------------------------------------------------------------------ fuzzy_graph.rs:335-454
def test_symmetrise_graph_fuzzy_union():
    ip, ind, w = O.symmetrise([0, 1], [1, 0], [0.8, 0.6], 2)
    assert len(w) == 2
    exp = 0.8 + 0.6 - 0.8 * 0.6          # 0.92 both ways
    assert all(abs(v - exp) < 1e-6 for v in w)


def test_symmetrise_graph_one_direction():
    ip, ind, w = O.symmetrise([0], [1], [0.7], 2)
    assert len(w) == 2 and all(abs(v - 0.7) < 1e-6 for v in w)
    assert ind.tolist() == [1, 0]


def test_symmetrise_drops_zero_and_keeps_last_duplicate():
    ip, ind, w = O.symmetrise([0, 0, 2, 1], [1, 1, 1, 2], [0.3, 0.5, 0.0, 0.0], 3)
    # (0,1) repeated: the last weight 0.5 stands; (2,1) and (1,2) both 0 -> dropped
    assert ip.tolist() == [0, 1, 2, 2] and ind.tolist() == [1, 0]
    assert np.allclose(w, [0.5, 0.5])


def test_knn_to_coo_explicit_weights():
    # knn_to_coo (fuzzy_graph.rs:114-159) with sigma and rho from smooth_knn_dist:
    # w = exp(-max(0, d - rho) / sigma); the first neighbour always weighs exactly 1
    idx = np.array([[1, 2], [0, 2], [0, 1]], np.int32)
    dist = np.array([[1.0, 2.0], [1.0, 1.5], [2.0, 1.5]], np.float32)
    ip, ind, w, sig, rho = O.fuzzy_graph(idx, dist, 2, symmetrise=False)
    assert rho.tolist() == [1.0, 1.0, 2.0]
    assert ip.tolist() == [0, 2, 4, 6] and ind.tolist() == [1, 2, 0, 2, 0, 1]
    for i in range(3):
        assert w[2 * i] == 1.0
        assert abs(w[2 * i + 1] - np.exp(-max(dist[i, 1] - rho[i], 0) / sig[i])) < 1e-6
    # log2(2) = 1 is the target of the row sums: 1 + exp(-x / sigma) = 1 needs sigma -> 0;
    # the bisection stops once the sum is within 1e-5 of it, so the second weight of rows
    # 0 and 1 is below 1e-5; row 2's second distance is below rho, so its weight is 1
    assert w[1] < 1e-5 and w[3] < 1e-5 and w[5] == 1.0


def test_coo_to_adjacency_list_unsymmetrised_order():
    # fuzzy_graph.rs:418-454: rows keep the kNN order when not symmetrised
    idx = np.array([[2, 1], [2, 0], [0, 1]], np.int32)
    dist = np.array([[0.5, 0.9], [0.2, 0.3], [0.4, 0.6]], np.float32)
    ip, ind, w, _, _ = O.fuzzy_graph(idx, dist, 2, symmetrise=False)
    assert ind.tolist() == [2, 1, 2, 0, 0, 1]
