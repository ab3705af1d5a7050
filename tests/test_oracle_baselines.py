"""The oracle's reference-complexity algorithms, used to time the CPU baseline, against the
parity references: the kd-tree Boruvka MST must be the dense Prim MST bit for bit (the MST
is unique under the canonical (w^2, min, max) order), kd-tree core distances must equal
brute force, and NN-Descent must find nearly all exact neighbours, with the exact search's
distance bits for the ones it finds."""
import numpy as np
import pytest

import datagen
import oracle as O


@pytest.mark.parametrize("n,dim,ms,seed", [(1500, 8, 5, 1), (3000, 4, 1, 2), (800, 16, 12, 3),
                                           (2, 8, 5, 4), (65, 3, 5, 5)])
def test_boruvka_equals_prim(n, dim, ms, seed):
    e, _ = datagen.gaussian_mixture(n, dim, 6, 3.0, 1.0, seed=seed)
    pu, pv, pw = O.mst(e, ms)
    bu, bv, bw = O.mst_boruvka(e, ms)
    assert np.array_equal(pu, bu) and np.array_equal(pv, bv)
    assert np.array_equal(pw.view(np.uint32), bw.view(np.uint32))
    np.testing.assert_array_equal(O.core_sq_kd(e, ms), O.core_sq(e, ms))


def test_boruvka_equals_prim_with_ties_and_duplicates():
    rng = np.random.default_rng(9)
    e = (np.round(rng.normal(size=(1200, 4)) * 2) / 2).astype(np.float32)
    e[100:160] = e[7]
    for ms in (1, 3, 7):
        pu, pv, pw = O.mst(e, ms)
        bu, bv, bw = O.mst_boruvka(e, ms)
        assert np.array_equal(pu, bu) and np.array_equal(pv, bv)
        assert np.array_equal(pw.view(np.uint32), bw.view(np.uint32))


def test_layers_identical_with_either_mst():
    e, _ = datagen.gaussian_mixture(2500, 8, 12, 3.0, 1.0, seed=6)
    ref = O.cluster_layers(e, 5, 5, 0.2, 10)
    old = O.set_mst_algo(1)
    try:
        fast = O.cluster_layers(e, 5, 5, 0.2, 10)
    finally:
        O.set_mst_algo(old)
    for a, b in zip(ref, fast):
        assert np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("metric", [0, 1])
def test_nndescent_recall_and_exact_distances(metric):
    # clustered data of moderate intrinsic dimension, where NN-Descent is near-exact (on
    # nearly uniform high-dimensional data its recall drops, as is known for the method)
    x, _ = datagen.gaussian_mixture(8000, 64, 30, 5.0, 1.0, seed=4)
    ei, ed = O.knn_exhaustive(x, 15, metric=metric)
    ni, nd, iters = O.knn_nndescent(x, 15, metric=metric, seed=3)
    assert 1 <= iters <= 12
    hits = sum(len(set(a) & set(b)) for a, b in zip(ni.tolist(), ei.tolist()))
    recall = hits / ei.size
    print(f"NN-Descent recall {recall:.4f} after {iters} iterations")
    assert recall > 0.98
    # a neighbour found at rank r has the exact search's distance for that pair
    same = ni == ei
    assert np.array_equal(nd[same].view(np.uint32), ed[same].view(np.uint32))
    assert (np.diff(nd, axis=1) >= 0).all()
    assert not (ni == np.arange(len(x))[:, None]).any()
