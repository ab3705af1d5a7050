"""Error of what the GPU pipeline stores against fp64 restatements of the same steps:
kNN distances, fuzzy weights (with sigma and rho), and the persistence scores
(stabilities) that select the layers. Each bound is written in its test; the figures are
printed.

- Distances: the library's f32 canonical distances of the returned pairs against the same
  pairs in fp64; the returned lists against the fp64 ranking, where a difference is allowed
  only between neighbours whose fp64 distances lie within the f32 error of each other.
- Fuzzy set: the GPU's f32 set from the GPU's kNN against the oracle's fp64 instantiation
  (oracle/evoc_oracle.cpp fuzzy_graph_impl<double>, the steps of
  src/graph/fuzzy_graph.rs:33-295) fed fp64 distances of the same pairs.
- Stabilities: the GPU's layers (fp64 sums of f32 lambdas) and the reference-order f32 sums
  (src/clustering/persistence.rs:135-173, the oracle's f32 mode) against the oracle's all-fp64
  clustering of the same embedding (fp64 core distances, MST and lambdas).
"""
import numpy as np
import pytest
from sklearn.metrics import adjusted_rand_score as ari

import datagen
import evoc_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def _fp64_dist(x, idx, metric):
    x64 = x.astype(np.float64)
    n, k = idx.shape
    rows = np.repeat(np.arange(n), k)
    cols = idx.ravel().astype(np.int64)
    if metric == "cosine":
        nrm = np.linalg.norm(x64, axis=1)
        dot = np.einsum("ij,ij->i", x64[rows], x64[cols])
        d = 1.0 - dot / (nrm[rows] * nrm[cols])
    else:
        d = np.linalg.norm(x64[rows] - x64[cols], axis=1)
    return d.reshape(n, k)


def _ranking_check(x, idx, metric, rows, tol):
    """Rows whose list differs from the fp64 top k (ties by index), and whether every
    difference is a near tie: the swapped neighbours lie within tol of the k-th fp64
    distance."""
    x64 = x.astype(np.float64)
    k = idx.shape[1]
    differ, near_ties = 0, True
    for i in rows:
        if metric == "cosine":
            nrm = np.linalg.norm(x64, axis=1)
            d = 1.0 - (x64 @ x64[i]) / (nrm * nrm[i])
        else:
            d = np.linalg.norm(x64 - x64[i], axis=1)
        d[i] = np.inf
        order = np.lexsort((np.arange(len(d)), d))[:k]
        if not np.array_equal(np.sort(order), np.sort(idx[i])):
            differ += 1
            dk = d[order[-1]]
            swapped = np.setxor1d(order, idx[i])
            near_ties &= bool(np.all(np.abs(d[swapped] - dk) <= tol))
    return differ, near_ties


@pytest.fixture(scope="module")
def cosine_case():
    # the 100,000 x 384 cosine workload's kind of data (unit mixture, f32), at 30,000 rows
    x, lab = datagen.unit_mixture(30000, 384, 50, 0.15, seed=21)
    idx, dist = H.knn(x, 15, metric="cosine")
    return x, lab, idx, dist


def test_knn_distance_error_vs_fp64(cosine_case):
    x, _, idx, dist = cosine_case
    d64 = _fp64_dist(x, idx, "cosine")
    abs_err = np.abs(dist.astype(np.float64) - d64).max()
    print(f"cosine 30000x384: max |d_f32 - d_fp64| = {abs_err:.3e}")
    assert abs_err <= 4e-7
    rows = np.random.default_rng(1).choice(len(x), 400, replace=False)
    differ, near = _ranking_check(x, idx, "cosine", rows, 2 * 4e-7)
    print(f"cosine: {differ} of 400 sampled rows differ from the fp64 ranking (near ties: {near})")
    assert near

    y, _ = datagen.gaussian_mixture(20000, 64, 20, 5.0, 1.0, seed=22)
    gi, gd = H.knn(y, 15, metric="euclidean")
    e64 = _fp64_dist(y, gi, "euclidean")
    rel = (np.abs(gd.astype(np.float64) - e64) / e64).max()
    print(f"euclidean 20000x64: max relative error {rel:.3e}")
    assert rel <= 1e-6
    differ, near = _ranking_check(y, gi, "euclidean", rows[rows < len(y)], 2e-6 * e64.max())
    print(f"euclidean: {differ} sampled rows differ from the fp64 ranking (near ties: {near})")
    assert near


def test_fuzzy_weight_error_vs_fp64(cosine_case):
    x, _, idx, dist = cosine_case
    d64 = _fp64_dist(x, idx, "cosine")
    gp, gi, gw, gs, gr = H.fuzzy_graph(idx, dist)
    op, oi, ow, os_, orh = O.fuzzy_graph(idx, d64, fp64=True)
    rho_rel = np.max(np.abs(gr - orh) / np.maximum(orh, 1e-30))
    sig_rel = np.max(np.abs(gs - os_) / os_)
    print(f"rho max rel {rho_rel:.3e}, sigma max rel {sig_rel:.3e}")
    assert np.array_equal(gp, op) and np.array_equal(gi, oi)
    w_abs = np.abs(gw.astype(np.float64) - ow).max()
    w_rel = (np.abs(gw.astype(np.float64) - ow) / ow).max()
    print(f"fuzzy weights: max abs error {w_abs:.3e}, max rel error {w_rel:.3e} over {len(gw)}")
    assert w_abs <= 1e-4 and w_rel <= 1e-3


def test_stability_error_vs_fp64(cosine_case):
    x, lab, _, _ = cosine_case
    r = H.evoc(x, H.EvocParams(embedding_dim=8), H.NearestNeighbourParams("cosine"), seed=42)
    e = np.asarray(r.embedding, np.float32)
    gl, gs, gp = H.cluster_layers(e, 5, 5, 0.2, 10)
    old = O.set_mst_algo(1)
    try:
        fl, fs, fp = O.cluster_layers(e.astype(np.float64), 5, 5, 0.2, 10)
        rl, rs, rp = O.cluster_layers(e, 5, 5, 0.2, 10, fp64=False)
    finally:
        O.set_mst_algo(old)
    print(f"layers: GPU {len(gp)}, reference f32 {len(rp)}, fp64 {len(fp)}")
    print("persistence GPU   ", gp)
    print("persistence f32   ", rp)
    print("persistence fp64  ", fp)
    assert len(gp) == len(fp) == len(rp)
    nz = fp != 0
    assert np.all(gp[~nz] == 0) and np.all(rp[~nz] == 0)
    g_rel = np.max(np.abs(gp - fp)[nz] / np.abs(fp[nz]))
    r_rel = np.max(np.abs(rp - fp)[nz] / np.abs(fp[nz]))
    agree = [ari(a, b) for a, b in zip(gl, fl)]
    print(f"persistence max rel error: GPU {g_rel:.3e}, reference f32 order {r_rel:.3e}; "
          f"ARI of GPU vs fp64 labels per layer {np.round(agree, 5)}")
    assert g_rel <= 1e-5
    assert min(agree) >= 0.999
