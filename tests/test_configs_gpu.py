"""The pipeline at the benchmark configurations' stated sizes.

- 10,000 x 64 Gaussian mixture (20 components, centres N(0, 5^2 I), sd 1), euclidean:
  the whole GPU pipeline against the whole CPU oracle, bit for bit.
- 100,000 x 384 unit-norm mixture (50 centres, sd 0.15, not bf16-representable), cosine:
  sampled kNN rows against the oracle's exact search, then the rest of the pipeline on a
  shared kNN graph against the oracle bit for bit (embedding, labels, strengths,
  persistence), as src/lib.rs:199-352 runs it.
- 1,000,000 x 512 bf16-valued unit-norm mixture (200 Zipf(1.1) clusters, 5% noise),
  cosine: sampled kNN rows against the oracle (the symmetric sweep runs at this size), the
  full fuzzy CSR against the oracle's on the GPU's kNN graph, and properties of the whole
  run (finite layout, layer structure, recovery of the planted clusters).
- 4,000,000 x 768 hierarchical mixture (100 parents ~ N(0, I), 10 children each at 0.3 x
  the parent spread, leaf spread 0.1, L2-normalised), cosine, approx_n_clusters = 500,
  learned warm start: sampled kNN rows against the oracle, the full fuzzy CSR against the
  oracle's, and properties of the whole run. It takes minutes, so it runs only with
  EVOC_HIP_FULL_SIZE=1.
"""
import os
import time

import numpy as np
import pytest
from sklearn.metrics import adjusted_rand_score as ari

import datagen
import evoc_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def _best(pers):
    return max(i for i in range(len(pers)) if pers[i] == pers.max()) if len(pers) > 1 else 0


def _sample_rows(n, m, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=m, replace=False))


def test_config_10k_x64_whole_pipeline_bit_exact():
    x, gt = datagen.gaussian_mixture(10_000, 64, 20, 5.0, 1.0, seed=42)
    r = H.evoc(x, params=H.EvocParams(embedding_dim=8), seed=42)
    o = O.evoc(x, O.params(embedding_dim=8), seed=42)
    assert np.array_equal(r.nn_indices, o["knn_idx"])
    assert np.array_equal(r.nn_distances.view(np.uint32), o["knn_dist"].view(np.uint32))
    assert np.array_equal(r.embedding.view(np.uint32), o["embedding"].view(np.uint32))
    assert np.array_equal(r.cluster_layers, o["labels"])
    assert np.array_equal(r.membership_strengths.view(np.uint32), o["strengths"].view(np.uint32))
    assert np.array_equal(r.persistence_scores.view(np.uint64), o["persistence"].view(np.uint64))
    assert r.best_layer == _best(o["persistence"])
    score = ari(r.best_labels(), gt)
    print(f"10k x 64: {len(r.cluster_layers)} layers, best {r.n_clusters()} clusters, "
          f"ARI vs planted {score:.4f}, stages {r.stage_ms}")


def test_config_100k_x384_cosine():
    x, gt = datagen.unit_mixture(100_000, 384, 50, 0.15, seed=42)
    t0 = time.perf_counter()
    idx, dist, st = H.knn(x, 15, metric="cosine", return_stats=True)
    t_knn = time.perf_counter() - t0
    rows = _sample_rows(len(x), 1024, 7)
    oi, od = O.knn_rows(x, rows, 15, metric=1)
    assert np.array_equal(idx[rows], oi)
    assert np.array_equal(dist[rows].view(np.uint32), od.view(np.uint32))
    nn = H.NearestNeighbourParams(dist_metric="cosine")
    p = H.EvocParams(embedding_dim=8)
    full = H.evoc(x, params=p, nn_params=nn, seed=42)
    assert np.array_equal(full.nn_indices, idx)
    shared = H.evoc(x, params=p, nn_params=nn, seed=42, precomputed_knn=(idx, dist))
    assert np.array_equal(shared.embedding.view(np.uint32), full.embedding.view(np.uint32))
    t0 = time.perf_counter()
    o = O.evoc(x, O.params(embedding_dim=8, metric=1), seed=42, pre_idx=idx, pre_dist=dist)
    t_cpu = time.perf_counter() - t0
    assert np.array_equal(shared.embedding.view(np.uint32), o["embedding"].view(np.uint32))
    assert np.array_equal(shared.cluster_layers, o["labels"])
    assert np.array_equal(shared.membership_strengths.view(np.uint32),
                          o["strengths"].view(np.uint32))
    assert np.array_equal(shared.persistence_scores.view(np.uint64),
                          o["persistence"].view(np.uint64))
    assert shared.best_layer == _best(o["persistence"])
    print(f"100k x 384: kNN {t_knn:.2f} s (fallback rows {st['fallback_rows']}), "
          f"{len(full.cluster_layers)} layers, best {full.n_clusters()} clusters, ARI vs planted "
          f"{ari(full.best_labels(), gt):.4f}; oracle on the shared kNN {t_cpu:.1f} s")


def test_config_1M_x512_bf16_cosine():
    import torch
    xd, gtd = datagen.unit_mixture_device(1_000_000, 512, seed=42)
    x = xd.cpu().numpy()
    gt = gtd.cpu().numpy()
    del xd, gtd
    torch.cuda.empty_cache()
    nn = H.NearestNeighbourParams(dist_metric="cosine")
    t0 = time.perf_counter()
    r = H.evoc(x, params=H.EvocParams(embedding_dim=8), nn_params=nn, seed=42)
    wall = time.perf_counter() - t0
    st = H.knn_last_stats()
    assert st["symmetric"]
    n = len(x)
    # kNN: 1024 sampled rows bit-exact against the oracle's exact f32 search
    rows = _sample_rows(n, 1024, 11)
    oi, od = O.knn_rows(x, rows, 15, metric=1)
    assert np.array_equal(r.nn_indices[rows], oi), \
        f"rows differ: {rows[(r.nn_indices[rows] != oi).any(1)][:10]}"
    assert np.array_equal(r.nn_distances[rows].view(np.uint32), od.view(np.uint32))
    # every row: sorted by (distance, index), self excluded, no repeats
    nd = r.nn_distances
    ni = r.nn_indices
    assert (ni >= 0).all() and (ni < n).all()
    assert not (ni == np.arange(n)[:, None]).any()
    assert ((nd[:, 1:] > nd[:, :-1]) | ((nd[:, 1:] == nd[:, :-1]) & (ni[:, 1:] > ni[:, :-1]))).all()
    # the full fuzzy CSR on the GPU's graph against the oracle's
    gp, gi, gw, _, _ = H.fuzzy_graph(ni, nd, 15.0)
    op, oi2, ow, _, _ = O.fuzzy_graph(ni, nd, 15.0)
    assert np.array_equal(gp, op) and np.array_equal(gi, oi2)
    np.testing.assert_allclose(gw, ow, rtol=1e-5, atol=0)
    # the rest of the run
    assert np.isfinite(r.embedding).all() and r.embedding.shape == (n, 8)
    nl = len(r.cluster_layers)
    assert 1 <= nl <= 10
    counts = [int(l.max()) + 1 for l in r.cluster_layers]
    assert counts == sorted(counts, reverse=True)
    for lab, s in zip(r.cluster_layers, r.membership_strengths):
        assert lab.min() >= -1
        assert ((s[lab >= 0] >= 0) & (s[lab >= 0] <= 1)).all()
    # Recovery of the planted clusters is a property of the data, not a parity bar: with
    # sd 0.15 in 512 dimensions the noise vector (norm ~ 0.15 * sqrt(512) = 3.4) outweighs
    # the unit centre, so the small Zipf clusters overlap and merge. A sanity floor only.
    clustered = gt >= 0
    scores = [ari(l[clustered], gt[clustered]) for l in r.cluster_layers]
    score = scores[r.best_layer]
    print(f"1M x 512: wall {wall:.2f} s with upload, stages {r.stage_ms}, fallback rows "
          f"{r.knn_fallback_rows}, layers {counts}, ARI vs planted (noise excluded) per layer "
          f"{[round(v, 4) for v in scores]}, best layer {score:.4f}")
    assert r.knn_fallback_rows >= 0
    assert max(scores) > 0.25


@pytest.mark.skipif(os.environ.get("EVOC_HIP_FULL_SIZE") != "1",
                    reason="4,000,000 x 768 takes minutes; set EVOC_HIP_FULL_SIZE=1")
def test_config_4M_x768_hierarchical_approx_500():
    import torch
    n, d = 4_000_000, 768
    xd, ld = datagen.unit_mixture_device(n, d, seed=5, hierarchical=(100, 10, 0.3, 0.1))
    leaves = ld.cpu().numpy()
    # EVOC_HIP_KNN_PRECISION=fp8 runs the FP8 first pass instead (auto by default)
    prec = os.environ.get("EVOC_HIP_KNN_PRECISION", "auto")
    nn = H.NearestNeighbourParams(dist_metric="cosine", knn_precision=prec, warm_start=True)
    p = H.EvocParams(embedding_dim=8, approx_n_clusters=500)
    t0 = time.perf_counter()
    r = H.evoc_device(xd.data_ptr(), n, d, params=p, nn_params=nn, seed=42, verbose=1)
    wall = time.perf_counter() - t0
    st = H.knn_last_stats()
    print(f"4M x 768: wall {wall:.1f} s, stages {r.stage_ms}, kNN tile kernel "
          f"{st['kernel_ms']:.0f} ms ({st['kernel_flop'] / st['kernel_ms'] / 1e9:.0f} TFLOP/s), "
          f"escalated rows {st['escalated_rows']}, exact fallback rows {r.knn_fallback_rows}, "
          f"warm start used {r.warm_start_used} (judge ARI {r.warm_start_ari:.4f}, "
          f"MST ratio {r.warm_start_mst_ratio:.4f})", flush=True)
    x = xd.cpu().numpy()
    del xd, ld
    torch.cuda.empty_cache()
    rows = _sample_rows(n, 256, 5)
    oi, od = O.knn_rows(x, rows, 15, metric=1)
    assert np.array_equal(r.nn_indices[rows], oi), \
        f"rows differ: {rows[(r.nn_indices[rows] != oi).any(1)][:10]}"
    assert np.array_equal(r.nn_distances[rows].view(np.uint32), od.view(np.uint32))
    nd, ni = r.nn_distances, r.nn_indices
    assert (ni >= 0).all() and (ni < n).all()
    assert not (ni == np.arange(n)[:, None]).any()
    assert ((nd[:, 1:] > nd[:, :-1]) | ((nd[:, 1:] == nd[:, :-1]) & (ni[:, 1:] > ni[:, :-1]))).all()
    del x
    gp, gi, gw, _, _ = H.fuzzy_graph(ni, nd, 15.0)
    op, oi2, ow, _, _ = O.fuzzy_graph(ni, nd, 15.0)
    assert np.array_equal(gp, op) and np.array_equal(gi, oi2)
    np.testing.assert_allclose(gw, ow, rtol=1e-5, atol=0)
    del gp, gi, gw, op, oi2, ow
    assert np.isfinite(r.embedding).all() and r.embedding.shape == (n, 8)
    assert len(r.cluster_layers) == 1
    lab = r.cluster_layers[0]
    k = int(lab.max()) + 1
    parents = leaves // 10
    print(f"approx_n_clusters=500: {k} clusters, noise fraction {(lab < 0).mean():.4f}, "
          f"ARI vs the 1,000 leaves {ari(lab, leaves):.4f}, vs the 100 parents "
          f"{ari(lab, parents):.4f}", flush=True)
    assert k >= 1 and lab.min() >= -1
