"""GPU kNN vs the oracle's exact f32 exhaustive search: indices and distances must be
bit-identical (ties broken by index, self excluded)."""
import numpy as np
import pytest

import datagen
import evoc_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def _check(x, k, metric, precision="auto"):
    gi, gd, st = H.knn(x, k, metric=metric, precision=precision, return_stats=True)
    oi, od = O.knn_exhaustive(x, k, metric=1 if metric == "cosine" else 0)
    assert np.array_equal(gi, oi), f"index mismatch rows {np.where((gi != oi).any(1))[0][:10]}"
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), "distance bits differ"
    return st


def test_knn_gaussian_mixture_euclidean():
    x, _ = datagen.gaussian_mixture(3000, 64, 20, 5.0, 1.0, seed=42)
    st = _check(x, 15, "euclidean")
    print("fallback rows", st)


def test_knn_unit_cosine_split():
    x, _ = datagen.unit_mixture(4000, 384, 50, 0.15, seed=1)
    _check(x, 15, "cosine")


def test_knn_bf16_exact_cosine():
    x, _ = datagen.unit_mixture(4000, 512, 40, 0.15, seed=2, bf16=True)
    _check(x, 15, "cosine", precision="auto")
    _check(x, 15, "cosine", precision="bf16")


def test_knn_ragged_shapes_and_ties():
    rng = np.random.default_rng(3)
    base = rng.normal(size=(37, 70)).astype(np.float32)
    x = np.concatenate([base, base[:20], base[:5]])  # exact duplicates -> distance ties
    _check(x, 15, "euclidean")
    _check(x, 15, "cosine")
    _check(x[:9], 15, "euclidean")   # k > n - 1: tail filled with (-1, inf)
    _check(x[:1], 3, "euclidean")
    _check(x[:130], 32, "euclidean")  # k = 32 uses 64 candidates per row


def test_knn_zero_rows_cosine():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(300, 33)).astype(np.float32)
    x[::17] = 0.0
    _check(x, 10, "cosine")
