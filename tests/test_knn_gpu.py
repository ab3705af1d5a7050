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


@pytest.mark.parametrize("n,d,metric", [
    (1000, 96, "euclidean"),     # d <= 128 tier, ragged n (not a multiple of 256)
    (2049, 200, "cosine"),       # d <= 256 tier, one row past a 256-row block
    (777, 512, "euclidean"),     # d == 512 tier
    (300, 13, "cosine"),         # tiny d padded to 128
])
def test_knn_resident_query_tiers(n, d, metric):
    """bf16-exact inputs with k <= 16 and d <= 512 run the resident-query tile kernel."""
    x, _ = datagen.gaussian_mixture(n, d, 12, 4.0, 1.0, seed=n + d)
    x = datagen.to_bf16_values(x)
    _check(x, 15, metric)
    _check(x, 7, metric, precision="bf16")


def test_knn_resident_query_ties_zero_rows_tiny():
    rng = np.random.default_rng(11)
    base = datagen.to_bf16_values(rng.normal(size=(180, 40)).astype(np.float32))
    x = np.concatenate([base, base[:60], base[:20]])  # exact duplicates -> distance ties
    _check(x, 15, "euclidean")
    _check(x, 16, "cosine")
    z = base.copy()
    z[::9] = 0.0
    _check(z, 10, "cosine")
    _check(base[:6], 15, "euclidean")   # k > n - 1
    _check(base[:1], 2, "cosine")


def test_knn_forced_bf16_on_inexact_input():
    """precision='bf16' on float data: the kernel rounds to bf16 and the certificate
    carries the input-rounding term; the exact re-rank keeps results bit-identical."""
    x, _ = datagen.unit_mixture(3000, 256, 30, 0.2, seed=5)
    _check(x, 15, "cosine", precision="bf16")
    _check(x, 12, "euclidean", precision="bf16")


@pytest.mark.parametrize("bf16,d", [(True, 300), (False, 64)])
def test_knn_exact_fallback_many_rows(bf16, d):
    """40 copies of one point: each copy has 39 neighbours at distance zero, more than the
    candidate lists hold, so its certificate fails and the exact scan (several batches of
    failing rows) decides it."""
    rng = np.random.default_rng(21)
    x = rng.normal(size=(1500, d)).astype(np.float32)
    x[100:140] = x[7]
    if bf16:
        x = datagen.to_bf16_values(x)
    for metric in ("euclidean", "cosine"):
        st = _check(x, 15, metric)
        assert st["fallback_rows"] >= 40, st


@pytest.mark.parametrize("n,d,metric", [
    (3000, 300, "cosine"),      # 12 row blocks: even count, the middle pair taken once
    (2800, 512, "euclidean"),   # 11 row blocks: odd count
    (777, 400, "cosine"),       # 4 blocks, ragged last block
    (513, 300, "euclidean"),    # 3 blocks, one row in the last
    (250, 300, "cosine"),       # one block: the diagonal only
])
def test_knn_symmetric_sweep(n, d, metric, monkeypatch):
    """The symmetric sweep (each pair of 256-row blocks once, transposed keys queued for
    the other block's rows, merged and certified) gives the oracle's exact kNN."""
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "1")
    x, _ = datagen.gaussian_mixture(n, d, 12, 4.0, 1.0, seed=n + d)
    x = datagen.to_bf16_values(x)
    st = _check(x, 15, metric)
    assert st["symmetric"], st
    _check(x, 7, metric, precision="bf16")


def test_knn_symmetric_sweep_ties_zero_rows(monkeypatch):
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "1")
    rng = np.random.default_rng(23)
    x = datagen.to_bf16_values(rng.normal(size=(1700, 320)).astype(np.float32))
    x[900:960] = x[5]           # 61 copies: certificates fail, exact rows decide
    x[1200:1210] = 0.0          # zero rows: cosine distance 1 to everything
    for metric in ("cosine", "euclidean"):
        st = _check(x, 15, metric)
        assert st["symmetric"] and st["fallback_rows"] >= 60, st


def test_knn_symmetric_equals_plain_at_300k(monkeypatch):
    """At a size the oracle cannot check quickly: the symmetric sweep and the plain sweep
    (each bit-exact against the oracle at the test sizes) return identical graphs."""
    x, _ = datagen.unit_mixture(300_000, 320, 60, 0.2, seed=8, bf16=True)
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "0")
    pi, pd, ps = H.knn(x, 15, metric="cosine", return_stats=True)
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "1")
    si, sd, ss = H.knn(x, 15, metric="cosine", return_stats=True)
    assert not ps["symmetric"] and ss["symmetric"]
    assert ss["kernel_flop"] < 0.51 * ps["kernel_flop"]
    assert np.array_equal(pi, si)
    assert np.array_equal(pd.view(np.uint32), sd.view(np.uint32))


def test_knn_escalation_of_concentrated_rows():
    """Leaves of a hierarchical mixture whose neighbour distances concentrate: the 32
    candidates of the first pass fail many certificates, those rows are searched again over
    four column shards (32 candidates per column half each), every candidate is re-ranked,
    and only the rest take the exact scan. Bit-exact either way; d = 768 takes the
    streaming kernel, d = 512 the resident-query one."""
    rng = np.random.default_rng(31)
    for d in (768, 512):
        par = rng.normal(size=(6, d))
        child = np.repeat(par, 5, 0) + 0.3 * rng.normal(size=(30, d))
        lab = np.arange(12000) % 30
        x = child[lab] + 0.1 * rng.normal(size=(12000, d))
        x = datagen.to_bf16_values((x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32))
        st = _check(x, 15, "cosine")
        esc = H.knn_last_stats()["escalated_rows"]
        print(d, "escalated", esc, "exact", st["fallback_rows"])
        assert esc >= 1024 and st["fallback_rows"] < esc


@pytest.mark.parametrize("n,d,metric", [(300, 300, "cosine"), (1500, 768, "cosine"),
                                        (2049, 64, "euclidean"), (700, 520, "euclidean")])
def test_knn_forced_escalation_small_inputs(n, d, metric, monkeypatch):
    """EVOC_HIP_KNN_ESCALATE=1 sends every failing row through the sharded search, also at
    sizes where a shard holds one column tile or none (300 rows: 3 tiles of 128 columns
    for 4 shards)."""
    monkeypatch.setenv("EVOC_HIP_KNN_ESCALATE", "1")
    # one tight leaf: the 15th and 33rd keys of a row lie closer than the certificate's
    # error term, so first-pass certificates fail
    rng = np.random.default_rng(n + d)
    x = rng.normal(size=d) + 0.02 * rng.normal(size=(n, d))
    if metric == "cosine":
        x = x / np.linalg.norm(x, axis=1, keepdims=True)
    x = datagen.to_bf16_values(x.astype(np.float32))
    st = _check(x, 15, metric)
    esc = H.knn_last_stats()["escalated_rows"]
    print(n, d, metric, "escalated", esc, "exact", st["fallback_rows"])
    assert esc > 0


def _hierarchical_rows(n, d, seed):
    import torch
    x, _ = datagen.unit_mixture_device(n, d, seed=seed, hierarchical=(10, 10, 0.3, 0.1))
    xh = x.cpu().numpy()
    del x
    torch.cuda.empty_cache()
    return xh


@pytest.mark.parametrize("sample", ["1", "0"])
def test_knn_sampled_certificate_rate_skips_first_pass(sample, monkeypatch):
    # 300,000 x 768 with 100 leaves of 3,000 points: neighbour distances inside a leaf are
    # concentrated below the first pass's error term, so the 4,096-row sample fails and
    # every row goes straight to the sharded escalation; with the sample test off the
    # first pass runs and the same rows are escalated after it. Both equal the exact search.
    monkeypatch.setenv("EVOC_HIP_KNN_SAMPLE", sample)
    n, d = 300_000, 768
    x = _hierarchical_rows(n, d, 7)
    gi, gd, st = H.knn(x, 15, metric="cosine", return_stats=True)
    rows = np.sort(np.random.default_rng(7).choice(n, size=512, replace=False))
    oi, od = O.knn_rows(x, rows, 15, metric=1)
    assert np.array_equal(gi[rows], oi)
    assert np.array_equal(gd[rows].view(np.uint32), od.view(np.uint32))
    print(f"sample test {'on' if sample == '1' else 'off'}: escalated {st['escalated_rows']}, "
          f"exact rows {st['fallback_rows']}, tile kernel {st['kernel_ms']:.0f} ms")
    if sample == "1":
        assert st["escalated_rows"] == n


@pytest.mark.parametrize("n,d,metric,escalate", [(2000, 1536, "cosine", False),
                                                 (1200, 3072, "euclidean", False),
                                                 (1500, 1536, "cosine", True)])
def test_knn_wide_rows(n, d, metric, escalate, monkeypatch):
    """Rows wider than 1024 (1536: OpenAI-size sentence embeddings, 3072) take the re-rank
    and exact kernels that hold 64 values per lane; f32 input (the split first pass), a
    tight leaf to force escalation, and 30 duplicates for the exact scan."""
    if escalate:
        monkeypatch.setenv("EVOC_HIP_KNN_ESCALATE", "1")
        rng = np.random.default_rng(d)
        x = (rng.normal(size=d) + 0.02 * rng.normal(size=(n, d))).astype(np.float32)
    else:
        x, _ = datagen.unit_mixture(n, d, 25, 0.15, seed=d)
        x = x.astype(np.float32)
        x[50:80] = x[3]  # 29 neighbours at distance zero: more than a first pass certifies
    st = _check(x, 15, metric)
    print(n, d, metric, st)
    if not escalate:
        assert st["fallback_rows"] >= 30


def test_knn_rejects_rows_wider_than_4096():
    x = np.zeros((10, 4097), np.float32)
    with pytest.raises(H.EvocError, match="4096"):
        H.knn(x, 3)


@pytest.mark.parametrize("n,d,k,metric", [(3000, 64, 40, "euclidean"), (2500, 384, 64, "cosine"),
                                          (1800, 1536, 48, "cosine")])
def test_knn_more_than_32_neighbours(n, d, k, metric):
    """33 to 64 neighbours: the first pass keeps 64 candidates, rows it cannot certify go to
    the sharded escalation (256 candidates) or the exact scan with 64-entry lists; 70
    duplicates make some rows need the exact scan."""
    x, _ = datagen.unit_mixture(n, d, 20, 0.2, seed=k)
    x = x.astype(np.float32)
    x[10:80] = x[1]
    st = _check(x, k, metric)
    print(n, d, k, metric, st)
    assert st["fallback_rows"] > 0  # the exact scan with 64-entry lists ran


def test_knn_rejects_more_than_64_neighbours():
    x = np.random.default_rng(0).normal(size=(200, 8)).astype(np.float32)
    with pytest.raises(H.EvocError, match="64"):
        H.knn(x, 65)


@pytest.mark.parametrize("n,d,metric", [(2999, 200, "euclidean"), (1025, 300, "cosine"),
                                        (4097, 512, "euclidean"), (129, 129, "cosine")])
def test_knn_split_resident_query(n, d, metric):
    """f32 inputs that are not bf16-exact, 128 < d <= 512, k <= 16: the split kernel with
    hi and lo A-fragments resident (rows padded to 256 or 512 dimensions), ragged n, zero
    rows and duplicates."""
    x, _ = datagen.unit_mixture(n, d, 20, 0.2, seed=n + d)
    x = x.astype(np.float32) * np.float32(1.7)
    if metric == "cosine":
        x[5] = 0.0
    x[20:24] = x[10]
    st = _check(x, 15, metric)
    print(n, d, metric, st)


@pytest.mark.parametrize("n,d,metric,bf16", [
    (3000, 300, "cosine", False),     # split, 24 row blocks of 128: the middle pair once
    (2700, 200, "euclidean", False),  # split, d padded to 256, 21 blocks: odd count
    (777, 400, "cosine", False),      # 7 blocks, ragged last block
    (129, 300, "euclidean", False),   # 2 blocks, one row in the last
    (100, 256, "cosine", False),      # one block: the diagonal only
    (2000, 700, "cosine", True),      # bf16 rows above 512: three stages of 256
    (1500, 1000, "euclidean", True),  # and four
])
def test_knn_resident_symmetric_sweep(n, d, metric, bf16, monkeypatch):
    """The resident kernel's symmetric sweep (each pair of 128-row blocks once, transposed
    keys queued, merged and certified), for f32 rows that are not bf16-exact (split) and
    bf16 rows above 512 dimensions, gives the oracle's exact kNN; zero rows and duplicates."""
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "1")
    x, _ = datagen.unit_mixture(n, d, 12, 0.2, seed=n + d, bf16=bf16)
    x = x.astype(np.float32)
    if not bf16:
        x *= np.float32(1.7)
    if metric == "cosine":
        x[5] = 0.0
    x[20:24] = x[10]
    st = _check(x, 15, metric)
    assert st["symmetric"], st
    # the matrix work names the kernel: 128-row blocks, each over half the blocks (the
    # middle pair of an even count once), d padded to 256, three products per pair if split
    m, dpad = -(-n // 128), -(-d // 256) * 256
    blocks = sum((m + 1) // 2 if m % 2 else m // 2 + (i < m // 2) for i in range(m))
    assert st["kernel_flop"] == 2.0 * blocks * 128 * 128 * dpad * (1 if bf16 else 3), st


def test_knn_resident_symmetric_equals_plain_at_200k(monkeypatch):
    """At a size the oracle cannot check quickly: the resident kernel's symmetric and plain
    sweeps of split rows return identical graphs, the symmetric one at about half the
    matrix work."""
    x, _ = datagen.unit_mixture(200_000, 300, 60, 0.2, seed=9)
    x = x.astype(np.float32)
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "0")
    pi, pd, ps = H.knn(x, 15, metric="cosine", return_stats=True)
    monkeypatch.setenv("EVOC_HIP_KNN_SYMMETRIC", "1")
    si, sd, ss = H.knn(x, 15, metric="cosine", return_stats=True)
    assert not ps["symmetric"] and ss["symmetric"]
    assert ss["kernel_flop"] < 0.51 * ps["kernel_flop"]
    assert np.array_equal(pi, si)
    assert np.array_equal(pd.view(np.uint32), sd.view(np.uint32))


def test_knn_sampled_certificate_rate_keeps_first_pass_split():
    """300,000 split f32 rows (d = 300): the certificate-rate sample (the resident kernel over
    column shards, merged with a bound per row) passes, so the symmetric first pass runs
    and few rows are escalated; sampled rows equal the exact search."""
    n, d = 300_000, 300
    x, _ = datagen.unit_mixture(n, d, 60, 0.2, seed=11)
    x = x.astype(np.float32)
    gi, gd, st = H.knn(x, 15, metric="cosine", return_stats=True)
    assert st["symmetric"] and st["escalated_rows"] < n // 100, st
    rows = np.sort(np.random.default_rng(11).choice(n, size=512, replace=False))
    oi, od = O.knn_rows(x, rows, 15, metric=1)
    assert np.array_equal(gi[rows], oi)
    assert np.array_equal(gd[rows].view(np.uint32), od.view(np.uint32))
