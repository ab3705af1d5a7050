"""FP8-e4m3 first pass of the kNN (EVOC_KNN_FP8, cosine).

The pass quantises each row to e4m3 with the per-row scale s = max|x| / 448, multiplies on
the fp8 MFMA with f32 accumulation and keeps, per row and column half, the 16 smallest
keys -(q_a . q_b) * s_b / ||b||. Its keys and its selection are checked against a numpy
statement of the same quantisation (torch's float8_e4m3fn rounding, nearest-even, pinned
on format values below) and dot products in f64; the certificate's error term against
the quantisation error measured on the host; and the whole search (FP8 pass, re-rank,
bf16 search of the rows it cannot certify, exact scan) against the oracle's exact f32
search, bit for bit."""
import numpy as np
import pytest
import torch

import datagen
import evoc_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def e4m3(y):
    """Round f32 values in [-448, 448] to e4m3 (OCP, nearest-even) and back to f32."""
    return torch.from_numpy(np.ascontiguousarray(y, np.float32)).to(torch.float8_e4m3fn).float().numpy()


def test_e4m3_rounding_reference():
    # format values: largest finite 448, smallest normal 2^-6, smallest subnormal 2^-9,
    # 3 mantissa bits (spacing 2^-3 in [1, 2), 32 in [256, 448]); ties to even
    v = np.array([448.0, 2 ** -6, 2 ** -9, 1.0625, 1.1875, 1.125 + 1e-3, 300.0, 304.0, -0.3],
                 np.float32)
    assert e4m3(v).tolist() == [448.0, 2 ** -6, 2 ** -9, 1.0, 1.25, 1.125, 288.0, 320.0,
                                -0.3125]


def quantise(x):
    """The tile pass's quantisation in f32 steps: scale, clamp, round to e4m3, subnormal
    results to zero."""
    amax = np.abs(x).max(axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        sc = np.where(amax > 0, (amax / np.float32(448.0)).astype(np.float32), 0).astype(np.float32)
        inv = np.where(amax > 0, (np.float32(448.0) / amax).astype(np.float32), 0).astype(np.float32)
    y = np.clip((x * inv[:, None]).astype(np.float32), -448.0, 448.0)
    q = e4m3(y).astype(np.float64)
    q[np.abs(q) < 2.0 ** -6] = 0.0  # subnormal results are stored as zero
    return q, sc


@pytest.mark.parametrize("n,d,seed", [(3000, 300, 11), (1500, 768, 12), (1000, 64, 13)])
def test_fp8_tile_pass_keys_and_selection(n, d, seed):
    x, _ = datagen.unit_mixture(n, d, 40, 0.15, seed=seed)
    x = x.astype(np.float32)
    idx, key, scale, gamma = H.knn_tile_pass(x, 15, metric="cosine", precision="fp8")
    q, sc = quantise(x)
    nrm = np.sqrt((x.astype(np.float64) ** 2).sum(1))
    # the key scale is s / ||x|| (||x|| in f32 summation order: agree to a few ulp)
    np.testing.assert_allclose(scale, sc / nrm, rtol=2e-6)
    # the certificate's error term covers the worst row's quantisation error
    rho = np.sqrt(((x - sc[:, None] * q) ** 2).sum(1)) / nrm
    assert gamma >= 2 * rho.max() + rho.max() ** 2 + (2 * ((d + 63) // 64 * 64) + 512) * 2.0 ** -24
    assert gamma < 2 * rho.max() + rho.max() ** 2 + 1e-3
    # keys of the returned candidates: -(Q_ab * scale_b), Q the dot of the e4m3 values
    rows = np.random.default_rng(seed).choice(n, size=min(n, 400), replace=False)
    u = 2.0 ** -24
    dp = (d + 63) // 64 * 64
    worst = 0.0
    for a in rows:
        c = idx[a]
        ok = c >= 0
        cj = c[ok]
        Q = q[cj] @ q[a]
        A = np.abs(q[cj]) @ np.abs(q[a])
        emu = -(Q * scale[cj])
        # sequential f32 adds give 2 d u; the fp8 MFMA's own sum of an instruction's 64
        # products measured up to ~160 u (of the sum of |products|), so 512 u is allowed,
        # as in the pass's error term
        tol = ((2 * dp + 512) * u) * A * scale[cj] + 4 * u * np.abs(emu) + 1e-30
        err = np.abs(key[a, ok] - emu)
        assert (err <= tol).all(), (a, (err / tol).max())
        worst = max(worst, float((err / tol).max()))
        # selection: each half keeps its 16 smallest keys (self excluded), to within the
        # keys' own rounding
        full = -(q @ q[a]) * scale
        full[a] = np.inf
        cols = np.arange(n)
        for h in range(2):
            sel = key[a, 16 * h:16 * h + 16]
            got = idx[a, 16 * h:16 * h + 16]
            hv = full[(cols % 128) // 64 == h]
            m = min(16, len(hv))
            assert (got[:m] >= 0).all()
            kth = np.sort(hv)[m - 1]
            slack = ((2 * dp + 512) * u) * float(np.abs(q).sum(1).max() * np.abs(q[a]).max()) * scale.max()
            assert sel[:m].max() <= kth + slack
            assert ((got[:m] % 128) // 64 == h).all()
    print(f"n={n} d={d}: worst key error / bound {worst:.3f}, gamma {gamma:.4f}, rho max {rho.max():.4f}")


@pytest.mark.parametrize("n,d,seed,escalate", [(3000, 300, 21, False), (1500, 768, 22, True),
                                                (2049, 64, 23, True), (700, 520, 24, False)])
def test_fp8_knn_equals_exact_search(n, d, seed, escalate, monkeypatch):
    if escalate:
        monkeypatch.setenv("EVOC_HIP_KNN_ESCALATE", "1")
    x, _ = datagen.unit_mixture(n, d, 30, 0.15, seed=seed)
    x = x.astype(np.float32)
    gi, gd, st = H.knn(x, 15, metric="cosine", precision="fp8", return_stats=True)
    oi, od = O.knn_exhaustive(x, 15, metric=1)
    assert np.array_equal(gi, oi), f"index mismatch rows {np.where((gi != oi).any(1))[0][:10]}"
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), "distance bits differ"
    print(f"n={n} d={d}: escalated {st['escalated_rows']}, exact rows {st['fallback_rows']}")


def test_fp8_bf16_valued_input_and_zero_rows():
    x, _ = datagen.unit_mixture(2000, 384, 20, 0.15, seed=31, bf16=True)
    x = x.astype(np.float32)
    x[::97] = 0.0
    gi, gd = H.knn(x, 15, metric="cosine", precision="fp8")
    oi, od = O.knn_exhaustive(x, 15, metric=1)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_fp8_rejects_euclidean():
    x = np.random.default_rng(0).normal(size=(100, 16)).astype(np.float32)
    with pytest.raises(H.EvocError, match="cosine"):
        H.knn(x, 5, metric="euclidean", precision="fp8")
