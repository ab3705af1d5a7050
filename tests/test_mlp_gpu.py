"""The warm-start MLP's hand-written forward/backward kernels against an fp64 numpy
restatement of the same network (gradient check), and what the learned warm start
gains over the plain label-propagation initialisation."""
import time

import numpy as np
import pytest
from scipy.special import erf
from sklearn.metrics import adjusted_rand_score as ari

import datagen
import evoc_hip as H

pytestmark = pytest.mark.gpu


def _gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def _gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def _ref(x, t, W, b):
    x, t = x.astype(np.float64), t.astype(np.float64)
    W = [w.astype(np.float64) for w in W]
    b = [v.astype(np.float64) for v in b]
    z1 = x @ W[0].T + b[0]
    a1 = _gelu(z1)
    z2 = a1 @ W[1].T + b[1]
    a2 = _gelu(z2)
    y = a2 @ W[2].T + b[2]
    loss = np.mean((y - t) ** 2)
    dy = 2.0 * (y - t) / y.size
    gw2, gb2 = dy.T @ a2, dy.sum(0)
    dz2 = (dy @ W[2]) * _gelu_grad(z2)
    gw1, gb1 = dz2.T @ a1, dz2.sum(0)
    dz1 = (dz2 @ W[1]) * _gelu_grad(z1)
    gw0, gb0 = dz1.T @ x, dz1.sum(0)
    return loss, y, [gw0, gw1, gw2], [gb0, gb1, gb2]


def _cos(a, b):
    a, b = a.ravel(), b.ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


@pytest.mark.parametrize("rows,d,out", [(256, 512, 8), (100, 70, 3)])
def test_mlp_gradients_match_fp64(rows, d, out):
    rng = np.random.default_rng(rows + d)
    x = rng.normal(size=(rows, d)).astype(np.float32) / np.sqrt(d)
    t = rng.normal(size=(rows, out)).astype(np.float32)
    dims = [(256, d), (256, 256), (out, 256)]
    W = [(rng.normal(size=s) * np.sqrt(2.0 / (s[0] + s[1]))).astype(np.float32) for s in dims]
    b = [rng.normal(scale=0.1, size=s[0]).astype(np.float32) for s in dims]
    loss, y, gw, gb = H.mlp_loss_grads(x, t, W, b)
    rl, ry, rgw, rgb = _ref(x, t, W, b)
    # bf16 operands, fp32 accumulation: a few bf16 ulps of relative error
    assert abs(loss - rl) / rl < 2e-2
    np.testing.assert_allclose(y, ry, atol=3e-2 * np.abs(ry).max())
    for g, r in zip(gw + gb, rgw + rgb):
        assert _cos(g, r) > 0.999
        assert abs(np.linalg.norm(g) - np.linalg.norm(r)) / np.linalg.norm(r) < 3e-2


def test_learned_warm_start_gate_and_fallback():
    """The learned layout is used only if, on its training subsample, the warm-started short
    run reproduces the plain run (best-layer ARI >= 0.98, MST total within 1%); otherwise the
    plain initial layout and epoch count run. Either way the result must meet the gate
    against the plain run on the full data: accepted runs by the check, rejected ones by
    being the plain run itself."""
    x, gt = datagen.gaussian_mixture(120_000, 64, 20, 5.0, 1.0, seed=42)
    p = H.EvocParams(embedding_dim=8)
    t0 = time.perf_counter()
    plain = H.evoc(x, params=p, seed=42)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    warm = H.evoc(x, params=p, nn_params=H.NearestNeighbourParams(warm_start=True,
                                                                  mlp_subsample=20000), seed=42)
    t_warm = time.perf_counter() - t0
    score = ari(warm.best_labels(), plain.best_labels())
    mst_p = H.mst(plain.embedding, 5)[2].astype(np.float64).sum()
    mst_w = H.mst(warm.embedding, 5)[2].astype(np.float64).sum()
    print(f"plain {t_plain:.2f}s warm {t_warm:.2f}s; subsample check ARI {warm.warm_start_ari:.4f} "
          f"MST ratio {warm.warm_start_mst_ratio:.4f} -> used {warm.warm_start_used}; full data "
          f"ARI(warm, plain) {score:.4f}, MST ratio {mst_w / mst_p:.4f}; stages plain "
          f"{plain.stage_ms} warm {warm.stage_ms}")
    assert 0.0 < warm.warm_start_mst_ratio and -1.0 <= warm.warm_start_ari <= 1.0
    if warm.warm_start_used:
        assert score >= 0.98 and abs(mst_w / mst_p - 1.0) <= 0.01
    else:
        assert np.array_equal(warm.embedding.view(np.uint32), plain.embedding.view(np.uint32))
        assert np.array_equal(warm.cluster_layers, plain.cluster_layers)
