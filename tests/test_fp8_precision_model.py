"""Why the kNN has no fp8-e4m3 path: a CPU model of an fp8 candidate pass with per-row
scales (max|x| / 448), an exact re-rank of its 32 best candidates and the certificate the
bf16 kernels use (knn.hip: knn_rerank_kernel), on the two data shapes the fp8 form was
meant for.

The certificate needs a bound on |key_fp8 - key_exact| for every pair. With per-row
quantisation errors e_x = q(x) - x, the product error x.e_y + e_x.y + e_x.e_y is bounded
by |x||e_y| + |e_x||y| + |e_x||e_y| (Cauchy-Schwarz), i.e. r_i + r_j + r_i r_j in cosine
key units with r = |e|/|x|. For e4m3 (3 mantissa bits) r is about 0.026, so the bound is
about 0.055, while the gap between a row's 15th and 33rd exact keys is 1e-4 to 1e-2 on
these shapes: no row can be certified, and every row would be searched again. Without the
certificate the fp8 top 32 misses most of the exact top 15 on the hierarchical shape
(1,000-leaf mixture at 0.1 leaf spread), where fp8 was meant to run.

e4m3 is emulated by torch's float8_e4m3fn conversion (round to nearest even), pinned
below against values fixed by the format (OCP FP8 E4M3: bias 7, max 448, no infinities,
subnormal step 2^-9), the format of gfx950's fp8 MFMA operands.
"""
import numpy as np
import pytest
import torch

import datagen

K = 15
C = 32


def _e4m3(v):
    return torch.tensor(v, dtype=torch.float32).to(torch.float8_e4m3fn).to(torch.float32).numpy()


def test_e4m3_emulation_known_values():
    v = np.array([1.0, 1.0625, 1.1875, 448.0, 2.0 ** -9, 2.0 ** -10 * 1.5, 0.3, -3.3],
                 np.float32)
    want = np.array([1.0, 1.0, 1.25, 448.0, 2.0 ** -9, 2.0 ** -9, 0.3125, -3.25],
                    np.float32)
    np.testing.assert_array_equal(_e4m3(v), want)


def _quantise(x):
    s = np.abs(x).max(axis=1, keepdims=True) / 448.0
    return _e4m3((x / s).astype(np.float32)) * s


def _hierarchical(n, d, n_par, n_child, spread, leaf_sd, seed):
    rng = np.random.default_rng(seed)
    par = rng.normal(size=(n_par, d))
    c = np.repeat(par, n_child, 0) + spread * rng.normal(size=(n_par * n_child, d))
    lab = np.arange(n) % (n_par * n_child)
    x = c[lab] + leaf_sd * rng.normal(size=(n, d))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return datagen.to_bf16_values(x.astype(np.float32))


def _model(x, xq, rows, eps_of):
    """Per sampled row: recall of the exact top K within the approximate top C, and whether
    the certificate (C+1-th approximate key - eps > exact K-th of the candidates) passes."""
    x64 = x.astype(np.float64)
    q64 = xq.astype(np.float64)
    nrm = np.linalg.norm(x64, axis=1)
    sel = np.random.default_rng(0).choice(len(x), rows, replace=False)
    exact = -(x64[sel] @ x64.T) / (nrm[sel, None] * nrm[None, :])
    approx = -(q64[sel] @ q64.T) / (nrm[sel, None] * nrm[None, :])
    exact[np.arange(rows), sel] = np.inf
    approx[np.arange(rows), sel] = np.inf
    recall, cert, gaps = [], [], []
    for t, i in enumerate(sel):
        order = np.argsort(approx[t], kind="stable")
        cand = order[:C]
        top = np.argsort(exact[t], kind="stable")[:K]
        recall.append(np.isin(top, cand).mean())
        dk = np.sort(exact[t, cand])[K - 1]
        cert.append(approx[t, order[C]] - eps_of(i) > dk)
        ek = np.sort(exact[t])
        gaps.append(ek[C] - ek[K - 1])
    return np.array(recall), np.array(cert), np.array(gaps)


def _fp8_case(x, rows):
    xq = _quantise(x)
    r = np.linalg.norm(xq.astype(np.float64) - x, axis=1) / np.linalg.norm(x, axis=1)
    rmax = r.max()
    rec, cert, gaps = _model(x, xq, rows, lambda i: r[i] + rmax + r[i] * rmax)
    return r, rec, cert, gaps


@pytest.mark.slow
def test_fp8_cannot_be_certified_and_misses_neighbours_on_the_hierarchical_shape():
    x = _hierarchical(12000, 768, 10, 10, 0.3, 0.1, seed=1)
    r, rec, cert, gaps = _fp8_case(x, 300)
    print(f"hierarchical: r median {np.median(r):.4f}, gap median {np.median(gaps):.2e}, "
          f"recall {rec.mean():.3f}, certified {cert.mean():.3f}")
    assert 0.02 < np.median(r) < 0.035
    assert np.median(gaps) < 1e-3            # neighbour gaps far below the fp8 bound
    assert cert.sum() == 0
    assert rec.mean() < 0.7                  # the fp8 top 32 misses exact neighbours


@pytest.mark.slow
def test_fp8_recovers_but_cannot_certify_the_unit_mixture_shape():
    x, _ = datagen.unit_mixture(12000, 512, 200, 0.15, 2, noise_frac=0.05, zipf=1.1,
                                bf16=True)
    r, rec, cert, gaps = _fp8_case(x, 300)
    print(f"unit mixture: recall {rec.mean():.3f}, certified {cert.mean():.3f}")
    assert rec.mean() > 0.99
    assert cert.sum() == 0


@pytest.mark.slow
def test_the_certificate_passes_with_exact_and_bf16_keys():
    """Controls: the same model certifies rows when the keys are exact, and with bf16-exact
    inputs through one bf16 product (the error is the fp32 accumulation only), so the fp8
    failures above come from the precision, not from the model."""
    x, _ = datagen.unit_mixture(12000, 512, 200, 0.15, 2, noise_frac=0.05, zipf=1.1,
                                bf16=True)
    _, cert_exact, _ = _model(x, x, 300, lambda i: 0.0)
    assert cert_exact.mean() == 1.0
    d = x.shape[1]
    eps_bf16 = 1.5 * (2 * d * 2.0 ** -24) + 64 * 2.0 ** -24
    _, cert_bf16, _ = _model(x, x, 300, lambda i: eps_bf16)
    assert cert_bf16.mean() > 0.95
