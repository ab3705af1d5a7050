"""The C-ABI library loads, exports every entry point include/evoc_hip.h declares, and
rejects bad arguments with status codes (no compute call is made here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import evoc_hip as H

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(REPO, "include", "evoc_hip.h")).read()
    return sorted(set(re.findall(r"\b(evoc_hip_\w+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = H.lib()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_front_end_covers_every_symbol():
    assert set(_declared()) <= set(H._SIGS)


def test_default_params_mirror_reference_defaults():
    # EvocParams::default, src/lib.rs:83-99
    p = H._Params()
    H.lib().evoc_hip_default_params(C.byref(p))
    assert (p.n_neighbours, p.n_epochs, p.embedding_dim, p.symmetrise) == (15, 50, 0, 1)
    assert abs(p.noise_level - 0.5) < 1e-7 and p.neighbour_scale == 1.0
    assert (p.min_samples, p.base_min_cluster_size, p.approx_n_clusters, p.max_layers) == (5, 5, 0, 10)
    assert abs(p.min_similarity_threshold - 0.2) < 1e-12
    q = H._NnParams()
    H.lib().evoc_hip_default_nn_params(C.byref(q))
    assert (q.metric, q.knn_precision, q.warm_start, q.device_id) == (0, 0, 0, 0)
    d = H.EvocParams()
    assert (d.n_neighbours, d.n_epochs, d.min_samples, d.max_layers) == (15, 50, 5, 10)


def test_invalid_arguments_are_status_codes():
    L = H.lib()
    x = np.zeros((4, 3), np.float32)
    idx = np.zeros((4, 2), np.int32)
    dist = np.zeros((4, 2), np.float32)
    st = L.evoc_hip_knn(x, 0, 3, 0, 2, 0, idx, dist, None, None)
    assert st == 1 and b"n and d" in L.evoc_hip_last_error()
    st = L.evoc_hip_knn(x, 4, 3, 7, 2, 0, idx, dist, None, None)
    assert st == 1 and b"metric" in L.evoc_hip_last_error()
    out = C.POINTER(H._Result)()
    p = H.EvocParams()._c()
    st = L.evoc_hip_run(x.reshape(-1), 1, 3, 3, 1, None, None, 0, C.byref(p), None, 1, 0, C.byref(out))
    assert st == 1 and not out


def test_condense_tree_rejects_rows_that_are_not_a_linkage():
    # rows for n = 4: node 4 = (0, 1), node 5 = (2, 3), node 6 = (4, 5)
    L = np.array([1, 3, 5], np.int64)
    R = np.array([0, 2, 4], np.int64)
    D = np.array([0.5, 0.7, 1.0], np.float32)
    S = np.array([2, 2, 4], np.int64)
    bad = [
        (np.array([1, 1, 5]), R, D, S, "binary tree"),   # point 1 has two parents
        (np.array([1, 6, 5]), R, D, S, "binary tree"),   # child is not an earlier node
        (L, R, D, np.array([2, 2, 5]), "size"),          # sizes do not add up
        (L, R, np.array([0.5, np.nan, 1.0]), S, "finite"),
        (L, R, np.array([0.5, -0.1, 1.0]), S, "finite"),
    ]
    for l, r, d, s, msg in bad:
        with pytest.raises(H.EvocError, match=msg):
            H.condense_tree(l, r, d, s, 4, 2)
    with pytest.raises(ValueError):
        H.condense_tree(L[:2], R, D, S, 4, 2)


def test_device_count_without_gpu_is_not_an_error():
    n = H.device_count()
    assert n >= 0


def test_non_finite_embeddings_rejected_before_any_device_call():
    """Every entry point that takes an embedding or distances rejects NaN / inf with
    EVOC_ERR_INVALID_ARGUMENT before it touches a device (so this runs without a GPU)."""
    L = H.lib()
    e = np.zeros((6, 2), np.float32)
    e[3, 1] = np.nan
    u = np.zeros(5, np.int32)
    w = np.zeros(5, np.float32)
    assert L.evoc_hip_mst(e, 6, 2, 2, u, u.copy(), w) == 1
    assert b"non-finite" in L.evoc_hip_last_error()
    assert L.evoc_hip_core_sq(e, 6, 2, 2, np.zeros(6, np.float32)) == 1
    lab = np.zeros((2, 6), np.int64)
    st = np.zeros((2, 6), np.float32)
    nl = C.c_int64(0)
    assert L.evoc_hip_cluster_layers(e, 6, 2, 2, 2, 0.2, 2, lab, st, np.zeros(2), C.byref(nl)) == 1
    e[3, 1] = np.inf
    assert L.evoc_hip_search_n_clusters(e, 6, 2, 2, 2, lab[0].copy(), st[0].copy()) == 1
    assert b"non-finite" in L.evoc_hip_last_error()
    x = np.ones((4, 3), np.float32)
    x[2, 2] = -np.inf
    idx = np.zeros((4, 2), np.int32)
    dist = np.zeros((4, 2), np.float32)
    assert L.evoc_hip_knn(x, 4, 3, 0, 2, 0, idx, dist, None, None) == 1
    dist[1, 1] = np.nan
    assert L.evoc_hip_fuzzy(idx, dist, 4, 2, 2.0, 1, np.zeros(4, np.float32),
                            np.zeros(4, np.float32), np.zeros(5, np.int64),
                            np.zeros(8, np.int32), np.zeros(8, np.float32),
                            C.byref(C.c_int64(0))) == 1


def test_precomputed_knn_shapes_and_values_checked():
    import pytest
    x = np.zeros((5, 3), np.float32)
    good_i = np.array([[1, 2], [0, 2], [0, 1], [4, -1], [3, -1]])
    good_d = np.array([[0.1, 0.2], [0.1, 0.3], [0.2, 0.3], [0.5, np.inf], [0.5, np.nan]],
                      np.float32)
    pi, pd = H._check_precomputed((good_i, good_d), 5)   # -1 pads: their distance is free
    assert pi.dtype == np.int32 and pd.dtype == np.float32 and pi.shape == (5, 2)
    bad = [
        (good_i[:4], good_d[:4]),                 # too few rows
        (good_i, good_d[:, :1]),                  # distance shape differs
        (good_i.ravel(), good_d.ravel()),         # not 2-d
        (good_i + 10, good_d),                    # index out of range
        (good_i.astype(np.float32), good_d),      # not integers
        (good_i, np.where(good_i >= 0, -1.0, 0.0).astype(np.float32)),  # negative distance
        (good_i, np.where(good_i >= 0, np.nan, 0.0).astype(np.float32)),  # NaN distance
    ]
    for pi, pd in bad:
        with pytest.raises(ValueError):
            H.evoc(x, precomputed_knn=(pi, pd))


def _struct_fields(src, name):
    """Field names of `typedef struct { ... } name;` in the C header."""
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", src).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return [re.findall(r"(\w+)(?:\[\d+\])?\s*;", line)[0] for line in body.split("\n")
            if re.search(r"\w+(?:\[\d+\])?\s*;", line)]


def test_rust_and_python_bindings_match_the_header():
    hdr = open(os.path.join(REPO, "include", "evoc_hip.h")).read()
    rs = open(os.path.join(REPO, "rust", "evoc-hip-sys", "src", "lib.rs")).read()
    # every function the header declares is bound in the -sys crate
    rust_fns = set(re.findall(r"pub fn (evoc_hip_\w+)\s*\(", rs))
    assert set(_declared()) <= rust_fns, set(_declared()) - rust_fns
    for cname, pyname in (("evoc_params_t", "_Params"), ("evoc_hip_nn_params_t", "_NnParams"),
                          ("evoc_result_t", "_Result")):
        cf = _struct_fields(hdr, cname)
        body = re.search(r"pub struct " + cname + r" \{(.*?)\n\}", rs, re.S).group(1)
        rf = re.findall(r"pub (\w+):", body)
        assert rf == cf, (cname, rf, cf)
        assert [f[0] for f in getattr(H, pyname)._fields_] == cf, cname
