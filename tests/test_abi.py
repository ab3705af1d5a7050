"""The C-ABI library loads, exports every entry point include/evoc_hip.h declares, and
rejects bad arguments with status codes (no compute call is made here)."""
import ctypes as C
import os
import re

import numpy as np

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


def test_device_count_without_gpu_is_not_an_error():
    n = H.device_count()
    assert n >= 0
