"""ctypes front-end of the C++ CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use this
module, and only as the checker. The product library never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
vp = C.c_void_p


class OrcParams(C.Structure):
    _fields_ = [
        ("n_neighbours", C.c_int32),
        ("noise_level", C.c_float),
        ("n_epochs", C.c_int32),
        ("embedding_dim", C.c_int32),
        ("neighbour_scale", C.c_float),
        ("symmetrise", C.c_int32),
        ("min_samples", C.c_int32),
        ("base_min_cluster_size", C.c_int32),
        ("approx_n_clusters", C.c_int32),
        ("min_similarity_threshold", C.c_double),
        ("max_layers", C.c_int32),
        ("metric", C.c_int32),
        ("persistence_fp64", C.c_int32),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_knn_exhaustive.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int, C.c_int, i32p, f32p]
        L.orc_knn_rows.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int, C.c_int, i64p, C.c_int64,
                                   i32p, f32p]
        L.orc_canon_dist.argtypes = [f32p, f32p, C.c_int64, C.c_int]
        L.orc_canon_dist.restype = C.c_float
        L.orc_fuzzy_graph.argtypes = [i32p, f32p, C.c_int64, C.c_int, C.c_float, C.c_int,
                                      f32p, f32p, i64p, i32p, f32p]
        L.orc_fuzzy_graph.restype = C.c_int64
        L.orc_fuzzy_graph_f64.argtypes = [i32p, f64p, C.c_int64, C.c_int, C.c_float, C.c_int,
                                          f64p, f64p, i64p, i32p, f64p]
        L.orc_fuzzy_graph_f64.restype = C.c_int64
        L.orc_embedding.argtypes = [i64p, i32p, f32p, C.c_int64, C.c_int, C.c_int, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    vp, C.c_uint64, f32p]
        L.orc_label_prop_init.argtypes = [i64p, i32p, f32p, C.c_int64, C.c_int, vp, C.c_int64,
                                          C.c_uint64, f32p]
        L.orc_core_sq.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, f32p]
        L.orc_core_sq_f64.argtypes = [f64p, C.c_int64, C.c_int, C.c_int, f64p]
        L.orc_mst.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, i32p, i32p, f32p]
        L.orc_mst.restype = C.c_int64
        L.orc_mst_f64.argtypes = [f64p, C.c_int64, C.c_int, C.c_int, i32p, i32p, f64p]
        L.orc_mst_f64.restype = C.c_int64
        L.orc_linkage.argtypes = [i32p, i32p, f32p, C.c_int64, C.c_int64, i64p, i64p, f32p, i64p]
        L.orc_linkage_f64.argtypes = [i32p, i32p, f64p, C.c_int64, C.c_int64, i64p, i64p, f64p,
                                      i64p]
        L.orc_condense.argtypes = [i64p, i64p, f32p, i64p, C.c_int64, C.c_int64, C.c_int64,
                                   i64p, i64p, f32p, i64p]
        L.orc_condense.restype = C.c_int64
        L.orc_condense_f64.argtypes = [i64p, i64p, f64p, i64p, C.c_int64, C.c_int64, C.c_int64,
                                       i64p, i64p, f64p, i64p]
        L.orc_condense_f64.restype = C.c_int64
        L.orc_extract_leaves.argtypes = [i64p, i64p, i64p, C.c_int64, C.c_int64, i64p]
        L.orc_extract_leaves.restype = C.c_int64
        L.orc_label_vector.argtypes = [i64p, i64p, C.c_int64, i64p, C.c_int64, C.c_int64, i64p]
        L.orc_strengths.argtypes = [i64p, i64p, f32p, i64p, C.c_int64, i64p, C.c_int64, i64p,
                                    C.c_int64, f32p]
        L.orc_find_peaks.argtypes = [f64p, C.c_int64, i64p]
        L.orc_find_peaks.restype = C.c_int64
        L.orc_persistence.argtypes = [i64p, i64p, f32p, i64p, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int, f64p, f64p, f64p, f64p, f64p]
        L.orc_persistence.restype = C.c_int64
        L.orc_select_diverse_peaks.argtypes = [i64p, C.c_int64, f64p, f64p, f64p, f64p,
                                               C.c_int64, C.c_double, C.c_int64, i64p, i64p]
        L.orc_cluster_layers.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_double,
                                         C.c_int64, C.c_int, i64p, f32p, f64p]
        L.orc_cluster_layers.restype = C.c_int64
        L.orc_cluster_layers_f64.argtypes = [f64p, C.c_int64, C.c_int, C.c_int, C.c_int64,
                                             C.c_double, C.c_int64, i64p, f64p, f64p]
        L.orc_cluster_layers_f64.restype = C.c_int64
        L.orc_search_n_clusters.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, C.c_int64, i64p,
                                            f32p]
        L.orc_search_n_clusters_f64.argtypes = [f64p, C.c_int64, C.c_int, C.c_int, C.c_int64,
                                                i64p, f64p]
        L.orc_evoc.argtypes = [f32p, C.c_int64, C.c_int64, vp, vp, C.POINTER(OrcParams),
                               C.c_uint64, i64p, f32p, f64p, vp, vp, vp, vp, vp]
        L.orc_evoc.restype = C.c_int64
        L.orc_philox_u32.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_philox_u32.restype = C.c_uint32
        L.orc_num_threads.restype = C.c_int
        L.orc_symmetrise.argtypes = [i64p, i32p, f32p, C.c_int64, C.c_int64, i64p, i32p, f32p]
        L.orc_symmetrise.restype = C.c_int64
        L.orc_mst_boruvka.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, i32p, i32p, f32p]
        L.orc_mst_boruvka.restype = C.c_int64
        L.orc_core_sq_kd.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, f32p]
        L.orc_set_mst_algo.argtypes = [C.c_int]
        L.orc_set_mst_algo.restype = C.c_int
        L.orc_knn_nndescent.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_uint64,
                                        C.c_int, C.c_float, C.c_float, i32p, f32p,
                                        C.POINTER(C.c_int)]
        csr3 = [i64p, i64p, f32p, C.c_int64, C.c_int64]
        L.orc_csr_op.argtypes = [C.c_int] + csr3 + csr3 + [i64p, i64p, f32p, C.c_int64, i64p]
        L.orc_csr_op.restype = C.c_int64
        L.orc_csr_from_coo.argtypes = [i64p, i64p, f32p, C.c_int64, C.c_int64, i64p, i64p, f32p]
        L.orc_csr_from_coo.restype = C.c_int64
        L.orc_csr_from_partition.argtypes = [i64p, C.c_int64, C.c_int64, i64p, i64p, f32p]
        L.orc_csr_matmul_dense.argtypes = csr3 + [f32p, C.c_int64, f32p]
        L.orc_lp_iteration.argtypes = [i64p, i64p, f32p, C.c_int64, i64p, i64p]
        L.orc_lp_outliers.argtypes = [i64p, i64p, f32p, C.c_int64, i64p, C.c_uint64]
        L.orc_lp_remap.argtypes = [i64p, C.c_int64]
        L.orc_lp_remap.restype = C.c_int64
        L.orc_lp_loop.argtypes = [i64p, i64p, f32p, C.c_int64, C.c_int64, C.c_int, C.c_uint64,
                                  i64p]
        L.orc_lp_loop.restype = C.c_int64
        L.orc_unit_range.argtypes = [f32p, C.c_int64, C.c_int64]
        L.orc_pca_init.argtypes = [f32p, C.c_int64, C.c_int64, C.c_int, f32p]
        L.orc_random_init.argtypes = [C.c_int64, C.c_int, C.c_uint64, f32p]
        L.orc_lp_expand.argtypes = [i64p, i64p, f32p, C.c_int64, i64p, C.c_int64, f32p, C.c_int,
                                    C.c_double, f32p]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def knn_exhaustive(x, k, metric=0):
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    lib().orc_knn_exhaustive(x, n, d, metric, k, idx, dist)
    return idx, dist


def knn_rows(x, rows, k, metric=0):
    x = _f32(x)
    n, d = x.shape
    rows = np.ascontiguousarray(rows, np.int64)
    idx = np.empty((len(rows), k), np.int32)
    dist = np.empty((len(rows), k), np.float32)
    lib().orc_knn_rows(x, n, d, metric, k, rows, len(rows), idx, dist)
    return idx, dist


def canon_dist(a, b, metric=0):
    a, b = _f32(a), _f32(b)
    return lib().orc_canon_dist(a, b, a.shape[0], metric)


def fuzzy_graph(idx, dist, effective_k=None, symmetrise=True, fp64=False):
    """build_fuzzy_simplicial_set; fp64=True runs the same steps in fp64 (distances taken
    as fp64) for error statements."""
    idx = np.ascontiguousarray(idx, np.int32)
    ft = np.float64 if fp64 else np.float32
    dist = _f64(dist) if fp64 else _f32(dist)
    n, k = idx.shape
    eff = float(k) if effective_k is None else float(effective_k)
    sig = np.empty(n, ft)
    rho = np.empty(n, ft)
    indptr = np.empty(n + 1, np.int64)
    ind = np.empty(2 * n * k, np.int32)
    w = np.empty(2 * n * k, ft)
    fn = lib().orc_fuzzy_graph_f64 if fp64 else lib().orc_fuzzy_graph
    nnz = fn(idx, dist, n, k, eff, int(symmetrise), sig, rho, indptr, ind, w)
    return indptr, ind[:nnz].copy(), w[:nnz].copy(), sig, rho


def symmetrise(rows, cols, vals, n):
    """symmetrise_graph on an explicit directed COO list -> (indptr, indices, weights)."""
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = _f32(vals)
    m = 2 * max(len(vals), 1)
    ip = np.empty(n + 1, np.int64)
    ix = np.empty(m, np.int32)
    w = np.empty(m, np.float32)
    nnz = lib().orc_symmetrise(rows, cols, vals, len(vals), n, ip, ix, w)
    return ip, ix[:nnz].copy(), w[:nnz].copy()


def embedding(indptr, ind, w, dim, n_epochs=50, noise_level=0.5, negative_sample_rate=1.0,
              initial_alpha=0.1, beta1=0.5, beta2=0.9, eps=1e-7, init=None, seed=42):
    indptr = np.ascontiguousarray(indptr, np.int64)
    n = indptr.shape[0] - 1
    ind = np.ascontiguousarray(ind, np.int32)
    w = _f32(w)
    out = np.empty((n, dim), np.float32)
    init_a = None if init is None else _f32(init)
    lib().orc_embedding(indptr, ind, w, n, dim, n_epochs, noise_level, negative_sample_rate,
                        initial_alpha, beta1, beta2, eps, _ptr(init_a), seed, out)
    return out


def label_prop_init(indptr, ind, w, dim, data=None, seed=42):
    indptr = np.ascontiguousarray(indptr, np.int64)
    n = indptr.shape[0] - 1
    d_a = None if data is None else _f32(data)
    out = np.empty((n, dim), np.float32)
    lib().orc_label_prop_init(indptr, np.ascontiguousarray(ind, np.int32), _f32(w), n, dim,
                              _ptr(d_a), 0 if d_a is None else d_a.shape[1], seed, out)
    return out


def core_sq(emb, min_samples):
    emb = np.ascontiguousarray(emb)
    n, dim = emb.shape
    if emb.dtype == np.float64:
        out = np.empty(n, np.float64)
        lib().orc_core_sq_f64(emb, n, dim, min_samples, out)
    else:
        emb = _f32(emb)
        out = np.empty(n, np.float32)
        lib().orc_core_sq(emb, n, dim, min_samples, out)
    return out


def mst(emb, min_samples):
    emb = np.ascontiguousarray(emb)
    n, dim = emb.shape
    m = max(n - 1, 0)
    u = np.empty(max(m, 1), np.int32)
    v = np.empty(max(m, 1), np.int32)
    if emb.dtype == np.float64:
        w = np.empty(max(m, 1), np.float64)
        ne = lib().orc_mst_f64(emb, n, dim, min_samples, u, v, w)
    else:
        emb = _f32(emb)
        w = np.empty(max(m, 1), np.float32)
        ne = lib().orc_mst(emb, n, dim, min_samples, u, v, w)
    return u[:ne].copy(), v[:ne].copy(), w[:ne].copy()


def linkage(u, v, w, n):
    u = np.array(u, np.int32)
    v = np.array(v, np.int32)
    ne = u.shape[0]
    L = np.empty(max(ne, 1), np.int64)
    R = np.empty(max(ne, 1), np.int64)
    S = np.empty(max(ne, 1), np.int64)
    if np.asarray(w).dtype == np.float64:
        w = np.array(w, np.float64)
        D = np.empty(max(ne, 1), np.float64)
        lib().orc_linkage_f64(u, v, w, ne, n, L, R, D, S)
    else:
        w = np.array(w, np.float32)
        D = np.empty(max(ne, 1), np.float32)
        lib().orc_linkage(u, v, w, ne, n, L, R, D, S)
    return L[:ne], R[:ne], D[:ne], S[:ne]


def condense(L, R, D, S, n, mcs):
    nr = len(L)
    cap = max(2 * n, 1)
    P = np.empty(cap, np.int64)
    Ch = np.empty(cap, np.int64)
    CS = np.empty(cap, np.int64)
    L = np.ascontiguousarray(L, np.int64)
    R = np.ascontiguousarray(R, np.int64)
    S = np.ascontiguousarray(S, np.int64)
    if np.asarray(D).dtype == np.float64:
        lam = np.empty(cap, np.float64)
        nn = lib().orc_condense_f64(L, R, _f64(D), S, nr, n, mcs, P, Ch, lam, CS)
    else:
        lam = np.empty(cap, np.float32)
        nn = lib().orc_condense(L, R, _f32(D), S, nr, n, mcs, P, Ch, lam, CS)
    return P[:nn], Ch[:nn], lam[:nn], CS[:nn]


def extract_leaves(P, Ch, CS, n):
    out = np.empty(max(len(P), 1), np.int64)
    m = lib().orc_extract_leaves(np.ascontiguousarray(P, np.int64),
                                 np.ascontiguousarray(Ch, np.int64),
                                 np.ascontiguousarray(CS, np.int64), len(P), n, out)
    return out[:m]


def label_vector(P, Ch, clusters, n):
    out = np.empty(n, np.int64)
    cl = np.ascontiguousarray(clusters, np.int64)
    lib().orc_label_vector(np.ascontiguousarray(P, np.int64), np.ascontiguousarray(Ch, np.int64),
                           len(P), cl if len(cl) else np.zeros(1, np.int64), len(cl), n, out)
    return out


def strengths(P, Ch, lam, CS, clusters, labels):
    n = len(labels)
    out = np.empty(n, np.float32)
    cl = np.ascontiguousarray(clusters, np.int64)
    lib().orc_strengths(np.ascontiguousarray(P, np.int64), np.ascontiguousarray(Ch, np.int64),
                        _f32(lam), np.ascontiguousarray(CS, np.int64), len(P),
                        cl if len(cl) else np.zeros(1, np.int64), len(cl),
                        np.ascontiguousarray(labels, np.int64), n, out)
    return out


def find_peaks(x):
    x = _f64(x)
    out = np.empty(max(len(x), 1), np.int64)
    m = lib().orc_find_peaks(x if len(x) else np.zeros(1), len(x), out)
    return out[:m]


def persistence(P, Ch, lam, CS, n, mcs, fp64=True):
    nn = len(P)
    cap = max(2 * n, 4)
    sizes = np.empty(cap, np.float64)
    pers = np.empty(cap, np.float64)
    b = np.empty(cap, np.float64)
    d = np.empty(cap, np.float64)
    ld = np.empty(cap, np.float64)
    r = lib().orc_persistence(np.ascontiguousarray(P, np.int64), np.ascontiguousarray(Ch, np.int64),
                              _f32(lam), np.ascontiguousarray(CS, np.int64), nn, n, mcs,
                              int(fp64), sizes, pers, b, d, ld)
    nbc, ns = r >> 32, r & 0xFFFFFFFF
    return sizes[:ns], pers[:ns], b[:nbc], d[:nbc], ld[:nbc]


def select_diverse_peaks(peaks, pers, sizes, births, deaths, min_sim, max_layers):
    peaks = np.ascontiguousarray(peaks, np.int64)
    sel = np.empty(max(len(peaks), 1), np.int64)
    ns = np.zeros(1, np.int64)
    lib().orc_select_diverse_peaks(peaks if len(peaks) else np.zeros(1, np.int64), len(peaks),
                                   _f64(pers) if len(pers) else np.zeros(1),
                                   _f64(sizes) if len(sizes) else np.zeros(1),
                                   _f64(births) if len(births) else np.zeros(1),
                                   _f64(deaths) if len(deaths) else np.zeros(1),
                                   len(births), min_sim, max_layers, sel, ns)
    return sel[:ns[0]]


def cluster_layers(emb, min_samples=5, base_mcs=5, min_sim=0.2, max_layers=10, fp64=True):
    emb = np.ascontiguousarray(emb)
    n, dim = emb.shape
    ml = max(max_layers, 1)
    labels = np.empty((ml, max(n, 1)), np.int64)
    pers = np.empty(ml, np.float64)
    if emb.dtype == np.float64:
        st = np.empty((ml, max(n, 1)), np.float64)
        nl = lib().orc_cluster_layers_f64(emb, n, dim, min_samples, base_mcs, min_sim,
                                          max_layers, labels, st, pers)
    else:
        emb = _f32(emb)
        st = np.empty((ml, max(n, 1)), np.float32)
        nl = lib().orc_cluster_layers(emb, n, dim, min_samples, base_mcs, min_sim, max_layers,
                                      int(fp64), labels, st, pers)
    return labels[:nl, :n], st[:nl, :n], pers[:nl]


def search_n_clusters(emb, min_samples, target_k):
    emb = np.ascontiguousarray(emb)
    n, dim = emb.shape
    labels = np.empty(n, np.int64)
    if emb.dtype == np.float64:
        st = np.empty(n, np.float64)
        lib().orc_search_n_clusters_f64(emb, n, dim, min_samples, target_k, labels, st)
    else:
        emb = _f32(emb)
        st = np.empty(n, np.float32)
        lib().orc_search_n_clusters(emb, n, dim, min_samples, target_k, labels, st)
    return labels, st


def params(n_neighbours=15, noise_level=0.5, n_epochs=50, embedding_dim=0, neighbour_scale=1.0,
           symmetrise=True, min_samples=5, base_min_cluster_size=5, approx_n_clusters=0,
           min_similarity_threshold=0.2, max_layers=10, metric=0, persistence_fp64=1):
    return OrcParams(n_neighbours, noise_level, n_epochs, embedding_dim, neighbour_scale,
                     int(symmetrise), min_samples, base_min_cluster_size, approx_n_clusters,
                     min_similarity_threshold, max_layers, metric, persistence_fp64)


def evoc(x, p=None, seed=42, pre_idx=None, pre_dist=None, want_intermediates=False):
    """Full CPU pipeline. Returns dict with labels (layers x n), strengths, persistence,
    knn_idx, knn_dist, embedding, init and per-stage timings (seconds)."""
    x = _f32(x)
    n, d = x.shape
    p = p or params()
    k = p.n_neighbours
    dim = p.embedding_dim if p.embedding_dim > 0 else min(max(k // 4, 4), 16)
    ml = max(p.max_layers, 1)
    labels = np.empty((ml, n), np.int64)
    st = np.empty((ml, n), np.float32)
    pers = np.empty(ml, np.float64)
    ki = np.empty((n, k), np.int32)
    kd = np.empty((n, k), np.float32)
    emb = np.empty((n, dim), np.float32)
    init = np.empty((n, dim), np.float32)
    tim = np.zeros(5, np.float64)
    pi = None if pre_idx is None else np.ascontiguousarray(pre_idx, np.int32)
    pd = None if pre_dist is None else _f32(pre_dist)
    nl = lib().orc_evoc(x, n, d, _ptr(pi), _ptr(pd), C.byref(p), seed, labels, st, pers,
                        _ptr(ki), _ptr(kd), _ptr(emb), _ptr(init), _ptr(tim))
    return dict(labels=labels[:nl], strengths=st[:nl], persistence=pers[:nl], knn_idx=ki,
                knn_dist=kd, embedding=emb, init=init, timings=tim)


# ---------------------------------------------------------------- CSR algebra and label-prop
class Csr:
    """A CSR matrix as the oracle stores it (int64 indices, f32 values)."""

    def __init__(self, indptr, indices, data, nrows, ncols):
        self.indptr = np.ascontiguousarray(indptr, np.int64)
        self.indices = np.ascontiguousarray(indices, np.int64)
        self.data = _f32(data)
        self.nrows, self.ncols = int(nrows), int(ncols)
        if len(self.indices) == 0:
            self.indices = np.zeros(1, np.int64)
            self.data = np.zeros(1, np.float32)
            self._nnz = 0
        else:
            self._nnz = int(self.indptr[-1])

    def nnz(self):
        return self._nnz

    def args(self):
        return (self.indptr, self.indices, self.data, self.nrows, self.ncols)

    def to_dense(self):
        d = np.zeros((self.nrows, self.ncols), np.float32)
        for i in range(self.nrows):
            for p in range(self.indptr[i], self.indptr[i + 1]):
                d[i, self.indices[p]] += self.data[p]
        return d

    def rows(self):
        return [[(int(self.indices[p]), float(self.data[p]))
                 for p in range(self.indptr[i], self.indptr[i + 1])] for i in range(self.nrows)]


_CSR_OPS = {"transpose": 0, "matmul": 1, "elementwise_mul": 2, "normalise_cols_l2": 3,
            "normalise_rows_l1": 4, "clip_values": 5}


def csr_op(op, a, b=None):
    cap = max(a.nrows * (b.ncols if b is not None else a.ncols), a.nnz(), 1) + 1
    bb = b if b is not None else a
    ip = np.empty(max(a.nrows, a.ncols, bb.ncols) + 2, np.int64)
    ix = np.empty(cap, np.int64)
    dv = np.empty(cap, np.float32)
    shape = np.zeros(2, np.int64)
    nnz = lib().orc_csr_op(_CSR_OPS[op], *a.args(), *bb.args(), ip, ix, dv, cap, shape)
    assert nnz >= 0
    return Csr(ip[:shape[0] + 1].copy(), ix[:nnz].copy(), dv[:nnz].copy(), shape[0], shape[1])


def csr_from_coo(rows, cols, vals, n):
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    vals = _f32(vals)
    m = max(len(vals), 1)
    ip = np.empty(n + 1, np.int64)
    ix = np.empty(m, np.int64)
    dv = np.empty(m, np.float32)
    nnz = lib().orc_csr_from_coo(rows if len(rows) else np.zeros(1, np.int64),
                                 cols if len(cols) else np.zeros(1, np.int64),
                                 vals if len(vals) else np.zeros(1, np.float32), len(vals), n,
                                 ip, ix, dv)
    return Csr(ip, ix[:nnz].copy(), dv[:nnz].copy(), n, n)


def csr_from_partition(part, n_parts):
    part = np.ascontiguousarray(part, np.int64)
    n = len(part)
    ip = np.empty(n + 1, np.int64)
    ix = np.empty(max(n, 1), np.int64)
    dv = np.empty(max(n, 1), np.float32)
    lib().orc_csr_from_partition(part, n, n_parts, ip, ix, dv)
    return Csr(ip, ix[:n], dv[:n], n, n_parts)


def csr_matmul_dense(a, rhs):
    rhs = _f32(rhs)
    out = np.empty((a.nrows, rhs.shape[1]), np.float32)
    lib().orc_csr_matmul_dense(*a.args(), rhs, rhs.shape[1], out)
    return out


def lp_iteration(g, labels):
    labels = np.ascontiguousarray(labels, np.int64)
    out = np.empty_like(labels)
    lib().orc_lp_iteration(g.indptr, g.indices, g.data, g.nrows, labels, out)
    return out


def lp_outliers(g, labels, seed):
    labels = np.array(labels, np.int64)
    lib().orc_lp_outliers(g.indptr, g.indices, g.data, g.nrows, labels, seed)
    return labels


def lp_remap(labels):
    labels = np.array(labels, np.int64)
    m = lib().orc_lp_remap(labels, len(labels))
    return m, labels


def lp_loop(g, approx_parts, n_iter, seed):
    part = np.empty(g.nrows, np.int64)
    m = lib().orc_lp_loop(g.indptr, g.indices, g.data, g.nrows, approx_parts, n_iter, seed, part)
    return part, m


def unit_range(m):
    m = np.array(m, np.float32)
    lib().orc_unit_range(m, m.shape[0], m.shape[1])
    return m


def pca_init(data, dim):
    data = _f32(data)
    out = np.empty((data.shape[0], dim), np.float32)
    lib().orc_pca_init(data, data.shape[0], data.shape[1], dim, out)
    return out


def random_init(n, dim, seed):
    out = np.empty((n, dim), np.float32)
    lib().orc_random_init(n, dim, seed, out)
    return out


def lp_expand(g, part, n_parts, layout, scaling=0.1):
    layout = _f32(layout)
    dim = layout.shape[1]
    out = np.empty((g.nrows, dim), np.float32)
    lib().orc_lp_expand(g.indptr, g.indices, g.data, g.nrows, np.ascontiguousarray(part, np.int64),
                        n_parts, layout, dim, scaling, out)
    return out


# ---------------------------------------------------------------- reference-complexity baselines
def mst_boruvka(emb, min_samples):
    """kd-tree Boruvka: the same unique MST as mst() (canonical tie order), O(n log n) rounds."""
    emb = _f32(emb)
    n, dim = emb.shape
    m = max(n - 1, 1)
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    w = np.empty(m, np.float32)
    ne = lib().orc_mst_boruvka(emb, n, dim, min_samples, u, v, w)
    return u[:ne].copy(), v[:ne].copy(), w[:ne].copy()


def core_sq_kd(emb, min_samples):
    emb = _f32(emb)
    n, dim = emb.shape
    out = np.empty(n, np.float32)
    lib().orc_core_sq_kd(emb, n, dim, min_samples, out)
    return out


def set_mst_algo(algo):
    """0: dense Prim (the parity reference), 1: kd-tree Boruvka. Returns the previous."""
    return lib().orc_set_mst_algo(int(algo))


def knn_nndescent(x, k, metric=0, seed=42, max_iters=12, delta=0.001, rho=1.0):
    """Textbook NN-Descent; returns (idx, dist, iterations)."""
    x = _f32(x)
    n, d = x.shape
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    it = C.c_int(0)
    lib().orc_knn_nndescent(x, n, d, metric, k, seed, max_iters, delta, rho, idx, dist,
                            C.byref(it))
    return idx, dist, it.value
