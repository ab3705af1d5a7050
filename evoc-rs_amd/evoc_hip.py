"""Python front-end of libevoc_hip.so, mirroring evoc-rs's public API.

``evoc()`` corresponds to ``evoc_rs::evoc`` (src/lib.rs:199-352): it takes an
``(n, d)`` float32 matrix, ``EvocParams`` and nearest-neighbour settings and returns an
``EvocResult`` with ``cluster_layers``, ``membership_strengths``, ``persistence_scores``,
``nn_indices`` and ``nn_distances``, plus ``best_labels()`` / ``best_strengths()`` /
``n_clusters()`` (src/lib.rs:120-161). All compute runs in the HIP library; there is no
CPU fallback: loading fails loudly if the library is missing.
"""
import ctypes as C
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libevoc_hip.so")

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
vp = C.c_void_p

METRICS = {"euclidean": 0, "cosine": 1}
PRECISIONS = {"auto": 0, "split": 1, "bf16": 2, "fp8": 3}


class EvocError(RuntimeError):
    """Raised when a library call returns a non-zero status."""


class _Params(C.Structure):
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
    ]


class _NnParams(C.Structure):
    _fields_ = [
        ("metric", C.c_int32),
        ("knn_precision", C.c_int32),
        ("warm_start", C.c_int32),
        ("warm_epochs", C.c_int32),
        ("mlp_subsample", C.c_int32),
        ("device_id", C.c_int32),
    ]


class _Result(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("k", C.c_int32),
        ("n_layers", C.c_int32),
        ("dim", C.c_int32),
        ("best_layer", C.c_int32),
        ("cluster_layers", C.POINTER(C.c_int64)),
        ("membership_strengths", C.POINTER(C.c_float)),
        ("persistence_scores", C.POINTER(C.c_double)),
        ("nn_indices", C.POINTER(C.c_int32)),
        ("nn_distances", C.POINTER(C.c_float)),
        ("embedding", C.POINTER(C.c_float)),
        ("stage_ms", C.c_double * 8),
        ("knn_fallback_rows", C.c_int64),
        ("knn_kernel_ms", C.c_double),
        ("embed_kernel_ms", C.c_double),
        ("warm_start_used", C.c_int32),
        ("warm_start_ari", C.c_double),
        ("warm_start_mst_ratio", C.c_double),
    ]


_lib = None

# symbol -> (restype, argtypes)
_SIGS = {
    "evoc_hip_last_error": (C.c_char_p, []),
    "evoc_hip_default_params": (None, [C.POINTER(_Params)]),
    "evoc_hip_default_nn_params": (None, [C.POINTER(_NnParams)]),
    "evoc_hip_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "evoc_hip_release": (C.c_int, []),
    "evoc_hip_run": (C.c_int, [f32p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, vp, vp,
                               C.c_int32, C.POINTER(_Params), C.POINTER(_NnParams), C.c_uint64,
                               C.c_int32, C.POINTER(C.POINTER(_Result))]),
    "evoc_hip_run_device": (C.c_int, [vp, C.c_int64, C.c_int64, C.c_int64, C.POINTER(_Params),
                                      C.POINTER(_NnParams), C.c_uint64, C.c_int32,
                                      C.POINTER(C.POINTER(_Result))]),
    "evoc_hip_result_free": (None, [C.POINTER(_Result)]),
    "evoc_hip_knn": (C.c_int, [f32p, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                               i32p, f32p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "evoc_hip_knn_last_stats": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "evoc_hip_knn_tile_pass": (C.c_int, [f32p, C.c_int64, C.c_int64, C.c_int32, C.c_int32,
                                         C.c_int32, i32p, f32p, f32p, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_float)]),
    "evoc_hip_fuzzy": (C.c_int, [i32p, f32p, C.c_int64, C.c_int32, C.c_float, C.c_int32, f32p,
                                 f32p, i64p, i32p, f32p, C.POINTER(C.c_int64)]),
    "evoc_hip_embed": (C.c_int, [i64p, i32p, f32p, C.c_int64, C.c_int32, C.c_int32, C.c_float,
                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, vp,
                                 C.c_uint64, f32p]),
    "evoc_hip_embed_noalloc": (C.c_int, [vp, vp, vp, C.c_int64, C.c_int64, C.c_int32,
                                         C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_int32, C.c_uint64, C.c_int32,
                                         vp]),
    "evoc_hip_label_prop_init": (C.c_int, [i64p, i32p, f32p, C.c_int64, C.c_int32, vp,
                                           C.c_int64, C.c_uint64, f32p]),
    "evoc_hip_core_sq": (C.c_int, [f32p, C.c_int64, C.c_int32, C.c_int32, f32p]),
    "evoc_hip_mst": (C.c_int, [f32p, C.c_int64, C.c_int32, C.c_int32, i32p, i32p, f32p]),
    "evoc_hip_condense_tree": (C.c_int, [i64p, i64p, f32p, i64p, C.c_int64, C.c_int64, i64p,
                                         i64p, f32p, i64p, C.POINTER(C.c_int64)]),
    "evoc_hip_cluster_layers": (C.c_int, [f32p, C.c_int64, C.c_int32, C.c_int32, C.c_int64,
                                          C.c_double, C.c_int64, i64p, f32p, f64p,
                                          C.POINTER(C.c_int64)]),
    "evoc_hip_search_n_clusters": (C.c_int, [f32p, C.c_int64, C.c_int32, C.c_int32, C.c_int64,
                                             i64p, f32p]),
    "evoc_hip_mlp_loss_grads": (C.c_int, [f32p, f32p, C.c_int32, C.c_int32, C.c_int32] + [f32p] * 12
                                + [f32p, C.POINTER(C.c_double)]),
}


def lib():
    """Loads libevoc_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EvocError(f"{LIB_PATH} not built: run `python evoc-rs_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        msg = lib().evoc_hip_last_error()
        raise EvocError(f"status {status}: {msg.decode() if msg else ''}")


def device_count():
    c = C.c_int32(0)
    _check(lib().evoc_hip_device_count(C.byref(c)))
    return c.value


def release():
    _check(lib().evoc_hip_release())


# ---------------------------------------------------------------- parameters
@dataclasses.dataclass
class EvocParams:
    """Mirrors ``EvocParams<T>`` (src/lib.rs:52-99) with its defaults."""
    n_neighbours: int = 15
    noise_level: float = 0.5
    n_epochs: int = 50
    embedding_dim: int = None
    neighbour_scale: float = 1.0
    symmetrise: bool = True
    min_samples: int = 5
    base_min_cluster_size: int = 5
    approx_n_clusters: int = None
    min_similarity_threshold: float = 0.2
    max_layers: int = 10

    def _c(self):
        return _Params(self.n_neighbours, self.noise_level, self.n_epochs,
                       self.embedding_dim or 0, self.neighbour_scale, int(self.symmetrise),
                       self.min_samples, self.base_min_cluster_size,
                       self.approx_n_clusters or 0, self.min_similarity_threshold,
                       self.max_layers)


@dataclasses.dataclass
class NearestNeighbourParams:
    """``dist_metric`` of NearestNeighbourParamsEvoc plus this library's device knobs."""
    dist_metric: str = "euclidean"
    knn_precision: str = "auto"
    warm_start: bool = False
    warm_epochs: int = 0
    mlp_subsample: int = 0
    device_id: int = 0

    def _c(self):
        return _NnParams(METRICS[self.dist_metric], PRECISIONS[self.knn_precision],
                         int(self.warm_start), self.warm_epochs, self.mlp_subsample,
                         self.device_id)


class _Owner:
    """Frees a library result once no array viewing its buffers is left."""

    def __init__(self, ptr):
        self.ptr = ptr
        self._free = lib().evoc_hip_result_free

    def __del__(self):
        if self.ptr:
            self._free(self.ptr)
            self.ptr = None


class _View:
    """Array interface over one buffer of a library result; keeps the result alive."""

    def __init__(self, owner, addr, shape, typestr):
        self._owner = owner
        self.__array_interface__ = {"data": (addr, False), "shape": shape, "typestr": typestr,
                                    "version": 3}


class EvocResult:
    """Mirrors ``EvocResult<T>`` (src/lib.rs:106-161). The arrays view the library's
    result buffers (no copy); the buffers are freed with the last array."""

    def __init__(self, ptr):
        r = ptr.contents
        n, k, nl, dim = r.n, r.k, r.n_layers, r.dim
        own = _Owner(ptr)

        def view(field, shape, typestr, empty_dtype):
            addr = C.cast(getattr(r, field), C.c_void_p).value
            if not addr or 0 in shape:
                return np.zeros(shape, empty_dtype)
            return np.asarray(_View(own, addr, shape, typestr))

        self.cluster_layers = view("cluster_layers", (nl, n), "<i8", np.int64)
        self.membership_strengths = view("membership_strengths", (nl, n), "<f4", np.float32)
        self.persistence_scores = np.ctypeslib.as_array(r.persistence_scores, (nl,)).copy() if nl else np.zeros(0)
        self.nn_indices = view("nn_indices", (n, k), "<i4", np.int32)
        self.nn_distances = view("nn_distances", (n, k), "<f4", np.float32)
        self.embedding = view("embedding", (n, dim), "<f4", np.float32)
        self.best_layer = r.best_layer
        self.stage_ms = dict(zip(["knn", "fuzzy", "init", "embed", "mst", "layers", "total",
                                  "upload"], list(r.stage_ms)))
        self.knn_fallback_rows = r.knn_fallback_rows
        self.knn_kernel_ms = r.knn_kernel_ms
        self.embed_kernel_ms = r.embed_kernel_ms
        self.warm_start_used = bool(r.warm_start_used)
        self.warm_start_ari = r.warm_start_ari
        self.warm_start_mst_ratio = r.warm_start_mst_ratio

    def best_labels(self):
        return self.cluster_layers[self.best_layer]

    def best_strengths(self):
        return self.membership_strengths[self.best_layer]

    def n_clusters(self):
        lab = self.best_labels()
        return int(max(lab.max(initial=-1) + 1, 0))


def _check_precomputed(precomputed_knn, n):
    """PreComputedKnn (src/prelude.rs:85): an (n, k) index array (-1 pads a short row) and
    an (n, k) distance array. The library reads n * k entries of each, so the shapes are
    checked here; real entries need finite, non-negative distances."""
    if len(precomputed_knn) != 2:
        raise ValueError("precomputed_knn must be a pair (indices, distances)")
    pi = np.asarray(precomputed_knn[0])
    pd = np.asarray(precomputed_knn[1])
    if pi.ndim != 2 or pi.shape[0] != n or pi.shape[1] < 1:
        raise ValueError(f"precomputed kNN indices must have shape ({n}, k), got {pi.shape}")
    if pd.shape != pi.shape:
        raise ValueError(f"precomputed kNN distances have shape {pd.shape}, indices {pi.shape}")
    if pi.shape[1] > 64:
        raise ValueError("precomputed kNN: at most 64 neighbours per row")
    if not np.issubdtype(pi.dtype, np.integer):
        raise ValueError("precomputed kNN indices must be integers")
    if pi.size and (pi.min() < -1 or pi.max() >= n):
        raise ValueError("precomputed kNN index out of range")
    pi = np.ascontiguousarray(pi, np.int32)
    pd = np.ascontiguousarray(pd, np.float32)
    real = pi >= 0
    if not np.all(np.isfinite(pd[real])) or np.any(pd[real] < 0):
        raise ValueError("precomputed kNN distances must be finite and non-negative")
    return pi, pd


def evoc(data, params=None, nn_params=None, seed=42, verbose=0, precomputed_knn=None):
    """evoc_rs::evoc (src/lib.rs:199-352) on the GPU."""
    x = np.asarray(data)
    if x.dtype != np.float32:
        x = x.astype(np.float32)
    n, d = x.shape
    rs, cs = x.strides[0] // 4, x.strides[1] // 4
    base = x
    if x.strides[0] % 4 or x.strides[1] % 4 or not (x.flags.c_contiguous or x.flags.f_contiguous):
        base = np.ascontiguousarray(x)
        rs, cs = d, 1
    elif x.flags.f_contiguous and not x.flags.c_contiguous:
        rs, cs = 1, n
    p = (params or EvocParams())._c()
    q = (nn_params or NearestNeighbourParams())._c()
    pi = pd = None
    pk = 0
    if precomputed_knn is not None:
        pi, pd = _check_precomputed(precomputed_knn, n)
        pk = pi.shape[1]
    out = C.POINTER(_Result)()
    flat = base.ravel(order="K")  # memory order; strides describe the layout
    _check(lib().evoc_hip_run(flat, n, d, rs, cs,
                              None if pi is None else pi.ctypes.data_as(vp),
                              None if pd is None else pd.ctypes.data_as(vp), pk,
                              C.byref(p), C.byref(q), seed, verbose, C.byref(out)))
    return EvocResult(out)


def evoc_device(dev_ptr, n, d, ld=None, params=None, nn_params=None, seed=42, verbose=0):
    """evoc() on an n x d row-major float32 matrix already resident in device memory."""
    p = (params or EvocParams())._c()
    q = (nn_params or NearestNeighbourParams())._c()
    out = C.POINTER(_Result)()
    _check(lib().evoc_hip_run_device(C.c_void_p(dev_ptr), n, d, ld or d, C.byref(p), C.byref(q),
                                     seed, verbose, C.byref(out)))
    return EvocResult(out)


# ---------------------------------------------------------------- stage entry points
def knn(x, k, metric="euclidean", precision="auto", return_stats=False):
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    idx = np.empty((n, k), np.int32)
    dist = np.empty((n, k), np.float32)
    fb = C.c_int64(0)
    ms = C.c_double(0)
    _check(lib().evoc_hip_knn(x, n, d, METRICS[metric], k, PRECISIONS[precision], idx, dist,
                              C.byref(fb), C.byref(ms)))
    if return_stats:
        st = knn_last_stats()
        return idx, dist, dict(fallback_rows=fb.value, kernel_ms=ms.value,
                               kernel_flop=st["kernel_flop"], symmetric=st["symmetric"],
                               escalated_rows=st["escalated_rows"])
    return idx, dist


def knn_tile_pass(x, k, metric="cosine", precision="fp8"):
    """The first tile pass alone (evoc_hip_knn_tile_pass): raw candidate lists (n x 32: two
    column halves of 16, ascending by key), their keys, the per-row key scales and the
    certificate's relative error term of the pass."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    idx = np.empty((n, 32), np.int32)
    key = np.empty((n, 32), np.float32)
    scale = np.empty(n, np.float32)
    nc = C.c_int32(0)
    gam = C.c_float(0)
    _check(lib().evoc_hip_knn_tile_pass(x, n, d, METRICS[metric], k, PRECISIONS[precision], idx,
                                        key, scale, C.byref(nc), C.byref(gam)))
    return idx, key, scale, gam.value


def knn_last_stats():
    """The calling thread's last kNN search: tile-kernel flop and time, symmetric sweep,
    rows searched again with 64 candidates after a failed certificate."""
    fl, ms, sym, esc = C.c_double(0), C.c_double(0), C.c_int32(0), C.c_int64(0)
    _check(lib().evoc_hip_knn_last_stats(C.byref(fl), C.byref(ms), C.byref(sym), C.byref(esc)))
    return dict(kernel_flop=fl.value, kernel_ms=ms.value, symmetric=bool(sym.value),
                escalated_rows=esc.value)


def fuzzy_graph(idx, dist, effective_k=None, symmetrise=True):
    idx = np.ascontiguousarray(idx, np.int32)
    dist = np.ascontiguousarray(dist, np.float32)
    n, k = idx.shape
    eff = float(k) if effective_k is None else float(effective_k)
    sig = np.empty(n, np.float32)
    rho = np.empty(n, np.float32)
    indptr = np.empty(n + 1, np.int64)
    ind = np.empty(max(2 * n * k, 1), np.int32)
    w = np.empty(max(2 * n * k, 1), np.float32)
    nnz = C.c_int64(0)
    _check(lib().evoc_hip_fuzzy(idx, dist, n, k, eff, int(symmetrise), sig, rho, indptr, ind, w,
                                C.byref(nnz)))
    return indptr, ind[:nnz.value].copy(), w[:nnz.value].copy(), sig, rho


def embedding(indptr, ind, w, dim, n_epochs=50, noise_level=0.5, negative_sample_rate=1.0,
              initial_alpha=0.1, beta1=0.5, beta2=0.9, eps=1e-7, init=None, seed=42):
    indptr = np.ascontiguousarray(indptr, np.int64)
    n = indptr.shape[0] - 1
    out = np.empty((n, dim), np.float32)
    ia = None if init is None else np.ascontiguousarray(init, np.float32)
    _check(lib().evoc_hip_embed(indptr, np.ascontiguousarray(ind, np.int32),
                                np.ascontiguousarray(w, np.float32), n, dim, n_epochs,
                                noise_level, negative_sample_rate, initial_alpha, beta1, beta2,
                                eps, None if ia is None else ia.ctypes.data_as(vp), seed, out))
    return out


def embedding_noalloc(d_indptr, d_indices, d_weights, n, nnz, dim, d_emb, n_epochs=50,
                      noise_level=0.5, negative_sample_rate=1.0, initial_alpha=0.1, beta1=0.5,
                      beta2=0.9, eps=1e-7, random_init=False, seed=42, device_id=0):
    """evoc_hip_embed_noalloc: the epochs on a symmetric column-sorted CSR already in device
    memory (integer addresses), d_emb updated in place; allocates no device memory."""
    _check(lib().evoc_hip_embed_noalloc(d_indptr, d_indices, d_weights, n, nnz, dim, n_epochs,
                                        noise_level, negative_sample_rate, initial_alpha, beta1,
                                        beta2, eps, int(random_init), seed, device_id, d_emb))


def label_prop_init(indptr, ind, w, dim, data=None, seed=42):
    indptr = np.ascontiguousarray(indptr, np.int64)
    n = indptr.shape[0] - 1
    da = None if data is None else np.ascontiguousarray(data, np.float32)
    out = np.empty((n, dim), np.float32)
    _check(lib().evoc_hip_label_prop_init(indptr, np.ascontiguousarray(ind, np.int32),
                                          np.ascontiguousarray(w, np.float32), n, dim,
                                          None if da is None else da.ctypes.data_as(vp),
                                          0 if da is None else da.shape[1], seed, out))
    return out


def core_sq(emb, min_samples):
    emb = np.ascontiguousarray(emb, np.float32)
    n, dim = emb.shape
    out = np.empty(n, np.float32)
    _check(lib().evoc_hip_core_sq(emb, n, dim, min_samples, out))
    return out


def mst(emb, min_samples):
    emb = np.ascontiguousarray(emb, np.float32)
    n, dim = emb.shape
    m = max(n - 1, 1)
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    w = np.empty(m, np.float32)
    _check(lib().evoc_hip_mst(emb, n, dim, min_samples, u, v, w))
    ne = max(n - 1, 0)
    return u[:ne], v[:ne], w[:ne]


def condense_tree(left, right, distance, size, n, min_cluster_size):
    """condense_tree (src/clustering/condensed_tree.rs:127-248) of n - 1 linkage rows on the
    device: (parent, child, lambda, child_size) in the reference's row order."""
    nr = max(n - 1, 0)
    left = np.ascontiguousarray(left, np.int64)
    right = np.ascontiguousarray(right, np.int64)
    distance = np.ascontiguousarray(distance, np.float32)
    size = np.ascontiguousarray(size, np.int64)
    if not (left.shape == right.shape == distance.shape == size.shape == (nr,)):
        raise ValueError(f"condense_tree: need {nr} rows in every array")
    if nr == 0:  # keep valid pointers for empty inputs
        left, right, size = (np.zeros(1, np.int64) for _ in range(3))
        distance = np.zeros(1, np.float32)
    cap = max(2 * n, 1)
    P = np.empty(cap, np.int64)
    Ch = np.empty(cap, np.int64)
    lam = np.empty(cap, np.float32)
    CS = np.empty(cap, np.int64)
    m = C.c_int64(0)
    _check(lib().evoc_hip_condense_tree(left, right, distance, size, n, min_cluster_size, P, Ch,
                                        lam, CS, C.byref(m)))
    k = m.value
    return P[:k], Ch[:k], lam[:k], CS[:k]


def cluster_layers(emb, min_samples=5, base_min_cluster_size=5, min_similarity_threshold=0.2,
                   max_layers=10):
    emb = np.ascontiguousarray(emb, np.float32)
    n, dim = emb.shape
    ml = max(max_layers, 1)
    labels = np.empty((ml, max(n, 1)), np.int64)
    st = np.empty((ml, max(n, 1)), np.float32)
    pers = np.empty(ml, np.float64)
    nl = C.c_int64(0)
    _check(lib().evoc_hip_cluster_layers(emb, n, dim, min_samples, base_min_cluster_size,
                                         min_similarity_threshold, max_layers, labels, st, pers,
                                         C.byref(nl)))
    k = nl.value
    return labels[:k, :n], st[:k, :n], pers[:k]


def search_for_n_clusters(emb, min_samples, target_k):
    emb = np.ascontiguousarray(emb, np.float32)
    n, dim = emb.shape
    labels = np.empty(n, np.int64)
    st = np.empty(n, np.float32)
    _check(lib().evoc_hip_search_n_clusters(emb, n, dim, min_samples, target_k, labels, st))
    return labels, st


def mlp_loss_grads(x, t, weights, biases):
    """One forward/backward pass of the warm-start MLP (bf16 MFMA, fp32 accumulation).
    weights[l]: (out_l, in_l) float32, biases[l]: (out_l,). Returns (loss, y, gw, gb)."""
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    rows, d = x.shape
    out = t.shape[1]
    ws = [np.ascontiguousarray(w, np.float32) for w in weights]
    bs = [np.ascontiguousarray(b, np.float32) for b in biases]
    gw = [np.empty_like(w) for w in ws]
    gb = [np.empty_like(b) for b in bs]
    y = np.empty((rows, out), np.float32)
    loss = C.c_double(0)
    _check(lib().evoc_hip_mlp_loss_grads(x, t, rows, d, out, ws[0], bs[0], ws[1], bs[1], ws[2],
                                         bs[2], gw[0], gb[0], gw[1], gb[1], gw[2], gb[2], y,
                                         C.byref(loss)))
    return loss.value, y, gw, gb
