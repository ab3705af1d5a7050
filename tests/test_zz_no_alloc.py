"""evoc_hip_embed_noalloc (the epochs with their scratch in a static device array) against
evoc_hip_embed on the same graphs: bit-identical layouts, on consecutive calls whose
sizes shrink (the scratch keeps the last call's contents), with random and given initial
layouts, on graphs with wavefront and workgroup hub nodes."""
import ctypes as C

import numpy as np
import pytest

import datagen
import evoc_hip as H
import oracle as O
from devbuf import DeviceArray

pytestmark = pytest.mark.gpu


def _fuzzy(n, d, seed):
    x, _ = datagen.gaussian_mixture(n, d, 10, 4.0, 1.0, seed=seed)
    idx, dist = O.knn_exhaustive(x, 15)
    ip, ind, w, _, _ = O.fuzzy_graph(idx, dist, 15.0)
    return ip, ind, w


def _hubs(n, hub, seed):
    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, size=4, replace=False):
            if int(j) != i:
                adj[i].add(int(j))
                adj[int(j)].add(i)
    for h in (0, 5):
        for j in rng.choice(n, size=hub, replace=False):
            if int(j) != h:
                adj[h].add(int(j))
                adj[int(j)].add(h)
    rows = [sorted(a) for a in adj]
    ip = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    ind = np.concatenate(rows).astype(np.int32)
    wt = {}
    w = np.empty(len(ind), np.float32)
    for i, r in enumerate(rows):
        for q, j in enumerate(r):
            key = (min(i, j), max(i, j))
            if key not in wt:
                wt[key] = np.float32(rng.uniform(0.05, 1.0))
            w[ip[i] + q] = wt[key]
    return ip, ind, w


def _noalloc(ip, ind, w, dim, init, **kw):
    n = len(ip) - 1
    d_ip, d_ind, d_w = DeviceArray(ip, np.int64), DeviceArray(ind, np.int32), DeviceArray(w)
    d_e = DeviceArray(init if init is not None else np.zeros((n, dim), np.float32))
    H.embedding_noalloc(d_ip.ptr.value, d_ind.ptr.value, d_w.ptr.value, n, len(ind), dim,
                        d_e.ptr.value, random_init=init is None, **kw)
    return d_e.to_host()


def test_noalloc_embedding_matches_allocating_entry_point():
    cases = [(_fuzzy(3000, 16, 1), 8, 30, 11, True), (_fuzzy(1200, 8, 2), 4, 20, 3, False),
             (_hubs(2500, 1500, 4), 8, 12, 7, True), (_hubs(700, 300, 5), 16, 15, 2, False),
             (_fuzzy(500, 8, 6), 5, 10, 9, True)]
    for (ip, ind, w), dim, epochs, seed, given in cases:
        n = len(ip) - 1
        init = np.random.default_rng(seed).uniform(-1, 1, (n, dim)).astype(np.float32) \
            if given else None
        ref = H.embedding(ip, ind, w, dim, n_epochs=epochs, init=init, seed=seed,
                          negative_sample_rate=2.0)
        got = _noalloc(ip, ind, w, dim, init, n_epochs=epochs, seed=seed,
                       negative_sample_rate=2.0)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (n, dim, given)


def test_noalloc_embedding_rejects_sizes_beyond_the_scratch():
    L = H.lib()
    for n, nnz, dim in ((65537, 10, 8), (100, 2097153, 8), (100, 10, 17)):
        st = L.evoc_hip_embed_noalloc(C.c_void_p(8), C.c_void_p(8), C.c_void_p(8), n, nnz, dim,
                                      10, 0.5, 1.0, 0.1, 0.5, 0.9, 1e-7, 0, 1, 0, C.c_void_p(8))
        assert st == 1, (n, nnz, dim)
