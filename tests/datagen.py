"""Seeded synthetic inputs for the tests and the benchmark (no external data)."""
import numpy as np


def make_blobs(n_per_cluster, n_clusters, dim, separation, spread, seed):
    """Restates tests/commons/mod.rs:11-40 of the reference (centres at
    (c * separation, 0, ...), Gaussian noise of std `spread`); numpy's generator
    replaces SmallRng, so the draws differ but the distribution is the same."""
    rng = np.random.default_rng(seed)
    data = rng.normal(0.0, spread, size=(n_per_cluster * n_clusters, dim))
    labels = np.repeat(np.arange(n_clusters), n_per_cluster)
    data[:, 0] += labels * separation
    return data, labels


def gaussian_mixture(n, dim, n_centres, centre_sd, sd, seed):
    """Isotropic Gaussian mixture with equal weights, shuffled."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(0.0, centre_sd, size=(n_centres, dim))
    lab = rng.integers(0, n_centres, size=n)
    x = centres[lab] + rng.normal(0.0, sd, size=(n, dim))
    return x.astype(np.float32), lab


def unit_mixture(n, dim, n_centres, sd, seed, noise_frac=0.0, zipf=None, bf16=False):
    """Centres uniform on the unit sphere; point = centre + N(0, sd^2 I), L2-normalised.
    Optional Zipf(zipf) cluster sizes, a fraction of uniform-on-sphere noise points, and
    rounding to bf16-representable values (round to nearest even)."""
    rng = np.random.default_rng(seed)
    centres = rng.normal(size=(n_centres, dim))
    centres /= np.linalg.norm(centres, axis=1, keepdims=True)
    n_noise = int(round(noise_frac * n))
    n_cl = n - n_noise
    if zipf is None:
        lab = rng.integers(0, n_centres, size=n_cl)
    else:
        p = 1.0 / np.arange(1, n_centres + 1) ** zipf
        lab = rng.choice(n_centres, size=n_cl, p=p / p.sum())
    x = np.empty((n, dim), np.float32)
    chunk = 200_000
    for s in range(0, n_cl, chunk):
        e = min(n_cl, s + chunk)
        pts = centres[lab[s:e]] + rng.normal(0.0, sd, size=(e - s, dim))
        x[s:e] = pts / np.linalg.norm(pts, axis=1, keepdims=True)
    if n_noise:
        z = rng.normal(size=(n_noise, dim))
        x[n_cl:] = z / np.linalg.norm(z, axis=1, keepdims=True)
    labels = np.concatenate([lab, -np.ones(n_noise, np.int64)])
    perm = rng.permutation(n)
    x, labels = x[perm], labels[perm]
    if bf16:
        x = to_bf16_values(x)
    return x, labels


def to_bf16_values(x):
    """Round float32 values to the nearest bf16-representable value (ties to even)."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32)


def unit_mixture_device(n, d, n_centres=200, zipf=1.1, sd=0.15, noise_frac=0.05, seed=42,
                        hierarchical=None, bf16=True):
    """The same kind of mixture as `unit_mixture`, generated on the GPU with torch (seeded;
    torch is only the allocator and RNG here). Values are rounded to bf16-representable
    float32 (bf16=False: left as computed in float32). Returns (x, labels) as device
    tensors; noise points are labelled -1.

    hierarchical=(n_parents, children_per_parent, child_spread, leaf_sd) instead draws
    parents ~ N(0, I), children = parent + child_spread * N(0, I), points =
    child + leaf_sd * N(0, I), all L2-normalised, equal-size leaves, no noise."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.empty(n, d, device="cuda", dtype=torch.float32)
    chunk = 1 << 18
    if hierarchical is not None:
        n_par, n_child, spread, leaf_sd = hierarchical
        par = torch.randn(n_par, d, device="cuda", generator=g)
        c = par.repeat_interleave(n_child, 0) + spread * torch.randn(n_par * n_child, d,
                                                                     device="cuda", generator=g)
        lab = torch.arange(n, device="cuda") % (n_par * n_child)
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            pts = c[lab[s:e]] + leaf_sd * torch.randn(e - s, d, device="cuda", generator=g)
            x[s:e] = pts / pts.norm(dim=1, keepdim=True)
        labels = lab
    else:
        c = torch.randn(n_centres, d, device="cuda", generator=g)
        c = c / c.norm(dim=1, keepdim=True)
        n_noise = int(round(noise_frac * n))
        n_cl = n - n_noise
        p = 1.0 / torch.arange(1, n_centres + 1, device="cuda", dtype=torch.float64) ** zipf
        lab = torch.multinomial(p / p.sum(), n_cl, replacement=True, generator=g)
        for s in range(0, n_cl, chunk):
            e = min(n_cl, s + chunk)
            pts = c[lab[s:e]] + sd * torch.randn(e - s, d, device="cuda", generator=g)
            x[s:e] = pts / pts.norm(dim=1, keepdim=True)
        if n_noise:
            z = torch.randn(n_noise, d, device="cuda", generator=g)
            x[n_cl:] = z / z.norm(dim=1, keepdim=True)
        labels = torch.cat([lab, torch.full((n_noise,), -1, device="cuda", dtype=lab.dtype)])
    perm = torch.randperm(n, device="cuda", generator=g)
    x = x[perm]
    x = (x.to(torch.bfloat16).to(torch.float32) if bf16 else x).contiguous()
    labels = labels[perm]
    torch.cuda.synchronize()
    return x, labels
