"""Seeded synthetic workloads of the exact k-NN path.

- ``mixture(n, d, seed)``: rows drawn from a mixture of 1,000 isotropic Gaussians (centres
  N(0, 1)^d, per-cluster sigma 0.3, uniform cluster labels), L2-normalised in float32. The
  centres depend only on ``centers_seed`` so base rows and held-out queries share clusters.
- ``unit_vectors(n, d, seed)``: i.i.d. Gaussian rows, normalised (the reference benchmark's
  data, bench/gpu_bench.py:25-30).
- ``zipf_groups(n, n_groups, s, seed, shift, cutoff)``: one group label per row with
  P(rank r) ~ (r + shift)^-s * exp(-(r - 1) / cutoff) (plain Zipf with shift 0, no cutoff).
- ``filtered_groups(n, seed)``: the filtered-scan workload's labels: 100 groups, Zipf exponent
  1.1 with shift 5.56 and cutoff 24.54, so the expected group sizes run from 12 % of the rows
  (rank 1) down to 0.01 % (rank 100). Plain Zipf(1.1) over 100 ranks spans 23.4 % .. 0.15 %.
- ``group_csr(groups, n_groups)``: per-group sorted row ids as CSR (offsets, ids).
- ``query_filters`` / ``query_ranges``: a random group and k per query, as ranges into the CSR.
Rows are generated in slices so float64 temporaries stay small.
"""
import numpy as np

N_CLUSTERS = 1000
SIGMA = 0.3


def _normalize_(x):
    nrm = np.sqrt(np.einsum("ij,ij->i", x, x, dtype=np.float32))
    x /= np.maximum(nrm, 1e-12)[:, None]
    return x


def centers(d, centers_seed=0, n_clusters=N_CLUSTERS):
    return np.random.default_rng([centers_seed, 7]).standard_normal((n_clusters, d), dtype=np.float32)


def mixture(n, d, seed, centers_seed=0, n_clusters=N_CLUSTERS, sigma=SIGMA, slice_rows=100_000,
            workers=1):
    """workers > 1 draws the slices in parallel threads, slice i from its own generator
    seeded (seed, 11, i): a different, equally deterministic stream than workers = 1."""
    c = centers(d, centers_seed, n_clusters)
    x = np.empty((n, d), np.float32)
    if workers <= 1:
        rng = np.random.default_rng([seed, 11])
        for i in range(0, n, slice_rows):
            m = min(slice_rows, n - i)
            lab = rng.integers(0, n_clusters, m)
            x[i:i + m] = c[lab] + np.float32(sigma) * rng.standard_normal((m, d), dtype=np.float32)
        return _normalize_(x)
    from concurrent.futures import ThreadPoolExecutor

    def fill(j):
        i = j * slice_rows
        m = min(slice_rows, n - i)
        rng = np.random.default_rng([seed, 11, j])
        lab = rng.integers(0, n_clusters, m)
        x[i:i + m] = c[lab] + np.float32(sigma) * rng.standard_normal((m, d), dtype=np.float32)
        _normalize_(x[i:i + m])

    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(fill, range((n + slice_rows - 1) // slice_rows)))
    return x


def unit_vectors(n, d, seed, slice_rows=100_000):
    rng = np.random.default_rng(seed)
    x = np.empty((n, d), np.float32)
    for i in range(0, n, slice_rows):
        x[i:i + slice_rows] = rng.standard_normal((min(slice_rows, n - i), d), dtype=np.float32)
    return _normalize_(x)


def zipf_weights(n_groups=100, s=1.1, shift=0.0, cutoff=None):
    """Probability of each rank 1..n_groups: (r + shift)^-s, times exp(-(r - 1) / cutoff)."""
    r = np.arange(1, n_groups + 1, dtype=np.float64)
    w = (r + shift) ** -s
    if cutoff:
        w *= np.exp(-(r - 1) / cutoff)
    return w / w.sum()


def zipf_groups(n, n_groups=100, s=1.1, seed=2, shift=0.0, cutoff=None):
    p = zipf_weights(n_groups, s, shift, cutoff)
    return np.random.default_rng([seed, 13]).choice(n_groups, size=n, p=p).astype(np.int32)


# The filtered workload's group law: the Zipf(1.1) rank exponent, with the rank shift and the
# exponential cutoff solved so that rank 1 holds 12 % of the rows and rank 100 holds 0.01 %.
F_GROUPS, F_S, F_SHIFT, F_CUTOFF = 100, 1.1, 5.5613, 24.5404


def filtered_groups(n, seed=2):
    return zipf_groups(n, F_GROUPS, F_S, seed, F_SHIFT, F_CUTOFF)


def group_csr(groups, n_groups):
    order = np.argsort(groups, kind="stable").astype(np.uint32)
    counts = np.bincount(groups, minlength=n_groups)
    off = np.zeros(n_groups + 1, np.uint64)
    np.cumsum(counts, out=off[1:])
    return off, order


def query_filters(nq, n_groups, seed=3, ks=(10, 100)):
    """Per-query random group (uniform over groups) and k (uniform over ks)."""
    rng = np.random.default_rng([seed, 17])
    g = rng.integers(0, n_groups, nq)
    k = np.asarray(ks, np.uint32)[rng.integers(0, len(ks), nq)]
    return k, g


def query_ranges(g, group_off):
    """[begin, end) of each query's group inside the group CSR's id array, shape (nq, 2):
    queries of one group share one range."""
    g = np.asarray(g)
    return np.stack([group_off[g], group_off[g + 1]], axis=1).astype(np.uint64)
