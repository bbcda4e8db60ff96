import numpy as np

from vecsearch_amd import datagen as dg


def test_mixture_is_seeded_unit_and_clustered():
    a = dg.mixture(2000, 32, seed=0)
    b = dg.mixture(2000, 32, seed=0)
    c = dg.mixture(2000, 32, seed=1)
    assert a.dtype == np.float32 and np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.allclose(np.linalg.norm(a, axis=1), 1, atol=1e-6)
    # rows of the same cluster are much closer than random unit vectors
    cen = dg.centers(32)
    cen /= np.linalg.norm(cen, axis=1, keepdims=True)
    assert np.max(a @ cen.T, axis=1).mean() > 0.8


def test_zipf_groups_and_ranges():
    g = dg.zipf_groups(50000, 100, 1.1, seed=2)
    frac = np.bincount(g, minlength=100) / len(g)
    assert frac[0] > frac[10] > frac[99] > 0
    off, ids = dg.group_csr(g, 100)
    assert off[-1] == len(g) and np.all(np.diff(off.astype(np.int64)) >= 0)
    for grp in (0, 57, 99):
        seg = ids[off[grp]:off[grp + 1]]
        assert np.all(g[seg] == grp) and np.all(np.diff(seg.astype(np.int64)) > 0)
    ks, qg = dg.query_filters(64, 100)
    r = dg.query_ranges(qg, off)
    assert r.shape == (64, 2) and set(np.unique(ks)) <= {10, 100}
    assert np.all(r[:, 1] - r[:, 0] == np.bincount(g, minlength=100)[qg])


def test_filtered_groups_span():
    """The filtered workload's allowed sets run from about 12 % to about 0.01 % of the rows."""
    p = dg.zipf_weights(dg.F_GROUPS, dg.F_S, dg.F_SHIFT, dg.F_CUTOFF)
    assert abs(p[0] - 0.12) < 1e-5 and abs(p[-1] - 1e-4) < 1e-8 and np.all(np.diff(p) < 0)
    # with no shift and no cutoff the law is plain Zipf(1.1): 23.4 % .. 0.15 % over 100 ranks
    z = dg.zipf_weights(100, 1.1)
    assert abs(z[0] - 0.2338) < 1e-3 and abs(z[-1] - 0.001475) < 1e-5
    frac = np.bincount(dg.filtered_groups(892_800), minlength=100) / 892_800
    assert 0.11 < frac.max() < 0.13 and 0.00004 < frac.min() < 0.0002
