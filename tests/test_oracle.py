"""Pins the CPU oracle (oracle/) to the reference.

Checks, in order of strength:
  1. golden vectors written by the compiled reference (tests/golden/make_golden.py): distances
     for every tail length, search_batch at 1 and 3 threads, padding, the exact filtered scan;
  2. the compiled reference itself (oracle/_ref), when present in this container;
  3. what the mathematics fixes: the known answer on a line (test_brute_force.cpp:9-20), an
     FP64 error bound, brute force on tiny inputs, thread invariance, filter == subset search;
  4. planted mistakes: plausible wrong orders of summation fail the golden vectors.
"""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


# ---- 1. golden vectors --------------------------------------------------------------------

def test_distances_match_golden():
    g = load("distances.npz")
    for d in g["dims"]:
        a, b = g[f"a{d}"], g[f"b{d}"]
        got_dot = [oracle.dot(a[i], b[i]) for i in range(len(a))]
        got_l2 = [oracle.l2sq(a[i], b[i]) for i in range(len(a))]
        got_ip = [oracle.distance("ip", a[i], b[i]) for i in range(len(a))]
        assert np.array_equal(bits(got_dot), bits(g[f"dot{d}"])), d
        assert np.array_equal(bits(got_l2), bits(g[f"l2{d}"])), d
        assert np.array_equal(bits(got_ip), bits(g[f"ipd{d}"])), d


SEARCH_FIXTURES = sorted(f for f in os.listdir(GOLDEN) if f.startswith("search_d"))


@pytest.mark.parametrize("name", SEARCH_FIXTURES)
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_search_matches_golden(name, metric):
    g = load(name)
    x, q = (g["x"], g["q"]) if metric == "l2" else (g["xn"], g["qn"])
    k = int(g["k"])
    for th in (1, 3):
        ids, dists = oracle.search(x, q, k, metric, threads=th)
        assert np.array_equal(ids, g[f"{metric}_ids_t{th}"])
        assert np.array_equal(bits(dists), bits(g[f"{metric}_dists_t{th}"]))
    ids, dists = oracle.search(x, q[:1], x.shape[0] + 3, metric, threads=2)
    assert np.array_equal(ids, g[f"{metric}_pad_ids"])
    assert np.array_equal(bits(dists), bits(g[f"{metric}_pad_dists"]))
    assert ids[0, -1] == oracle.NONE and np.isinf(dists[0, -3:]).all()


@pytest.mark.parametrize("mask_name", ["third", "sparse", "two"])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_filtered_matches_golden(mask_name, metric):
    g = load("filtered.npz")
    ids, dists = oracle.search(g["x"], g["q"], 10, metric, threads=2, mask=g[f"mask_{mask_name}"])
    assert np.array_equal(ids, g[f"{mask_name}_{metric}_ids"])
    assert np.array_equal(bits(dists), bits(g[f"{mask_name}_{metric}_dists"]))


def test_known_answer_on_a_line():
    g = load("line.npz")
    ids, dists = oracle.search(g["x"], g["q"], 3, "l2")
    assert list(ids[0]) == [3, 4, 2] == list(g["ids"])
    assert np.array_equal(bits(dists[0]), bits(g["dists"]))


# ---- 2. the compiled reference (development container only) -------------------------------

@pytest.mark.skipif(oracle.ref() is None, reason="oracle/_ref not built (reference tree absent)")
def test_matches_compiled_reference():
    ref = oracle.ref()
    rng = np.random.default_rng(5)
    for d in list(range(1, 40)) + [64, 128, 384, 385, 768]:
        a = rng.standard_normal((20, d)).astype(np.float32)
        b = rng.standard_normal((20, d)).astype(np.float32)
        for m in ("l2", "ip"):
            got = [oracle.distance(m, a[i], b[i]) for i in range(20)]
            exp = [ref.distance(m, a[i], b[i]) for i in range(20)]
            assert np.array_equal(bits(got), bits(exp)), (d, m)
    x = rng.standard_normal((2000, 45)).astype(np.float32)
    q = rng.standard_normal((9, 45)).astype(np.float32)
    for m in ("l2", "ip"):
        i1, d1 = oracle.search(x, q, 25, m, threads=4)
        i2, d2 = ref.search(x, q, 25, m, threads=4)
        assert np.array_equal(i1, i2) and np.array_equal(bits(d1), bits(d2))


# ---- 3. mathematics -----------------------------------------------------------------------

@pytest.mark.parametrize("d", [37, 384, 768])
def test_fp64_error_bound(d):
    # Every term passes through at most m = ceil(d/16) + 4 + 8 roundings, so
    # |fl(s) - s| <= gamma_m * sum |a_i b_i| with gamma_m = m u / (1 - m u), u = 2^-24.
    rng = np.random.default_rng(d)
    a = rng.standard_normal((50, d)).astype(np.float32)
    b = rng.standard_normal((50, d)).astype(np.float32)
    m = -(-d // 16) + 12
    u = 2.0 ** -24
    gamma = m * u / (1 - m * u)
    for i in range(50):
        s = float(np.dot(a[i].astype(np.float64), b[i].astype(np.float64)))
        bound = gamma * float(np.abs(a[i].astype(np.float64) * b[i]).sum())
        assert abs(oracle.dot(a[i], b[i]) - s) <= bound
        assert abs(oracle.distance_f64("ip", a[i], b[i]) - (1 - s)) < 1e-12


def test_topk_equals_brute_force_on_tiny_inputs():
    rng = np.random.default_rng(1)
    for d in (3, 16, 19):
        x = rng.standard_normal((40, d)).astype(np.float32)
        x[7] = x[3]  # an exact tie: the lower id must come first
        q = rng.standard_normal((6, d)).astype(np.float32)
        q[0] = x[3] + 0.25
        for metric in ("l2", "ip"):
            ids, dists = oracle.search(x, q, 12, metric)
            for j in range(len(q)):
                dd = np.array([oracle.distance(metric, q[j], x[i]) for i in range(len(x))], np.float32)
                order = np.lexsort((np.arange(len(x)), dd))[:12]
                assert list(ids[j]) == list(order)
                assert np.array_equal(bits(dists[j]), bits(dd[order]))
            assert list(ids[0]).index(3) < list(ids[0]).index(7)


def test_threads_agree():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3000, 64)).astype(np.float32)
    q = rng.standard_normal((50, 64)).astype(np.float32)
    i1, d1 = oracle.search(x, q, 10, "l2", threads=1)
    for th in (2, 4, 7):
        i2, d2 = oracle.search(x, q, 10, "l2", threads=th)
        assert np.array_equal(i1, i2) and np.array_equal(bits(d1), bits(d2))


def test_filter_equals_search_over_subset():
    # test_bindings.py:76-90: exact filtered results == brute force over x[mask] mapped back.
    rng = np.random.default_rng(3)
    x = rng.standard_normal((3000, 64)).astype(np.float32)
    q = rng.standard_normal((30, 64)).astype(np.float32)
    mask = np.zeros(len(x), bool)
    mask[::3] = True
    ids, dists = oracle.search(x, q, 10, "l2", mask=mask)
    gt, gd = oracle.search(x[mask], q, 10, "l2")
    assert np.array_equal(ids, np.flatnonzero(mask)[gt])
    assert np.array_equal(bits(dists), bits(gd))
    ids, _ = oracle.search(x, q, 10, "l2", mask=np.zeros(len(x), bool))
    assert (ids == oracle.NONE).all()


def test_lists_equal_masked_search():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2000, 24)).astype(np.float32)
    q = rng.standard_normal((12, 24)).astype(np.float32)
    masks = [rng.random(2000) < p for p in (0.5, 0.05, 0.001, 0.0)]
    ids_all = np.concatenate([np.flatnonzero(m) for m in masks]).astype(np.uint32)
    starts = np.cumsum([0] + [int(m.sum()) for m in masks])
    qm = rng.integers(0, len(masks), len(q))
    ranges = np.stack([starts[qm], starts[qm + 1]], 1)
    ks = rng.choice([1, 5, 17], len(q)).astype(np.uint32)
    ids, dists = oracle.search_lists(x, q, ks, ranges, ids_all, "ip", threads=3, kmax=17)
    for j in range(len(q)):
        ei, ed = oracle.search(x, q[j], int(ks[j]), "ip", mask=masks[qm[j]])
        assert np.array_equal(ids[j, :ks[j]], ei[0]) and np.array_equal(bits(dists[j, :ks[j]]), bits(ed[0]))
        assert (ids[j, ks[j]:] == oracle.NONE).all() and np.isinf(dists[j, ks[j]:]).all()


def test_lists_unfiltered_rows_equal_plain_search():
    # a range beginning with ALL_ROWS scans every row: the same answer as the unfiltered search
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1500, 20)).astype(np.float32)
    q = rng.standard_normal((6, 20)).astype(np.float32)
    allow = np.array([2, 40, 41, 900], np.uint32)
    ranges = np.array([[oracle.ALL_ROWS, 0], [0, 4], [oracle.ALL_ROWS, 7], [1, 3], [oracle.ALL_ROWS, 0],
                       [0, 0]], np.uint64)
    ks = np.array([10, 3, 1, 2, 0, 5], np.uint32)
    ids, dists = oracle.search_lists(x, q, ks, ranges, allow, "l2", threads=2, kmax=10)
    for j in (0, 2):
        ei, ed = oracle.search(x, q[j], int(ks[j]), "l2")
        assert np.array_equal(ids[j, :ks[j]], ei[0]) and np.array_equal(bits(dists[j, :ks[j]]), bits(ed[0]))
    assert (ids[4] == oracle.NONE).all() and (ids[5] == oracle.NONE).all()
    assert set(ids[1, :3]) <= {2, 40, 41, 900}


# ---- 4. planted mistakes fail the golden vectors ------------------------------------------

def _model_dot(a, b, mistake=None):
    """numpy model of the oracle's dot with one optional planted mistake."""
    f32 = np.float32
    d = len(a)
    acc = np.zeros(16, f32)
    i = 0

    def fma(x, y, z):
        return f32(np.float64(x) * np.float64(y) + np.float64(z))

    if mistake == "sequential":
        r = f32(0)
        for t in range(d):
            r = fma(a[t], b[t], r)
        return r
    while i + 16 <= d:
        for j in range(16):
            acc[j] = fma(a[i + j], b[i + j], acc[j])
        i += 16
    while i + 8 <= d:
        for j in range(8):
            acc[j] = fma(a[i + j], b[i + j], acc[j])
        i += 8
    if mistake == "pairwise_tree":
        v = acc[:8] + acc[8:]
        r = f32(((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7])))
    else:
        v = acc[:8] + acc[8:]
        w = v[:4] + v[4:]
        r = f32(f32(w[0] + w[1]) + f32(w[2] + w[3]))
    if mistake != "fused_tail" and i + 4 <= d:
        for _ in range(4):
            r = f32(r + f32(a[i] * b[i]))
            i += 1
    while i < d:
        r = fma(a[i], b[i], r)
        i += 1
    return r


def test_model_reproduces_golden_and_mistakes_are_caught():
    g = load("distances.npz")
    caught = {"sequential": False, "pairwise_tree": False, "fused_tail": False}
    for d in g["dims"]:
        a, b, exp = g[f"a{d}"], g[f"b{d}"], bits(g[f"dot{d}"])
        assert np.array_equal(bits([_model_dot(a[i], b[i]) for i in range(len(a))]), exp), d
        for m in caught:
            if not np.array_equal(bits([_model_dot(a[i], b[i], m) for i in range(len(a))]), exp):
                caught[m] = True
    assert all(caught.values()), caught
