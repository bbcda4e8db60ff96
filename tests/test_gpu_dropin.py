"""The drop-in boundary on the GPU: the reference's call forms, concurrent callers, and the
reference's own Python module built over this engine."""
import os
import subprocess
import sys
import threading

import numpy as np
import pytest

import oracle
import vecsearch_amd as vs
from vecsearch_amd import datagen as dg

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN_DIR = os.path.join(ROOT, "oracle", "_ref", "dropin")


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(bits(a[1]), bits(b[1]))


def test_reference_module_over_engine():
    """The reference's unchanged python/bindings.cpp, linked to this engine, passes the
    reference's GPU test cases bit for bit against the reference's CPU index."""
    mods = [f for f in os.listdir(DROPIN_DIR) if f.startswith("_vecsearch")] if os.path.isdir(DROPIN_DIR) else []
    if not mods:
        pytest.skip("oracle/_ref/dropin was not built (needs the reference sources)")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "dropin", "child.py"),
                        os.path.join(DROPIN_DIR, mods[0])], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0 and "dropin ok" in r.stdout, r.stderr[-3000:]


@pytest.fixture(scope="module")
def small():
    x = dg.mixture(60_000, 128, seed=40)
    q = dg.mixture(64, 128, seed=41)
    idx = vs.GpuBruteForceIndex(128, "cosine")
    idx.add(x)
    return x, q, idx


def test_positional_method_and_filter(small):
    x, q, idx = small
    mask = np.random.default_rng(42).random(len(x)) < 0.1
    same(idx.search(q, 10, "skinny"), idx.search(q, 10, method="skinny"))  # reference order
    same(idx.search(q, 10, "tiled"), oracle.search(x, q, 10, "ip", threads=8))
    want = oracle.search(x, q, 10, "ip", threads=8, mask=mask)
    got = idx.search(q, 10, mask)  # search(queries, k, filter)
    np.testing.assert_array_equal(got[0], want[0].astype(np.int64))
    same(got, idx.search(q, 10, filter=mask))
    same(got, idx.search(q, 10, "skinny", mask))
    with pytest.raises(ValueError):
        idx.search(q, 10, mask, filter=mask)
    with pytest.raises(ValueError):
        idx.search(q, 10, "fast")


def test_concurrent_callers_overlap_and_match_alone(small):
    """Callers are admitted side by side (each on its own lane), not serialised, and each gets
    the bits it gets alone."""
    x, q, idx = small
    mask = np.random.default_rng(43).random(len(x)) < 0.4
    jobs = [(q[i:i + 8], m, f) for i in range(0, 64, 8) for m, f in (("skinny", None), ("auto", None),
                                                                     ("naive", mask), ("tiled", None))]
    alone = [idx.search(qb, 10, m, f) for qb, m, f in jobs]
    out = [None] * len(jobs)
    errors = []
    start = threading.Barrier(4)

    def worker(t):
        try:
            start.wait()
            for rep in range(5):
                for j in range(t, len(jobs), 4):
                    qb, m, f = jobs[j]
                    out[j] = idx.search(qb, 10, m, f)
        except Exception as e:  # noqa: BLE001 (reported below)
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors, errors
    for a, b in zip(out, alone):
        same(a, b)
    assert idx.get_param("search_lanes_peak") >= 2


def test_search_device_on_two_streams(small):
    """Two device searches enqueued back to back on different streams, with no wait between
    them, use one workspace in turn: the second stream waits for the first's work."""
    torch = pytest.importorskip("torch")
    x, q, idx = small
    qa = torch.from_numpy(q[:32]).cuda()
    qb = torch.from_numpy(q[32:]).cuda()
    outs = []
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for qq, s in ((qa, sa), (qb, sb)):
        ids = torch.empty((32, 10), dtype=torch.int32, device="cuda")
        dd = torch.empty((32, 10), dtype=torch.float32, device="cuda")
        idx.search_device(qq.data_ptr(), 32, 10, ids.data_ptr(), dd.data_ptr(), stream=s.cuda_stream,
                          method="skinny")
        outs.append((ids, dd))
    torch.cuda.synchronize()
    for (ids, dd), lo in zip(outs, (0, 32)):
        want = idx.search(q[lo:lo + 32], 10, "skinny")
        np.testing.assert_array_equal(ids.cpu().numpy().astype(np.int64), want[0])
        np.testing.assert_array_equal(bits(dd.cpu().numpy()), bits(want[1]))
