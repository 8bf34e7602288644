"""GPU: device-resident searches enqueued on caller streams.

Bar: searches from two host threads on two different streams, with different
batch sizes and without waiting between calls, return exactly what the
host-buffer search returns for the same queries (ids, counts, distance bits).
A workspace whose last batch is still queued on one stream must not be handed
to a batch on another stream.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, PREC_EXACT, PREC_MIXED

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.12):
    centers = np.random.default_rng(99).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[np.arange(n) % n_centers] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("precision", [PREC_EXACT, PREC_MIXED])
def test_two_threads_two_streams_match_host_search(precision):
    import torch
    x = clustered(20000, 64, 40, seed=1)
    ix = build_index(x, "l2", None, n_postings=96, rerank_size=200, seed=0)
    dev = DeviceIndex(ix)
    k, n_probe, iters = 10, 12, 6
    sizes = {0: [700, 90, 1500, 33, 900, 250], 1: [64, 2000, 5, 1200, 300, 1700]}
    queries = {t: [clustered(n, 64, 40, seed=100 + 10 * t + i) for i, n in enumerate(sizes[t])]
               for t in sizes}
    want = {t: [dev.search(q, k, n_probe, precision=precision, rerank_size=200 if precision else 0)
                for q in queries[t]] for t in sizes}
    got = {t: [None] * iters for t in sizes}
    errors = []
    print(f"[case] two streams precision={precision}", flush=True)

    def worker(t):
        try:
            s = torch.cuda.Stream()
            outs = []
            with torch.cuda.stream(s):
                for i, q in enumerate(queries[t]):
                    nq = q.shape[0]
                    qd = torch.from_numpy(q).cuda(non_blocking=False)
                    ids = torch.empty((nq, k), dtype=torch.int64, device="cuda")
                    dd = torch.empty((nq, k), dtype=torch.float32, device="cuda")
                    cnt = torch.empty((nq,), dtype=torch.int32, device="cuda")
                    s.wait_stream(torch.cuda.default_stream())
                    dev.search_device(qd.data_ptr(), nq, 64, k, n_probe, ids.data_ptr(),
                                      dd.data_ptr(), cnt.data_ptr(), stream=s.cuda_stream,
                                      precision=precision,
                                      rerank_size=200 if precision else 0)
                    outs.append((qd, ids, dd, cnt))
            s.synchronize()
            for i, (_, ids, dd, cnt) in enumerate(outs):
                got[t][i] = (ids.cpu().numpy().view(np.uint64), dd.cpu().numpy(),
                             cnt.cpu().numpy().astype(np.uint32))
        except Exception as e:  # reported by the main thread
            errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in sizes]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for t in sizes:
        for i in range(iters):
            w_ids, w_d, w_c = want[t][i]
            g_ids, g_d, g_c = got[t][i]
            np.testing.assert_array_equal(g_c, w_c, err_msg=f"thread {t} call {i}")
            np.testing.assert_array_equal(g_ids, w_ids, err_msg=f"thread {t} call {i}")
            np.testing.assert_array_equal(g_d.view(np.uint32), w_d.view(np.uint32),
                                          err_msg=f"thread {t} call {i}")
