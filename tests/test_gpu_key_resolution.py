"""GPU: what the 16-bit candidate keys of mixed precision cost in recall on
data whose distances crowd together.

The mixed scan keys each candidate by the top 16 bits of the order-preserving
f32 key of its f16 distance (sign, 8 exponent bits, 7 mantissa bits, so
distances within about 2^-7 relative of each other can share a key), and the
pool is the `want` smallest keys, ties in slot order; the pool is then
reranked exactly.  A true neighbour is lost only when `want` or more
candidates have keys at or below its own.  On isotropic Gaussian data in 128
dimensions (no cluster structure: the squared distances of a query's
candidates lie within a few percent of each other), probing every posting so
that each query has all 20,000 points as candidates, the test measures recall@10
against FP64 for pools from k to 2,048 and checks that recall does not fall
as the pool grows (a larger pool is a superset), and that pools of 10 k and
more stay within 0.02 of the exact f32 path.  It prints the table.
"""
from __future__ import annotations

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, brute_force_f64, PREC_MIXED, PREC_EXACT

pytestmark = pytest.mark.gpu

POOLS = [10, 20, 50, 100, 200, 500, 1000, 2048]


def recall(ids, gt, k=10):
    return float(np.mean([len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist())) / k
                          for i in range(gt.shape[0])]))


@pytest.mark.parametrize("metric", ["l2", "cosine"])
def test_recall_against_pool_size_on_crowded_distances(metric):
    rng = np.random.default_rng(11)
    x = rng.normal(size=(20000, 128)).astype(np.float32)
    q = rng.normal(size=(300, 128)).astype(np.float32)
    ix = build_index(x, metric, None, n_postings=64, seed=0)
    dev = DeviceIndex(ix)
    n_probe, k = 64, 10   # every posting: recall is decided by the pool alone
    print(f"[case] crowded distances, metric={metric}", flush=True)
    gt, _ = brute_force_f64(x, q, k, ix.metric)
    e_ids, _, _ = dev.search(q, k, n_probe, precision=PREC_EXACT)
    r_exact = recall(e_ids, gt)
    rows = []
    for pool in POOLS:
        ids, _, cnt = dev.search(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool)
        assert (cnt == k).all()
        rows.append((pool, recall(ids, gt)))
    print(f"recall@10 vs fp64, exact f32 path {r_exact:.4f}; mixed by pool: "
          + ", ".join(f"{p}: {r:.4f}" for p, r in rows), flush=True)
    for (p0, r0), (p1, r1) in zip(rows, rows[1:]):
        assert r1 >= r0 - 0.005, (p0, r0, p1, r1)
    for p, r in rows:
        if p >= 10 * k:
            assert r >= r_exact - 0.02, (p, r, r_exact)
    dev.close()
