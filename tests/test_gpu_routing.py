"""GPU: insert routing, `nearest(v, 1)` per vector (src/spfresh.rs:437-444,
nearest :358-372) through spf_nearest_postings.

Bar: the block index equals the oracle's nearest posting (ties -> lower block
index) and the distance is bit-identical to the oracle's probe distance, for all
metrics; on integer data with many exact ties the winner is the lowest block.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, build_index

pytestmark = pytest.mark.gpu


def clustered(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 2, size=(30, d))
    return (c[rng.integers(0, 30, n)] + rng.normal(0, 1, size=(n, d))).astype(np.float32)


def check(ix, v):
    dev = DeviceIndex(ix)
    blocks, dd = dev.nearest_postings(v)
    o = oracle.Oracle(ix)
    for i in range(v.shape[0]):
        posts, pd = o.probe(v[i], 1)
        assert blocks[i] == ix.block_ids[posts[0]], i
        assert np.float32(dd[i]).view(np.uint32) == np.float32(pd[0]).view(np.uint32), i


@pytest.mark.parametrize("metric", ["l2", "l2_root", "cosine", "dot"])
@pytest.mark.parametrize("dim", [17, 64, 128])
def test_nearest_posting_matches_oracle(metric, dim):
    x = clustered(3000, dim, 1)
    ix = build_index(x, metric, None, n_postings=60, seed=0)
    ix.block_ids = (ix.block_ids * 7 + 3).astype(np.uint32)   # sparse block indices
    print(f"[case] routing metric={metric} dim={dim}", flush=True)
    check(ix, clustered(500, dim, 2))


def test_integer_ties_take_the_lower_block():
    rng = np.random.default_rng(3)
    x = rng.integers(0, 4, size=(2000, 8)).astype(np.float32)
    ix = build_index(x, "l2", None, n_postings=40, seed=0)
    # duplicate centroids in reversed block order: every tie must go to the lower block
    ix.centroids[1::2] = ix.centroids[0::2][: ix.centroids[1::2].shape[0]]
    ix.block_ids = ix.block_ids[::-1].copy()
    check(ix, rng.integers(0, 4, size=(300, 8)).astype(np.float32))


def test_many_vectors_against_fp64():
    x = clustered(20000, 32, 4)
    ix = build_index(x, "l2", "int8", n_postings=256, seed=0)
    v = clustered(50000, 32, 5)
    blocks, dd = DeviceIndex(ix).nearest_postings(v)
    c = ix.centroids.astype(np.float64)
    v64 = v.astype(np.float64)
    d2 = np.maximum((v64 ** 2).sum(1)[:, None] + (c ** 2).sum(1)[None, :] - 2.0 * v64 @ c.T, 0.0)
    best = d2.argmin(1)
    srt = np.sort(d2, 1)
    clear = (srt[:, 1] - srt[:, 0]) > 1e-4 * srt[:, 0] + 1e-6
    assert (blocks[clear] == ix.block_ids[best[clear]]).all()
    np.testing.assert_allclose(dd, srt[:, 0], rtol=1e-5, atol=1e-5)
