"""GPU parity: the HIP search path against the CPU oracle on the same index.

Bar (exact precision): identical ids in identical order, identical counts, and
distances bit-identical (the oracle shares the canonical f32 reduction
order); the stated tolerance 1e-5 relative is also asserted so a failure names
the size of a difference.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, build_index
from pydiskann_amd._lib import SpfError, SPF_EDIM

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def clustered(n, d, n_centers, seed, spread=0.12, center_seed=99):
    crng = np.random.default_rng(center_seed)
    centers = crng.uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    lab = np.arange(n) % n_centers
    return (centers[lab] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def compare(dev, ix, qs, k, n_probe, threads=8):
    print(f"[case] nq={qs.shape[0]} d={qs.shape[1]} k={k} n_probe={n_probe} metric={ix.metric} "
          f"quant={ix.quant} rerank={ix.rerank_size}", flush=True)
    ids, dd, cnt = dev.search(qs, k, n_probe)
    rid, rd, rc = oracle.Oracle(ix).search_batch(qs, k, n_probe, threads=threads)
    np.testing.assert_array_equal(cnt, rc)
    for i in range(qs.shape[0]):
        c = int(cnt[i])
        np.testing.assert_array_equal(ids[i, :c], rid[i, :c], err_msg=f"query {i}")
        np.testing.assert_allclose(dd[i, :c], rd[i, :c], rtol=REL_TOL, atol=0,
                                   err_msg=f"query {i}")
        assert np.array_equal(dd[i, :c].view(np.uint32), rd[i, :c].view(np.uint32)), \
            f"query {i}: distances not bit-identical"
        assert (ids[i, c:] == np.iinfo(np.uint64).max).all()
        assert np.isinf(dd[i, c:]).all()
    return ids, dd, cnt


@pytest.mark.parametrize("metric", ["l2", "l2_root", "cosine", "dot"])
@pytest.mark.parametrize("dim", [32, 17, 128])
def test_plain_metrics(metric, dim):
    x = clustered(3000, dim, 25, seed=1)
    q = clustered(40, dim, 25, seed=2)
    ix = build_index(x, metric, None, n_postings=48, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 8)
    compare(dev, ix, q, 10, ix.n_post)       # exhaustive probe


@pytest.mark.parametrize("quant", ["f16", "int8"])
@pytest.mark.parametrize("metric", ["l2", "cosine", "dot", "l2_root"])
@pytest.mark.parametrize("rerank", [0, 64])
def test_quantized(quant, metric, rerank):
    x = clustered(2500, 48, 25, seed=14)
    q = clustered(30, 48, 25, seed=15)
    ix = build_index(x, metric, quant, n_postings=40, rerank_size=rerank, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 16)


def test_deleted_hidden_and_matching():
    x = clustered(3000, 32, 25, seed=3)
    q = x[:40:2].copy()
    ix = build_index(x, "l2", None, n_postings=50, seed=0)
    dev = DeviceIndex(ix)
    dead = np.arange(0, 3000, 3, dtype=np.uint64)
    ix.set_deleted(dead)
    dev.delete(dead)
    ids, _, cnt = compare(dev, ix, q, 10, ix.n_post)
    for i in range(q.shape[0]):
        assert not (ids[i, : cnt[i]] % 3 == 0).any()
    assert dev.info().n_deleted == dead.shape[0]


def test_deleted_at_create_and_quantized_rerank():
    x = clustered(2000, 32, 25, seed=4)
    ix = build_index(x, "l2", "int8", n_postings=30, rerank_size=40, seed=0)
    ix.set_deleted(np.arange(1, 2000, 7))
    dev = DeviceIndex(ix)
    compare(dev, ix, x[:25], 10, 8)


def test_short_rows_and_padding():
    x = clustered(30, 16, 5, seed=5)
    ix = build_index(x, "l2", None, n_postings=6, seed=0)
    dev = DeviceIndex(ix)
    ids, dd, cnt = compare(dev, ix, x[:5], 50, 2)
    assert (cnt < 50).all()


def test_probe_ratio_cut():
    x = clustered(3000, 24, 25, seed=6)
    ix = build_index(x, "l2", None, n_postings=60, probe_ratio=1.3, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, clustered(30, 24, 25, seed=7), 10, 32)


def test_integer_data_ties_and_closed_form():
    """Integer-valued data (the 1M x 128 descriptor shape): every f32 L2 sum is
    exact, so distances equal the int64 closed form and ties are common."""
    rng = np.random.default_rng(8)
    c = rng.uniform(0, 200, size=(16, 32))
    x = np.clip(np.round(c[rng.integers(0, 16, 4000)] + rng.normal(0, 20, (4000, 32))), 0, 255)
    x = np.concatenate([x, x[:500]]).astype(np.float32)      # exact duplicates -> ties
    q = np.clip(np.round(c[rng.integers(0, 16, 30)] + rng.normal(0, 20, (30, 32))), 0, 255)
    q = q.astype(np.float32)
    ix = build_index(x, "l2", None, n_postings=40, seed=0)
    dev = DeviceIndex(ix)
    ids, dd, cnt = compare(dev, ix, q, 20, 40)
    xi, qi = x.astype(np.int64), q.astype(np.int64)
    for i in range(q.shape[0]):
        exact = ((xi[ids[i, : cnt[i]].astype(np.int64)] - qi[i]) ** 2).sum(1)
        np.testing.assert_array_equal(dd[i, : cnt[i]].astype(np.int64), exact)
        full = ((xi - qi[i]) ** 2).sum(1)
        order = np.lexsort((np.arange(x.shape[0]), full))[:20]
        np.testing.assert_array_equal(ids[i, : cnt[i]], order)   # exhaustive: (dist, id) order


@pytest.mark.parametrize("k,rerank", [(1000, 0), (100, 1000), (2048, 0)])
def test_large_pools(k, rerank):
    x = clustered(6000, 32, 25, seed=9)
    quant = "f16" if rerank else None
    ix = build_index(x, "l2", quant, n_postings=24, rerank_size=rerank, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, clustered(12, 32, 25, seed=10), k, 12)


def test_n_probe_edges_and_empty_posting():
    x = clustered(1000, 16, 10, seed=11)
    ix = build_index(x, "l2", None, n_postings=20, seed=0)
    # append an empty posting far away
    ix.block_ids = np.append(ix.block_ids, np.uint32(ix.n_post))
    ix.centroids = np.vstack([ix.centroids, np.full((1, 16), 50, np.float32)])
    ix.post_off = np.append(ix.post_off, ix.post_off[-1])
    dev = DeviceIndex(ix)
    q = clustered(10, 16, 10, seed=12)
    compare(dev, ix, q, 10, 0)
    compare(dev, ix, q, 10, 1000)
    compare(dev, ix, np.full((1, 16), 50, np.float32), 10, 1)   # probes only the empty posting


def test_block_id_order_breaks_centroid_ties():
    # two postings with identical centroids: the probe keeps the lower block id
    x = clustered(400, 8, 4, seed=13)
    ix = build_index(x, "l2", None, n_postings=4, seed=0)
    ix.centroids[1] = ix.centroids[0]
    ix.block_ids = np.array([7, 3, 9, 1], dtype=np.uint32)
    dev = DeviceIndex(ix)
    compare(dev, ix, ix.centroids[:1].copy(), 10, 1)


def test_config1_shape_parity():
    """100k x 128 Gaussian mixture, 256 postings, 1000 queries, n_probe 16, k 10."""
    from pydiskann_amd.datasets import config1
    base, queries = config1()
    ix = build_index(base, "l2", None, n_postings=256, seed=0, kmeans_iters=4)
    dev = DeviceIndex(ix)
    compare(dev, ix, queries, 10, 16, threads=16)


def test_batch_matches_single_and_errors():
    x = clustered(2000, 32, 25, seed=16)
    ix = build_index(x, "l2", None, n_postings=32, seed=0)
    dev = DeviceIndex(ix)
    q = clustered(25, 32, 25, seed=18)
    ids, dd, cnt = dev.search(q, 10, 8)
    for i in range(q.shape[0]):
        i1, d1, c1 = dev.search(q[i:i + 1], 10, 8)
        assert c1[0] == cnt[i]
        np.testing.assert_array_equal(i1[0, : c1[0]], ids[i, : cnt[i]])
        np.testing.assert_array_equal(d1[0, : c1[0]], dd[i, : cnt[i]])
    with pytest.raises(SpfError) as e:
        dev.search(np.zeros((1, 31), np.float32), 10, 8)
    assert e.value.code == SPF_EDIM and "Query dim 31 != index dim 32" in str(e.value)
    with pytest.raises(ValueError):
        dev.search(q, 0, 8)


@pytest.mark.parametrize("quant,k,rerank", [(None, 10, 0), (None, 1000, 0), ("int8", 50, 400)])
def test_long_candidate_lists_select_in_chunks(quant, k, rerank):
    """Candidate lists far above the 24,576-key chunks of the long-list select."""
    rng = np.random.default_rng(11)
    x = rng.normal(size=(300000, 16)).astype(np.float32)
    cents = np.concatenate([rng.normal(size=(3, 16)) * 0.01, rng.normal(size=(40, 16)) * 3])
    ix = build_index(x, "l2", quant, centroids=cents.astype(np.float32), rerank_size=rerank, seed=0)
    assert int(ix.posting_lengths().max()) > 65536
    dev = DeviceIndex(ix)
    q = np.concatenate([x[:12], (rng.normal(size=(12, 16)) * 3).astype(np.float32)])
    compare(dev, ix, q, k, 3)
    compare(dev, ix, q, k, ix.n_post)


@pytest.mark.parametrize("n_probe", [1, 64, 700])
def test_exact_probe_above_4096_postings(n_probe):
    """Exact probe over 9,000 centroids (rows selected with 96 keys per thread
    in registers), including integer centroids with equal distances
    (block-order ties)."""
    rng = np.random.default_rng(12)
    x = rng.integers(-4, 5, size=(60000, 8)).astype(np.float32)
    cents = x[rng.choice(60000, 9000, replace=False)] + rng.integers(0, 2, size=(9000, 8))
    ix = build_index(x, "l2", None, centroids=cents.astype(np.float32), seed=0)
    assert ix.n_post > 4096
    dev = DeviceIndex(ix)
    q = rng.integers(-4, 5, size=(40, 8)).astype(np.float32)
    compare(dev, ix, q, 10, n_probe)


@pytest.mark.parametrize("nq", [40, 300])
def test_pair_ranks_above_16384_postings(nq):
    """30,000 postings: above the 16,384 LDS counters of the block-wise pair
    ranking, so ranks come from device-wide atomics, and above the 24,576-key
    rows the probe select holds in registers (chunks, then a merge); 300
    queries span two ranking blocks on the smaller index (8,000 postings)."""
    rng = np.random.default_rng(13)
    x = rng.normal(size=(60000, 8)).astype(np.float32)
    for n_post in (30000, 8000):
        cents = x[rng.choice(60000, n_post, replace=False)]
        ix = build_index(x, "l2", None, centroids=cents, seed=0)
        dev = DeviceIndex(ix)
        q = rng.normal(size=(nq, 8)).astype(np.float32)
        compare(dev, ix, q, 10, 64)


@pytest.mark.parametrize("k", [10, 100])
def test_candidate_lists_in_register_band(k):
    """About 10,000 candidates per query: the exact select's 4,097 - 24,576
    band (96 keys per thread in registers)."""
    rng = np.random.default_rng(14)
    x = rng.normal(size=(20000, 24)).astype(np.float32)
    ix = build_index(x, "l2", None, n_postings=16, seed=0)
    dev = DeviceIndex(ix)
    q = rng.normal(size=(24, 24)).astype(np.float32)
    compare(dev, ix, q, k, 8)
