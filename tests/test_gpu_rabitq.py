"""GPU: RaBitQ postings (src/rabitq.rs, src/spfresh.rs:697-703, 731-770).

Bar (exact precision): the device search_rabitq (per-pair query prep, estimate
+ error bound, exact distance only when the lower bound beats the current k-th,
heap updates in entry order) returns the oracle's ids in the oracle's order with
bit-identical distances; cosine / dot postings search the decoded vectors
(decode_with on the device) and match the oracle the same way, with and without
rerank_size.  Mixed precision (estimator keys -> pool -> exact rerank) returns
exact distances, reaches the stated recall, and equals the exact f32 search of
the same postings when the pool covers every candidate.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, build_index, brute_force_f64, PREC_MIXED, PREC_EXACT

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.12):
    centers = np.random.default_rng(99).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[np.arange(n) % n_centers] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def compare(dev, ix, qs, k, n_probe):
    print(f"[case] rabitq nq={qs.shape[0]} d={qs.shape[1]} k={k} n_probe={n_probe} "
          f"metric={ix.metric} rerank={ix.rerank_size}", flush=True)
    ids, dd, cnt = dev.search(qs, k, n_probe)
    rid, rd, rc = oracle.Oracle(ix).search_batch(qs, k, n_probe, threads=8)
    np.testing.assert_array_equal(cnt, rc)
    for i in range(qs.shape[0]):
        c = int(cnt[i])
        np.testing.assert_array_equal(ids[i, :c], rid[i, :c], err_msg=f"query {i}")
        assert np.array_equal(dd[i, :c].view(np.uint32), rd[i, :c].view(np.uint32)), \
            f"query {i}: distances not bit-identical"
        assert (ids[i, c:] == np.iinfo(np.uint64).max).all()
    return ids, dd, cnt


def recall(ids, gt, k=10):
    return float(np.mean([len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist())) / k
                          for i in range(gt.shape[0])]))


@pytest.mark.parametrize("metric", ["l2", "l2_root"])
@pytest.mark.parametrize("dim", [17, 64, 100, 128, 300])
def test_search_rabitq_parity(metric, dim):
    x = clustered(3000, dim, 25, seed=21)
    q = clustered(40, dim, 25, seed=22)
    ix = build_index(x, metric, "rabitq", n_postings=40, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 8)
    compare(dev, ix, q, 10, ix.n_post)


@pytest.mark.parametrize("k", [1, 100, 1024])
def test_search_rabitq_heap_sizes(k):
    x = clustered(4000, 48, 20, seed=23)
    q = clustered(16, 48, 20, seed=24)
    ix = build_index(x, "l2", "rabitq", n_postings=24, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, k, 12)


def test_search_rabitq_k_limit():
    x = clustered(500, 16, 5, seed=25)
    ix = build_index(x, "l2", "rabitq", n_postings=8, seed=0)
    dev = DeviceIndex(ix)
    with pytest.raises(ValueError, match="k above 1024"):
        dev.search(x[:2], 1025, 4)


@pytest.mark.parametrize("metric", ["cosine", "dot"])
@pytest.mark.parametrize("rerank", [0, 64])
def test_decoded_postings_parity(metric, rerank):
    x = clustered(2500, 40, 25, seed=26)
    q = clustered(30, 40, 25, seed=27)
    ix = build_index(x, metric, "rabitq", n_postings=30, rerank_size=rerank, seed=0)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 10)


def test_rabitq_deleted_hidden():
    x = clustered(3000, 32, 25, seed=28)
    q = x[:60:3].copy()
    ix = build_index(x, "l2", "rabitq", n_postings=30, seed=0)
    dev = DeviceIndex(ix)
    dead = np.arange(0, 3000, 4, dtype=np.uint64)
    ix.set_deleted(dead)
    dev.delete(dead)
    ids, _, cnt = compare(dev, ix, q, 10, ix.n_post)
    dead_set = set(dead.tolist())
    for i in range(q.shape[0]):
        assert not (set(ids[i, :cnt[i]].tolist()) & dead_set)


@pytest.mark.parametrize("metric", ["l2", "l2_root"])
@pytest.mark.parametrize("dim", [32, 128])
def test_rabitq_mixed_exact_distances_and_recall(metric, dim):
    x = clustered(6000, dim, 30, seed=29)
    q = clustered(64, dim, 30, seed=30)
    ix = build_index(x, metric, "rabitq", n_postings=48, seed=0)
    dev = DeviceIndex(ix)
    print(f"[case] rabitq mixed metric={metric} dim={dim}", flush=True)
    ids, dd, cnt = dev.search(q, 10, 16, precision=PREC_MIXED, rerank_size=200)
    assert (cnt == 10).all()
    for i in range(q.shape[0]):
        for j in range(10):
            want = oracle.eval_metric(ix.metric, q[i], ix.raw[ids[i, j]])
            assert np.float32(dd[i, j]).view(np.uint32) == np.float32(want).view(np.uint32)
        keys = list(zip(dd[i].tolist(), ids[i].tolist()))
        assert keys == sorted(keys)
    gt, _ = brute_force_f64(x, q, 10, ix.metric)
    r = recall(ids, gt)
    print(f"recall@10 vs fp64 = {r:.4f}")
    assert r >= 0.95


def test_rabitq_mixed_equals_exact_f32_when_pool_covers_all():
    x = clustered(3000, 64, 25, seed=31)
    q = clustered(40, 64, 25, seed=32)
    rq = build_index(x, "l2", "rabitq", n_postings=30, seed=0)
    f32 = build_index(x, "l2", None, n_postings=30, seed=0)
    np.testing.assert_array_equal(rq.post_off, f32.post_off)
    np.testing.assert_array_equal(rq.ids, f32.ids)
    m_ids, m_d, m_c = DeviceIndex(rq).search(q, 10, 4, precision=PREC_MIXED, rerank_size=2048)
    e_ids, e_d, e_c = DeviceIndex(f32).search(q, 10, 4, precision=PREC_EXACT)
    np.testing.assert_array_equal(e_c, m_c)
    np.testing.assert_array_equal(e_ids, m_ids)
    np.testing.assert_array_equal(e_d.view(np.uint32), m_d.view(np.uint32))


@pytest.mark.parametrize("metric", ["l2", "l2_root", "cosine", "dot"])
@pytest.mark.parametrize("dim", [32, 128, 300, 768])
def test_rabitq_mixed_every_metric_f16_stage_equals_full_pool(metric, dim):
    """Mixed precision over 1-bit codes for every metric (the L2 estimate
    turned into the metric's estimate with exact per-entry norms), then the
    f16-row stage ordering the pool by a proven bound and the exact f32
    rerank in that order: bit-identical to reranking the whole pool in f32,
    every distance the exact distance of its id, and recall stated against
    FP64."""
    from pydiskann_amd._lib import FLAG_FULL_RERANK
    x = clustered(6000, dim, 30, seed=33)
    q = clustered(48, dim, 30, seed=34)
    if metric in ("cosine", "dot"):
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    ix = build_index(x, metric, "rabitq", n_postings=48, seed=0)
    dev = DeviceIndex(ix)
    print(f"[case] rabitq mixed f16 stage metric={metric} dim={dim}", flush=True)
    for k, pool in ((10, 200), (100, 1000)):
        a = dev.search(q, k, 16, precision=PREC_MIXED, rerank_size=pool)
        b = dev.search(q, k, 16, precision=PREC_MIXED, rerank_size=pool, flags=FLAG_FULL_RERANK)
        np.testing.assert_array_equal(a[2], b[2])
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    ids, dd, cnt = dev.search(q, 10, 16, precision=PREC_MIXED, rerank_size=200)
    assert (cnt == 10).all()
    for i in range(0, q.shape[0], 4):
        for j in range(10):
            want = oracle.eval_metric(ix.metric, q[i], ix.raw[ids[i, j]])
            assert np.float32(dd[i, j]).view(np.uint32) == np.float32(want).view(np.uint32)
        keys = list(zip(dd[i].tolist(), ids[i].tolist()))
        assert keys == sorted(keys)
    gt, _ = brute_force_f64(x, q, 10, ix.metric)
    r = recall(ids, gt)
    print(f"recall@10 vs fp64 = {r:.4f}")
    assert r >= 0.9


@pytest.mark.parametrize("metric", ["l2", "l2_root", "cosine", "dot"])
@pytest.mark.parametrize("dim", [32, 100, 300, 768, 1500, 2100])
def test_rabitq_matrix_core_estimator_equals_popcount(metric, dim):
    """The mixed scan's estimator sums S = sum_i bit_i * level_i on int8
    matrix cores (bits unpacked to bytes, one 64-dimension K step per code
    word) instead of four bit-plane popcounts (src/rabitq.rs:46-58).  S is
    the same integer, so the 16-bit keys, the pool and the reranked results
    equal the popcount scan's bit for bit: every padded width (64 .. 4096),
    every metric, with deletes and label filters, pools from k to 10 k."""
    from pydiskann_amd._lib import FLAG_RQ_POPCOUNT
    n = 6000 if dim <= 768 else 2500
    x = clustered(n, dim, 30, seed=41)
    q = clustered(40, dim, 30, seed=42)
    if metric in ("cosine", "dot"):
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    ix = build_index(x, metric, "rabitq", n_postings=40, seed=0)
    ix.labels = (np.arange(ix.n_raw) % 11).astype(np.uint8)
    ix.set_deleted(np.arange(0, n, 7, dtype=np.uint64))
    dev = DeviceIndex(ix)
    print(f"[case] rabitq matrix-core estimator metric={metric} dim={dim} "
          f"padded={ix.rq_padded}", flush=True)
    masks = np.full(q.shape[0], np.uint64(0b10110101101), dtype=np.uint64)
    for k, pool, m in ((10, 10, None), (10, 200, None), (100, 1000, None), (10, 100, masks)):
        a = dev.search(q, k, 12, precision=PREC_MIXED, rerank_size=pool, masks=m)
        b = dev.search(q, k, 12, precision=PREC_MIXED, rerank_size=pool, masks=m,
                       flags=FLAG_RQ_POPCOUNT)
        msg = f"k={k} pool={pool} filtered={m is not None}"
        np.testing.assert_array_equal(a[2], b[2], err_msg=msg)
        np.testing.assert_array_equal(a[0], b[0], err_msg=msg)
        np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32), err_msg=msg)
        assert (a[2] > 0).all()
