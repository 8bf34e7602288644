"""GPU: mixed-precision search (f16 matrix-core scan, exact f32 rerank of a pool).

Bar: the matrix-core fragment layout matches the host product exactly; every
returned distance is the exact f32 distance of its id in the canonical order
(bit-identical to the oracle's metric); rows are sorted by (dist, id); recall
against the FP64 brute-force oracle is stated per case; with a pool at least as
large as every query's candidate set the result equals the exact path.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, build_index, brute_force_f64, PREC_MIXED, PREC_EXACT
from pydiskann_amd import _lib

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.12):
    centers = np.random.default_rng(99).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[np.arange(n) % n_centers] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def test_mfma_fragment_layout_selftest():
    rc = _lib.lib().spf_selftest(0)
    assert rc == 0, _lib.lib().spf_last_error().decode()


def check_rows(ix, q, ids, dd, cnt):
    for i in range(q.shape[0]):
        c = int(cnt[i])
        for j in range(c):
            want = oracle.eval_metric(ix.metric, q[i], ix.raw[ids[i, j]])
            assert np.float32(dd[i, j]).view(np.uint32) == np.float32(want).view(np.uint32), (i, j)
        keys = list(zip(dd[i, :c].tolist(), ids[i, :c].tolist()))
        assert keys == sorted(keys)


def recall(ids, gt, k=10):
    return float(np.mean([len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist())) / k
                          for i in range(gt.shape[0])]))


@pytest.mark.parametrize("quant", [None, "f16", "int8"])
@pytest.mark.parametrize("metric", ["l2", "cosine", "dot", "l2_root"])
@pytest.mark.parametrize("dim", [32, 64, 96, 100, 128, 200])   # 1, 2, 3, 4, 4, 8 K steps of 32
def test_mixed_exact_rerank_and_recall(quant, metric, dim):
    x = clustered(4000, dim, 25, seed=1)
    q = clustered(64, dim, 25, seed=2)
    ix = build_index(x, metric, quant, n_postings=40, rerank_size=200, seed=0)
    dev = DeviceIndex(ix)
    print(f"[case] mixed quant={quant} metric={metric} dim={dim}", flush=True)
    ids, dd, cnt = dev.search(q, 10, 16, precision=PREC_MIXED)
    assert (cnt == 10).all()
    check_rows(ix, q, ids, dd, cnt)
    gt, _ = brute_force_f64(x, q, 10, ix.metric)
    r = recall(ids, gt)
    print(f"recall@10 vs fp64 = {r:.4f}")
    assert r >= 0.95


@pytest.mark.parametrize("metric", ["l2", "cosine", "dot"])
@pytest.mark.parametrize("dim", [300, 512])   # int8 scan K steps of 64: 5, 8; long-row probe
def test_mixed_int8_long_rows(metric, dim):
    x = clustered(3000, dim, 25, seed=11)
    q = clustered(40, dim, 25, seed=12)
    ix = build_index(x, metric, "int8", n_postings=40, rerank_size=200, seed=0)
    dev = DeviceIndex(ix)
    print(f"[case] mixed int8 metric={metric} dim={dim}", flush=True)
    ids, dd, cnt = dev.search(q, 10, 16, precision=PREC_MIXED)
    assert (cnt == 10).all()
    check_rows(ix, q, ids, dd, cnt)
    gt, _ = brute_force_f64(x, q, 10, ix.metric)
    r = recall(ids, gt)
    print(f"recall@10 vs fp64 = {r:.4f}")
    assert r >= 0.95


def test_mixed_equals_exact_when_pool_covers_all_candidates():
    x = clustered(3000, 64, 25, seed=3)
    q = clustered(40, 64, 25, seed=4)
    ix = build_index(x, "l2", None, n_postings=30, seed=0)
    dev = DeviceIndex(ix)
    e_ids, e_d, e_c = dev.search(q, 10, 4, precision=PREC_EXACT)
    m_ids, m_d, m_c = dev.search(q, 10, 4, precision=PREC_MIXED, rerank_size=2048)
    np.testing.assert_array_equal(e_c, m_c)
    np.testing.assert_array_equal(e_ids, m_ids)
    np.testing.assert_array_equal(e_d.view(np.uint32), m_d.view(np.uint32))


def test_mixed_integer_data_is_exact():
    """Integer data in [0, 255] (the 1M x 128 descriptor shape): f16 holds every
    value and every f32 sum is exact, so the pool is exact and mixed = exact."""
    rng = np.random.default_rng(5)
    c = rng.uniform(0, 200, size=(16, 128))
    x = np.clip(np.round(c[rng.integers(0, 16, 5000)] + rng.normal(0, 20, (5000, 128))), 0, 255)
    x = x.astype(np.float32)
    q = np.clip(np.round(c[rng.integers(0, 16, 50)] + rng.normal(0, 20, (50, 128))), 0, 255)
    q = q.astype(np.float32)
    ix = build_index(x, "l2", None, n_postings=48, seed=0)
    dev = DeviceIndex(ix)
    e = dev.search(q, 100, 16, precision=PREC_EXACT)
    m = dev.search(q, 100, 16, precision=PREC_MIXED, rerank_size=1000)
    for i in range(q.shape[0]):
        c_ = int(e[2][i])
        assert int(m[2][i]) == c_
        np.testing.assert_array_equal(m[1][i, :c_], e[1][i, :c_])
        # equal distances may be ordered differently only if the pool boundary
        # cut a tie group; the (dist, id) order makes ids equal too
        np.testing.assert_array_equal(m[0][i, :c_], e[0][i, :c_])


def test_mixed_deleted_hidden():
    x = clustered(3000, 32, 25, seed=6)
    ix = build_index(x, "l2", None, n_postings=30, rerank_size=100, seed=0)
    dev = DeviceIndex(ix)
    dead = np.arange(0, 3000, 2, dtype=np.uint64)
    dev.delete(dead)
    ids, dd, cnt = dev.search(x[:20], 10, 30, precision=PREC_MIXED)
    for i in range(20):
        assert not (ids[i, : cnt[i]] % 2 == 0).any()
    live = np.ones(3000, bool)
    live[dead.astype(np.int64)] = False
    gt, _ = brute_force_f64(x, x[:20], 10, 1, live=live)
    assert recall(ids, gt) >= 0.99


def test_mixed_short_rows():
    x = clustered(40, 16, 4, seed=7)
    ix = build_index(x, "l2", None, n_postings=4, seed=0)
    dev = DeviceIndex(ix)
    ids, dd, cnt = dev.search(x[:3], 50, 1, precision=PREC_MIXED, rerank_size=64)
    e_ids, e_dd, e_cnt = dev.search(x[:3], 50, 1)
    np.testing.assert_array_equal(cnt, e_cnt)
    for i in range(3):
        np.testing.assert_array_equal(ids[i, : cnt[i]], e_ids[i, : cnt[i]])


@pytest.mark.parametrize("metric", ["l2", "l2_root"])
@pytest.mark.parametrize("data", ["float", "integer", "wide"])
@pytest.mark.parametrize("k,pool", [(1, 64), (10, 200), (100, 1000)])
def test_pruned_rerank_equals_full_pool_rerank(metric, data, k, pool):
    """The bound-pruned rerank returns exactly what reranking every pool member
    returns, while gathering fewer raw rows."""
    rng = np.random.default_rng(11)
    if data == "integer":
        c = rng.uniform(0, 200, size=(40, 64))
        x = np.clip(np.rint(c[rng.integers(0, 40, 8000)] + rng.normal(0, 20, (8000, 64))), 0, 255)
        q = np.clip(np.rint(c[rng.integers(0, 40, 64)] + rng.normal(0, 20, (64, 64))), 0, 255)
    elif data == "wide":   # large magnitudes: f16 rounding error matters
        x = clustered(8000, 48, 30, seed=12) * 3000.0
        q = clustered(64, 48, 30, seed=13) * 3000.0
    else:
        x = clustered(8000, 100, 30, seed=12)
        q = clustered(64, 100, 30, seed=13)
    x, q = x.astype(np.float32), q.astype(np.float32)
    ix = build_index(x, metric, None, n_postings=64, seed=0)
    dev = DeviceIndex(ix)
    print(f"[case] prune metric={metric} data={data} k={k} pool={pool}", flush=True)
    a_ids, a_d, a_c = dev.search(q, k, 24, precision=PREC_MIXED, rerank_size=pool)
    b_ids, b_d, b_c = dev.search(q, k, 24, precision=PREC_MIXED, rerank_size=pool,
                                 flags=_lib.FLAG_FULL_RERANK)
    np.testing.assert_array_equal(a_c, b_c)
    np.testing.assert_array_equal(a_ids, b_ids)
    np.testing.assert_array_equal(a_d.view(np.uint32), b_d.view(np.uint32))


def test_pruned_rerank_gathers_fewer_rows():
    x = clustered(20000, 64, 40, seed=14)
    q = clustered(256, 64, 40, seed=15)
    ix = build_index(x, "l2", None, n_postings=128, seed=0)
    dev = DeviceIndex(ix)
    dev.search(q, 10, 16, precision=PREC_MIXED, rerank_size=500, flags=_lib.FLAG_PROFILE)
    pr = _lib.last_profile()
    print(f"rerank rows {pr['rerank_rows']} of {pr['pool_members']} pool members")
    assert pr["pool_members"] == 256 * 500
    assert 256 * 10 <= pr["rerank_rows"] < pr["pool_members"] // 2


def integer_mixture(n, nq, d, comps, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 200, size=(comps, d))
    x = np.clip(np.rint(c[rng.integers(0, comps, n)] + rng.normal(0, 20, (n, d))), 0, 255)
    q = np.clip(np.rint(c[rng.integers(0, comps, nq)] + rng.normal(0, 20, (nq, d))), 0, 255)
    return x.astype(np.float32), q.astype(np.float32)


@pytest.mark.parametrize("pool", [1000, 2048])
def test_mixed_long_candidate_lists(pool):
    """Every query scans all 60,000 entries: above the 24,576 keys a pool
    select workgroup holds in registers, so later passes re-read the rest.
    Integer data: mixed = exact."""
    x, q = integer_mixture(60000, 48, 32, 12, seed=21)
    ix = build_index(x, "l2", None, n_postings=16, seed=0)
    dev = DeviceIndex(ix)
    print(f"[case] long lists pool={pool}", flush=True)
    e = dev.search(q, 100, 16, precision=PREC_EXACT)
    m = dev.search(q, 100, 16, precision=PREC_MIXED, rerank_size=pool)
    f = dev.search(q, 100, 16, precision=PREC_MIXED, rerank_size=pool, flags=_lib.FLAG_FULL_RERANK)
    for a in (m, f):
        np.testing.assert_array_equal(a[2], e[2])
        np.testing.assert_array_equal(a[0], e[0])
        np.testing.assert_array_equal(a[1].view(np.uint32), e[1].view(np.uint32))


def test_mixed_probe_above_4096_postings():
    """The matrix-core probe over 5,000 postings (probe pools above the 4,096
    keys a wave holds in registers) with its exact recheck: the probe list,
    and so the result, equals the exact path's on integer data."""
    x, q = integer_mixture(40000, 64, 32, 64, seed=22)
    ix = build_index(x, "l2", None, n_postings=5000, seed=0, kmeans_iters=2)
    dev = DeviceIndex(ix)
    print("[case] mixed probe P=5000", flush=True)
    for n_probe in (8, 64):
        e = dev.search(q, 10, n_probe, precision=PREC_EXACT)
        m = dev.search(q, 10, n_probe, precision=PREC_MIXED, rerank_size=2048)
        np.testing.assert_array_equal(m[2], e[2])
        np.testing.assert_array_equal(m[0], e[0])
        np.testing.assert_array_equal(m[1].view(np.uint32), e[1].view(np.uint32))
