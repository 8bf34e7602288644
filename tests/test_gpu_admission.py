"""GPU: two-phase mixed scans with admission.

The matrix-core scan first covers each query's nearest probes until they hold
`rerank_size` entries and stores their keys; the `rerank_size`-th smallest live
key among them bounds the pool's largest key (the pool is the smallest keys of
a superset), so the second phase lists only keys within the bound and the
pool is selected from both.  This is the reference's top-`want` heap inside
the scan loop (src/spfresh.rs:703-717) turned batch-wide.

Bar: bit-identical ids, distances and counts (SPF_FLAG_ADMIT) to the one-phase
scan that stores every key (the default), for every code type and metric, with
label filters, deletes, a probe_ratio cut, ties at the pool boundary, pools
larger than the candidate sets, repeated and replayed calls; the second phase
lists fewer keys than it scans.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, PREC_MIXED
from pydiskann_amd import _lib

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.12):
    centers = np.random.default_rng(99).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[np.arange(n) % n_centers] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def same(a, b, msg):
    np.testing.assert_array_equal(a[2], b[2], err_msg=msg)
    np.testing.assert_array_equal(a[0], b[0], err_msg=msg)
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32), err_msg=msg)


def both(dev, q, k, n_probe, pool, masks=None):
    a = dev.search(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool, masks=masks,
                   flags=_lib.FLAG_ADMIT | _lib.FLAG_PROFILE)
    pa = _lib.last_profile()
    b = dev.search(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool, masks=masks)
    return a, b, pa


@pytest.mark.parametrize("quant", [None, "f16", "int8"])
@pytest.mark.parametrize("metric", ["l2", "l2_root", "cosine", "dot"])
def test_admission_equals_all_keys(quant, metric):
    x = clustered(12000, 64, 40, seed=31)
    q = clustered(96, 64, 40, seed=32)
    ix = build_index(x, metric, quant, n_postings=96, seed=0)
    dev = DeviceIndex(ix)
    # pools sorted in registers (up to 1,024 and 2,048 candidates) and, at pool
    # 2,048, the radix select over more candidates than LDS holds
    for k, pool, n_probe in ((10, 100, 16), (10, 400, 24), (100, 1000, 32), (1, 64, 8),
                             (100, 2048, 40)):
        print(f"[case] admission quant={quant} metric={metric} k={k} pool={pool} "
              f"n_probe={n_probe}", flush=True)
        a, b, pa = both(dev, q, k, n_probe, pool)
        same(a, b, f"k={k} pool={pool} n_probe={n_probe}")
        assert (a[2] == k).all()
        # the second phase lists fewer keys than the whole scan computes
        print(f"admitted {pa['admitted']} of {pa['candidates']} candidates", flush=True)
        assert 0 < pa["admitted"] < pa["candidates"], pa


def test_admission_filters_deletes_and_ratio_cut():
    x = clustered(10000, 48, 30, seed=33)
    q = clustered(80, 48, 30, seed=34)
    ix = build_index(x, "l2", "f16", n_postings=80, seed=0)
    ix.labels = (np.arange(ix.n_raw) % 7).astype(np.uint8)
    ix.set_deleted(np.arange(0, ix.n_raw, 3, dtype=np.uint64))
    dev = DeviceIndex(ix)
    rng = np.random.default_rng(5)
    masks = rng.integers(1, 1 << 7, size=q.shape[0]).astype(np.uint64)
    masks[::5] = np.uint64(1)   # one label in seven: first phases short of `want` live keys
    for k, pool in ((10, 200), (50, 600)):
        print(f"[case] admission filtered k={k} pool={pool}", flush=True)
        a, b, _ = both(dev, q, k, 20, pool, masks=masks)
        same(a, b, f"filtered k={k} pool={pool}")
        a, b, _ = both(dev, q, k, 20, pool)
        same(a, b, f"deleted k={k} pool={pool}")
    cut = DeviceIndex(dataclasses.replace(ix, probe_ratio=1.3))
    a, b, _ = both(cut, q, 10, 40, 300)
    same(a, b, "probe_ratio cut")


def test_admission_ties_at_the_pool_boundary():
    """Each row repeated 16 times on integer data: many candidates share the
    boundary key, and the pool takes them in slot order in both paths."""
    rng = np.random.default_rng(35)
    c = rng.uniform(0, 200, size=(24, 32))
    base = np.clip(np.rint(c[rng.integers(0, 24, 600)] + rng.normal(0, 20, (600, 32))), 0, 255)
    x = np.repeat(base, 16, axis=0).astype(np.float32)
    q = np.clip(np.rint(c[rng.integers(0, 24, 64)] + rng.normal(0, 20, (64, 32))), 0, 255)
    q = q.astype(np.float32)
    ix = build_index(x, "l2", None, n_postings=48, seed=0)
    dev = DeviceIndex(ix)
    for k, pool in ((10, 100), (100, 500), (16, 160)):
        print(f"[case] admission ties k={k} pool={pool}", flush=True)
        a, b, _ = both(dev, q, k, 12, pool)
        same(a, b, f"ties k={k} pool={pool}")


def test_admission_pool_above_candidates_and_reuse():
    """Pools larger than some queries' candidate sets (no second phase for
    them), then smaller batches on the same workspace (stale keys beyond the
    first phase) and replayed calls: every result equals the one-phase scan."""
    x = clustered(6000, 32, 30, seed=36)
    q = clustered(128, 32, 30, seed=37)
    ix = build_index(x, "l2", None, n_postings=60, seed=0)
    dev = DeviceIndex(ix)
    print("[case] admission pool above candidates", flush=True)
    a, b, _ = both(dev, q, 10, 4, 2048)
    same(a, b, "pool above candidates")
    ref = {}
    for nq in (128, 40, 7, 128):
        ref[nq] = dev.search(q[:nq], 10, 12, precision=PREC_MIXED, rerank_size=300)
    for rep in range(3):   # the second and later calls replay a captured chain
        for nq in (128, 40, 7):
            got = dev.search(q[:nq], 10, 12, precision=PREC_MIXED, rerank_size=300,
                             flags=_lib.FLAG_ADMIT)
            same(got, ref[nq], f"nq={nq} call {rep}")


def test_admission_off_for_long_segments():
    """Segments beyond 2^16 slots keep the one-phase scan (16-bit list slots)."""
    x = clustered(90000, 16, 4, seed=38)
    q = clustered(16, 16, 4, seed=39)
    ix = build_index(x, "l2", None, n_postings=8, seed=0)
    dev = DeviceIndex(ix)
    print("[case] admission long segments", flush=True)
    a, b, pa = both(dev, q, 10, 8, 500)
    same(a, b, "long segments")
    assert pa["admitted"] == 0
