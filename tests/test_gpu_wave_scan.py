"""GPU: the wave scan (one wave per work item of 32 or 64 queries) against the
workgroup scan (128-query items) of mixed precision.

Bar: both scans write the same 16-bit key for every (query, entry), so every
search returns the same ids, distances and counts bit for bit, whichever scan
ran: f32 / f16 / int8 postings, every metric, K steps 1 to 8 (dims 17 to 300),
postings longer than one 256-entry chunk, short and empty postings, tombstones,
label filters, and pools above the candidate count.  The scan is chosen through
the launch-shape cache (mix_g 32 / 64: wave scan; 128: workgroup scan).
"""
from __future__ import annotations

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, PREC_MIXED

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.15):
    centers = np.random.default_rng(5).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[rng.integers(0, n_centers, n)] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def with_mix_g(ix, q, k, n_probe, pool, m, tmp_path):
    """A fresh device mirror whose mixed scan takes `m` queries per item."""
    cache = str(tmp_path / f"tune_{m}.txt")
    dev = DeviceIndex(ix)
    dev.autotune(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool, cache_path=cache)
    key = open(cache).read().split(" scan_g=")[0]
    open(cache, "w").write(f"{key} scan_g=8 scan_chunk=1024 mix_g={m}\n")
    dev.close()
    dev = DeviceIndex(ix)
    r = dev.autotune(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool, cache_path=cache)
    assert r["from_cache"] == 1 and r["mix_g"] == m
    return dev


def same(a, b, label):
    np.testing.assert_array_equal(a[2], b[2], err_msg=label)
    np.testing.assert_array_equal(a[0], b[0], err_msg=label)
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32), err_msg=label)


def compare(ix, q, k, n_probe, pool, tmp_path, label, masks=None):
    devs = {m: with_mix_g(ix, q, k, n_probe, pool, m, tmp_path) for m in (128, 64, 32)}
    print(f"[case] {label}", flush=True)
    base = devs[128].search(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool, masks=masks)
    assert base[2].sum() > 0
    for m in (64, 32):
        got = devs[m].search(q, k, n_probe, precision=PREC_MIXED, rerank_size=pool, masks=masks)
        same(got, base, f"{label}: mix_g {m}")
    for d in devs.values():
        d.close()
    return base


@pytest.mark.parametrize("quant", [None, "f16", "int8"])
@pytest.mark.parametrize("metric", ["l2", "l2_root", "cosine", "dot"])
@pytest.mark.parametrize("dim", [17, 96, 200])   # K steps 1, 3, 7 (f16) / 1, 2, 4 (int8)
def test_wave_scan_equals_workgroup_scan(tmp_path, quant, metric, dim):
    x = clustered(6000, dim, 30, seed=1)
    q = clustered(150, dim, 30, seed=2)
    ix = build_index(x, metric, quant, n_postings=48, rerank_size=64, seed=0)
    compare(ix, q, 10, 9, 64, tmp_path, f"quant={quant} metric={metric} dim={dim}")


def test_long_postings_many_chunks_and_tombstones(tmp_path):
    # 20 postings of ~1,500 entries: every item walks 16 tiles, postings span
    # several 256-entry chunks, and the posting ends fall inside tiles
    x = clustered(30011, 64, 20, seed=3)
    q = clustered(200, 64, 20, seed=4)
    ix = build_index(x, "l2", None, n_postings=20, seed=0)
    rng = np.random.default_rng(0)
    ix.set_deleted(rng.choice(x.shape[0], 4000, replace=False))
    compare(ix, q, 50, 6, 500, tmp_path, "long postings, tombstones")


def test_int8_wide_rows_and_filter(tmp_path):
    # int8 at 300 dims: 64-dim K steps, 5 of 8 in use
    x = clustered(8000, 300, 40, seed=5)
    q = clustered(100, 300, 40, seed=6)
    ix = build_index(x, "dot", "int8", n_postings=64, rerank_size=100, seed=0)
    ix.labels = (np.arange(ix.n_raw) % 64).astype(np.uint8)
    rng = np.random.default_rng(1)
    masks = np.zeros(q.shape[0], dtype=np.uint64)
    for i in range(q.shape[0]):
        for l in rng.choice(64, int(rng.integers(1, 9)), replace=False):
            masks[i] |= np.uint64(1) << np.uint64(int(l))
    compare(ix, q, 10, 12, 100, tmp_path, "int8 300-d dot, unfiltered")
    compare(ix, q, 10, 12, 100, tmp_path, "int8 300-d dot, filtered", masks=masks)


def test_f16_filter_small_postings_and_big_pool(tmp_path):
    # many tiny postings (ragged tails everywhere) and a pool above the candidates
    x = clustered(3000, 128, 12, seed=7)
    q = clustered(97, 128, 12, seed=8)
    ix = build_index(x, "cosine", "f16", n_postings=300, rerank_size=64, seed=0)
    ix.labels = (np.arange(ix.n_raw) * 7 % 64).astype(np.uint8)
    masks = np.full(q.shape[0], np.uint64(0x0f0f0f0f0f0f0f0f))
    compare(ix, q, 10, 40, 2048, tmp_path, "f16 cosine, 300 postings, filtered", masks=masks)
    compare(ix, q, 10, 40, 2048, tmp_path, "f16 cosine, 300 postings")
