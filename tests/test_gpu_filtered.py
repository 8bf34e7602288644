"""GPU: label-filtered search (Filter::LabelIn on one label field,
src/filtered.rs:120-135, applied where the scan skips tombstones,
src/spfresh.rs:707,745).

Bar: the exact path returns the oracle's filtered result bit for bit (ids,
order, counts, distances); the mixed path returns exact distances of matching
ids only, equals the exact path when its pool covers every candidate, and
keeps recall against the FP64 brute force over the matching ids; filtering
equals tombstoning the non-matching ids; a query's result does not depend on
the other queries of its batch.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, build_index, brute_force_f64, PREC_MIXED, PREC_EXACT

pytestmark = pytest.mark.gpu


def labelled(n, d, quant, metric, seed=0, n_post=40, rerank=0, n_labels=64, nq=48):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=(20, d))
    x = (c[rng.integers(0, 20, n)] + rng.normal(0, 0.15, size=(n, d))).astype(np.float32)
    ix = build_index(x, metric, quant, n_postings=n_post, rerank_size=rerank, seed=0)
    ix.labels = rng.integers(0, n_labels, size=ix.n_raw).astype(np.uint8)
    q = (c[rng.integers(0, 20, nq)] + rng.normal(0, 0.15, size=(nq, d))).astype(np.float32)
    masks = np.zeros(nq, dtype=np.uint64)
    for i in range(nq):
        for b in rng.choice(64, size=rng.integers(1, 12), replace=False):
            masks[i] |= np.uint64(1) << np.uint64(b)
    return x, ix, q, masks


def matching(ix, mask):
    lab = ix.labels.astype(np.uint64)
    return ((np.uint64(mask) >> np.minimum(lab, np.uint64(63))) & np.uint64(1)).astype(bool) \
        & (lab < np.uint64(64))


def compare(dev, ix, q, masks, k, n_probe):
    print(f"[case] filtered nq={q.shape[0]} d={q.shape[1]} k={k} n_probe={n_probe} "
          f"metric={ix.metric} quant={ix.quant} rerank={ix.rerank_size}", flush=True)
    ids, dd, cnt = dev.search(q, k, n_probe, masks=masks)
    rid, rd, rc = oracle.Oracle(ix).search_batch(q, k, n_probe, masks=masks, threads=8)
    np.testing.assert_array_equal(cnt, rc)
    for i in range(q.shape[0]):
        c = int(cnt[i])
        np.testing.assert_array_equal(ids[i, :c], rid[i, :c], err_msg=f"query {i}")
        assert np.array_equal(dd[i, :c].view(np.uint32), rd[i, :c].view(np.uint32)), \
            f"query {i}: distances not bit-identical"
        assert matching(ix, masks[i])[ids[i, :c].astype(np.int64)].all()
    return ids, dd, cnt


@pytest.mark.parametrize("quant,metric,rerank", [
    (None, "l2", 0), (None, "cosine", 0), (None, "dot", 0), ("f16", "l2", 0),
    ("int8", "l2", 64), ("int8", "cosine", 0), ("rabitq", "l2", 0), ("rabitq", "l2_root", 64),
    ("rabitq", "dot", 0)])
def test_filtered_exact_parity(quant, metric, rerank):
    _, ix, q, masks = labelled(3000, 48, quant, metric, rerank=rerank)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, masks, 10, 8)
    compare(dev, ix, q, masks, 10, ix.n_post)      # exhaustive: every matching id is a candidate


def test_filtered_edge_masks_and_labels():
    _, ix, q, masks = labelled(2500, 32, None, "l2", seed=1, n_labels=80)  # labels 64..79 never match
    dev = DeviceIndex(ix)
    edge = masks.copy()
    edge[0] = 0                                    # matches nothing
    edge[1] = np.uint64((1 << 64) - 1)             # every label below 64
    edge[2] = np.uint64(1) << np.uint64(63)
    ids, dd, cnt = compare(dev, ix, q, edge, 10, ix.n_post)
    assert cnt[0] == 0 and (ids[0] == np.iinfo(np.uint64).max).all() and np.isinf(dd[0]).all()
    assert cnt[1] == min(10, int(matching(ix, edge[1]).sum()))
    compare(dev, ix, q, edge, 100, 6)


def test_filter_equals_tombstoning_the_rest():
    _, ix, q, masks = labelled(3000, 32, "int8", "l2", seed=2, rerank=32)
    dev = DeviceIndex(ix)
    for i in (0, 9, 31):
        a = dev.search(q[i:i + 1], 10, 8, masks=masks[i:i + 1])
        _, ix2, _, _ = labelled(3000, 32, "int8", "l2", seed=2, rerank=32)
        dev2 = DeviceIndex(ix2)
        dev2.delete(np.nonzero(~matching(ix2, masks[i]))[0].astype(np.uint64))
        b = dev2.search(q[i:i + 1], 10, 8)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)


@pytest.mark.parametrize("precision", [PREC_EXACT, PREC_MIXED])
def test_filtered_query_alone_equals_in_batch(precision):
    _, ix, q, masks = labelled(3000, 64, None, "l2", seed=3, rerank=200)
    dev = DeviceIndex(ix)
    ids, dd, cnt = dev.search(q, 10, 8, precision=precision, masks=masks)
    for i in (0, 17, 47):
        a = dev.search(q[i:i + 1], 10, 8, precision=precision, masks=masks[i:i + 1])
        assert int(a[2][0]) == int(cnt[i])
        np.testing.assert_array_equal(a[0][0], ids[i])
        np.testing.assert_array_equal(a[1][0].view(np.uint32), dd[i].view(np.uint32))


@pytest.mark.parametrize("quant,metric", [(None, "l2"), (None, "cosine"), ("f16", "dot"),
                                          ("rabitq", "l2")])
def test_filtered_mixed_exact_rows_and_recall(quant, metric):
    x, ix, q, masks = labelled(4000, 64, quant, metric, seed=4, rerank=200)
    dev = DeviceIndex(ix)
    print(f"[case] filtered mixed quant={quant} metric={metric}", flush=True)
    ids, dd, cnt = dev.search(q, 10, 16, precision=PREC_MIXED, masks=masks)
    hits = 0
    total = 0
    for i in range(q.shape[0]):
        c = int(cnt[i])
        m = matching(ix, masks[i])
        assert m[ids[i, :c].astype(np.int64)].all()
        for j in range(c):
            want = oracle.eval_metric(ix.metric, q[i], ix.raw[ids[i, j]])
            assert np.float32(dd[i, j]).view(np.uint32) == np.float32(want).view(np.uint32)
        keys = list(zip(dd[i, :c].tolist(), ids[i, :c].tolist()))
        assert keys == sorted(keys)
        gt, gd = brute_force_f64(x, q[i:i + 1], 10, ix.metric, live=m)
        g = [int(v) for v, dv in zip(gt[0], gd[0]) if np.isfinite(dv)]
        hits += len(set(ids[i, :c].tolist()) & set(g))
        total += len(g)
    r = hits / max(total, 1)
    print(f"filtered recall@10 vs fp64 = {r:.4f}")
    assert r >= 0.9


@pytest.mark.parametrize("dim", [64, 128, 200])   # filtered matrix-core scan: 2, 4, 8 K steps
def test_filtered_mixed_equals_exact_when_pool_covers_all(dim):
    _, ix, q, masks = labelled(3000, dim, None, "l2", seed=5)
    dev = DeviceIndex(ix)
    e = dev.search(q, 10, 4, precision=PREC_EXACT, masks=masks)
    m = dev.search(q, 10, 4, precision=PREC_MIXED, rerank_size=2048, masks=masks)
    np.testing.assert_array_equal(e[2], m[2])
    np.testing.assert_array_equal(e[0], m[0])
    np.testing.assert_array_equal(e[1].view(np.uint32), m[1].view(np.uint32))


def test_filtered_device_form_and_errors():
    """Device-resident buffers from the HIP runtime the library itself links."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    _, ix, q, masks = labelled(2000, 32, None, "l2", seed=6)
    dev = DeviceIndex(ix)
    want = dev.search(q, 10, 8, masks=masks)
    nq = q.shape[0]
    host = {"q": q, "m": masks, "ids": np.empty((nq, 10), np.uint64),
            "dd": np.empty((nq, 10), np.float32), "cnt": np.empty(nq, np.uint32)}
    ptr = {}
    try:
        for name, a in host.items():
            ptr[name] = C.c_void_p()
            assert hip.hipMalloc(C.byref(ptr[name]), a.nbytes) == 0
        for name in ("q", "m"):
            assert hip.hipMemcpy(ptr[name], host[name].ctypes.data, host[name].nbytes, 1) == 0
        dev.search_device(ptr["q"].value, nq, 32, 10, 8, ptr["ids"].value, ptr["dd"].value,
                          ptr["cnt"].value, masks_ptr=ptr["m"].value)
        for name in ("ids", "dd", "cnt"):
            assert hip.hipMemcpy(host[name].ctypes.data, ptr[name], host[name].nbytes, 2) == 0
    finally:
        for p_ in ptr.values():
            hip.hipFree(p_)
    np.testing.assert_array_equal(host["cnt"], want[2])
    np.testing.assert_array_equal(host["ids"], want[0])
    np.testing.assert_array_equal(host["dd"].view(np.uint32), want[1].view(np.uint32))
    ix.labels = None
    plain = DeviceIndex(ix)
    with pytest.raises(ValueError, match="without labels"):
        plain.search(q, 10, 8, masks=masks)


def test_spfresh_search_batch_filtered_surface():
    from pydiskann_amd import SPFresh
    x, ix, q, masks = labelled(2000, 32, None, "l2", seed=7)
    labels = ix.labels.copy()
    ix.labels = None
    s = SPFresh.from_arrays(ix)
    with pytest.raises(RuntimeError, match="no labels"):
        s.search_batch_filtered(q, [1, 2])
    s.set_labels(labels)
    vals = [2, 3, 40]
    ids, dd, cnt = s.search_batch_filtered(q, vals, k=10, n_probe=8)
    o = oracle.Oracle(s.arrays).search_batch(q, 10, 8, masks=np.full(
        q.shape[0], np.uint64((1 << 2) | (1 << 3) | (1 << 40))))
    np.testing.assert_array_equal(cnt, o[2])
    for i in range(q.shape[0]):
        np.testing.assert_array_equal(ids[i, :cnt[i]], o[0][i, :cnt[i]])
        np.testing.assert_array_equal(dd[i, :cnt[i]].view(np.uint32),
                                      o[1][i, :cnt[i]].view(np.uint32))
    with pytest.raises(ValueError):
        s.search_batch_filtered(q, [64])
