"""GPU: candidate pools of long segments (the chunked select).

A query whose candidate segment holds more than 24,576 slots and more than
half the batch's slots per concurrent pool-select workgroup (1,024 of them)
has its pool
selected by many workgroups: each 24,576-slot chunk contributes its smallest
(key, slot) values in slot order, and the pool and phase A are selected from
the concatenated chunk lists with the same rule.  Bar: the results equal the
two-phase admission path's bit for bit (SPF_FLAG_ADMIT: its pool is selected
by another kernel from per-query lists), for float and tie-heavy integer data,
L2 / cosine / dot, f32 and int8 postings, pools from 100 to 2,048 and k 10 /
100, with deletes; and equal the full-pool rerank of the same pool.
"""
from __future__ import annotations

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, PREC_MIXED
from pydiskann_amd import _lib

pytestmark = pytest.mark.gpu


def data(kind, n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 60, size=(8, d))
    x = c[rng.integers(0, 8, n)] + rng.normal(0, 6, size=(n, d))
    if kind == "int":
        x = np.clip(np.round(x), 0, 255)
    return x.astype(np.float32)


def same(a, b, label):
    np.testing.assert_array_equal(a[2], b[2], err_msg=label)
    np.testing.assert_array_equal(a[0], b[0], err_msg=label)
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32), err_msg=label)


@pytest.mark.parametrize("kind", ["float", "int"])
@pytest.mark.parametrize("metric,quant", [("l2", None), ("cosine", None), ("dot", "int8"),
                                          ("l2", "int8")])
def test_chunked_pools_equal_admission_pools(kind, metric, quant):
    x = data(kind, 50000, 48, seed=3)
    q = data(kind, 160, 48, seed=4)
    ix = build_index(x, metric, quant, n_postings=12, rerank_size=100, seed=0)
    ix.set_deleted(np.arange(0, 50000, 7))
    dev = DeviceIndex(ix)
    for pool, k in ((100, 10), (400, 10), (1000, 100), (2048, 100)):
        label = f"{kind} {metric} {quant} pool {pool} k {k}"
        print(f"[case] {label}", flush=True)
        one = dev.search(q, k, 12, precision=PREC_MIXED, rerank_size=pool)
        adm = dev.search(q, k, 12, precision=PREC_MIXED, rerank_size=pool, flags=_lib.FLAG_ADMIT)
        full = dev.search(q, k, 12, precision=PREC_MIXED, rerank_size=pool,
                          flags=_lib.FLAG_FULL_RERANK)
        assert (one[2] == k).all()
        same(one, adm, label + " vs admission")
        same(one, full, label + " vs full rerank")
    dev.close()
