"""GPU: launch-shape autotuner (spf_autotune).

Bar: every shape it may pick returns the same result bit for bit (each shape
is forced through a cache line and compared with the default); the chosen
shape is one of the candidates; the cache is read back without timing.
"""
from __future__ import annotations

import numpy as np
import pytest

from pydiskann_amd import DeviceIndex, build_index, PREC_EXACT, PREC_MIXED

pytestmark = pytest.mark.gpu


def index(quant, metric="l2"):
    rng = np.random.default_rng(7)
    c = rng.normal(size=(40, 48))
    x = (c[rng.integers(0, 40, 30000)] + 0.3 * rng.normal(size=(30000, 48))).astype(np.float32)
    q = (c[rng.integers(0, 40, 300)] + 0.3 * rng.normal(size=(300, 48))).astype(np.float32)
    return build_index(x, metric, quant, n_postings=300, rerank_size=64 if quant else 0, seed=0), q


def same(a, b):
    np.testing.assert_array_equal(a[2], b[2])
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@pytest.mark.parametrize("quant,precision,shapes", [
    ("int8", PREC_EXACT, [(4, 256, 128), (16, 2048, 128), (8, 512, 128)]),
    (None, PREC_EXACT, [(4, 2048, 128), (16, 256, 128)]),
    (None, PREC_MIXED, [(8, 1024, 32), (8, 1024, 64)])])
def test_every_shape_gives_the_same_result(tmp_path, quant, precision, shapes):
    ix, q = index(quant)
    dev = DeviceIndex(ix)
    print(f"[case] autotune quant={quant} precision={precision}", flush=True)
    base = dev.search(q, 10, 12, precision=precision)
    cache = str(tmp_path / "tune.txt")
    r = dev.autotune(q, 10, 12, precision=precision, cache_path=cache)
    assert r["from_cache"] == 0 and r["tried"] >= 1 and r["best_ms"] <= r["default_ms"] * 1.0001
    assert r["scan_g"] in (4, 8, 16) and r["scan_chunk"] in (256, 512, 1024, 2048)
    assert r["mix_g"] in (32, 64, 128)
    same(dev.search(q, 10, 12, precision=precision), base)
    key = open(cache).read().split(" scan_g=")[0]
    for g, c, m in shapes:   # force each shape through the cache
        open(cache, "w").write(f"{key} scan_g={g} scan_chunk={c} mix_g={m}\n")
        d2 = DeviceIndex(ix)
        r2 = d2.autotune(q, 10, 12, precision=precision, cache_path=cache)
        assert r2["from_cache"] == 1 and (r2["scan_g"], r2["scan_chunk"], r2["mix_g"]) == (g, c, m)
        same(d2.search(q, 10, 12, precision=precision), base)
    # a malformed line is ignored (tuned afresh)
    open(cache, "w").write(f"{key} scan_g=5 scan_chunk=1000 mix_g=7\n")
    assert DeviceIndex(ix).autotune(q, 10, 12, precision=precision, cache_path=cache)["from_cache"] == 0
