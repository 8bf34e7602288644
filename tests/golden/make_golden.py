"""Writes the committed golden fixtures of the SPFresh batch search.

The reference commits no golden vectors for this path, so these fixtures are
produced here: each holds a small index (the arrays the device mirror is built
from), its queries, and the oracle's expected (ids, dists, counts) for a
shallow probe and for an exhaustive probe.  Before a fixture is written its
exhaustive-probe result is cross-checked against FP64 brute force over the live
vectors (ids equal wherever the FP64 gap between neighbours is clear of f32
rounding; distances within 1e-5 relative); reranked and RaBitQ-L2 fixtures are
checked to carry the exact distance of every returned id, and every fixture to
reach recall@k >= 0.9.  A fixture therefore cannot freeze an oracle mistake
that changes the answer.

    python tests/golden/make_golden.py          # rewrites tests/golden/*.npz
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from pydiskann_amd.index import build_index, brute_force_f64, IndexArrays  # noqa: E402

# name: (n, d, n_postings, metric, quantizer, rerank, k, n_probe, deleted every)
CASES = {
    "c1_small_f32_l2": (2000, 32, 16, "l2", None, 0, 10, 4, 0),
    "f32_l2root_deletes": (1500, 24, 12, "l2_root", None, 0, 10, 3, 7),
    "f16_cosine_rerank": (1500, 40, 12, "cosine", "f16", 48, 10, 4, 0),
    "int8_l2_rerank": (2000, 32, 16, "l2", "int8", 64, 10, 4, 0),
    "int8_dot": (1500, 48, 12, "dot", "int8", 0, 10, 4, 0),
    "rabitq_l2": (2000, 48, 16, "l2", "rabitq", 0, 10, 4, 0),
    "rabitq_dot_rerank": (1500, 40, 12, "dot", "rabitq", 32, 10, 4, 0),
}

ARRAY_FIELDS = ("block_ids", "centroids", "post_off", "ids", "codes", "raw", "deleted")


def dataset(n, d, seed):
    """config-1 shape scaled down: centres N(0, 4I), x = centre + N(0, I)."""
    ss = np.random.SeedSequence(seed).spawn(3)
    centres = np.random.default_rng(ss[0]).normal(0, 2.0, size=(max(n // 100, 4), d))
    lab = np.random.default_rng(ss[1]).integers(0, centres.shape[0], size=n)
    x = centres[lab] + np.random.default_rng(ss[2]).normal(0, 1.0, size=(n, d))
    q = centres[lab[:40]] + np.random.default_rng(ss[2]).normal(0, 1.0, size=(40, d))
    return x.astype(np.float32), q.astype(np.float32)


def fixture_arrays(ix: IndexArrays) -> dict:
    out = {f: getattr(ix, f) for f in ARRAY_FIELDS}
    out.update(dim=ix.dim, metric=ix.metric, quant=ix.quant, int8=np.array(
        [ix.int8_scale, ix.int8_offset], np.float32), rq_padded=ix.rq_padded,
        rq_signs=ix.rq_signs if ix.rq_signs is not None else np.zeros((0,), np.float32),
        probe_ratio=np.float32(ix.probe_ratio), rerank_size=ix.rerank_size)
    return out


def load(path: str):
    """(IndexArrays, fixture dict) from a fixture file."""
    z = dict(np.load(path, allow_pickle=False))
    ix = IndexArrays(dim=int(z["dim"]), metric=int(z["metric"]), quant=int(z["quant"]),
                     block_ids=z["block_ids"], centroids=z["centroids"], post_off=z["post_off"],
                     ids=z["ids"], codes=z["codes"], raw=z["raw"], deleted=z["deleted"],
                     int8_scale=float(z["int8"][0]), int8_offset=float(z["int8"][1]),
                     rq_padded=int(z["rq_padded"]),
                     rq_signs=z["rq_signs"].reshape(3, -1) if z["rq_signs"].size else None,
                     probe_ratio=float(z["probe_ratio"]), rerank_size=int(z["rerank_size"]))
    return ix, z


def check_against_fp64(ix, q, k, ids, dd, cnt):
    live = live_mask(ix)
    gt, gd = brute_force_f64(ix.raw, q, k + 1, ix.metric, live=live)
    for i in range(q.shape[0]):
        c = int(cnt[i])
        assert c == min(k, int(live.sum())), (i, c)
        np.testing.assert_allclose(dd[i, :c], gd[i, :c], rtol=1e-5, atol=1e-6)
        for j in range(c):
            lo = gd[i, j - 1] if j else -math.inf
            hi = gd[i, j + 1]
            clear = (gd[i, j] - lo) > 1e-5 * abs(gd[i, j]) + 1e-6 and \
                    (hi - gd[i, j]) > 1e-5 * abs(gd[i, j]) + 1e-6
            if clear:
                assert int(ids[i, j]) == int(gt[i, j]), (i, j)


def live_mask(ix):
    bits = np.unpackbits(ix.deleted, bitorder="little")[: ix.n_raw]
    return bits == 0


def check_exact_distances(ix, q, ids, dd, cnt):
    for i in range(q.shape[0]):
        c = int(cnt[i])
        _, gd = brute_force_f64(ix.raw[ids[i, :c].astype(np.int64)], q[i:i + 1], c, ix.metric)
        want = np.sort(gd[0])
        np.testing.assert_allclose(dd[i, :c], want, rtol=1e-5, atol=1e-5)


def check_recall(ix, q, k, ids, floor):
    gt, _ = brute_force_f64(ix.raw, q, k, ix.metric, live=live_mask(ix))
    r = np.mean([len(set(ids[i, :k].tolist()) & set(gt[i].tolist())) / k
                 for i in range(q.shape[0])])
    assert r >= floor, r


def make(name, spec, seed):
    n, d, P, metric, quant, rerank, k, n_probe, every = spec
    x, q = dataset(n, d, seed)
    ix = build_index(x, metric, quant, n_postings=P, rerank_size=rerank, seed=0)
    if every:
        ix.set_deleted(np.arange(0, n, every, dtype=np.uint64))
    o = oracle.Oracle(ix)
    out = fixture_arrays(ix)
    out.update(queries=q, k=k, n_probe=n_probe)
    for tag, npr in (("shallow", n_probe), ("full", ix.n_post)):
        ids, dd, cnt = o.search_batch(q, k, npr, threads=4)
        out[f"{tag}_ids"], out[f"{tag}_dists"], out[f"{tag}_counts"] = ids, dd, cnt
    # exhaustive probe of an unquantized index is exact k-NN; reranked and
    # RaBitQ-L2 results carry the exact distance of each returned id; every
    # case reaches recall@k >= 0.9 against FP64
    if quant is None:
        check_against_fp64(ix, q, k, out["full_ids"], out["full_dists"], out["full_counts"])
    elif rerank or (quant == "rabitq" and metric in ("l2", "l2_root")):
        check_exact_distances(ix, q, out["full_ids"], out["full_dists"], out["full_counts"])
    check_recall(ix, q, k, out["full_ids"], 0.9)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


def main():
    for seed, (name, spec) in enumerate(CASES.items()):
        p = make(name, spec, 1000 + seed)
        print(f"wrote {os.path.relpath(p)} ({os.path.getsize(p)} bytes)")


if __name__ == "__main__":
    main()
