"""GPU: the routing-graph probe (Inner::nearest above 4,096 live postings,
src/spfresh.rs:20,374-378) against the oracle's restatement of the reference's
beam search.

Bar: with a routing graph the device search returns the oracle's result bit
for bit (ids, order, counts, distances) for every metric, quantizer and probe
count tried, with and without tombstones and a delta layer; insert routing
agrees with the oracle's graph answer for nearest(v, 1); mixed precision
returns exact distances and equals the exact path when its pool covers every
candidate.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from graph_util import graph_index
from pydiskann_amd import DeviceIndex, PREC_EXACT, PREC_MIXED

pytestmark = pytest.mark.gpu


def compare(dev, ix, q, k, n_probe, masks=None):
    print(f"[case] graph nq={q.shape[0]} k={k} n_probe={n_probe} metric={ix.metric} "
          f"quant={ix.quant} tombstones={ix.graph.n_tombstones} nodes={ix.graph.n_nodes}",
          flush=True)
    ids, dd, cnt = dev.search(q, k, n_probe, masks=masks)
    rid, rd, rc = oracle.Oracle(ix).search_batch(q, k, n_probe, threads=8, masks=masks)
    np.testing.assert_array_equal(cnt, rc)
    for i in range(q.shape[0]):
        c = int(cnt[i])
        np.testing.assert_array_equal(ids[i, :c], rid[i, :c], err_msg=f"query {i}")
        assert np.array_equal(dd[i, :c].view(np.uint32), rd[i, :c].view(np.uint32)), \
            f"query {i}: distances not bit-identical"


@pytest.mark.parametrize("delta,dead", [(0, 0), (300, 0), (0, 40), (300, 40)])
def test_graph_probe_parity_layers_and_tombstones(delta, dead):
    x, ix, q = graph_index(n=30000, n_post=4200, delta=delta, dead=dead, seed=1)
    dev = DeviceIndex(ix)
    for n_probe in (1, 8, 32, 100):
        compare(dev, ix, q, 10, n_probe)
    compare(dev, ix, q, 100, 64)


# dead 0: no tombstones, beam 64 (the wave-uniform walk); dead 25: tombstones
# (the walk with a result list)
@pytest.mark.parametrize("dead", [0, 25])
@pytest.mark.parametrize("quant,metric,rerank", [(None, "cosine", 0), (None, "dot", 0),
                                                 ("int8", "l2", 64), ("f16", "l2", 0),
                                                 ("rabitq", "l2", 0)])
def test_graph_probe_parity_metrics_and_quantizers(quant, metric, rerank, dead):
    x, ix, q = graph_index(n=24000, n_post=4150, quant=quant, metric=metric, rerank=rerank,
                           delta=200, dead=dead, seed=2)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 16)


def test_graph_probe_wide_beam_and_probe_count_above_postings():
    x, ix, q = graph_index(n=20000, n_post=4100, delta=100, dead=10, beam=200, seed=3)
    dev = DeviceIndex(ix)
    compare(dev, ix, q[:16], 10, 150)            # beam = max(200, n_probe)
    compare(dev, ix, q[:8], 10, 5000)            # n_probe above the posting count


def test_graph_probe_filtered():
    x, ix, q = graph_index(n=30000, n_post=4200, delta=300, dead=40, seed=4)
    rng = np.random.default_rng(4)
    ix.labels = rng.integers(0, 64, size=ix.n_raw).astype(np.uint8)
    masks = rng.integers(0, 2 ** 63, size=q.shape[0], dtype=np.uint64)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 32, masks=masks)


def test_routing_uses_the_graph():
    x, ix, q = graph_index(n=30000, n_post=4200, delta=300, dead=40, seed=5)
    dev = DeviceIndex(ix)
    v = x[::97][:200]
    blocks, dd = dev.nearest_postings(v)
    o = oracle.Oracle(ix)
    for i in range(v.shape[0]):
        posts, pd = o.probe(v[i], 1)
        assert blocks[i] == ix.block_ids[posts[0]], i
        assert np.float32(dd[i]).view(np.uint32) == np.float32(pd[0]).view(np.uint32), i


def test_graph_probe_mixed_precision():
    x, ix, q = graph_index(n=30000, n_post=4200, delta=300, dead=40, seed=6)
    dev = DeviceIndex(ix)
    print("[case] graph mixed", flush=True)
    e = dev.search(q, 10, 8, precision=PREC_EXACT)
    m = dev.search(q, 10, 8, precision=PREC_MIXED, rerank_size=2048)
    np.testing.assert_array_equal(e[2], m[2])
    np.testing.assert_array_equal(e[0], m[0])
    np.testing.assert_array_equal(e[1].view(np.uint32), m[1].view(np.uint32))


def test_graph_probe_frontier_outgrows_first_share():
    # Along a 20,000-node line with 128 links per node (two 64-lane batches
    # per expansion) and beam 32, queries near the entry end walk with a small
    # frontier, queries 1,000-3,000 along it outgrow the first LDS share
    # (640 entries) and walk again with the full one, and queries beyond
    # 12,000 outgrow that too and walk in global scratch; one batch holds all
    # three kinds, interleaved.
    from graph_util import line_index
    x, ix = line_index()
    rng = np.random.default_rng(8)
    at = np.concatenate([rng.uniform(-5, 40, 8), rng.uniform(1000, 3000, 8),
                         rng.uniform(12000, 20050, 8)])
    q = np.zeros((24, 8), np.float32)
    q[:, 0] = at[np.argsort(rng.random(24))]
    q[:, 1:] = rng.normal(0, 0.3, size=(24, 7))
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 8)
    compare(dev, ix, q, 10, 40)


@pytest.mark.parametrize("metric", ["l2", "cosine"])
def test_graph_walk_wide_rows(metric):
    # 200 dims: 50 four-dim units, each lane of a pair loads them in 3 batches
    x, ix, q = graph_index(n=20000, dim=200, n_post=4300, metric=metric, delta=0, dead=0, r=24,
                           seed=9)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 16)
    compare(dev, ix, q, 50, 64)


@pytest.mark.parametrize("r,repeat", [(12, 3), (24, 24)])
def test_graph_walk_neighbours_listed_twice(r, repeat):
    # Each node's list followed by a repeat of its first `repeat` neighbours:
    # 12 + 3 repeats fall in one 32-neighbour pass (the repeat found by the
    # visited bits' atomics, then its first copy by a shuffle scan); 24 + 24
    # spill into a second pass, where the repeats are already visited.
    from pydiskann_amd.index import CentroidGraph
    x, ix, q = graph_index(n=30000, n_post=4200, r=r, delta=0, dead=0, seed=12)
    g = ix.graph
    lists = [g.adj[g.adj_off[i]:g.adj_off[i + 1]] for i in range(g.n_nodes)]
    lists = [np.concatenate([v, v[:repeat]]).astype(np.uint32) for v in lists]
    adj_off = np.zeros(g.n_nodes + 1, dtype=np.uint32)
    adj_off[1:] = np.cumsum([len(v) for v in lists])
    ix.graph = CentroidGraph(vectors=g.vectors, adj_off=adj_off, adj=np.concatenate(lists),
                             live=g.live, seeds=g.seeds, node_post=g.node_post,
                             n_tombstones=g.n_tombstones, centroid_beam=g.centroid_beam,
                             base_n=g.base_n)
    dev = DeviceIndex(ix)
    compare(dev, ix, q, 10, 16)
    compare(dev, ix, q, 10, 64)
