"""Routing graph (Inner::nearest above 4,096 live postings, src/spfresh.rs:
20,374-378): the reference's .centroids format and the oracle's restatement
of IncrementalDiskANN::search_with_dists / beam_search.

Pinned by: hand-assembled file bytes laid out by the format rules of
src/incremental.rs:1213-1327 and src/lib.rs:90-98,436-489 (decoded, and
re-encoded byte for byte); hand-traced searches on a chain graph whose
outcome the termination rule (`best >= worst`) and the beam width fix;
exhaustive equivalence (a beam as wide as the graph visits every reachable
node, so the result is the brute-force top-k of the reachable live nodes);
planted mistakes (termination rule, delta-layer entry point) that these
checks catch.  Tie order among exactly equal distances follows the restated
std BinaryHeap and is unpinned by a reference run.
"""
from __future__ import annotations

import struct

import numpy as np
import pytest

import oracle
from graph_util import graph_index, knn_graph
from pydiskann_amd import refformat
from pydiskann_amd.index import CentroidGraph, DELTA_ID_OFFSET, build_index
from pydiskann_amd.refformat import GraphFile, decode_centroids, encode_centroids


def hand_centroids_bytes() -> bytes:
    """2 base nodes (dim 2, max_degree 2), 1 delta node, 1 tombstone."""
    name = b"L2"
    md = (struct.pack("<QQQ", 2, 2, 2) + struct.pack("<I", 1) +
          struct.pack("<QQ", 1 << 20, (1 << 20) + 16) + struct.pack("<Q", len(name)) + name)
    dann = bytearray((1 << 20) + 16 + 16)
    head = struct.pack("<IIQ", 0x44414E4E, 1, len(md)) + md
    dann[:len(head)] = head
    dann[1 << 20:(1 << 20) + 16] = struct.pack("<4f", 0.0, 0.0, 1.0, 0.0)   # node 0, node 1
    dann[(1 << 20) + 16:] = struct.pack("<4I", 1, 0xFFFFFFFF, 0, 0xFFFFFFFF)  # 0 -> [1], 1 -> [0]
    b = struct.pack("<II", 0x494E4352, 2) + b"\x01" + struct.pack("<Q", len(dann)) + bytes(dann)
    b += struct.pack("<Q", 2)                                   # dim
    b += struct.pack("<Q", 1) + struct.pack("<2f", 5.0, 5.0)    # one delta vector
    b += struct.pack("<Q", 1) + struct.pack("<I", 0)            # its (empty) neighbor list
    b += struct.pack("<q", 0) + struct.pack("<Q", 2)            # entry point 0, max_degree 2
    b += struct.pack("<Q", 1) + struct.pack("<Q", 1)            # tombstone: base node 1
    b += b"\x00\x00"                                            # no labels, no quantizer
    return b


def test_centroids_file_decodes_and_reencodes_byte_for_byte():
    raw = hand_centroids_bytes()
    gf = decode_centroids(raw)
    assert gf.dim == 2 and gf.has_base and gf.medoid == 1 and gf.distance_name == "L2"
    np.testing.assert_array_equal(gf.base_vectors, [[0, 0], [1, 0]])
    np.testing.assert_array_equal(gf.base_adj, [[1, 0xFFFFFFFF], [0, 0xFFFFFFFF]])
    np.testing.assert_array_equal(gf.delta_vectors, [[5, 5]])
    assert [list(x) for x in gf.delta_graph] == [[]]
    assert gf.entry_point == 0 and gf.max_degree == 2 and gf.tombstones == [1]
    assert encode_centroids(gf) == raw
    # unified view: cid of the delta node is DELTA_ID_OFFSET + 0; postings by cid
    g = refformat.graph_from_file(gf, [0, DELTA_ID_OFFSET], centroid_beam=8)
    assert g.seeds == [1, 2] and list(g.live) == [1, 0, 1] and list(g.node_post) == [0, -1, 1]
    assert list(g.adj_off) == [0, 1, 2, 2] and list(g.adj) == [1, 0]
    assert g.n_tombstones == 1 and g.cid(2) == DELTA_ID_OFFSET
    with pytest.raises(ValueError):
        decode_centroids(raw + b"\x00")
    bad = bytearray(raw)
    bad[4:8] = struct.pack("<I", 3)
    with pytest.raises(ValueError):
        decode_centroids(bytes(bad))


def chain(n=6, positions=None):
    """Nodes on a line (dim 4), each linked to its neighbors on the line."""
    pos = np.arange(n, dtype=np.float32) if positions is None else np.asarray(positions, np.float32)
    vec = np.zeros((n, 4), np.float32)
    vec[:, 0] = pos
    lists = [[i - 1] if i == n - 1 else ([i + 1] if i == 0 else [i - 1, i + 1]) for i in range(n)]
    adj_off = np.zeros(n + 1, np.uint32)
    adj_off[1:] = np.cumsum([len(x) for x in lists])
    return CentroidGraph(vectors=vec, adj_off=adj_off, adj=np.concatenate(lists).astype(np.uint32),
                         live=np.ones(n, np.uint8), seeds=[0], node_post=np.arange(n),
                         n_tombstones=0, centroid_beam=1, base_n=n)


def graph_search(g, q, k, beam, metric=1):
    o, keep = oracle.graph_struct(g)
    qv = np.ascontiguousarray(q, np.float32)
    nodes = np.zeros(max(k, 1), np.uint32)
    dd = np.zeros(max(k, 1), np.float32)
    import ctypes as C
    c = oracle.lib().spf_o_graph_search(C.byref(o), metric, g.vectors.shape[1],
                                        qv.ctypes.data_as(C.c_void_p), k, beam,
                                        nodes.ctypes.data_as(C.c_void_p),
                                        dd.ctypes.data_as(C.c_void_p))
    return list(nodes[:c]), list(dd[:c])


def test_chain_walk_hand_traced():
    g = chain()
    q = np.array([3.2, 0, 0, 0], np.float32)
    # beam 1: the seed is both the best frontier entry and the worst of a full
    # w, and best >= worst ends the walk at once
    assert graph_search(g, q, 1, 1)[0] == [0]
    # beam 2: 0 -> 1 -> 2 -> 3 -> 4, then best (4, 0.64) >= worst (0.64) stops it
    nodes, dd = graph_search(g, q, 2, 2)
    assert nodes == [3, 4]
    np.testing.assert_allclose(dd, [0.2 ** 2, 0.8 ** 2], rtol=1e-6)
    oracle.set_fault(6)      # best > worst: the beam-1 walk no longer stops at the seed
    try:
        assert graph_search(g, q, 1, 1)[0] == [3]
    finally:
        oracle.set_fault(0)


def test_tombstoned_nodes_route_but_are_not_returned():
    g = chain()
    g.live[3] = 0
    g.n_tombstones = 1
    q = np.array([3.2, 0, 0, 0], np.float32)
    nodes, dd = graph_search(g, q, 2, 2)
    # expanded beam max(2*2, 2+1) = 4; node 3 is walked through but dropped
    assert nodes == [4, 2]
    np.testing.assert_allclose(dd, [0.8 ** 2, 1.2 ** 2], rtol=1e-6)


def reachable(g):
    seen, stack = set(g.seeds), list(g.seeds)
    while stack:
        u = stack.pop()
        for v in g.adj[g.adj_off[u]:g.adj_off[u + 1]]:
            if int(v) not in seen:
                seen.add(int(v))
                stack.append(int(v))
    return seen


@pytest.mark.parametrize("delta", [0, 40])
def test_full_beam_equals_brute_force_over_reachable_nodes(delta):
    """No tombstones: a beam wider than the graph never fills, so the walk
    visits every reachable node and returns their exact top-k."""
    rng = np.random.default_rng(3)
    cents = rng.normal(size=(300, 8)).astype(np.float32)
    g = knn_graph(cents, r=10, delta=delta, beam=1000, seed=1)
    reach = reachable(g)
    for t in range(10):
        q = rng.normal(size=8).astype(np.float32)
        nodes, dd = graph_search(g, q, 10, g.n_nodes + 10)
        cand = sorted(reach)
        d = np.array([oracle.eval_metric(1, q, g.vectors[u]) for u in cand], np.float32)
        order = np.lexsort((np.array(cand), d))[:10]
        assert nodes == [cand[i] for i in order]
        np.testing.assert_array_equal(np.array(dd, np.float32), d[order])


@pytest.mark.parametrize("dead,delta", [(15, 0), (10, 40)])
def test_tombstone_mode_returns_live_nodes_in_order(dead, delta):
    """With tombstones the walk may stop early (best > 1.5 x worst result), so
    only the result's form is fixed: live nodes, exact distances, ascending,
    and mostly the true top-k."""
    rng = np.random.default_rng(4)
    cents = rng.normal(size=(300, 8)).astype(np.float32)
    g = knn_graph(cents, r=10, delta=delta, dead=dead, beam=64, seed=1)
    live = [u for u in sorted(reachable(g)) if g.live[u]]
    hits = 0
    for t in range(10):
        q = rng.normal(size=8).astype(np.float32)
        nodes, dd = graph_search(g, q, 10, 64)
        assert len(nodes) == 10 and all(g.live[u] for u in nodes)
        assert dd == sorted(dd)
        for u, dv in zip(nodes, dd):
            assert np.float32(dv) == oracle.eval_metric(1, q, g.vectors[u])
        d = np.array([oracle.eval_metric(1, q, g.vectors[u]) for u in live], np.float32)
        hits += len(set(nodes) & {live[i] for i in np.argsort(d, kind="stable")[:10]})
    assert hits >= 80


def test_delta_entry_point_is_searched():
    # base: a chain at x in [0, 5]; delta: a chain at x in [100, 105]
    g0 = chain(6)
    d1 = chain(6, positions=np.arange(100, 106))
    vec = np.concatenate([g0.vectors, d1.vectors])
    adj = np.concatenate([g0.adj, d1.adj + 6]).astype(np.uint32)
    adj_off = np.concatenate([g0.adj_off, d1.adj_off[1:] + g0.adj_off[-1]]).astype(np.uint32)
    g = CentroidGraph(vectors=vec, adj_off=adj_off, adj=adj, live=np.ones(12, np.uint8),
                      seeds=[0, 6], node_post=np.arange(12), n_tombstones=0, centroid_beam=4,
                      base_n=6)
    q = np.array([103.1, 0, 0, 0], np.float32)
    assert graph_search(g, q, 1, 4)[0] == [9]
    oracle.set_fault(7)      # only the base medoid as a seed
    try:
        assert graph_search(g, q, 1, 4)[0] != [9]
    finally:
        oracle.set_fault(0)


def test_search_uses_the_graph_above_4096_postings_only():
    x, ix, q = graph_index(n=12000, n_post=4200, delta=200, dead=20, seed=2)
    o = oracle.Oracle(ix)
    g = ix.graph
    for i in range(8):
        posts, dd = o.probe(q[i], 16)
        nodes, gd = o.graph_search(q[i], 16, max(g.centroid_beam, 16))
        want = [int(g.node_post[u]) for u in nodes if g.node_post[u] >= 0]
        assert list(posts) == want                     # the probe is the graph's answer
        np.testing.assert_array_equal(dd, np.array(gd, np.float32)[[u >= 0 for u in
                                                                   g.node_post[nodes]]])
    # below the threshold the same graph is ignored: exact probe
    small = build_index(x[:3000], "l2", None, n_postings=100, seed=0)
    small.graph = knn_graph(small.centroids, r=8, beam=4)
    a = oracle.Oracle(small).search_batch(q, 10, 8)
    small.graph = None
    b = oracle.Oracle(small).search_batch(q, 10, 8)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_index_files_round_trip_with_the_graph(tmp_path):
    x, ix, q = graph_index(n=12000, n_post=4200, delta=150, dead=30, seed=4)
    ix.block_ids = (ix.block_ids * 2 + 3).astype(np.uint32)
    path = str(tmp_path / "gidx")
    refformat.write_index(ix, path)
    back, meta = refformat.read_index(path)
    g, h = ix.graph, back.graph
    for f in ("vectors", "adj_off", "adj", "live", "node_post"):
        np.testing.assert_array_equal(getattr(h, f), getattr(g, f), err_msg=f)
    assert (h.seeds, h.n_tombstones, h.base_n, h.centroid_beam) == \
        (g.seeds, g.n_tombstones, g.base_n, g.centroid_beam)
    a = oracle.Oracle(ix).search_batch(q, 10, 12)
    b = oracle.Oracle(back).search_batch(q, 10, 12)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
