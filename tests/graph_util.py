"""Synthetic routing graphs for tests (the reference builds its own with
Vamana, src/lib.rs:436-489, which this path only reads)."""
from __future__ import annotations

import numpy as np

from pydiskann_amd.index import CentroidGraph, build_index


def _knn_lists(v: np.ndarray, r: int) -> list:
    d = ((v[:, None, :].astype(np.float64) - v[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    order = np.argsort(d, axis=1, kind="stable")[:, :r]
    return [order[i].astype(np.uint32) for i in range(v.shape[0])]


def _medoid(v: np.ndarray) -> int:
    c = v.mean(0)
    return int(np.argmin(((v - c) ** 2).sum(1)))


def knn_graph(centroids: np.ndarray, r: int = 12, delta: int = 0, dead: int = 0,
              beam: int = 64, seed: int = 0) -> CentroidGraph:
    """kNN routing graph over `centroids` (node i = posting i for the first
    len(centroids) nodes): the last `delta` postings form the delta layer
    (own kNN lists, own entry point); `dead` extra tombstoned nodes (no
    posting) are added to the base, linked from their nearest base nodes."""
    rng = np.random.default_rng(seed)
    p, dim = centroids.shape
    bn_live = p - delta
    extra = (centroids[rng.integers(0, bn_live, dead)] +
             0.01 * rng.normal(size=(dead, dim))).astype(np.float32)
    base = np.concatenate([centroids[:bn_live], extra]).astype(np.float32)
    blists = _knn_lists(base, min(r, base.shape[0] - 1)) if base.shape[0] > 1 else [np.zeros(0, np.uint32)]
    dv = centroids[bn_live:].astype(np.float32)
    dlists = _knn_lists(dv, min(r, delta - 1)) if delta > 1 else [np.zeros(0, np.uint32)] * delta
    bn = base.shape[0]
    vec = np.concatenate([base, dv]).astype(np.float32)
    lists = blists + [x + np.uint32(bn) for x in dlists]
    adj_off = np.zeros(len(lists) + 1, dtype=np.uint32)
    adj_off[1:] = np.cumsum([len(x) for x in lists])
    adj = np.concatenate(lists).astype(np.uint32)
    live = np.ones(vec.shape[0], dtype=np.uint8)
    live[bn_live:bn] = 0
    node_post = np.full(vec.shape[0], -1, dtype=np.int64)
    node_post[:bn_live] = np.arange(bn_live)
    node_post[bn:] = np.arange(bn_live, p)
    seeds = [_medoid(base)] + ([bn + _medoid(dv)] if delta else [])
    return CentroidGraph(vectors=vec, adj_off=adj_off, adj=adj, live=live, seeds=seeds,
                         node_post=node_post, n_tombstones=dead, centroid_beam=beam, base_n=bn)


def graph_index(n=30000, dim=16, n_post=4200, quant=None, metric="l2", r=12, delta=300, dead=40,
                beam=64, rerank=0, seed=0):
    """Clustered data, postings seeded from sampled points (as the reference's
    build seeds them), and a kNN routing graph over the centroids."""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(200, dim))
    x = (c[rng.integers(0, 200, n)] + 0.4 * rng.normal(size=(n, dim))).astype(np.float32)
    if metric in ("dot", "cosine"):
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    cents = x[rng.choice(n, n_post, replace=False)]
    ix = build_index(x, metric, quant, centroids=cents, rerank_size=rerank, seed=seed)
    ix.graph = knn_graph(ix.centroids, r=r, delta=delta, dead=dead, beam=beam, seed=seed)
    q = (c[rng.integers(0, 200, 64)] + 0.4 * rng.normal(size=(64, dim))).astype(np.float32)
    if metric in ("dot", "cosine"):
        q /= np.linalg.norm(q, axis=1, keepdims=True)
    return x, ix, q


def line_index(n=20000, r=64, beam=32, seed=0):
    """Postings whose centroids lie on a line (1.0 apart along axis 0), each
    linked to the r nodes on either side, entered at one end.  A query beyond
    the far end makes every expansion reveal r nodes nearer than all seen so
    far, each of which enters the w heap and the frontier: the frontier grows
    by about r - 1 per expansion, so the walk's frontier size is set by how
    far along the line the query lies."""
    rng = np.random.default_rng(seed)
    cents = np.zeros((n, 8), np.float32)
    cents[:, 0] = np.arange(n, dtype=np.float32)
    x = (np.repeat(cents, 2, axis=0) + rng.normal(0, 0.05, size=(2 * n, 8))).astype(np.float32)
    ix = build_index(x, "l2", None, centroids=cents, seed=seed)
    c = ix.centroids.astype(np.float32)
    order = np.argsort(c[:, 0], kind="stable")
    p = c.shape[0]
    lists = [None] * p
    for j in range(p):
        nb = [order[t] for t in range(j + 1, min(p, j + r + 1))] + \
             [order[t] for t in range(max(0, j - r), j)]
        lists[order[j]] = np.array(nb, dtype=np.uint32)
    adj_off = np.zeros(p + 1, dtype=np.uint32)
    adj_off[1:] = np.cumsum([len(v) for v in lists])
    ix.graph = CentroidGraph(vectors=c, adj_off=adj_off, adj=np.concatenate(lists),
                             live=np.ones(p, np.uint8), seeds=[int(order[0])],
                             node_post=np.arange(p, dtype=np.int64), n_tombstones=0,
                             centroid_beam=beam, base_n=p)
    return x, ix
