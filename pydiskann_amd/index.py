"""Host-side SPFresh index arrays and a deterministic index builder.

``IndexArrays`` is the in-memory form of the state that ``Inner::search``
reads (src/spfresh.rs:165-195): live postings with their centroids
(``block_centroid``), each posting's ``(id, code)`` entries
(src/spfresh.rs:258-270), the raw vector store by id (src/spfresh.rs:271-273),
the tombstone set, the quantizer state and the two search knobs of
``SPFreshConfig`` (probe_ratio, rerank_size; src/spfresh.rs:25-39).  The
postings are CSR (offsets + entries) instead of fixed 2*max_posting_size
blocks; the entry order inside a posting is kept.

``build_index`` partitions vectors with seeded k-means and one nearest-centroid
assignment.  The reference grows its partition by inserts and splits under a
thread RNG (src/spfresh.rs:786-817), so its partitions are not reproducible;
search parity is defined on a given index, which this builder provides.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import codecs

# Python "l2" is squared L2 (python/src/lib.rs:13,28); "l2_root" is DistL2.
METRICS = {"l2_root": 0, "l2": 1, "l2sq": 1, "cosine": 2, "dot": 3}
METRIC_NAMES = {0: "l2_root", 1: "l2", 2: "cosine", 3: "dot"}
QUANTS = {None: 0, "none": 0, "f16": 1, "int8": 2, "rabitq": 3}
QUANT_NAMES = {0: None, 1: "f16", 2: "int8", 3: "rabitq"}


def metric_code(metric: str) -> int:
    if metric not in METRICS:
        raise ValueError(f"unknown metric: {metric}")
    return METRICS[metric]


def quant_code(quantizer) -> int:
    if quantizer == "pq":
        raise NotImplementedError(
            "pq postings have no device path; the reference CPU search serves them")
    if quantizer not in QUANTS:
        raise ValueError(f"unknown quantizer: {quantizer}")
    return QUANTS[quantizer]


def code_bytes_for(quant: int, dim: int, rq_padded: int = 0) -> int:
    """code_size (src/spfresh.rs:99-107)."""
    if quant == 0:
        return 4 * dim
    if quant == 1:
        return 2 * dim
    if quant == 2:
        return dim
    if quant == 3:
        return codecs.rabitq_code_bytes(rq_padded)
    raise ValueError(quant)


DELTA_ID_OFFSET = 1 << 48   # global id of delta node i = DELTA_ID_OFFSET + i (src/incremental.rs:334)


@dataclass
class CentroidGraph:
    """The reference's centroid routing graph (an IncrementalDiskANN over the
    posting centroids: DANN base + delta layer + tombstones), which
    Inner::nearest searches above 4,096 live postings (src/spfresh.rs:20,
    374-378), in unified node ids: base nodes 0..base_n-1, then delta nodes
    (UnifiedView, src/incremental.rs:357-487)."""
    vectors: np.ndarray            # f32 [n, d] node vectors (base, then delta)
    adj_off: np.ndarray            # u32 [n + 1]
    adj: np.ndarray                # u32 [E] neighbor lists in stored order, PAD removed
    live: np.ndarray               # u8 [n] 0 = tombstoned
    seeds: list                    # entry points: base medoid, then base_n + delta entry
    node_post: np.ndarray          # i64 [n] posting index of the node's cid, -1 = none
    n_tombstones: int = 0          # size of the tombstone set
    centroid_beam: int = 64        # SPFreshConfig::centroid_beam
    base_n: int = 0

    @property
    def n_nodes(self) -> int:
        return int(self.vectors.shape[0])

    def cid(self, node: int) -> int:
        """Global u64 id of a unified node (UnifiedView::to_global_u64)."""
        return node if node < self.base_n else DELTA_ID_OFFSET + node - self.base_n


@dataclass
class IndexArrays:
    dim: int
    metric: int
    quant: int
    block_ids: np.ndarray          # u32 [P]  reference block index of each live posting
    centroids: np.ndarray          # f32 [P, d]
    post_off: np.ndarray           # u64 [P + 1]
    ids: np.ndarray                # u64 [N]
    codes: np.ndarray              # u8  [N, code_bytes]
    raw: np.ndarray                # f32 [n_raw, d] by id
    deleted: np.ndarray = None     # u8 bitmap over ids < n_raw
    int8_scale: float = 1.0
    int8_offset: float = 0.0
    rq_padded: int = 0
    rq_signs: np.ndarray = None    # f32 [3, padded]
    probe_ratio: float = math.inf
    rerank_size: int = 0
    labels: np.ndarray = None      # optional u8 [n_raw] label per id (label-filtered search)
    graph: CentroidGraph = None    # routing graph above 4,096 postings (None = exact probe)

    def __post_init__(self):
        self.block_ids = np.ascontiguousarray(self.block_ids, dtype=np.uint32)
        self.centroids = np.ascontiguousarray(self.centroids, dtype=np.float32)
        self.post_off = np.ascontiguousarray(self.post_off, dtype=np.uint64)
        self.ids = np.ascontiguousarray(self.ids, dtype=np.uint64)
        self.codes = np.ascontiguousarray(self.codes, dtype=np.uint8)
        self.raw = np.ascontiguousarray(self.raw, dtype=np.float32)
        if self.deleted is None:
            self.deleted = np.zeros((self.n_raw + 7) // 8, dtype=np.uint8)
        self.deleted = np.ascontiguousarray(self.deleted, dtype=np.uint8)
        if self.rq_signs is not None:
            self.rq_signs = np.ascontiguousarray(self.rq_signs, dtype=np.float32)

    @property
    def n_post(self) -> int:
        return int(self.block_ids.shape[0])

    @property
    def n_entries(self) -> int:
        return int(self.ids.shape[0])

    @property
    def n_raw(self) -> int:
        return int(self.raw.shape[0])

    @property
    def code_bytes(self) -> int:
        return int(self.codes.shape[1]) if self.codes.ndim == 2 else 0

    def posting_lengths(self) -> np.ndarray:
        return np.diff(self.post_off.astype(np.int64))

    def set_deleted(self, ids) -> int:
        """Tombstone ids < n_raw (SPFresh::delete, src/spfresh.rs:867-874).
        Returns how many were newly deleted."""
        new = 0
        for i in np.asarray(ids, dtype=np.uint64).ravel():
            i = int(i)
            if i < self.n_raw and not (self.deleted[i >> 3] >> (i & 7)) & 1:
                self.deleted[i >> 3] |= np.uint8(1 << (i & 7))
                new += 1
        return new

    def is_deleted(self, i: int) -> bool:
        return i < self.n_raw and bool((self.deleted[i >> 3] >> (i & 7)) & 1)

    def n_deleted(self) -> int:
        return int(np.unpackbits(self.deleted, bitorder="little")[: self.n_raw].sum())


# ------------------------------------------------------------------ k-means
def _sqnorm(x: np.ndarray) -> np.ndarray:
    return np.einsum("ij,ij->i", x, x, dtype=np.float32)


def assign_nearest(x: np.ndarray, c: np.ndarray, chunk: int = 16384, device=None) -> np.ndarray:
    """Index of the L2-nearest centroid for each row (ties -> lowest index)."""
    n = x.shape[0]
    out = np.empty(n, dtype=np.int64)
    if device is not None:
        import torch
        ct = torch.from_numpy(c).to(device)
        cn = (ct * ct).sum(1)
        for s in range(0, n, chunk):
            xt = torch.from_numpy(x[s:s + chunk]).to(device)
            d = cn[None, :] - 2.0 * (xt @ ct.T)
            out[s:s + chunk] = d.argmin(1).cpu().numpy()
        return out
    cn = _sqnorm(c)
    for s in range(0, n, chunk):
        xs = x[s:s + chunk]
        d = cn[None, :] - 2.0 * (xs @ c.T)
        out[s:s + chunk] = d.argmin(1)
    return out


def kmeans(x: np.ndarray, k: int, iters: int = 10, seed: int = 0, sample: int = 262144,
           device=None) -> np.ndarray:
    """Seeded Lloyd's k-means on a sample; returns [k, d] f32 centroids."""
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    if k >= n:
        return x[:k].copy()
    xs = x[rng.choice(n, size=min(n, max(sample, 4 * k)), replace=False)] if n > sample else x
    c = xs[rng.choice(xs.shape[0], size=k, replace=False)].copy()
    for _ in range(iters):
        a = assign_nearest(xs, c, device=device)
        order = np.argsort(a, kind="stable")
        cnt = np.bincount(a, minlength=k)
        sums = np.zeros((k, x.shape[1]), dtype=np.float64)
        nz = np.nonzero(cnt)[0]
        starts = np.concatenate([[0], np.cumsum(cnt)[:-1]])[nz]
        sums[nz] = np.add.reduceat(xs[order].astype(np.float64), starts, axis=0)
        empty = cnt == 0
        c = (sums / np.maximum(cnt, 1)[:, None]).astype(np.float32)
        if empty.any():
            c[empty] = xs[rng.choice(xs.shape[0], size=int(empty.sum()), replace=False)]
    return c


# ------------------------------------------------------------------ builder
def build_index(vectors, metric: str = "l2", quantizer=None, n_postings: int | None = None,
                max_posting_size: int = 128, rerank_size: int = 0,
                probe_ratio: float = math.inf, seed: int = 0, kmeans_iters: int = 10,
                device=None, centroids: np.ndarray | None = None,
                assign: np.ndarray | None = None) -> IndexArrays:
    """Partition `vectors` (ids 0..n-1) into postings and encode them.

    n_postings defaults to the reference's initial posting count
    n / (max_posting_size / 2) (src/spfresh.rs:807)."""
    m = metric_code(metric)
    q = quant_code(quantizer)
    x = np.ascontiguousarray(vectors, dtype=np.float32)
    if x.ndim != 2 or x.shape[0] == 0:
        raise ValueError("No vectors provided")
    n, d = x.shape
    if n_postings is None:
        n_postings = max(1, n // max(1, max_posting_size // 2))
    n_postings = min(n_postings, n)
    if centroids is None:
        centroids = kmeans(x, n_postings, iters=kmeans_iters, seed=seed, device=device)
    centroids = np.ascontiguousarray(centroids, dtype=np.float32)
    if assign is None:
        assign = assign_nearest(x, centroids, device=device)
    # drop empty postings (a live posting in the reference always holds entries
    # after build); keep block ids 0..P-1 in centroid order
    cnt = np.bincount(assign, minlength=centroids.shape[0])
    keep = np.nonzero(cnt > 0)[0]
    remap = np.full(centroids.shape[0], -1, dtype=np.int64)
    remap[keep] = np.arange(keep.shape[0])
    assign = remap[assign]
    centroids = centroids[keep]
    P = centroids.shape[0]
    order = np.argsort(assign, kind="stable")     # id order inside each posting
    counts = np.bincount(assign, minlength=P)
    post_off = np.zeros(P + 1, dtype=np.uint64)
    post_off[1:] = np.cumsum(counts)
    ids = order.astype(np.uint64)
    ent = x[order]
    scale, offset, padded, signs = 1.0, 0.0, 0, None
    if q == 0:
        codes = ent.view(np.uint8).reshape(n, 4 * d)
    elif q == 1:
        codes = codecs.f16_encode(ent)
    elif q == 2:
        scale, offset = codecs.int8_train(x)
        codes = codecs.int8_encode(ent, scale, offset)
    else:
        padded = codecs.rabitq_padded(d)
        signs = codecs.rabitq_signs(padded)
        codes = codecs.rabitq_encode(ent, centroids[assign[order]], padded, signs)
    return IndexArrays(dim=d, metric=m, quant=q, block_ids=np.arange(P, dtype=np.uint32),
                       centroids=centroids, post_off=post_off, ids=ids, codes=codes, raw=x,
                       int8_scale=scale, int8_offset=offset, rq_padded=padded, rq_signs=signs,
                       probe_ratio=probe_ratio, rerank_size=rerank_size)


def changed_blocks(old: IndexArrays, new: IndexArrays) -> list:
    """Reference block indices whose posting differs between two states of an
    index: blocks present in only one of them (created or freed) and blocks
    whose centroid or (id, code) entries changed -- the set Inner::mark_dirty
    records for a device mirror (src/spfresh.rs:283-289)."""
    def table(ix):
        t = {}
        for i, b in enumerate(ix.block_ids.tolist()):
            s, e = int(ix.post_off[i]), int(ix.post_off[i + 1])
            t[b] = (ix.centroids[i], ix.ids[s:e], ix.codes[s:e])
        return t
    a, b = table(old), table(new)
    out = []
    for blk in sorted(set(a) | set(b)):
        if blk not in a or blk not in b:
            out.append(blk)
            continue
        (ca, ia, xa), (cb, ib, xb) = a[blk], b[blk]
        if not (np.array_equal(ca.view(np.uint32), cb.view(np.uint32)) and np.array_equal(ia, ib)
                and np.array_equal(xa, xb)):
            out.append(blk)
    return out


def brute_force_f64(raw: np.ndarray, queries: np.ndarray, k: int, metric: int,
                    live: np.ndarray | None = None, chunk: int = 256):
    """Exact k-NN in float64 (ids, dists); ties by id.  Used for recall."""
    x = np.asarray(raw, dtype=np.float64)
    qs = np.asarray(queries, dtype=np.float64)
    ids_out = np.zeros((qs.shape[0], k), dtype=np.int64)
    d_out = np.zeros((qs.shape[0], k), dtype=np.float64)
    xn = np.einsum("ij,ij->i", x, x)
    for s in range(0, qs.shape[0], chunk):
        q = qs[s:s + chunk]
        ip = q @ x.T
        if metric in (0, 1):
            dd = np.maximum(xn[None, :] - 2 * ip + np.einsum("ij,ij->i", q, q)[:, None], 0)
            if metric == 0:
                dd = np.sqrt(dd)
        elif metric == 3:
            dd = 1.0 - ip
        else:
            qn = np.sqrt(np.einsum("ij,ij->i", q, q))[:, None]
            dd = 1.0 - ip / np.maximum(qn * np.sqrt(xn)[None, :], 1e-300)
        if live is not None:
            dd[:, ~live] = np.inf
        part = np.argpartition(dd, min(k, dd.shape[1] - 1), axis=1)[:, :k]
        pd = np.take_along_axis(dd, part, 1)
        o = np.lexsort((part, pd), axis=1)
        ids_out[s:s + chunk] = np.take_along_axis(part, o, 1)
        d_out[s:s + chunk] = np.take_along_axis(pd, o, 1)
    return ids_out, d_out
