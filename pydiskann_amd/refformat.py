"""Reader and writer of the reference's SPFresh files, so a device mirror can be
built straight from an index the reference saved.

Files (src/spfresh.rs:785-860):
  {path}.spf       bincode 1.x of `Meta` (src/spfresh.rs:62-72), legacy fixed-int
                   little-endian encoding: usize as u64, Vec as u64 length + items,
                   Option as a u8 tag, enum as a u32 variant index + payload,
                   struct fields in declaration order.
  {path}.postings  fixed blocks of 4 + 2 * max_posting_size * (8 + code_bytes)
                   bytes (src/spfresh.rs:109-111); block b = [len u32][len x (id u64,
                   code)] (src/spfresh.rs:258-270); the file may be longer than
                   block_cid.len() blocks (it grows by doubling).
  {path}.vectors   dense f32 rows by id (src/spfresh.rs:271-273), next_id rows used.
  {path}.centroids IncrementalDiskANN::to_bytes (src/incremental.rs:1213-1327):
                   [magic "INCR" u32][version 2 u32][has_base u8][base_len u64][DANN
                   bytes][dim u64][n_delta u64][delta vectors f32][n_graph u64]
                   [(deg u32, deg x u32 delta-local ids) per delta node][entry_point
                   i64, -1 = none][max_degree u64][n_tombstones u64][tombstone u64
                   ...][has_labels u8 ...][has_quantizer u8 ...].  DANN
                   (src/lib.rs:90-98,436-489): [magic "DANN" u32][version 1 u32]
                   [md_len u64][bincode Metadata {dim, num_vectors, max_degree,
                   medoid_id u32, vectors_offset, adjacency_offset, distance_name}];
                   vectors (f32 rows) at vectors_offset, adjacency (max_degree u32
                   per node, PAD = u32::MAX) at adjacency_offset.

Meta field order: dim, cfg {max_posting_size, min_posting_size, probe_ratio (f32),
rerank_size, reassign_neighbors, centroid_beam} (src/spfresh.rs:25-39),
quantizer: Option<QuantizerState> with variants PQ=0, F16=1, Int8=2, RaBitQ=3
(src/quantized.rs:150-155) and payloads F16Quantizer{dim} (src/sq.rs:36-38),
Int8Quantizer{dim, scale, offset} (src/sq.rs:121-125), RaBitQ{dim, padded,
centroid: Vec<f32>, signs: Vec<f32>} (src/rabitq.rs:12-18); block_cid: Vec<u64>
(u64::MAX = free block), block_centroid: Vec<Vec<f32>>, next_id, deleted:
Vec<u64>, live.
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from .index import CentroidGraph, DELTA_ID_OFFSET, IndexArrays, METRICS, QUANTS, code_bytes_for

FREE = (1 << 64) - 1
INCR_MAGIC, INCR_VERSION = 0x494E4352, 2
CORE_MAGIC, CORE_VERSION = 0x44414E4E, 1
PAD_U32 = 0xFFFFFFFF
QUANT_VARIANT = {"pq": 0, "f16": 1, "int8": 2, "rabitq": 3}


@dataclass
class SPFreshConfig:
    """src/spfresh.rs:25-51 (defaults as there)."""
    max_posting_size: int = 128
    min_posting_size: int = 16
    probe_ratio: float = math.inf
    rerank_size: int = 0
    reassign_neighbors: int = 16
    centroid_beam: int = 64


@dataclass
class Meta:
    dim: int
    cfg: SPFreshConfig
    quantizer: tuple | None          # (name, payload dict) or None
    block_cid: list = field(default_factory=list)
    block_centroid: list = field(default_factory=list)
    next_id: int = 0
    deleted: list = field(default_factory=list)
    live: int = 0


class _Reader:
    def __init__(self, buf: bytes):
        self.b, self.o = buf, 0

    def take(self, fmt: str):
        v = struct.unpack_from("<" + fmt, self.b, self.o)
        self.o += struct.calcsize("<" + fmt)
        return v[0] if len(v) == 1 else v

    def u8(self):
        return self.take("B")

    def u32(self):
        return self.take("I")

    def u64(self):
        return self.take("Q")

    def f32(self):
        return self.take("f")

    def vec_f32(self) -> np.ndarray:
        n = self.u64()
        a = np.frombuffer(self.b, dtype="<f4", count=n, offset=self.o).astype(np.float32)
        self.o += 4 * n
        return a

    def vec_u64(self) -> np.ndarray:
        n = self.u64()
        a = np.frombuffer(self.b, dtype="<u8", count=n, offset=self.o).astype(np.uint64)
        self.o += 8 * n
        return a


def decode_meta(buf: bytes) -> Meta:
    r = _Reader(buf)
    dim = r.u64()
    cfg = SPFreshConfig(r.u64(), r.u64(), r.f32(), r.u64(), r.u64(), r.u64())
    quant = None
    if r.u8():
        v = r.u32()
        if v == 1:
            quant = ("f16", {"dim": r.u64()})
        elif v == 2:
            quant = ("int8", {"dim": r.u64(), "scale": r.f32(), "offset": r.f32()})
        elif v == 3:
            quant = ("rabitq", {"dim": r.u64(), "padded": r.u64(), "centroid": r.vec_f32(),
                                "signs": r.vec_f32()})
        elif v == 0:
            raise NotImplementedError("PQ postings have no device path")
        else:
            raise ValueError(f"unknown quantizer variant {v}")
    block_cid = r.vec_u64()
    nb = r.u64()
    cents = [r.vec_f32() for _ in range(nb)]
    next_id = r.u64()
    deleted = r.vec_u64()
    live = r.u64()
    if r.o != len(buf):
        raise ValueError(f"{len(buf) - r.o} trailing bytes after Meta")
    return Meta(dim, cfg, quant, block_cid, cents, next_id, deleted, live)


def encode_meta(m: Meta) -> bytes:
    out = bytearray()
    p = lambda fmt, *v: out.extend(struct.pack("<" + fmt, *v))  # noqa: E731

    def vec_f32(a):
        a = np.ascontiguousarray(a, dtype="<f4")
        p("Q", a.size)
        out.extend(a.tobytes())

    p("Q", m.dim)
    c = m.cfg
    p("QQfQQQ", c.max_posting_size, c.min_posting_size, c.probe_ratio, c.rerank_size,
      c.reassign_neighbors, c.centroid_beam)
    if m.quantizer is None:
        p("B", 0)
    else:
        name, q = m.quantizer
        p("B", 1)
        p("I", QUANT_VARIANT[name])
        if name == "f16":
            p("Q", q["dim"])
        elif name == "int8":
            p("Qff", q["dim"], q["scale"], q["offset"])
        elif name == "rabitq":
            p("QQ", q["dim"], q["padded"])
            vec_f32(q["centroid"])
            vec_f32(q["signs"])
        else:
            raise NotImplementedError(name)
    bc = np.ascontiguousarray(m.block_cid, dtype="<u8")
    p("Q", bc.size)
    out.extend(bc.tobytes())
    p("Q", len(m.block_centroid))
    for cen in m.block_centroid:
        vec_f32(cen)
    p("Q", m.next_id)
    dl = np.ascontiguousarray(m.deleted, dtype="<u8")
    p("Q", dl.size)
    out.extend(dl.tobytes())
    p("Q", m.live)
    return bytes(out)


def _code_bytes(m: Meta) -> int:
    if m.quantizer is None:
        return 4 * m.dim
    name, q = m.quantizer
    return code_bytes_for(QUANTS[name], m.dim, q.get("padded", 0))


def block_bytes(cfg: SPFreshConfig, code_bytes: int) -> int:
    """src/spfresh.rs:109-111."""
    return 4 + 2 * cfg.max_posting_size * (8 + code_bytes)


def read_index(path: str, metric: str = "l2") -> tuple[IndexArrays, Meta]:
    """IndexArrays of the live postings (block order) of a reference index at
    `path` (files {path}.spf / .postings / .vectors) and its Meta."""
    if metric not in METRICS:
        raise ValueError(f"Unknown metric '{metric}'. Use 'l2', 'cosine', or 'dot'.")
    with open(path + ".spf", "rb") as f:
        m = decode_meta(f.read())
    d, cb = m.dim, _code_bytes(m)
    blk = block_bytes(m.cfg, cb)
    entry = 8 + cb
    post = np.memmap(path + ".postings", dtype=np.uint8, mode="r")
    n_blocks = len(m.block_cid)
    if post.size < n_blocks * blk:
        raise ValueError("postings file shorter than its block table")
    live = [b for b in range(n_blocks) if int(m.block_cid[b]) != FREE]
    lens = [int(np.frombuffer(post[b * blk:b * blk + 4], dtype="<u4")[0]) for b in live]
    if any(n > 2 * m.cfg.max_posting_size for n in lens):
        raise ValueError("posting length beyond block capacity")
    off = np.zeros(len(live) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens, dtype=np.uint64)
    n = int(off[-1])
    ids = np.empty(n, dtype=np.uint64)
    codes = np.empty((n, cb), dtype=np.uint8)
    for i, b in enumerate(live):
        if lens[i] == 0:
            continue
        rows = np.asarray(post[b * blk + 4:b * blk + 4 + lens[i] * entry]).reshape(lens[i], entry)
        ids[off[i]:off[i + 1]] = rows[:, :8].copy().view("<u8").ravel()
        codes[off[i]:off[i + 1]] = rows[:, 8:]
    vec = np.memmap(path + ".vectors", dtype="<f4", mode="r")
    if vec.size < m.next_id * d:
        raise ValueError("vectors file shorter than next_id rows")
    raw = np.array(vec[:m.next_id * d], dtype=np.float32).reshape(m.next_id, d)
    cents = np.stack([np.asarray(m.block_centroid[b], dtype=np.float32) for b in live]) \
        if live else np.zeros((0, d), np.float32)
    quant = QUANTS[m.quantizer[0]] if m.quantizer else 0
    qp = m.quantizer[1] if m.quantizer else {}
    ix = IndexArrays(dim=d, metric=METRICS[metric], quant=quant,
                     block_ids=np.asarray(live, dtype=np.uint32), centroids=cents, post_off=off,
                     ids=ids, codes=codes, raw=raw,
                     int8_scale=float(qp.get("scale", 1.0)), int8_offset=float(qp.get("offset", 0.0)),
                     rq_padded=int(qp.get("padded", 0)),
                     rq_signs=np.asarray(qp["signs"], np.float32).reshape(3, -1) if "signs" in qp else None,
                     probe_ratio=float(m.cfg.probe_ratio), rerank_size=int(m.cfg.rerank_size))
    ix.set_deleted(np.asarray(m.deleted, dtype=np.uint64))
    if os.path.exists(path + ".centroids"):
        with open(path + ".centroids", "rb") as f:
            gf = decode_centroids(f.read())
        if gf.dim != d:
            raise ValueError("centroid graph dim differs from the index dim")
        ix.graph = graph_from_file(gf, [int(m.block_cid[b]) for b in live], int(m.cfg.centroid_beam))
    return ix, m


def write_index(ix: IndexArrays, path: str, cfg: SPFreshConfig | None = None,
                rabitq_centroid: np.ndarray | None = None) -> Meta:
    """Write IndexArrays in the reference's layout ({path}.spf / .postings /
    .vectors).  Posting i goes to block ix.block_ids[i]; other blocks below the
    largest id are free.  max_posting_size grows to hold the longest posting."""
    cfg = SPFreshConfig(**vars(cfg)) if cfg else SPFreshConfig(
        probe_ratio=ix.probe_ratio, rerank_size=ix.rerank_size)
    lens = ix.posting_lengths()
    if lens.size:
        cfg.max_posting_size = max(cfg.max_posting_size, int((lens.max() + 1) // 2), 2)
    cb = ix.code_bytes
    blk = block_bytes(cfg, cb)
    n_blocks = int(ix.block_ids.max()) + 1 if ix.n_post else 0
    block_cid = np.full(n_blocks, FREE, dtype=np.uint64)
    cents = [np.zeros(0, np.float32)] * n_blocks
    buf = np.zeros(max(n_blocks, 1) * blk, dtype=np.uint8)
    post_cid = list(range(ix.n_post))
    g = getattr(ix, "graph", None)
    if g is not None:   # cids are the routing graph's global node ids
        for u in range(g.n_nodes):
            if g.node_post[u] >= 0:
                post_cid[int(g.node_post[u])] = g.cid(u)
        if len(set(int(x) for x in g.node_post if x >= 0)) != ix.n_post:
            raise ValueError("every posting needs exactly one routing-graph node")
        cfg.centroid_beam = g.centroid_beam
        with open(path + ".centroids", "wb") as f:
            f.write(encode_centroids(file_from_graph(g, ix.dim)))
    for i in range(ix.n_post):
        b = int(ix.block_ids[i])
        block_cid[b] = post_cid[i]
        cents[b] = ix.centroids[i]
        s, e = int(ix.post_off[i]), int(ix.post_off[i + 1])
        base = b * blk
        buf[base:base + 4] = np.frombuffer(struct.pack("<I", e - s), dtype=np.uint8)
        rows = np.zeros((e - s, 8 + cb), dtype=np.uint8)
        rows[:, :8] = ix.ids[s:e].astype("<u8").view(np.uint8).reshape(-1, 8)
        rows[:, 8:] = ix.codes[s:e]
        buf[base + 4:base + 4 + rows.size] = rows.ravel()
    buf.tofile(path + ".postings")
    np.ascontiguousarray(ix.raw, dtype="<f4").tofile(path + ".vectors")
    quant = None
    if ix.quant == QUANTS["f16"]:
        quant = ("f16", {"dim": ix.dim})
    elif ix.quant == QUANTS["int8"]:
        quant = ("int8", {"dim": ix.dim, "scale": ix.int8_scale, "offset": ix.int8_offset})
    elif ix.quant == QUANTS["rabitq"]:
        quant = ("rabitq", {"dim": ix.dim, "padded": ix.rq_padded,
                            "centroid": rabitq_centroid if rabitq_centroid is not None
                            else np.zeros(ix.dim, np.float32),
                            "signs": ix.rq_signs.ravel()})
    deleted = [i for i in range(ix.n_raw) if ix.is_deleted(i)]
    m = Meta(dim=ix.dim, cfg=cfg, quantizer=quant, block_cid=list(block_cid), block_centroid=cents,
             next_id=ix.n_raw, deleted=deleted, live=ix.n_raw - len(deleted))
    with open(path + ".spf", "wb") as f:
        f.write(encode_meta(m))
    return m


# ---------------------------------------------------------------- centroid graph
@dataclass
class GraphFile:
    """Decoded {path}.centroids: base (DANN) part, delta layer, tombstones."""
    dim: int
    base_vectors: np.ndarray          # f32 [base_n, dim]
    base_adj: np.ndarray              # u32 [base_n, max_degree], PAD = u32::MAX
    medoid: int
    distance_name: str
    delta_vectors: np.ndarray         # f32 [n_delta, dim]
    delta_graph: list                 # per delta node: u32 array of delta-local ids
    entry_point: int | None
    max_degree: int
    tombstones: list                  # global u64 ids
    has_base: bool = True


def _decode_dann(buf: bytes):
    r = _Reader(buf)
    off = 0
    if len(buf) >= 8 and r.u32() == CORE_MAGIC:
        if r.u32() != CORE_VERSION:
            raise ValueError("unsupported DANN version")
        off = 8
    else:
        r.o = 0
    md_len = r.u64()
    md_end = off + 8 + md_len
    dim, n, max_deg = r.u64(), r.u64(), r.u64()
    medoid = r.u32()
    v_off, a_off = r.u64(), r.u64()
    name_len = r.u64()
    name = buf[r.o:r.o + name_len].decode()
    r.o += name_len
    if r.o != md_end:
        raise ValueError("DANN metadata length mismatch")
    if v_off + n * dim * 4 > len(buf) or a_off + n * max_deg * 4 > len(buf):
        raise ValueError("DANN buffer shorter than its vectors / adjacency")
    vec = np.frombuffer(buf, dtype="<f4", count=n * dim, offset=v_off).astype(np.float32)
    adj = np.frombuffer(buf, dtype="<u4", count=n * max_deg, offset=a_off).astype(np.uint32)
    return dim, vec.reshape(n, dim), adj.reshape(n, max_deg), medoid, name


def _encode_dann(vectors: np.ndarray, adj: np.ndarray, medoid: int, name: str) -> bytes:
    """DANN bytes laid out as DiskANN::build writes them (vectors at 1 MiB)."""
    n, dim = vectors.shape
    max_deg = adj.shape[1] if adj.ndim == 2 else 0
    v_off = 1 << 20
    a_off = v_off + n * dim * 4
    nb = name.encode()
    md = struct.pack("<QQQIQQQ", dim, n, max_deg, medoid, v_off, a_off, len(nb)) + nb
    head = struct.pack("<IIQ", CORE_MAGIC, CORE_VERSION, len(md)) + md
    out = bytearray(a_off + n * max_deg * 4)
    out[:len(head)] = head
    out[v_off:a_off] = np.ascontiguousarray(vectors, dtype="<f4").tobytes()
    out[a_off:] = np.ascontiguousarray(adj, dtype="<u4").tobytes()
    return bytes(out)


def decode_centroids(buf: bytes) -> GraphFile:
    r = _Reader(buf)
    if r.u32() != INCR_MAGIC:
        raise NotImplementedError("pre-INCR (legacy) centroid graph files are not read")
    if r.u32() != INCR_VERSION:
        raise ValueError("unsupported INCR version")
    has_base = r.u8() == 1
    base = None
    if has_base:
        blen = r.u64()
        base = _decode_dann(buf[r.o:r.o + blen])
        r.o += blen
    dim = r.u64()
    nd = r.u64()
    dvec = np.frombuffer(buf, dtype="<f4", count=nd * dim, offset=r.o).astype(np.float32)
    r.o += nd * dim * 4
    ng = r.u64()
    dgraph = []
    for _ in range(ng):
        deg = r.u32()
        dgraph.append(np.frombuffer(buf, dtype="<u4", count=deg, offset=r.o).astype(np.uint32))
        r.o += 4 * deg
    ep = r.take("q")
    max_deg = r.u64()
    nt = r.u64()
    tomb = [int(x) for x in np.frombuffer(buf, dtype="<u8", count=nt, offset=r.o)]
    r.o += 8 * nt
    if r.u8():
        raise NotImplementedError("labelled centroid graphs are not read")
    if r.u8():
        raise NotImplementedError("quantized centroid graphs are not read")
    if r.o != len(buf):
        raise ValueError(f"{len(buf) - r.o} trailing bytes after the centroid graph")
    if base is not None:
        bdim, bvec, badj, medoid, name = base
        if bdim != dim:
            raise ValueError("DANN dim differs from the graph dim")
    else:
        bvec, badj, medoid, name = np.zeros((0, dim), np.float32), np.zeros((0, 0), np.uint32), 0, ""
    return GraphFile(dim, bvec, badj, medoid, name, dvec.reshape(nd, dim), dgraph,
                     None if ep < 0 else int(ep), max_deg, tomb, has_base)


def encode_centroids(gf: GraphFile) -> bytes:
    out = bytearray(struct.pack("<II", INCR_MAGIC, INCR_VERSION))
    if gf.has_base:
        b = _encode_dann(gf.base_vectors, gf.base_adj, gf.medoid, gf.distance_name)
        out += struct.pack("<BQ", 1, len(b)) + b
    else:
        out += b"\x00"
    out += struct.pack("<QQ", gf.dim, gf.delta_vectors.shape[0])
    out += np.ascontiguousarray(gf.delta_vectors, dtype="<f4").tobytes()
    out += struct.pack("<Q", len(gf.delta_graph))
    for nb in gf.delta_graph:
        out += struct.pack("<I", len(nb)) + np.asarray(nb, dtype="<u4").tobytes()
    out += struct.pack("<qQQ", -1 if gf.entry_point is None else gf.entry_point, gf.max_degree,
                       len(gf.tombstones))
    out += np.asarray(gf.tombstones, dtype="<u8").tobytes()
    out += b"\x00\x00"   # no labels, no quantizer
    return bytes(out)


def graph_from_file(gf: GraphFile, posting_cids, centroid_beam: int = 64) -> CentroidGraph:
    """Unified routing graph (UnifiedView, src/incremental.rs:357-487) with each
    node's posting: posting_cids[i] = cid (block_cid) of posting i."""
    bn = gf.base_vectors.shape[0] if gf.has_base else 0
    nd = gf.delta_vectors.shape[0]
    n = bn + nd
    vec = np.concatenate([gf.base_vectors, gf.delta_vectors]).astype(np.float32) \
        if n else np.zeros((0, gf.dim), np.float32)
    lists = []
    for u in range(bn):
        row = gf.base_adj[u]
        lists.append(row[row != PAD_U32])
    for j in range(nd):
        lists.append(gf.delta_graph[j] + np.uint32(bn) if j < len(gf.delta_graph)
                     else np.zeros(0, np.uint32))
    adj_off = np.zeros(n + 1, dtype=np.uint32)
    adj_off[1:] = np.cumsum([len(x) for x in lists], dtype=np.uint64).astype(np.uint32)
    adj = np.concatenate(lists).astype(np.uint32) if lists else np.zeros(0, np.uint32)
    if adj.size and int(adj.max()) >= n:
        raise ValueError("graph neighbor beyond the nodes")
    tomb = set(int(t) for t in gf.tombstones)
    cid = lambda u: u if u < bn else DELTA_ID_OFFSET + u - bn
    live = np.array([cid(u) not in tomb for u in range(n)], dtype=np.uint8)
    node_of = {}
    for u in range(n):
        node_of[cid(u)] = u
    node_post = np.full(n, -1, dtype=np.int64)
    for i, c in enumerate(posting_cids):
        u = node_of.get(int(c))
        if u is not None:
            node_post[u] = i
    seeds = []
    if gf.has_base:
        seeds.append(int(gf.medoid))
    if gf.entry_point is not None:
        seeds.append(bn + int(gf.entry_point))
    return CentroidGraph(vectors=vec, adj_off=adj_off, adj=adj, live=live, seeds=seeds,
                         node_post=node_post, n_tombstones=len(tomb), centroid_beam=centroid_beam,
                         base_n=bn)


def file_from_graph(g: CentroidGraph, dim: int,
                    distance_name: str = "diskann_rs::DistL2") -> GraphFile:
    """GraphFile of a unified routing graph (inverse of graph_from_file)."""
    bn, n = g.base_n, g.n_nodes
    lists = [g.adj[g.adj_off[u]:g.adj_off[u + 1]] for u in range(n)]
    maxd = max([len(x) for x in lists[:bn]] + [0])
    badj = np.full((bn, maxd), PAD_U32, dtype=np.uint32)
    for u in range(bn):
        if len(lists[u]) and int(lists[u].max()) >= bn:
            raise ValueError("a base node links to a delta node")
        badj[u, :len(lists[u])] = lists[u]
    dgraph = []
    for u in range(bn, n):
        if len(lists[u]) and int(lists[u].min()) < bn:
            raise ValueError("a delta node links to a base node")
        dgraph.append((lists[u] - np.uint32(bn)).astype(np.uint32))
    seeds = list(g.seeds)
    medoid = seeds.pop(0) if bn else 0
    ep = seeds[0] - bn if seeds else None
    dead = [g.cid(u) for u in range(n) if not g.live[u]]
    extra = g.n_tombstones - len(dead)
    dead += [DELTA_ID_OFFSET + (n - bn) + i for i in range(max(extra, 0))]
    return GraphFile(dim, g.vectors[:bn], badj, medoid, distance_name, g.vectors[bn:], dgraph, ep,
                     maxd, dead, bn > 0)
