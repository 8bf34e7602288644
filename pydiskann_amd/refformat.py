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
The `.centroids` graph file is not read: the device probe is exact.

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
import struct
from dataclasses import dataclass, field

import numpy as np

from .index import IndexArrays, METRICS, QUANTS, code_bytes_for

FREE = (1 << 64) - 1
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
    for i in range(ix.n_post):
        b = int(ix.block_ids[i])
        block_cid[b] = i
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
