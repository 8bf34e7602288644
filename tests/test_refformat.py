"""Reference file formats (.spf bincode Meta, .postings blocks, .vectors rows).

The reference commits no saved index, so the layout is pinned by (1) a Meta
byte string assembled by hand from bincode 1.x's legacy encoding rules
(little-endian fixed-width integers, u64 lengths, u8 Option tags, u32 enum
variant indices) and the field order of src/spfresh.rs:25-39,62-72, which the
decoder must read and the encoder must reproduce byte for byte; (2) block
offsets from src/spfresh.rs:109-111,258-270 checked on the written file;
(3) write -> read round trips that give back the same arrays and the same
search results.  Parity against a file the reference itself wrote is
unpinned (no such file exists here).
"""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

import oracle
from pydiskann_amd import build_index, refformat
from pydiskann_amd.refformat import FREE, Meta, SPFreshConfig, decode_meta, encode_meta


def hand_meta_bytes() -> bytes:
    b = b""
    b += (3).to_bytes(8, "little")                      # dim: usize
    b += (128).to_bytes(8, "little")                    # cfg.max_posting_size
    b += (16).to_bytes(8, "little")                     # cfg.min_posting_size
    b += bytes.fromhex("0000803f")                      # cfg.probe_ratio = 1.0f32
    b += (7).to_bytes(8, "little")                      # cfg.rerank_size
    b += (16).to_bytes(8, "little")                     # cfg.reassign_neighbors
    b += (64).to_bytes(8, "little")                     # cfg.centroid_beam
    b += b"\x01"                                        # quantizer: Some
    b += (2).to_bytes(4, "little")                      # variant Int8
    b += (3).to_bytes(8, "little")                      # Int8Quantizer.dim
    b += bytes.fromhex("0000003f")                      # scale = 0.5f32
    b += bytes.fromhex("000080bf")                      # offset = -1.0f32
    b += (2).to_bytes(8, "little")                      # block_cid: len 2
    b += (0).to_bytes(8, "little") + b"\xff" * 8        # [0, FREE]
    b += (2).to_bytes(8, "little")                      # block_centroid: len 2
    b += (3).to_bytes(8, "little") + bytes.fromhex("0000803f" "00000040" "00004040")  # [1,2,3]
    b += (0).to_bytes(8, "little")                      # [] (free block)
    b += (5).to_bytes(8, "little")                      # next_id
    b += (1).to_bytes(8, "little") + (4).to_bytes(8, "little")   # deleted = [4]
    b += (4).to_bytes(8, "little")                      # live
    return b


def test_meta_decodes_hand_built_bytes_and_reencodes_them():
    raw = hand_meta_bytes()
    m = decode_meta(raw)
    assert m.dim == 3
    assert (m.cfg.max_posting_size, m.cfg.min_posting_size, m.cfg.rerank_size) == (128, 16, 7)
    assert m.cfg.probe_ratio == 1.0 and m.cfg.reassign_neighbors == 16 and m.cfg.centroid_beam == 64
    assert m.quantizer == ("int8", {"dim": 3, "scale": 0.5, "offset": -1.0})
    assert list(m.block_cid) == [0, FREE]
    assert [list(c) for c in m.block_centroid] == [[1.0, 2.0, 3.0], []]
    assert (m.next_id, list(m.deleted), m.live) == (5, [4], 4)
    assert encode_meta(m) == raw


def test_meta_variants_and_errors():
    for q in (None, ("f16", {"dim": 4}),
              ("rabitq", {"dim": 4, "padded": 64, "centroid": np.arange(4, dtype=np.float32),
                          "signs": np.linspace(-1, 1, 192).astype(np.float32)})):
        m = Meta(4, SPFreshConfig(), q, [0], [np.ones(4, np.float32)], 1, [], 1)
        back = decode_meta(encode_meta(m))
        assert back.dim == 4 and back.cfg.probe_ratio == math.inf
        if q and q[0] == "rabitq":
            np.testing.assert_array_equal(back.quantizer[1]["signs"], q[1]["signs"])
            assert back.quantizer[1]["padded"] == 64
        else:
            assert back.quantizer == q
    raw = hand_meta_bytes()
    with pytest.raises(ValueError):
        decode_meta(raw + b"\x00")                         # trailing bytes
    pq = bytearray(raw)
    pq[8 + 44 + 1:8 + 44 + 5] = (0).to_bytes(4, "little")  # variant PQ (after dim, cfg, tag)
    with pytest.raises(NotImplementedError):
        decode_meta(bytes(pq))
    with pytest.raises(struct.error):
        decode_meta(raw[:-3])                              # truncated


def clustered(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.normal(0, 2, size=(12, d))
    return (c[rng.integers(0, 12, n)] + rng.normal(0, 1, size=(n, d))).astype(np.float32)


@pytest.mark.parametrize("quant,metric", [(None, "l2"), ("f16", "cosine"), ("int8", "l2"),
                                          ("rabitq", "l2"), ("rabitq", "dot")])
def test_write_read_round_trip(tmp_path, quant, metric):
    x = clustered(1500, 24, 1)
    ix = build_index(x, metric, quant, n_postings=20, rerank_size=16 if quant else 0, seed=0)
    # non-contiguous block indices (free blocks between) and tombstones
    ix.block_ids = (ix.block_ids * 3 + 1).astype(np.uint32)
    ix.set_deleted(np.arange(0, 1500, 11, dtype=np.uint64))
    path = str(tmp_path / "idx")
    refformat.write_index(ix, path)
    back, meta = refformat.read_index(path, metric)
    assert len(meta.block_cid) == int(ix.block_ids.max()) + 1
    assert sum(int(c) != FREE for c in meta.block_cid) == ix.n_post
    for f in ("block_ids", "centroids", "post_off", "ids", "codes", "raw", "deleted"):
        np.testing.assert_array_equal(getattr(back, f), getattr(ix, f), err_msg=f)
    assert (back.quant, back.metric, back.rerank_size) == (ix.quant, ix.metric, ix.rerank_size)
    assert (back.int8_scale, back.int8_offset, back.rq_padded) == \
        (ix.int8_scale, ix.int8_offset, ix.rq_padded)
    if ix.rq_signs is not None:
        np.testing.assert_array_equal(back.rq_signs, ix.rq_signs)
    q = clustered(20, 24, 2)
    a = oracle.Oracle(ix).search_batch(q, 10, 6)
    b = oracle.Oracle(back).search_batch(q, 10, 6)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)


def test_block_layout_on_disk(tmp_path):
    x = clustered(300, 8, 3)
    ix = build_index(x, "l2", "int8", n_postings=5, seed=0)
    path = str(tmp_path / "idx")
    m = refformat.write_index(ix, path)
    cb = ix.code_bytes
    blk = 4 + 2 * m.cfg.max_posting_size * (8 + cb)
    data = open(path + ".postings", "rb").read()
    assert len(data) == blk * len(m.block_cid)
    i = 2
    b = int(ix.block_ids[i])
    s, e = int(ix.post_off[i]), int(ix.post_off[i + 1])
    assert int.from_bytes(data[b * blk:b * blk + 4], "little") == e - s
    first = data[b * blk + 4:b * blk + 4 + 8 + cb]
    assert int.from_bytes(first[:8], "little") == int(ix.ids[s])
    assert first[8:] == ix.codes[s].tobytes()
    vec = np.fromfile(path + ".vectors", dtype="<f4").reshape(-1, 8)
    np.testing.assert_array_equal(vec, ix.raw)


def test_reader_errors(tmp_path):
    x = clustered(200, 8, 4)
    ix = build_index(x, "l2", None, n_postings=4, seed=0)
    path = str(tmp_path / "idx")
    refformat.write_index(ix, path)
    with pytest.raises(ValueError):
        refformat.read_index(path, "hamming")
    open(path + ".vectors", "wb").write(b"\x00" * 16)
    with pytest.raises(ValueError):
        refformat.read_index(path)
