"""GPU: incremental sync of the device mirror (spf_index_update).

The mutations the reference's control plane makes -- insert into the nearest
posting (src/spfresh.rs:416-476), split of a full posting into itself plus a
new block (:568-611), freeing a small posting and reassigning its entries
(gc, :613-643), deletes (:867-874) and purging deleted entries (rewrite,
:304-313) -- are applied to a host copy of the index; the mirror is then
updated with the changed blocks only.  Bar: after each round the updated
mirror returns the oracle's results on the mutated index bit for bit (exact
precision), and exactly what a mirror created from the mutated state returns
(exact and mixed precision, filtered and unfiltered).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from pydiskann_amd import DeviceIndex, build_index, codecs, PREC_EXACT, PREC_MIXED
from pydiskann_amd.index import IndexArrays, changed_blocks

pytestmark = pytest.mark.gpu


def clustered(n, d, n_centers, seed, spread=0.12):
    centers = np.random.default_rng(99).uniform(-1, 1, size=(n_centers, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    return (centers[np.arange(n) % n_centers] + rng.normal(0, spread, size=(n, d))).astype(np.float32)


def encode(ix, rows, cent):
    if ix.quant == 0:
        return np.ascontiguousarray(rows, dtype=np.float32).view(np.uint8).reshape(rows.shape[0], -1)
    if ix.quant == 1:
        return codecs.f16_encode(rows)
    if ix.quant == 2:
        return codecs.int8_encode(rows, ix.int8_scale, ix.int8_offset)
    return codecs.rabitq_encode(rows, cent, ix.rq_padded, ix.rq_signs)


def to_blocks(ix):
    out = {}
    for i, b in enumerate(ix.block_ids.tolist()):
        s, e = int(ix.post_off[i]), int(ix.post_off[i + 1])
        out[b] = [ix.centroids[i].copy(), ix.ids[s:e].copy(), ix.codes[s:e].copy()]
    return out


def from_blocks(ix, blocks, raw, deleted, labels):
    keys = sorted(blocks)
    ids = [blocks[b][1] for b in keys]
    codes = [blocks[b][2] for b in keys]
    lens = [len(x) for x in ids]
    post_off = np.zeros(len(keys) + 1, dtype=np.uint64)
    post_off[1:] = np.cumsum(lens)
    out = IndexArrays(dim=ix.dim, metric=ix.metric, quant=ix.quant,
                      block_ids=np.asarray(keys, dtype=np.uint32),
                      centroids=np.stack([blocks[b][0] for b in keys]),
                      post_off=post_off, ids=np.concatenate(ids).astype(np.uint64),
                      codes=np.concatenate(codes), raw=raw, deleted=deleted,
                      int8_scale=ix.int8_scale, int8_offset=ix.int8_offset,
                      rq_padded=ix.rq_padded, rq_signs=ix.rq_signs, probe_ratio=ix.probe_ratio,
                      rerank_size=ix.rerank_size)
    out.labels = labels
    return out


def mutate(ix, rng, n_new=90):
    """One round of control-plane mutations on a copy of `ix` (inserts near a
    few existing points, so that only some postings change)."""
    blocks = to_blocks(ix)
    raw = ix.raw
    cent = lambda b: blocks[b][0]

    def nearest(v, skip=()):
        ks = [b for b in blocks if b not in skip]
        c = np.stack([cent(b) for b in ks])
        return ks[int(((c - v) ** 2).sum(1).argmin())]

    def push(b, id_, v):
        blocks[b][1] = np.append(blocks[b][1], np.uint64(id_))
        blocks[b][2] = np.concatenate([blocks[b][2], encode(ix, v[None], cent(b)[None])])

    # insert: new ids at the end of the raw store, each into its nearest posting
    seeds = ix.raw[rng.choice(ix.n_raw, size=3, replace=False)]
    new = (np.repeat(seeds, n_new // 3, 0) +
           rng.normal(0, 0.05, size=(n_new // 3 * 3, ix.dim))).astype(np.float32)
    base = raw.shape[0]
    raw = np.concatenate([raw, new])
    for t, v in enumerate(new):
        push(nearest(v), base + t, v)
    # split the largest posting: half stays (new centroid), half to a new block
    big = max(blocks, key=lambda b: len(blocks[b][1]))
    ids = blocks[big][1]
    rows = raw[ids.astype(np.int64)]
    proj = (rows - rows.mean(0)) @ rng.normal(size=ix.dim).astype(np.float32)
    a, bmask = proj <= np.median(proj), proj > np.median(proj)
    nb = max(blocks) + 1
    for blk, m in ((big, a), (nb, bmask)):
        c = rows[m].mean(0).astype(np.float32)
        blocks[blk] = [c, ids[m], encode(ix, rows[m], np.repeat(c[None], int(m.sum()), 0))]
    # free the smallest posting, each entry to its nearest other posting
    small = min(blocks, key=lambda b: len(blocks[b][1]))
    sids = blocks.pop(small)[1]
    for id_ in sids.tolist():
        push(nearest(raw[id_]), id_, raw[id_])
    # deletes, then a purge of the deleted entries of a few postings
    deleted = np.zeros((raw.shape[0] + 7) // 8, dtype=np.uint8)
    deleted[: ix.deleted.shape[0]] = ix.deleted
    dead = rng.choice(raw.shape[0], size=300, replace=False)
    for i in dead.tolist():
        deleted[i >> 3] |= np.uint8(1 << (i & 7))
    isdead = np.unpackbits(deleted, bitorder="little")[: raw.shape[0]].astype(bool)
    for blk in rng.choice(sorted(blocks), size=5, replace=False).tolist():
        keep = ~isdead[blocks[blk][1].astype(np.int64)]
        blocks[blk][1], blocks[blk][2] = blocks[blk][1][keep], blocks[blk][2][keep]
    labels = None
    if ix.labels is not None:
        labels = np.concatenate([ix.labels, (np.arange(base, raw.shape[0]) % 13).astype(np.uint8)])
    return from_blocks(ix, blocks, raw, deleted, labels)


def same(a, b, msg):
    np.testing.assert_array_equal(a[2], b[2], err_msg=msg)
    np.testing.assert_array_equal(a[0], b[0], err_msg=msg)
    np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32), err_msg=msg)


@pytest.mark.parametrize("quant,metric", [(None, "l2"), ("f16", "cosine"), ("int8", "dot"),
                                          ("rabitq", "l2"), ("rabitq", "cosine"), (None, "l2_root")])
def test_update_equals_oracle_and_fresh_mirror(quant, metric):
    rng = np.random.default_rng(7)
    x = clustered(6000, 32, 25, seed=1)
    q = clustered(48, 32, 25, seed=2)
    ix = build_index(x, metric, quant, n_postings=40, rerank_size=64 if quant else 0, seed=0)
    ix.labels = (np.arange(ix.n_raw) % 13).astype(np.uint8)
    dev = DeviceIndex(ix)
    for rnd in range(2):
        ix = mutate(ix, rng)
        sent = dev.update(ix)
        print(f"[case] update quant={quant} metric={metric} round={rnd}: {sent} blocks of "
              f"{ix.n_post}, {ix.n_entries} entries", flush=True)
        assert 0 < sent < ix.n_post
        inf = dev.info()
        assert (inf.n_postings, inf.n_entries, inf.n_raw) == (ix.n_post, ix.n_entries, ix.n_raw)
        got = dev.search(q, 10, 8)
        want = oracle.Oracle(ix).search_batch(q, 10, 8, threads=8)
        same(got, want, f"oracle, round {rnd}")
        fresh = DeviceIndex(ix)
        same(got, fresh.search(q, 10, 8), f"fresh exact, round {rnd}")
        masks = np.full(q.shape[0], np.uint64(0b1010101), dtype=np.uint64)
        same(dev.search(q, 10, 8, masks=masks), fresh.search(q, 10, 8, masks=masks),
             f"fresh filtered, round {rnd}")
        same(dev.search(q, 10, 8, precision=PREC_MIXED, rerank_size=100),
             fresh.search(q, 10, 8, precision=PREC_MIXED, rerank_size=100),
             f"fresh mixed, round {rnd}")
        fresh.close()


def test_update_without_changes_and_errors():
    x = clustered(2000, 16, 10, seed=3)
    ix = build_index(x, "l2", "f16", n_postings=16, seed=0)
    dev = DeviceIndex(ix)
    q = x[:20]
    before = dev.search(q, 5, 4)
    assert dev.update(ix) == 0   # nothing changed
    same(dev.search(q, 5, 4), before, "no-op update")
    # an entry whose id has no raw row on the device is refused
    bad = to_blocks(ix)
    b0 = sorted(bad)[0]
    bad[b0][1] = np.append(bad[b0][1], np.uint64(ix.n_raw + 5))
    bad[b0][2] = np.concatenate([bad[b0][2], bad[b0][2][:1]])
    with pytest.raises(ValueError):
        dev.update(from_blocks(ix, bad, ix.raw, ix.deleted, None))
    same(dev.search(q, 5, 4), before, "after a refused update")


@pytest.mark.parametrize("quant,metric", [(None, "l2"), ("f16", "cosine"), ("int8", "l2"),
                                          ("rabitq", "l2"), ("rabitq", "dot")])
def test_open_reference_files_searches_like_the_oracle(tmp_path, quant, metric):
    """An index in the reference's on-disk layout ({path}.spf / .postings /
    .vectors, SPFresh::save / open, src/spfresh.rs:819-859): SPFresh.open
    reads it into a device mirror whose search equals the oracle on the same
    files bit for bit; then a mutated state is written, and the mirror synced
    from the re-read files equals the oracle on them too."""
    from pydiskann_amd import SPFresh
    from pydiskann_amd import refformat
    x = clustered(4000, 24, 12, seed=5)
    q = clustered(32, 24, 12, seed=6)
    ix = build_index(x, metric, quant, n_postings=24, rerank_size=16 if quant else 0, seed=0)
    ix.block_ids = (ix.block_ids * 3 + 1).astype(np.uint32)   # free blocks between
    ix.set_deleted(np.arange(0, 4000, 11, dtype=np.uint64))
    path = str(tmp_path / "idx")
    refformat.write_index(ix, path)
    s = SPFresh.open(path, metric)
    print(f"[case] open quant={quant} metric={metric}", flush=True)
    got = s._dev.search(q, 10, 6)
    same(got, oracle.Oracle(s.arrays).search_batch(q, 10, 6), "opened files vs oracle")
    for i, row in enumerate(s.search_batch(q, k=10, n_probe=6)):   # the reference's surface
        assert row == got[0][i, :int(got[2][i])].tolist()
    # a mutated state on disk, synced into the open mirror
    mutated = mutate(s.arrays, np.random.default_rng(3), n_new=30)
    mutated.labels = None
    refformat.write_index(mutated, path)
    back, _ = refformat.read_index(path, metric)
    assert s.sync(back) > 0
    same(s._dev.search(q, 10, 6), oracle.Oracle(back).search_batch(q, 10, 6),
         "synced files vs oracle")
