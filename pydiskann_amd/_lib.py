"""ctypes binding of libspf_hip.so (the C ABI declared in include/spf.h).

The product search path: there is no fallback.  If the library is missing the
import of a search entry point raises, naming the build step.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspf_hip.so")

SPF_OK, SPF_EINVAL, SPF_EDIM, SPF_EUNSUPPORTED, SPF_EDEVICE, SPF_ENOMEM = 0, -1, -2, -3, -4, -5
SPF_EAGAIN = -6
PREC_EXACT, PREC_MIXED = 0, 1

# every entry point of include/spf.h
EXPORTS = ("spf_index_create", "spf_index_destroy", "spf_index_get_info", "spf_index_update",
           "spf_delete",
           "spf_search_batch", "spf_search_batch_device", "spf_search_batch_filtered",
           "spf_search_batch_filtered_device", "spf_search_batch_device_noalloc",
           "spf_nearest_postings", "spf_server_create",
           "spf_server_submit", "spf_server_wait", "spf_server_get_stats", "spf_server_destroy",
           "spf_autotune",
           "spf_last_profile", "spf_selftest", "spf_last_error", "spf_version")
FLAG_PROFILE = 1
FLAG_FULL_RERANK = 2
FLAG_SYNC_PLAN = 4
FLAG_NO_GRAPH = 8
FLAG_RQ_POPCOUNT = 16
FLAG_ADMIT = 32
STAGES = ("prep", "probe", "probe_select", "plan", "scan", "select", "rerank", "finalize")


class GraphDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("n_seeds", C.c_uint32), ("seeds", C.c_uint32 * 2),
        ("centroid_beam", C.c_uint32), ("reserved", C.c_uint32), ("n_tombstones", C.c_uint64),
        ("vectors", C.c_void_p), ("adj_off", C.c_void_p), ("adj", C.c_void_p),
        ("live", C.c_void_p), ("node_post", C.c_void_p),
    ]


class IndexDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_uint32), ("metric", C.c_uint32), ("quantizer", C.c_uint32),
        ("n_postings", C.c_uint32),
        ("block_ids", C.c_void_p), ("centroids", C.c_void_p), ("posting_offsets", C.c_void_p),
        ("entry_ids", C.c_void_p), ("codes", C.c_void_p), ("code_bytes", C.c_uint32),
        ("raw", C.c_void_p), ("n_raw", C.c_uint64),
        ("deleted_ids", C.c_void_p), ("n_deleted", C.c_uint64),
        ("int8_scale", C.c_float), ("int8_offset", C.c_float),
        ("rabitq_padded", C.c_uint32), ("rabitq_signs", C.c_void_p),
        ("probe_ratio", C.c_float), ("rerank_size", C.c_uint32), ("device", C.c_int32),
        ("labels", C.c_void_p), ("graph", C.POINTER(GraphDesc)),
    ]


class BlockUpdate(C.Structure):
    _fields_ = [("block", C.c_uint32), ("live", C.c_uint32), ("centroid", C.c_void_p),
                ("n_entries", C.c_uint64), ("entry_ids", C.c_void_p), ("codes", C.c_void_p)]


class UpdateDesc(C.Structure):
    _fields_ = [("blocks", C.POINTER(BlockUpdate)), ("n_blocks", C.c_uint32),
                ("reserved0", C.c_uint32), ("n_raw", C.c_uint64), ("raw_rows", C.c_void_p),
                ("labels", C.c_void_p), ("deleted_ids", C.c_void_p), ("n_deleted", C.c_uint64),
                ("graph", C.POINTER(GraphDesc)), ("reserved", C.c_uint32 * 4)]


class SearchOpts(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("rerank_size", C.c_uint32), ("flags", C.c_uint32),
                ("plan_bound_mib", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class Profile(C.Structure):
    _fields_ = [("stage_ms", C.c_float * 8), ("candidates", C.c_uint64),
                ("scan_code_bytes", C.c_uint64), ("probe_pairs", C.c_uint64),
                ("scan_items", C.c_uint32), ("rerank_rows", C.c_uint32),
                ("pool_members", C.c_uint32), ("admitted", C.c_uint32),
                ("scan_bytes", C.c_uint64), ("scan_flops", C.c_uint64),
                ("probe_flops", C.c_uint64), ("scan_hbm_bytes", C.c_uint64)]


class IndexInfo(C.Structure):
    _fields_ = [("dim", C.c_uint32), ("metric", C.c_uint32), ("quantizer", C.c_uint32),
                ("n_postings", C.c_uint32), ("n_entries", C.c_uint64), ("n_raw", C.c_uint64),
                ("n_deleted", C.c_uint64), ("device_bytes", C.c_uint64),
                ("max_posting", C.c_uint32), ("reserved", C.c_uint32),
                ("graph_replays", C.c_uint64)]


class ServerOpts(C.Structure):
    _fields_ = [("n_slots", C.c_uint32), ("max_batch", C.c_uint32),
                ("hbm_budget_bytes", C.c_uint64), ("reserved", C.c_uint32 * 4)]


class RequestInfo(C.Structure):
    _fields_ = [("submit_ns", C.c_uint64), ("start_ns", C.c_uint64), ("done_ns", C.c_uint64),
                ("batches", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("rows_done", C.c_uint64),
                ("batches", C.c_uint64), ("rejected", C.c_uint64),
                ("peak_reserved_bytes", C.c_uint64), ("budget_overruns", C.c_uint64),
                ("admission_waits", C.c_uint64), ("trims", C.c_uint64),
                ("max_concurrent", C.c_uint32), ("requeues", C.c_uint32),
                ("splits", C.c_uint32), ("reserved", C.c_uint32 * 1)]


class TuneResult(C.Structure):
    _fields_ = [("scan_g", C.c_uint32), ("scan_chunk", C.c_uint32), ("mix_g", C.c_uint32),
                ("from_cache", C.c_uint32), ("best_ms", C.c_float), ("default_ms", C.c_float),
                ("tried", C.c_uint32), ("reserved", C.c_uint32 * 1)]


class SpfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Load libspf_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (or make -C pydiskann_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        P, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
        L.spf_index_create.argtypes = [C.POINTER(IndexDesc), C.POINTER(C.c_void_p)]
        L.spf_index_create.restype = C.c_int
        L.spf_index_destroy.argtypes = [P]
        L.spf_index_destroy.restype = None
        L.spf_index_get_info.argtypes = [P, C.POINTER(IndexInfo)]
        L.spf_index_get_info.restype = C.c_int
        L.spf_index_update.argtypes = [P, C.POINTER(UpdateDesc)]
        L.spf_index_update.restype = C.c_int
        L.spf_delete.argtypes = [P, P, u64]
        L.spf_delete.restype = C.c_int
        L.spf_search_batch.argtypes = [P, P, u32, u32, u32, u32, C.POINTER(SearchOpts), P, P, P]
        L.spf_search_batch.restype = C.c_int
        L.spf_search_batch_device.argtypes = [P, P, u32, u32, u32, u32, C.POINTER(SearchOpts),
                                              P, P, P, P]
        L.spf_search_batch_device.restype = C.c_int
        L.spf_search_batch_filtered.argtypes = [P, P, u32, u32, u32, u32, C.POINTER(SearchOpts),
                                                P, P, P, P]
        L.spf_search_batch_filtered.restype = C.c_int
        L.spf_search_batch_filtered_device.argtypes = [P, P, u32, u32, u32, u32,
                                                       C.POINTER(SearchOpts), P, P, P, P, P]
        L.spf_search_batch_filtered_device.restype = C.c_int
        L.spf_search_batch_device_noalloc.argtypes = [P, P, u32, u32, u32, u32,
                                                      C.POINTER(SearchOpts), P, P, P, P, P]
        L.spf_search_batch_device_noalloc.restype = C.c_int
        L.spf_nearest_postings.argtypes = [P, P, u32, u32, P, P]
        L.spf_nearest_postings.restype = C.c_int
        L.spf_server_create.argtypes = [P, C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]
        L.spf_server_create.restype = C.c_int
        L.spf_server_submit.argtypes = [P, P, u32, u32, u32, u32, C.POINTER(SearchOpts), P, P, P,
                                        P, C.POINTER(C.c_uint64)]
        L.spf_server_submit.restype = C.c_int
        L.spf_server_wait.argtypes = [P, u64, C.c_int64, C.POINTER(RequestInfo)]
        L.spf_server_wait.restype = C.c_int
        L.spf_server_get_stats.argtypes = [P, C.POINTER(ServerStats)]
        L.spf_server_get_stats.restype = C.c_int
        L.spf_server_destroy.argtypes = [P]
        L.spf_server_destroy.restype = None
        L.spf_autotune.argtypes = [P, P, u32, u32, u32, u32, C.POINTER(SearchOpts), C.c_char_p,
                                   C.POINTER(TuneResult)]
        L.spf_autotune.restype = C.c_int
        L.spf_last_profile.argtypes = [C.POINTER(Profile)]
        L.spf_last_profile.restype = C.c_int
        L.spf_selftest.argtypes = [C.c_int]
        L.spf_selftest.restype = C.c_int
        L.spf_last_error.restype = C.c_char_p
        L.spf_version.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SPF_OK:
        msg = lib().spf_last_error().decode()
        if rc == SPF_EINVAL:
            raise ValueError(msg)
        raise SpfError(rc, msg)


def last_profile() -> dict:
    pr = Profile()
    check(lib().spf_last_profile(C.byref(pr)))
    return {"stage_ms": dict(zip(STAGES, [float(x) for x in pr.stage_ms])),
            "candidates": int(pr.candidates), "scan_code_bytes": int(pr.scan_code_bytes),
            "probe_pairs": int(pr.probe_pairs), "scan_items": int(pr.scan_items),
            "rerank_rows": int(pr.rerank_rows), "pool_members": int(pr.pool_members),
            "admitted": int(pr.admitted),
            "scan_bytes": int(pr.scan_bytes), "scan_flops": int(pr.scan_flops),
            "probe_flops": int(pr.probe_flops), "scan_hbm_bytes": int(pr.scan_hbm_bytes)}


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _graph_desc(g, node_post):
    keep = [np.ascontiguousarray(g.vectors, dtype=np.float32),
            np.ascontiguousarray(g.adj_off, dtype=np.uint32),
            np.ascontiguousarray(g.adj, dtype=np.uint32),
            np.ascontiguousarray(g.live, dtype=np.uint8),
            np.ascontiguousarray(node_post, dtype=np.int64)]
    gd = GraphDesc()
    gd.n_nodes, gd.n_seeds = keep[0].shape[0], len(g.seeds)
    for i, sd in enumerate(g.seeds):
        gd.seeds[i] = int(sd)
    gd.centroid_beam, gd.n_tombstones = g.centroid_beam, g.n_tombstones
    gd.vectors, gd.adj_off, gd.adj, gd.live, gd.node_post = (_p(a) for a in keep)
    return gd, keep


class DeviceIndex:
    """Device mirror of an ``IndexArrays`` (one GPU)."""

    def __init__(self, ix, device: int = 0):
        self._src = ix
        d = IndexDesc()
        d.dim, d.metric, d.quantizer, d.n_postings = ix.dim, ix.metric, ix.quant, ix.n_post
        d.block_ids, d.centroids, d.posting_offsets = _p(ix.block_ids), _p(ix.centroids), _p(ix.post_off)
        d.entry_ids, d.codes, d.code_bytes = _p(ix.ids), _p(ix.codes), ix.code_bytes
        d.raw, d.n_raw = _p(ix.raw), ix.n_raw
        dead = np.nonzero(np.unpackbits(ix.deleted, bitorder="little")[: ix.n_raw])[0].astype(np.uint64)
        d.deleted_ids, d.n_deleted = _p(dead), dead.shape[0]
        d.int8_scale, d.int8_offset = ix.int8_scale, ix.int8_offset
        d.rabitq_padded, d.rabitq_signs = ix.rq_padded, _p(ix.rq_signs)
        d.probe_ratio = ix.probe_ratio if ix.probe_ratio is not None else math.inf
        d.rerank_size, d.device = ix.rerank_size, device
        self._labels = None
        if getattr(ix, "labels", None) is not None:
            self._labels = np.ascontiguousarray(ix.labels, dtype=np.uint8)
            if self._labels.shape[0] != ix.n_raw:
                raise ValueError("labels must hold one u8 per id (n_raw)")
            d.labels = _p(self._labels)
        self._graph = None
        g = getattr(ix, "graph", None)
        if g is not None:
            self._graph = _graph_desc(g, np.asarray(g.node_post, dtype=np.int64))
            d.graph = C.pointer(self._graph[0])
        h = C.c_void_p()
        check(lib().spf_index_create(C.byref(d), C.byref(h)))
        self.h = h
        self.dim = ix.dim

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().spf_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> IndexInfo:
        inf = IndexInfo()
        check(lib().spf_index_get_info(self.h, C.byref(inf)))
        return inf

    def delete(self, ids) -> None:
        a = np.ascontiguousarray(np.asarray(ids, dtype=np.uint64).ravel())
        check(lib().spf_delete(self.h, _p(a), a.shape[0]))

    def update(self, new, blocks=None) -> int:
        """Sync the mirror to the mutated index `new` (spf_index_update): the
        blocks written since the mirror's state (every block whose centroid
        or entries differ when `blocks` is None; the reference tracks them
        with mark_dirty), the appended raw rows and labels, the new
        tombstones and the routing graph.  Returns the number of blocks sent."""
        from .index import changed_blocks
        old = self._src
        if new.n_raw < old.n_raw:
            raise ValueError("the raw store cannot shrink")
        if blocks is None:
            blocks = changed_blocks(old, new)
        pos = {int(b): i for i, b in enumerate(new.block_ids.tolist())}
        arr = (BlockUpdate * max(1, len(blocks)))()
        keep = []
        for j, b in enumerate(blocks):
            u = arr[j]
            u.block = int(b)
            i = pos.get(int(b))
            if i is None:
                u.live = 0
                continue
            s, e = int(new.post_off[i]), int(new.post_off[i + 1])
            cent = np.ascontiguousarray(new.centroids[i])
            ids = np.ascontiguousarray(new.ids[s:e])
            codes = np.ascontiguousarray(new.codes[s:e])
            keep += [cent, ids, codes]
            u.live, u.centroid, u.n_entries = 1, _p(cent), e - s
            u.entry_ids, u.codes = _p(ids), _p(codes)
        d = UpdateDesc()
        d.blocks, d.n_blocks, d.n_raw = arr, len(blocks), new.n_raw
        raw = np.ascontiguousarray(new.raw[old.n_raw:])
        d.raw_rows = _p(raw) if raw.shape[0] else None
        lab = None
        if getattr(new, "labels", None) is not None and new.n_raw > old.n_raw:
            lab = np.ascontiguousarray(new.labels[old.n_raw:], dtype=np.uint8)
            d.labels = _p(lab)
        was = np.unpackbits(old.deleted, bitorder="little")[: old.n_raw]
        now = np.unpackbits(new.deleted, bitorder="little")[: new.n_raw]
        dead = np.nonzero(now[: old.n_raw] > was)[0].tolist() + \
            (np.nonzero(now[old.n_raw:])[0] + old.n_raw).tolist()
        dead = np.ascontiguousarray(np.asarray(dead, dtype=np.uint64))
        d.deleted_ids, d.n_deleted = (_p(dead) if dead.size else None), dead.size
        g = getattr(new, "graph", None)
        gkeep = None
        if g is not None and g is not getattr(old, "graph", None):
            npost = np.asarray(g.node_post, dtype=np.int64)
            node_block = np.where(npost >= 0, new.block_ids[np.maximum(npost, 0)].astype(np.int64), -1)
            gkeep = _graph_desc(g, node_block)
            d.graph = C.pointer(gkeep[0])
        check(lib().spf_index_update(self.h, C.byref(d)))
        self._src = new
        self._labels = None if getattr(new, "labels", None) is None else new.labels
        self._graph = gkeep or self._graph
        return len(blocks)

    def search(self, queries, k: int, n_probe: int, precision: int = PREC_EXACT,
               rerank_size: int = 0, flags: int = 0, masks=None):
        """Batch search; ``masks`` (u64 [nq], bit l = label l allowed) makes it
        the label-filtered search."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        nq, dim = q.shape
        ids = np.empty((nq, max(k, 1)), dtype=np.uint64)
        dd = np.empty((nq, max(k, 1)), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.uint32)
        o = SearchOpts()
        o.precision, o.rerank_size, o.flags = precision, rerank_size, flags
        if masks is None:
            check(lib().spf_search_batch(self.h, _p(q), nq, dim, k, n_probe, C.byref(o),
                                         _p(ids), _p(dd), _p(cnt)))
        else:
            m = np.ascontiguousarray(np.broadcast_to(np.asarray(masks, dtype=np.uint64), (nq,)))
            check(lib().spf_search_batch_filtered(self.h, _p(q), nq, dim, k, n_probe,
                                                  C.byref(o), _p(m), _p(ids), _p(dd), _p(cnt)))
        return ids, dd, cnt

    def autotune(self, queries, k: int, n_probe: int, precision: int = PREC_EXACT,
                 rerank_size: int = 0, cache_path: str | None = None) -> dict:
        """Time the launch shapes on these sample queries and keep the fastest
        (spf_autotune); with cache_path, reuse / record the choice."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        o = SearchOpts()
        o.precision, o.rerank_size = precision, rerank_size
        r = TuneResult()
        check(lib().spf_autotune(self.h, _p(q), q.shape[0], q.shape[1], k, n_probe, C.byref(o),
                                 cache_path.encode() if cache_path else None, C.byref(r)))
        return {f: (float(getattr(r, f)) if f.endswith("_ms") else int(getattr(r, f)))
                for f, _ in TuneResult._fields_ if f != "reserved"}

    def nearest_postings(self, vectors):
        """(block index, distance) of each vector's nearest live posting:
        the routing step of Inner::insert (src/spfresh.rs:437-444)."""
        v = np.ascontiguousarray(vectors, dtype=np.float32)
        if v.ndim == 1:
            v = v[None, :]
        n, dim = v.shape
        blocks = np.empty(n, dtype=np.uint32)
        dd = np.empty(n, dtype=np.float32)
        check(lib().spf_nearest_postings(self.h, _p(v), n, dim, _p(blocks), _p(dd)))
        return blocks, dd

    def search_device(self, q_ptr: int, nq: int, dim: int, k: int, n_probe: int, ids_ptr: int,
                      dists_ptr: int, counts_ptr: int, stream: int = 0,
                      precision: int = PREC_EXACT, rerank_size: int = 0,
                      profile: bool = False, masks_ptr: int = 0, flags: int = 0) -> None:
        """Device-resident form: raw device pointers (e.g. torch .data_ptr());
        ``masks_ptr`` (u64 [nq] on the device) makes it the filtered search."""
        o = SearchOpts()
        o.precision, o.rerank_size = precision, rerank_size
        o.flags = flags | (FLAG_PROFILE if profile else 0)
        st = C.c_void_p(stream) if stream else None
        if masks_ptr:
            check(lib().spf_search_batch_filtered_device(
                self.h, C.c_void_p(q_ptr), nq, dim, k, n_probe, C.byref(o),
                C.c_void_p(masks_ptr), C.c_void_p(ids_ptr), C.c_void_p(dists_ptr),
                C.c_void_p(counts_ptr), st))
        else:
            check(lib().spf_search_batch_device(self.h, C.c_void_p(q_ptr), nq, dim, k, n_probe,
                                                C.byref(o), C.c_void_p(ids_ptr),
                                                C.c_void_p(dists_ptr), C.c_void_p(counts_ptr),
                                                st))


    def search_device_noalloc(self, q_ptr: int, nq: int, dim: int, k: int, n_probe: int,
                              ids_ptr: int, dists_ptr: int, counts_ptr: int, stream: int = 0,
                              precision: int = PREC_EXACT, rerank_size: int = 0,
                              masks_ptr: int = 0, flags: int = 0) -> None:
        """spf_search_batch_device_noalloc: the device-resident search with
        its working buffers carved from the library's fixed device scratch."""
        o = SearchOpts()
        o.precision, o.rerank_size, o.flags = precision, rerank_size, flags
        check(lib().spf_search_batch_device_noalloc(
            self.h, C.c_void_p(q_ptr), nq, dim, k, n_probe, C.byref(o),
            C.c_void_p(masks_ptr) if masks_ptr else None, C.c_void_p(ids_ptr),
            C.c_void_p(dists_ptr), C.c_void_p(counts_ptr),
            C.c_void_p(stream) if stream else None))


class Ticket:
    """A submitted request: holds its input and output arrays until waited."""

    def __init__(self, tid: int, q, masks, ids, dists, counts):
        self.id, self.q, self.masks = tid, q, masks
        self.ids, self.dists, self.counts = ids, dists, counts


class Server:
    """Continuous-batching search server over one DeviceIndex (spf_server_*):
    requests from any thread are merged into device batches that run on
    several streams at once, admitted by a device-memory budget."""

    def __init__(self, dev: "DeviceIndex", n_slots: int = 0, max_batch: int = 0,
                 hbm_budget: int = 0):
        o = ServerOpts()
        o.n_slots, o.max_batch, o.hbm_budget_bytes = n_slots, max_batch, hbm_budget
        h = C.c_void_p()
        check(lib().spf_server_create(dev.h, C.byref(o), C.byref(h)))
        self.h, self.dev, self.dim = h, dev, dev.dim
        self._live = {}   # submitted tickets: the device writes into their arrays
        self._mu = threading.Lock()

    def submit(self, queries, k: int, n_probe: int, precision: int = PREC_EXACT,
               rerank_size: int = 0, flags: int = 0, masks=None) -> Ticket:
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim == 1:
            q = q[None, :]
        nq, dim = q.shape
        m = None
        if masks is not None:
            m = np.ascontiguousarray(np.broadcast_to(np.asarray(masks, dtype=np.uint64), (nq,)))
        ids = np.empty((nq, max(k, 1)), dtype=np.uint64)
        dd = np.empty((nq, max(k, 1)), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.uint32)
        o = SearchOpts()
        o.precision, o.rerank_size, o.flags = precision, rerank_size, flags
        t = C.c_uint64()
        check(lib().spf_server_submit(self.h, _p(q), nq, dim, k, n_probe, C.byref(o), _p(m),
                                      _p(ids), _p(dd), _p(cnt), C.byref(t)))
        tk = Ticket(t.value, q, m, ids, dd, cnt)
        with self._mu:
            self._live[tk.id] = tk
        return tk

    def wait(self, ticket: Ticket, timeout_us: int = -1):
        """(ids, dists, counts, info) once the request is done; None if
        timeout_us passes first."""
        info = RequestInfo()
        rc = lib().spf_server_wait(self.h, ticket.id, timeout_us, C.byref(info))
        if rc == SPF_EAGAIN:
            return None
        with self._mu:
            self._live.pop(ticket.id, None)
        check(rc)
        return ticket.ids, ticket.dists, ticket.counts, {
            "submit_ns": int(info.submit_ns), "start_ns": int(info.start_ns),
            "done_ns": int(info.done_ns), "batches": int(info.batches)}

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().spf_server_get_stats(self.h, C.byref(st)))
        return {f: int(getattr(st, f)) for f, _ in ServerStats._fields_ if f != "reserved"}

    def close(self) -> None:
        """Completes every submitted request, then frees the server."""
        if getattr(self, "h", None):
            lib().spf_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
