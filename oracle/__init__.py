"""ctypes front of the CPU oracle (libspf_oracle.so, built from spf_oracle.c).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, and only to check or time against it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libspf_oracle.so")


class OIndex(C.Structure):
    _fields_ = [
        ("dim", C.c_uint32), ("metric", C.c_uint32), ("quant", C.c_uint32),
        ("n_post", C.c_uint32), ("code_bytes", C.c_uint32), ("rerank_size", C.c_uint32),
        ("probe_ratio", C.c_float), ("int8_scale", C.c_float), ("int8_offset", C.c_float),
        ("rq_padded", C.c_uint32),
        ("block_ids", C.c_void_p), ("centroids", C.c_void_p), ("post_off", C.c_void_p),
        ("ids", C.c_void_p), ("codes", C.c_void_p), ("raw", C.c_void_p),
        ("n_raw", C.c_uint64), ("deleted", C.c_void_p), ("rq_signs", C.c_void_p),
        ("labels", C.c_void_p), ("graph", C.c_void_p),
    ]


class OGraph(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32), ("n_seeds", C.c_uint32), ("seeds", C.c_uint32 * 2),
        ("centroid_beam", C.c_uint32), ("n_tombstones", C.c_uint64),
        ("vectors", C.c_void_p), ("adj_off", C.c_void_p), ("adj", C.c_void_p),
        ("live", C.c_void_p), ("node_post", C.c_void_p),
    ]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        u32, f32 = C.c_uint32, C.c_float
        L.spf_o_f32_to_f16.restype = C.c_uint16
        L.spf_o_f32_to_f16.argtypes = [f32]
        L.spf_o_f16_to_f32.restype = f32
        L.spf_o_f16_to_f32.argtypes = [C.c_uint16]
        for n in ("spf_o_l2sq", "spf_o_dot"):
            getattr(L, n).restype = f32
            getattr(L, n).argtypes = [P, P, u32]
        L.spf_o_eval.restype = f32
        L.spf_o_eval.argtypes = [u32, P, P, u32]
        L.spf_o_int8_encode.argtypes = [P, u32, f32, f32, P]
        L.spf_o_int8_l2sq.restype = f32
        L.spf_o_int8_l2sq.argtypes = [P, P, u32, f32]
        L.spf_o_fht.argtypes = [P, u32]
        L.spf_o_rotate.argtypes = [P, u32, P]
        L.spf_o_unrotate.argtypes = [P, u32, P]
        L.spf_o_rabitq_encode.argtypes = [P, P, u32, u32, P, P]
        L.spf_o_rabitq_decode.argtypes = [P, P, u32, u32, P, P]
        L.spf_o_rabitq_query.argtypes = [P, P, u32, u32, P, P, P]
        L.spf_o_rabitq_dist.argtypes = [P, P, P, u32, P, P]
        L.spf_o_probe.restype = u32
        L.spf_o_probe.argtypes = [P, P, u32, P, P]
        L.spf_o_search.restype = u32
        L.spf_o_search.argtypes = [P, P, u32, u32, P, P]
        L.spf_o_search_batch.argtypes = [P, P, u32, u32, u32, u32, P, P, P]
        L.spf_o_search_batch_filtered.argtypes = [P, P, P, u32, u32, u32, u32, P, P, P]
        L.spf_o_set_fault.argtypes = [C.c_int]
        L.spf_o_graph_search.restype = u32
        L.spf_o_graph_search.argtypes = [P, u32, u32, P, u32, u32, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def graph_struct(g):
    """(OGraph, kept arrays) for a pydiskann_amd.index.CentroidGraph."""
    keep = [np.ascontiguousarray(g.vectors, dtype=np.float32),
            np.ascontiguousarray(g.adj_off, dtype=np.uint32),
            np.ascontiguousarray(g.adj, dtype=np.uint32),
            np.ascontiguousarray(g.live, dtype=np.uint8),
            np.ascontiguousarray(g.node_post, dtype=np.int64)]
    o = OGraph()
    o.n_nodes = keep[0].shape[0]
    o.n_seeds = len(g.seeds)
    for i, sd in enumerate(g.seeds):
        o.seeds[i] = int(sd)
    o.centroid_beam = g.centroid_beam
    o.n_tombstones = g.n_tombstones
    o.vectors, o.adj_off, o.adj, o.live, o.node_post = (_p(a) for a in keep)
    return o, keep


class Oracle:
    """Oracle view of an ``pydiskann_amd.index.IndexArrays``."""

    def __init__(self, ix):
        self.ix = ix
        self._keep = ix
        s = OIndex()
        s.dim, s.metric, s.quant = ix.dim, ix.metric, ix.quant
        s.n_post, s.code_bytes, s.rerank_size = ix.n_post, ix.code_bytes, ix.rerank_size
        s.probe_ratio = ix.probe_ratio
        s.int8_scale, s.int8_offset = ix.int8_scale, ix.int8_offset
        s.rq_padded = ix.rq_padded
        s.block_ids, s.centroids, s.post_off = _p(ix.block_ids), _p(ix.centroids), _p(ix.post_off)
        s.ids, s.codes, s.raw = _p(ix.ids), _p(ix.codes), _p(ix.raw)
        s.n_raw = ix.n_raw
        s.deleted = _p(ix.deleted)
        s.rq_signs = _p(ix.rq_signs)
        self._labels = None
        if getattr(ix, "labels", None) is not None:
            self._labels = np.ascontiguousarray(ix.labels, dtype=np.uint8)
            assert self._labels.shape[0] == ix.n_raw
            s.labels = _p(self._labels)
        self._graph = None
        if getattr(ix, "graph", None) is not None:
            self._graph = graph_struct(ix.graph)
            s.graph = C.cast(C.byref(self._graph[0]), C.c_void_p)
        self.s = s

    def graph_search(self, q, k: int, beam: int):
        """IncrementalDiskANN::search_with_dists over the routing graph:
        (unified node ids, distances)."""
        q = np.ascontiguousarray(q, dtype=np.float32)
        nodes = np.zeros(max(k, 1), dtype=np.uint32)
        dd = np.zeros(max(k, 1), dtype=np.float32)
        c = lib().spf_o_graph_search(C.byref(self._graph[0]), self.ix.metric, self.ix.dim, _p(q),
                                     k, beam, _p(nodes), _p(dd))
        return nodes[:c].copy(), dd[:c].copy()

    def search(self, q, k: int, n_probe: int):
        q = np.ascontiguousarray(q, dtype=np.float32)
        ids = np.zeros(max(k, 1), dtype=np.uint64)
        dd = np.zeros(max(k, 1), dtype=np.float32)
        n = lib().spf_o_search(C.byref(self.s), _p(q), k, n_probe, _p(ids), _p(dd))
        return ids[:n].copy(), dd[:n].copy()

    def probe(self, q, n_probe: int):
        q = np.ascontiguousarray(q, dtype=np.float32)
        n = max(1, n_probe)
        posts = np.zeros(n, dtype=np.uint32)
        dd = np.zeros(n, dtype=np.float32)
        c = lib().spf_o_probe(C.byref(self.s), _p(q), n_probe, _p(posts), _p(dd))
        return posts[:c].copy(), dd[:c].copy()

    def search_batch(self, qs, k: int, n_probe: int, threads: int = 1, masks=None):
        """masks: per-query u64 label masks (filtered search) or None."""
        qs = np.ascontiguousarray(qs, dtype=np.float32)
        nq = qs.shape[0]
        ids = np.zeros((nq, k), dtype=np.uint64)
        dd = np.zeros((nq, k), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.uint32)
        if masks is None:
            lib().spf_o_search_batch(C.byref(self.s), _p(qs), nq, k, n_probe, threads, _p(ids),
                                     _p(dd), _p(cnt))
        else:
            m = np.ascontiguousarray(masks, dtype=np.uint64)
            assert m.shape == (nq,)
            lib().spf_o_search_batch_filtered(C.byref(self.s), _p(qs), _p(m), nq, k, n_probe,
                                              threads, _p(ids), _p(dd), _p(cnt))
        return ids, dd, cnt


def set_fault(f: int) -> None:
    lib().spf_o_set_fault(f)


# ---- scalar helpers -------------------------------------------------------
def f32_to_f16(x: float) -> int:
    return lib().spf_o_f32_to_f16(x)


def f16_to_f32(h: int) -> float:
    return lib().spf_o_f16_to_f32(h)


def eval_metric(metric: int, a, b) -> float:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return lib().spf_o_eval(metric, _p(a), _p(b), a.shape[0])


def int8_encode(v, scale, offset):
    v = np.ascontiguousarray(v, dtype=np.float32)
    out = np.zeros(v.shape[0], dtype=np.uint8)
    lib().spf_o_int8_encode(_p(v), v.shape[0], scale, offset, _p(out))
    return out


def int8_l2sq(a, b, scale):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    return lib().spf_o_int8_l2sq(_p(a), _p(b), a.shape[0], scale)


def rabitq_encode(v, centroid, padded, signs):
    v = np.ascontiguousarray(v, dtype=np.float32)
    centroid = np.ascontiguousarray(centroid, dtype=np.float32)
    code = np.zeros(padded // 8 + 8, dtype=np.uint8)
    lib().spf_o_rabitq_encode(_p(v), _p(centroid), v.shape[0], padded, _p(signs), _p(code))
    return code


def rabitq_decode(code, centroid, d, padded, signs):
    out = np.zeros(d, dtype=np.float32)
    centroid = np.ascontiguousarray(centroid, dtype=np.float32)
    code = np.ascontiguousarray(code, dtype=np.uint8)
    lib().spf_o_rabitq_decode(_p(code), _p(centroid), d, padded, _p(signs), _p(out))
    return out


def rabitq_query(q, centroid, padded, signs):
    q = np.ascontiguousarray(q, dtype=np.float32)
    centroid = np.ascontiguousarray(centroid, dtype=np.float32)
    planes = np.zeros(4 * (padded // 64), dtype=np.uint64)
    consts = np.zeros(5, dtype=np.float32)
    lib().spf_o_rabitq_query(_p(q), _p(centroid), q.shape[0], padded, _p(signs), _p(planes),
                             _p(consts))
    return planes, consts


def rabitq_dist(planes, consts, code, padded):
    est = np.zeros(1, dtype=np.float32)
    err = np.zeros(1, dtype=np.float32)
    code = np.ascontiguousarray(code, dtype=np.uint8)
    lib().spf_o_rabitq_dist(_p(planes), _p(consts), _p(code), padded, _p(est), _p(err))
    return float(est[0]), float(err[0])


def rotate(v, padded, signs, inverse=False):
    v = np.array(v, dtype=np.float32, copy=True)
    (lib().spf_o_unrotate if inverse else lib().spf_o_rotate)(_p(v), padded, _p(signs))
    return v
