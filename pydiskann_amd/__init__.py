"""pydiskann_amd -- MI355X batch search for SPFresh indexes.

``SPFresh`` mirrors the search surface of the reference binding
``pydiskann.SPFresh`` (python/src/lib.rs:303-361): ``build``, ``search``,
``search_with_dists``, ``search_batch``, ``delete``, ``is_deleted``,
``get_vector``, ``stats``, ``dim``, with the same argument names, defaults and
error types.  Searches run on the GPU through libspf_hip.so.  Index mutation
beyond tombstones (``insert``, ``gc``, ``compact``) belongs to the reference's
CPU index and is not part of this path.
"""
from __future__ import annotations

import math

import numpy as np

from .index import IndexArrays, build_index, brute_force_f64, METRICS, QUANTS  # noqa: F401
from ._lib import DeviceIndex, Server, PREC_EXACT, PREC_MIXED, SpfError, lib  # noqa: F401
from .refformat import read_index, write_index  # noqa: F401

__all__ = ["SPFresh", "IndexArrays", "build_index", "DeviceIndex", "Server", "PREC_EXACT",
           "PREC_MIXED", "label_masks"]

MAX_LABEL = 64   # labels 0..63 are filterable; an id whose label is >= 64 never matches


def label_masks(values, nq: int) -> np.ndarray:
    """u64 [nq] masks for Filter::LabelIn (src/filtered.rs:65,124-126): bit l
    set = label l allowed.  ``values``: one iterable of labels for every query,
    or a list of nq iterables.  Values must lie in [0, 64)."""
    vals = list(values)
    per_query = len(vals) == nq and all(not np.isscalar(v) for v in vals) and nq > 0
    groups = vals if per_query else [vals] * nq
    out = np.zeros(nq, dtype=np.uint64)
    for i, g in enumerate(groups):
        for v in g:
            v = int(v)
            if not 0 <= v < MAX_LABEL:
                raise ValueError(f"label value {v} outside [0, {MAX_LABEL})")
            out[i] |= np.uint64(1) << np.uint64(v)
    return out


class SPFresh:
    """Search-side mirror of ``pydiskann.SPFresh``."""

    def __init__(self, arrays: IndexArrays, path: str | None = None, device: int = 0):
        self.arrays = arrays
        self.path = path
        self._dev_ordinal = device
        self._dev = DeviceIndex(arrays, device=device)

    # -- construction (python/src/lib.rs:308-319)
    @staticmethod
    def build(vectors, path: str, metric: str = "l2", quantizer=None, rerank_size: int = 0,
              pq_subspaces: int = 8, max_posting_size: int = 128, min_posting_size: int = 16,
              reassign_neighbors: int = 16, probe_ratio: float = math.inf,
              n_postings: int | None = None, seed: int = 0) -> "SPFresh":
        arr = build_index(np.asarray(vectors, dtype=np.float32), metric=metric,
                          quantizer=quantizer, n_postings=n_postings,
                          max_posting_size=max_posting_size, rerank_size=rerank_size,
                          probe_ratio=probe_ratio, seed=seed)
        return SPFresh(arr, path=path)

    @staticmethod
    def from_arrays(arrays: IndexArrays, device: int = 0) -> "SPFresh":
        return SPFresh(arrays, device=device)

    def save(self) -> None:
        """Write {path}.spf / .postings / .vectors in the reference's layout
        (SPFresh::save, src/spfresh.rs:852-858; the centroid graph file is not
        written: this path probes centroids exactly)."""
        if not self.path:
            raise RuntimeError("index has no path")
        write_index(self.arrays, self.path)

    @staticmethod
    def open(path: str, metric: str = "l2") -> "SPFresh":
        """Open an index the reference saved at `path` (SPFresh::open,
        src/spfresh.rs:819-850; python/src/lib.rs:315-319)."""
        arr, _ = read_index(path, metric)
        return SPFresh(arr, path=path)

    # -- search (python/src/lib.rs:326-337)
    def _check_dim(self, n: int) -> None:
        if n != self.arrays.dim:
            raise RuntimeError(f"Query dim {n} != index dim {self.arrays.dim}")

    def search_with_dists(self, query, k: int = 10, n_probe: int = 8) -> list[tuple[int, float]]:
        q = np.asarray(query, dtype=np.float32).ravel()
        self._check_dim(q.shape[0])
        ids, dd, cnt = self._dev.search(q[None, :], k, n_probe)
        return [(int(ids[0, i]), float(dd[0, i])) for i in range(int(cnt[0]))]

    def search(self, query, k: int = 10, n_probe: int = 8) -> list[int]:
        return [i for i, _ in self.search_with_dists(query, k, n_probe)]

    def search_batch(self, queries, k: int = 10, n_probe: int = 8) -> list[list[int]]:
        q = np.asarray(queries, dtype=np.float32)
        if q.size == 0:
            return []
        if q.ndim != 2:
            raise RuntimeError("queries must be a list of equal-length vectors")
        self._check_dim(q.shape[1])
        ids, _, cnt = self._dev.search(q, k, n_probe)
        return [[int(x) for x in ids[i, : cnt[i]]] for i in range(q.shape[0])]

    def search_batch_with_dists(self, queries, k: int = 10, n_probe: int = 8,
                                precision: int = PREC_EXACT, rerank_size: int = 0):
        """numpy form: (ids [nq,k] u64, dists [nq,k] f32, counts [nq])."""
        q = np.asarray(queries, dtype=np.float32)
        self._check_dim(q.shape[1])
        return self._dev.search(q, k, n_probe, precision=precision, rerank_size=rerank_size)

    # -- label filter (Filter::LabelIn semantics, src/filtered.rs:65,124-126;
    #    the reference's SPFresh has none, so this surface is new)
    def set_labels(self, labels) -> None:
        """One label (u8) per id; re-creates the device mirror."""
        lab = np.asarray(labels)
        if lab.shape != (self.arrays.n_raw,) or lab.min(initial=0) < 0 or lab.max(initial=0) > 255:
            raise ValueError("labels must be one value in [0, 255] per id")
        self.arrays.labels = lab.astype(np.uint8)
        dev = DeviceIndex(self.arrays, device=self._dev_ordinal)
        self._dev.close()
        self._dev = dev

    def search_batch_filtered(self, queries, values, k: int = 10, n_probe: int = 8,
                              precision: int = PREC_EXACT, rerank_size: int = 0):
        """Search only ids whose label is in ``values`` (one set for every
        query, or one per query): (ids, dists, counts) as
        search_batch_with_dists."""
        q = np.asarray(queries, dtype=np.float32)
        if q.ndim != 2:
            raise RuntimeError("queries must be a list of equal-length vectors")
        self._check_dim(q.shape[1])
        if self.arrays.labels is None:
            raise RuntimeError("index has no labels (set_labels first)")
        return self._dev.search(q, k, n_probe, precision=precision, rerank_size=rerank_size,
                                masks=label_masks(values, q.shape[0]))

    # -- tombstones (python/src/lib.rs:324-325, 338)
    def delete(self, ids) -> None:
        ids = [int(i) for i in ids]
        self.arrays.set_deleted(ids)
        self._dev.delete(ids)

    def is_deleted(self, id: int) -> bool:
        return self.arrays.is_deleted(int(id))

    def get_vector(self, id: int):
        id = int(id)
        if id >= self.arrays.n_raw or self.arrays.is_deleted(id):
            return None
        return [float(x) for x in self.arrays.raw[id]]

    def stats(self) -> dict:
        a = self.arrays
        lens = a.posting_lengths()
        dead = a.n_deleted()
        return {"live": a.n_raw - dead, "deleted": dead, "postings": a.n_post,
                "max_posting": int(lens.max()) if lens.size else 0}

    @property
    def dim(self) -> int:
        return self.arrays.dim

    def sync(self, arrays: IndexArrays) -> int:
        """Adopt a mutated state of the index (the reference's insert / gc /
        compact run on its CPU index): the device mirror is updated in place
        with the postings that changed (spf_index_update).  Returns the
        number of blocks sent."""
        n = self._dev.update(arrays)
        self.arrays = arrays
        return n

    def insert(self, vectors):
        raise NotImplementedError("insert runs on the reference's CPU index")

    def gc(self):
        raise NotImplementedError("gc runs on the reference's CPU index")

    def compact(self):
        raise NotImplementedError("compact runs on the reference's CPU index")
