"""Seeded synthetic workloads (inputs only; no labels of truth).

Streams come from ``np.random.SeedSequence(0).spawn(n)`` so every workload is
reproducible on any host.

* config1  100k x 128 f32, 256-component Gaussian mixture, centres N(0, 4I),
           points = centre + N(0, I); 1,000 queries from the same mixture.
* config2  1M x 128 integer-valued f32 in [0, 255]: clamp(round(mixture of
           1,024 Gaussians, centres U[0, 200], sigma 20)); 10k queries.
* config3  10M x 96 unit vectors from a 16,384-component vMF-like mixture
           (concentration 50), component weights Zipf(1.1).
* config4  5M x 768 unit vectors, covariance eigenvalues i^-0.8 under a random
           rotation.
* labels   one of 64 uniform labels per base vector, and per-query allow masks
           of 1-8 random labels (the filtered serving workload over config3).
"""
from __future__ import annotations

import numpy as np


def _streams(n: int, seed: int = 0):
    return [np.random.default_rng(s) for s in np.random.SeedSequence(seed).spawn(n)]


def config1(n: int = 100_000, nq: int = 1_000, dim: int = 128, comps: int = 256):
    r_c, r_b, r_q = _streams(3)
    centres = r_c.normal(0.0, 2.0, size=(comps, dim)).astype(np.float32)

    def draw(r, m):
        lab = r.integers(0, comps, size=m)
        return (centres[lab] + r.normal(0.0, 1.0, size=(m, dim))).astype(np.float32)

    return draw(r_b, n), draw(r_q, nq)


def config2(n: int = 1_000_000, nq: int = 10_000, dim: int = 128, comps: int = 1_024,
            chunk: int = 250_000):
    r_c, r_b, r_q = _streams(3)
    centres = r_c.uniform(0.0, 200.0, size=(comps, dim)).astype(np.float32)

    def draw(r, m):
        out = np.empty((m, dim), dtype=np.float32)
        for s in range(0, m, chunk):
            e = min(m, s + chunk)
            lab = r.integers(0, comps, size=e - s)
            v = centres[lab] + r.normal(0.0, 20.0, size=(e - s, dim)).astype(np.float32)
            out[s:e] = np.clip(np.rint(v), 0, 255)
        return out

    return draw(r_b, n), draw(r_q, nq)


def _normalize(x):
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def config3(n: int = 10_000_000, nq: int = 10_000, dim: int = 96, comps: int = 16_384,
            kappa: float = 50.0, zipf: float = 1.1, chunk: int = 500_000,
            return_means: bool = False):
    """(base, queries[, component mean directions]); with one posting per
    component the posting sizes follow the Zipf weights."""
    r_c, r_w, r_b, r_q = _streams(4)
    mu = _normalize(r_c.normal(size=(comps, dim)))
    w = 1.0 / np.arange(1, comps + 1) ** zipf
    w /= w.sum()
    perm = r_w.permutation(comps)
    w = w[perm]

    def draw(r, m, weights):
        out = np.empty((m, dim), dtype=np.float32)
        for s in range(0, m, chunk):
            e = min(m, s + chunk)
            lab = r.choice(comps, size=e - s, p=weights)
            out[s:e] = _normalize(mu[lab] + r.normal(size=(e - s, dim)) / np.sqrt(kappa))
        return out

    # queries uniform over components (each component equally likely)
    base, queries = draw(r_b, n, w), draw(r_q, nq, np.full(comps, 1.0 / comps))
    return (base, queries, mu) if return_means else (base, queries)


def labels(n: int, n_labels: int = 64):
    """One uniform label in [0, n_labels) per base vector (u8)."""
    return _streams(5)[4].integers(0, n_labels, size=n).astype(np.uint8)


def label_masks(nq: int, seed: int = 1, lo: int = 1, hi: int = 8, n_labels: int = 64):
    """Per-query u64 allow masks of lo..hi distinct random labels."""
    r = _streams(6 + seed)[5 + seed]
    out = np.zeros(nq, dtype=np.uint64)
    for i in range(nq):
        for b in r.choice(n_labels, size=int(r.integers(lo, hi + 1)), replace=False):
            out[i] |= np.uint64(1) << np.uint64(int(b))
    return out


def config4(n: int = 5_000_000, nq: int = 10_000, dim: int = 768, chunk: int = 250_000):
    r_q0, r_b, r_q = _streams(3)
    lam = np.arange(1, dim + 1, dtype=np.float64) ** -0.8
    Q, _ = np.linalg.qr(r_q0.normal(size=(dim, dim)))
    A = (Q * np.sqrt(lam)[None, :]).astype(np.float32)

    def draw(r, m):
        out = np.empty((m, dim), dtype=np.float32)
        for s in range(0, m, chunk):
            e = min(m, s + chunk)
            z = r.normal(size=(e - s, dim)).astype(np.float32)
            out[s:e] = _normalize(z @ A.T)
        return out

    return draw(r_b, n), draw(r_q, nq)
