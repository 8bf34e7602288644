"""Host-side posting codecs, used when an index is built from raw vectors.

Each encoder reproduces the reference's arithmetic element by element in
float32 numpy (one rounding per operation, no fused multiply-add), so codes
built here are the codes the reference would store for the same parameters:

* F16   -- ``F16Quantizer::encode`` (src/sq.rs:89-95): IEEE round-to-nearest-even.
* Int8  -- ``Int8Quantizer::train`` / ``encode`` (src/sq.rs:128-157, 208-220):
  one global scale/offset, u8 codes.
* RaBitQ -- ``RaBitQ::encode_with`` / ``pack`` (src/rabitq.rs:136-149, 218-256):
  residual against the posting centroid, 3 rounds of sign flip + unnormalised
  fast Hadamard transform, 1 sign bit per padded dimension, then ``|r|^2`` and
  the estimator factor as two little-endian f32.
"""
from __future__ import annotations

import numpy as np

F32_EPS = np.float32(np.finfo(np.float32).eps)


# ---------------------------------------------------------------- F16
def f16_encode(v: np.ndarray) -> np.ndarray:
    """[n, d] f32 -> [n, 2d] u8 (little-endian IEEE half)."""
    v = np.ascontiguousarray(v, dtype=np.float32)
    return v.astype(np.float16).view(np.uint8).reshape(v.shape[0], -1)


# ---------------------------------------------------------------- Int8
def int8_train(vectors: np.ndarray) -> tuple[float, float]:
    """Global (scale, offset) as in src/sq.rs:128-157."""
    v = np.asarray(vectors, dtype=np.float32)
    if v.size == 0:
        raise ValueError("No vectors to train on")
    lo = np.float32(v.min())
    hi = np.float32(v.max())
    rng = np.float32(hi - lo)
    scale = np.float32(1.0) if abs(rng) < F32_EPS else np.float32(rng / np.float32(255.0))
    return float(scale), float(lo)


def int8_encode(v: np.ndarray, scale: float, offset: float) -> np.ndarray:
    """[n, d] f32 -> [n, d] u8: sat_round(v * (1/scale) + (-offset/scale))."""
    v = np.asarray(v, dtype=np.float32)
    scale = np.float32(scale)
    alpha = np.float32(np.float32(1.0) / scale)
    beta = np.float32(np.float32(-np.float32(offset)) / scale)
    s = (v * alpha).astype(np.float32)
    s = (s + beta).astype(np.float32)
    with np.errstate(invalid="ignore"):
        r = np.rint(s)
        r = np.where(s > 0, r, 0.0)
        r = np.minimum(r, 255.0)
    return r.astype(np.uint8)


# ---------------------------------------------------------------- RaBitQ
def rabitq_padded(dim: int) -> int:
    p = max(dim, 64)
    return 1 << (p - 1).bit_length()


def rabitq_code_bytes(padded: int) -> int:
    return padded // 8 + 8


def rabitq_signs(padded: int, seed: int = 0x7AB17) -> np.ndarray:
    """3 x padded random signs, each +-1/sqrt(padded) (src/rabitq.rs:96-100).

    The reference draws them from its own RNG; any draw is valid and the
    signs travel with the index, so this one is seeded numpy."""
    rng = np.random.default_rng(seed)
    scale = np.float32(np.float32(1.0) / np.sqrt(np.float32(padded)))
    b = rng.integers(0, 2, size=(3, padded)).astype(bool)
    return np.where(b, scale, -scale).astype(np.float32)


def fht_rows(r: np.ndarray) -> np.ndarray:
    """Unnormalised fast Hadamard transform of each row, same butterfly order
    as src/rabitq.rs:31-43 (each output element is the same IEEE sum)."""
    n, p = r.shape
    h = 1
    while h < p:
        a = r.reshape(n, p // (2 * h), 2, h)
        x = a[:, :, 0, :].copy()
        y = a[:, :, 1, :]
        a[:, :, 0, :] = x + y
        a[:, :, 1, :] = x - y
        h *= 2
    return r


def _seq_sum(cols: np.ndarray) -> np.ndarray:
    """Sequential left fold over axis 1 (Rust's Iterator::sum for f32)."""
    acc = np.zeros(cols.shape[0], dtype=np.float32)
    for i in range(cols.shape[1]):
        acc = (acc + cols[:, i]).astype(np.float32)
    return acc


def rabitq_prep(v: np.ndarray, centroid: np.ndarray, padded: int, signs: np.ndarray):
    """prep_with (src/rabitq.rs:136-149) for many rows; centroid is [n, d] or [d]."""
    v = np.asarray(v, dtype=np.float32)
    n, d = v.shape
    r = np.zeros((n, padded), dtype=np.float32)
    r[:, :d] = v - centroid
    sq = (r * r).astype(np.float32)
    norm = np.sqrt(_seq_sum(sq)).astype(np.float32)
    nz = norm > 0
    r[nz] = r[nz] / norm[nz, None]
    for rr in range(3):
        r *= signs[rr][None, :]
        fht_rows(r)
    return r, norm


def rabitq_encode(v: np.ndarray, centroid: np.ndarray, padded: int, signs: np.ndarray,
                  chunk: int = 65536) -> np.ndarray:
    """[n, d] f32 -> [n, padded/8 + 8] u8 codes (residual vs centroid rows)."""
    v = np.asarray(v, dtype=np.float32)
    n = v.shape[0]
    out = np.zeros((n, rabitq_code_bytes(padded)), dtype=np.uint8)
    cent = np.asarray(centroid, dtype=np.float32)
    sd = np.sqrt(np.float32(padded))
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        c = cent[s:e] if cent.ndim == 2 else cent
        r, norm = rabitq_prep(v[s:e], c, padded, signs)
        bits = np.packbits(r > 0, axis=1, bitorder="little")
        ip = _seq_sum(np.abs(r)) / sd
        ip = ip.astype(np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            fac = np.where(ip > 0, norm / ip, np.float32(0)).astype(np.float32)
        n2 = (norm * norm).astype(np.float32)
        out[s:e, : padded // 8] = bits
        out[s:e, padded // 8: padded // 8 + 4] = n2.view(np.uint8).reshape(-1, 4)
        out[s:e, padded // 8 + 4:] = fac.view(np.uint8).reshape(-1, 4)
    return out
