"""CPU restatement of the dual-quantisation 3-D Lorenzo front end --
TEST INFRASTRUCTURE ONLY (the checker; only tests/, __graft_entry__.smoke()
and bench.py's CPU leg may use it).

The reference has no Lorenzo kernel (its mapping programs read uint16 codes,
mapping/cpu/mappingcpuV1.cpp:33-193); this restates the definition the
library's kernel implements, in numpy:
  q = rint(f * float32(1 / (2 eb)))               float32 product, ties to even
  r = 3-D Lorenzo residual of q (q = 0 outside the field)
  code = 32768 + r if |r| < radius else 0, and (index, r) is an outlier;
outliers are listed in the kernel's tile order: tiles of 256 x-columns x 16
rows x 16 planes numbered x-tile fastest, then y, then z; within a tile by x,
then plane, then row.  The inverse is the 3-D prefix sum of the residuals
times float32(2 eb).
"""
from __future__ import annotations

import numpy as np

TX, TY, TZ = 256, 16, 16


def quantize(field: np.ndarray, error_bound: float) -> np.ndarray:
    inv = np.float32(1.0 / (2.0 * error_bound))
    return np.rint(np.ascontiguousarray(field, dtype=np.float32) * inv)


def residual(q: np.ndarray) -> np.ndarray:
    """3-D Lorenzo residual of integer-valued q (int64)."""
    qi = q.astype(np.int64)
    p = np.pad(qi, ((1, 0), (1, 0), (1, 0)))
    return (p[1:, 1:, 1:] - p[:-1, 1:, 1:] - p[1:, :-1, 1:] - p[1:, 1:, :-1]
            + p[:-1, :-1, 1:] + p[:-1, 1:, :-1] + p[1:, :-1, :-1] - p[:-1, :-1, :-1])


def lorenzo_quantize(field: np.ndarray, error_bound: float, radius: int = 512):
    """(codes uint16 [n], outlier indices uint32, outlier residuals int32)."""
    nz, ny, nx = field.shape
    q = quantize(field, error_bound)
    if not (np.abs(q) < 2 ** 28).all():
        raise ValueError("|f / (2 eb)| must stay below 2^28")
    r = residual(q)
    inside = np.abs(r) < radius
    codes = np.where(inside, r + 32768, 0).astype(np.uint16).reshape(-1)
    zz, yy, xx = np.nonzero(~inside)
    nxb, nyb = -(-nx // TX), -(-ny // TY)
    tile = (zz // TZ) * (nyb * nxb) + (yy // TY) * nxb + xx // TX
    order = np.lexsort((yy % TY, zz % TZ, xx % TX, tile))
    idx = ((zz * ny + yy) * nx + xx)[order].astype(np.uint32)
    val = r[zz, yy, xx][order].astype(np.int32)
    return codes, idx, val


def lorenzo_reconstruct(codes: np.ndarray, idx: np.ndarray, val: np.ndarray, shape, error_bound: float) -> np.ndarray:
    c = np.asarray(codes).reshape(-1).astype(np.int64)
    r = np.where(c != 0, c - 32768, 0)
    r[np.asarray(idx, dtype=np.int64)] = np.asarray(val, dtype=np.int64)
    q = r.reshape(shape).cumsum(0).cumsum(1).cumsum(2)
    return q.astype(np.float32) * np.float32(2.0 * error_bound)
