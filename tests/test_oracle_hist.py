"""Pins of the CPU histogram restatement (reference
include/ans/GpuANSStatistics.h:23-151): every count equals a brute-force
np.bincount of the same bytes, and the counts sum to the number of symbols,
for strided and offset byte selections as the multi-plane frames use them."""
import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("stride", [1, 2, 4])
def test_histogram_equals_bincount(stride):
    rng = np.random.default_rng(40 + stride)
    for n in [0, 1, 7, 255, 256, 4097, 100_003]:
        raw = rng.integers(0, 256, n * stride + 8).astype(np.uint8)
        raw[: n * stride: 5] = 200  # a hot bin
        for off in range(stride):
            sel = raw[off: off + n * stride]
            h = oracle.histogram(sel, stride=stride)
            ref = np.bincount(sel[::stride][:n], minlength=256).astype(np.uint32)
            assert h.dtype == np.uint32 and h.shape == (256,)
            assert (h == ref).all(), (n, off)
            assert int(h.sum()) == n


def test_histogram_offset_argument():
    rng = np.random.default_rng(7)
    raw = rng.integers(0, 256, 10_000).astype(np.uint8)
    for off, stride in [(3, 1), (1, 2), (2, 4), (9, 4)]:
        h = oracle.histogram(raw, stride=stride, offset=off)
        ref = np.bincount(raw[off::stride], minlength=256)
        assert (h == ref).all() and int(h.sum()) == raw[off::stride].size


def test_histogram_single_symbol_and_extremes():
    for v in (0, 255):
        x = np.full(12_345, v, dtype=np.uint8)
        h = oracle.histogram(x)
        assert h[v] == 12_345 and int(h.sum()) == 12_345
