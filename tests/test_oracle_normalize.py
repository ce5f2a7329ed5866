"""Known answers for the normalisation (reference GpuANSStatistics.h:186-385).

Each expected table below was derived by hand from the reference's steps
(fp32 quantisation, symbol-id "+1" rule when the sum is short, rank-order
decrements when it is over); the comments show the arithmetic.  They pin
oracle.normalize, which the GPU normalisation is then compared with.
"""
import numpy as np
import pytest

import oracle


def hist_of(d):
    h = np.zeros(256, dtype=np.uint32)
    for s, c in d.items():
        h[s] = c
    return h


def test_single_symbol():
    # q = trunc(1024 * 1.0) = 1024 = 2^p; shift = 32 - clz(1023) = 10;
    # magic = 2^32 * (1024 - 1024) / 1024 + 1 = 1
    t = oracle.normalize(hist_of({65: 1000}), 1000, 10)
    assert t[65].tolist() == [1024, 0, 1, 10]
    assert t[:, 0].sum() == 1024
    assert (t[66:, 1] == 1024).all() and (t[:65, 1] == 0).all()


def test_uniform():
    # q = trunc(1024 / 256) = 4 for all; cdf = 4 s; shift = 32 - clz(3) = 2
    t = oracle.normalize(np.ones(256, dtype=np.uint32), 256, 10)
    assert (t[:, 0] == 4).all()
    assert (t[:, 1] == 4 * np.arange(256)).all()
    assert (t[:, 3] == 2).all() and (t[:, 2] == 1).all()


def test_short_sum_goes_to_low_symbol_ids():
    # three symbols of 1/3: fp32(1/3) * 512 = 170.66.. -> 170 each, sum 510,
    # diff 2 -> symbol ids 0 and 1 (absent!) get +1 (GpuANSStatistics.h:276)
    t = oracle.normalize(hist_of({5: 1, 9: 1, 250: 1}), 3, 9)
    pdf = t[:, 0]
    assert pdf[5] == pdf[9] == pdf[250] == 170
    assert pdf[0] == 1 and pdf[1] == 1
    assert pdf.sum() == 512 and np.count_nonzero(pdf) == 5


def test_over_sum_single_large():
    # symbol 0: trunc(512 * 1e6/1000100) = 511; 100 tiny symbols -> 1 each;
    # sum 611, diff -99, G = 1 each round -> symbol 0 loses 99 -> 412
    h = hist_of({0: 1_000_000, **{s: 1 for s in range(1, 101)}})
    t = oracle.normalize(h, int(h.sum()), 9)
    assert t[0, 0] == 412
    assert (t[1:101, 0] == 1).all() and t[:, 0].sum() == 512


def test_over_sum_tie_decrements_smaller_id_first():
    # ids 10 and 20 at 500000 each, 99 tiny ids 100..198; T = 1000099.
    # q = trunc(512 * 0.4999505) = 255 each; sum 510 + 99 = 609, diff -97.
    # Sorted keys: (255,20) rank 0, (255,10) rank 1.  48 rounds of G = 2 take
    # 96; the last round takes ranks [1, 2) -> id 10.  20 -> 207, 10 -> 206.
    h = hist_of({10: 500_000, 20: 500_000, **{s: 1 for s in range(100, 199)}})
    t = oracle.normalize(h, int(h.sum()), 9)
    assert t[20, 0] == 207 and t[10, 0] == 206
    assert t[:, 0].sum() == 512


def test_over_sum_partial_round_hits_lower_ranks():
    # A=3:600000 B=7:300000 C=9:99000, 55 tiny ids 200..254; T = 999055.
    # q: 512*0.600567 = 307.49 -> 307, 153.74 -> 153, 50.74 -> 50, tiny 1.
    # sum 565, diff -53: 17 rounds of G = 3 (-51), then it = 2 at ranks
    # [1, 3) = B, C.  A = 290, B = 135, C = 32.
    h = hist_of({3: 600_000, 7: 300_000, 9: 99_000, **{s: 1 for s in range(200, 255)}})
    t = oracle.normalize(h, int(h.sum()), 9)
    assert (t[3, 0], t[7, 0], t[9, 0]) == (290, 135, 32)
    assert t[:, 0].sum() == 512


def test_empty_input_leaves_zero_table():
    t = oracle.normalize(np.zeros(256, dtype=np.uint32), 0, 10)
    assert not t.any()


@pytest.mark.parametrize("p", [9, 10, 11])
def test_invariants_random(p):
    rng = np.random.default_rng(p)
    for trial in range(200):
        k = int(rng.integers(1, 257))
        syms = rng.choice(256, size=k, replace=False)
        counts = rng.integers(1, 10 ** int(rng.integers(1, 7)), size=k)
        h = np.zeros(256, dtype=np.uint32)
        h[syms] = counts
        t = oracle.normalize(h, int(h.sum()), p)
        pdf = t[:, 0].astype(np.int64)
        assert pdf.sum() == 1 << p
        assert (pdf[h > 0] >= 1).all()
        assert (t[:, 1] == np.concatenate([[0], np.cumsum(pdf)[:-1]])).all()


def test_magic_division_is_exact():
    # the table's (magic, shift) must give floor(x / pdf) for every state the
    # encoder can see (x < 2^31): brute force over sampled x for every pdf that
    # random tables produce, plus the edges x = k*pdf - 1, k*pdf.
    rng = np.random.default_rng(7)
    seen = {}
    # counts (c, 2048 - c) out of 2048 at p = 11 quantise exactly to pdf c
    for c in range(1, 2048):
        h = np.zeros(256, dtype=np.uint32)
        h[3], h[200] = c, 2048 - c
        t = oracle.normalize(h, 2048, 11)
        assert t[3, 0] == c
        seen[c] = (int(t[3, 2]), int(t[3, 3]))
    t = oracle.normalize(np.eye(1, 256, 9, dtype=np.uint32)[0], 1, 11)
    seen[2048] = (int(t[9, 2]), int(t[9, 3]))
    assert sorted(seen) == list(range(1, 2049))
    xs = np.concatenate([rng.integers(0, 2 ** 31, 200), [0, 1, 2 ** 15, 2 ** 31 - 1]])
    for pdf, (magic, shift) in seen.items():
        for x in list(xs) + [pdf * 1000 - 1, pdf * 1000, (2 ** 31 - 1) // pdf * pdf]:
            x = int(x)
            assert oracle.magic_div(x, magic, shift) == x // pdf, (x, pdf)
