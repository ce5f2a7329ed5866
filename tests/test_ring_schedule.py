"""CPU model of the 64-lane coder's LDS ring schedules.

The 64-lane decoder keeps a ring of R words per chain plus a 64-word mirror
of the ring head, refilled half a ring at a time from the unit's words in HBM
(mans_amd/csrc/decode.hip, `refill`); the 64-lane encoder keeps a ring of R
words per chain and moves each full half to a spill area
(mans_amd/csrc/encode.hip, `spill_full`).  These tests replay the same index
arithmetic on random per-step word counts (0..64 words per step, as a
wavefront can take or emit) and check that every word a lane reads or writes
is the right one and every access stays inside the chain's LDS region.
"""
import numpy as np
import pytest


def decode_ring(words: np.ndarray, counts: list[int], R: int) -> None:
    H = R // 2
    W = words.size
    lds = np.full(R + 64, -1, dtype=np.int64)
    lo = 0
    if W > 0:
        jt = (W - 1) // H
        j0 = jt - 1 if jt > 0 else 0
        lo = j0 * H
        for j in range(j0, jt + 1):
            n = min(W - j * H, H)
            n8 = -(-n // 8) * 8
            for i in range(n8):  # whole uint4s: the padded tail reads as data too
                v = words[j * H + i] if j * H + i < W else 0
                lds[(j & 1) * H + i] = v
                if (j & 1) == 0 and i < 64:
                    lds[R + i] = v
    ptr, pend, pending = W, False, None
    for cnt in counts:
        base = max(ptr - cnt, 0)
        assert ptr >= cnt, "schedule consumed more words than the unit holds"
        for rank in range(64):  # every lane reads (base & (R-1)) + rank
            addr = (base & (R - 1)) + rank
            assert 0 <= addr < R + 64
            if rank < cnt:  # lanes that take a word
                assert lds[addr] == words[base + rank], (base, rank)
        ptr = base
        # refill(): trigger, then commit
        if not pend and lo > 0 and ptr <= lo + H:
            j = lo // H - 1
            pending = words[j * H:(j + 1) * H].copy()
            assert pending.size == H
            pend = True
        if pend and ptr < lo + 64:
            j = lo // H - 1
            # the half being overwritten holds chunk j + 2: all taken already
            assert ptr <= (j + 2) * H
            lds[(j & 1) * H:(j & 1) * H + H] = pending
            if (j & 1) == 0:
                lds[R:R + 64] = pending[:64]
            lo -= H
            pend = False
    assert ptr == 0


def encode_ring(counts: list[int], R: int, cap: int) -> None:
    H = R // 2
    lds = np.full(R, -1, dtype=np.int64)
    spill = np.full(cap, -1, dtype=np.int64)
    cnt = fl = 0
    for add in counts:
        for rank in range(add):
            pos = (cnt + rank) & (R - 1)
            # never overwrite a word that is not yet spilled
            assert lds[pos] < fl or lds[pos] == -1
            lds[pos] = cnt + rank  # the word's index stands for its value
        cnt += add
        if cnt - fl >= H:  # spill_full()
            assert fl + H <= cap
            spill[fl:fl + H] = lds[(fl & (R - 1)):(fl & (R - 1)) + H]
            fl += H
    padded = -(-cnt // 8) * 8
    for k in range(cnt, padded):
        lds[k & (R - 1)] = k
    out = np.array([spill[j] if j < fl else lds[j & (R - 1)] for j in range(padded)])
    assert (out == np.arange(padded)).all()


def counts_for(rng, total: int, hi: int) -> list[int]:
    out, left = [], total
    while left > 0:
        c = min(int(rng.integers(0, hi + 1)), left)
        out.append(c)
        left -= c
    out += [0] * int(rng.integers(0, 3))
    return out


@pytest.mark.parametrize("R", [512, 1024])
@pytest.mark.parametrize("seed", range(6))
def test_decoder_ring_reads_the_right_words(R, seed):
    rng = np.random.default_rng(seed)
    for W in [0, 1, 8, 63, 64, 65, R // 2 - 1, R // 2, R // 2 + 1, R - 1, R, R + 1, 3 * R + 17, 5760]:
        words = rng.integers(0, 1 << 16, W)
        for hi in (1, 7, 40, 64):
            decode_ring(words, counts_for(rng, W, hi), R)


@pytest.mark.parametrize("R", [512, 1024])
@pytest.mark.parametrize("seed", range(6))
def test_encoder_ring_spills_in_order(R, seed):
    rng = np.random.default_rng(100 + seed)
    cap = 5760
    for total in [0, 1, 9, R // 2 - 1, R // 2, R // 2 + 63, R, 2 * R + 5, cap - 7, cap]:
        for hi in (1, 13, 64):
            encode_ring(counts_for(rng, total, hi), R, cap)
