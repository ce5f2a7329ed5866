"""CPU model of the decoders' word rings (no GPU).  A chain's unit of `words`
16-bit words is read backwards, at most 64 words per step; LDS holds a ring
of two halves of H words (plus a mirror of its first 64 words), the half
below is fetched into registers and written over the half just consumed.
The model replays the upkeep schedules of the kernels and checks that every
word a step reads sits in the slot (ring index, or mirror) it reads:

* batched decode, one upkeep per step (KB = 1): fetch once ptr <= lo + H,
  write when ptr < lo + 64 (batch_kernels.hpp, decode_ticket);
* batched decode, one upkeep every KB steps (KB = 4, 8, 16, H = max(512,
  64 KB)): fetch right after the previous write, write once ptr <= lo + H;
* single-buffer decode, checks at most RK = 4 steps apart (decode.hip; the
  model checks only every RK steps, the sparsest case): fetch once
  ptr <= lo + H, write when ptr < lo + L RK, H = 512 (or 256 for four
  chains), L = 64 lanes, or L = 32 for the reference format's units (at most
  32 words per step, one ring per unit of a half-wave)."""
import random

import pytest


def run(cons, full, kb, early, h, commit_gap):
    """cons[s]: words taken by step s (steps run from len(cons) - 1 down to
    0); steps below `full` are grouped kb at a time.  Raises on a bad read."""
    words = sum(cons)
    r = 2 * h
    mask = r - 1
    ring = [None] * (r + 64)
    top = (words + 7) // 8 * 8

    def load_half(j):
        for w in range(j * h, min(j * h + h, top)):
            ring[w & mask] = w
            if (j & 1) == 0 and w - j * h < 64:
                ring[r + w - j * h] = w

    st = {"lo": 0, "pend": False, "ptr": words, "pj": None}
    if words > 0:
        jt = (words - 1) // h
        j0 = jt - 1 if jt > 0 else 0
        st["lo"] = j0 * h
        for j in range(j0, jt + 1):
            load_half(j)

    def fetch():
        st["pj"] = st["lo"] // h - 1
        st["pend"] = True

    def commit():
        assert st["pj"] == st["lo"] // h - 1
        load_half(st["pj"])
        st["lo"] -= h
        st["pend"] = False

    def upkeep():
        if early:
            if not st["pend"] and st["lo"] > 0:
                fetch()
            if st["pend"] and st["ptr"] <= st["lo"] + h:
                commit()
                if st["lo"] > 0:
                    fetch()
        else:
            if not st["pend"] and st["lo"] > 0 and st["ptr"] <= st["lo"] + h:
                fetch()
            if st["pend"] and st["ptr"] < st["lo"] + commit_gap:
                commit()

    def step(s):
        base = st["ptr"] - cons[s]
        for k in range(cons[s]):
            slot = (base & mask) + k
            assert ring[slot] == base + k, (s, base + k, slot, ring[slot], st)
        st["ptr"] = base

    s = len(cons)
    while s > full or s % kb:
        s -= 1
        step(s)
        if s % kb == 0:
            upkeep()
    while s > 0:
        for _ in range(kb):
            s -= 1
            step(s)
        upkeep()
    assert st["ptr"] == 0


def cases(seed, n, lanes=64):
    rng = random.Random(seed)
    for _ in range(n):
        mode = rng.random()
        if mode < 0.3:
            cons = [lanes] * 128
        elif mode < 0.6:
            p = rng.random()
            cons = [sum(rng.random() < p for _ in range(lanes)) for _ in range(128)]
        else:
            cons = [rng.choice([0, lanes, rng.randrange(lanes + 1)]) for _ in range(128)]
        yield cons, rng.choice([0, 128, rng.randrange(129)])


@pytest.mark.parametrize("kb", [1, 4, 8, 16])
def test_batched_ring_schedules(kb):
    h = max(512, 64 * kb)
    for cons, full in cases(kb, 150):
        run(cons, full, kb, early=kb > 1, h=h, commit_gap=64)


def test_single_buffer_ring_schedule():
    for cons, full in cases(99, 150):
        run(cons, full, 4, early=False, h=512, commit_gap=64 * 4)


@pytest.mark.parametrize("h", [512, 256])
def test_single_buffer_ring_schedule_32_lanes(h):
    for cons, full in cases(31 + h, 150, lanes=32):
        run(cons, full, 4, early=False, h=h, commit_gap=32 * 4)


def test_model_catches_a_late_write():
    # the same schedule writing the next half one check too late reads a
    # word that is not in the ring yet
    cons = [64] * 128
    with pytest.raises(AssertionError):
        run(cons, 128, 4, early=False, h=512, commit_gap=0)
