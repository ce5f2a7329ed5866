"""Pins of the CPU rANS stream / frame restatement (reference
include/ans/GpuANSEncode.h:24-401, GpuANSDecode.h:27-451, GpuANSCodec.h).

Hand-traced known answers (no reference vectors exist: the reference's GPU
codec cannot be built here), container-layout checks written from the
reference's struct definitions, and invariants a wrong lane/step/word order
breaks.
"""
import numpy as np
import pytest

import oracle

REF, NAT = oracle.FORMAT_REF, oracle.FORMAT_NATIVE


def u32(b, off):
    return int(np.frombuffer(bytes(b[off:off + 4]), dtype="<u4")[0])


def parse(c, fmt=REF):
    """Container fields by the reference layout (GpuANSCodec.h:68-209)."""
    lanes = fmt[0]
    nb = u32(c, 4)
    states = np.frombuffer(bytes(c[544:544 + 4 * lanes * nb]), dtype="<u4").reshape(nb, lanes)
    bw_off = 544 + 4 * lanes * nb
    bw = np.frombuffer(bytes(c[bw_off:bw_off + 8 * nb]), dtype="<u4").reshape(nb, 2)
    data_off = bw_off + 8 * (nb + (nb & 1))
    return dict(magic=u32(c, 0), nb=nb, n=u32(c, 8), words=u32(c, 12), options=u32(c, 16),
                rest=[u32(c, 20), u32(c, 24), u32(c, 28)],
                probs=np.frombuffer(bytes(c[32:544]), dtype="<u2"), states=states, bw=bw,
                data=np.frombuffer(bytes(c[data_off:]), dtype="<u2"), data_off=data_off)


def test_single_symbol_emits_nothing():
    # pdf = 2^p: x' = floor(x/2^p) 2^p + x mod 2^p + 0 = x, never >= 2^31
    x = np.full(10000, 77, dtype=np.uint8)
    c = oracle.stream_encode(x, 10, REF)
    f = parse(c)
    assert f["magic"] == 0xd00d0001 and f["nb"] == 3 and f["n"] == 10000
    assert f["words"] == 0 and f["options"] == 10 and f["rest"] == [0, 0, 0]
    assert f["probs"][77] == 1024 and f["probs"].sum() == 1024
    assert (f["states"] == 1 << 15).all()
    assert f["bw"][:, 0].tolist() == [4096 << 16, 4096 << 16, 1808 << 16]
    assert f["bw"][:, 1].tolist() == [0, 0, 0]
    assert len(c) == 544 + 128 * 3 + 8 * 4  # odd nb: one pad entry
    assert (oracle.stream_decode(c, 10) == x).all()


def test_one_step_per_lane():
    # 32 symbols, 16 x 'a'(97) and 16 x 'b'(98) -> pdf 256 each at p = 9,
    # cdf(a) = 0, cdf(b) = 256.  One step per lane from x = 2^15:
    # threshold 256 * 2^22 not reached, q = 2^15 / 256 = 128,
    # x = 128 * 512 + 0 + cdf -> 65536 (a) or 65792 (b).  No words.
    x = np.array([97, 98] * 16, dtype=np.uint8)
    f = parse(oracle.stream_encode(x, 9, REF))
    assert f["words"] == 0 and f["nb"] == 1
    assert f["states"][0].tolist() == [65536, 65792] * 16


def test_hand_traced_emit():
    # n = 1024 at p = 9: lane 0 gets id 2 at step 0 and id 1 at step 1; the
    # other 1022 symbols are id 3.  Counts 1, 1, 1022 -> q = 0->1, 0->1,
    # trunc(511.0) = 511; sum 513 -> id 3 (rank 0) loses one: pdf 1, 1, 510,
    # cdf(1) = 0, cdf(2) = 1, cdf(3) = 2.
    # lane 0 step 0 (id 2): x = 2^15 * 512 + 0 + 1 = 2^24 + 1
    # lane 0 step 1 (id 1): 2^24 + 1 >= 1 * 2^22 -> emit 0x0001, x = 256,
    #   then x = 256 * 512 + 0 + 0 = 131072.  No other lane ever reaches
    #   510 * 2^22.  So the block holds exactly one word, 0x0001.
    x = np.full(1024, 3, dtype=np.uint8)
    x[0], x[32] = 2, 1
    c = oracle.stream_encode(x, 9, REF)
    f = parse(c)
    assert f["probs"][1] == 1 and f["probs"][2] == 1 and f["probs"][3] == 510
    assert f["bw"][0].tolist() == [(1024 << 16) | 1, 0]
    assert f["words"] == 8
    assert f["data"].tolist() == [1, 0, 0, 0, 0, 0, 0, 0]
    # lanes 1..31 each encoded 32 x id 3 from the same start
    assert len(set(f["states"][0, 1:].tolist())) == 1
    assert (oracle.stream_decode(c, 9) == x).all()


def test_word_order_within_step_is_lane_order():
    # two emitting lanes in one step: lanes 4 and 9 get a pdf-1 symbol at
    # steps 0 and 1 (x = 2^24 + cdf after step 0); their step-1 words are
    # their low halves (cdf of the step-0 symbol), lane 4 first.
    x = np.full(4096, 5, dtype=np.uint8)
    x[4], x[9] = 1, 2          # step 0 (cdf 0 and 1)
    x[32 + 4], x[32 + 9] = 3, 3  # step 1, any rare symbol
    c = oracle.stream_encode(x, 9, REF)
    f = parse(c)
    t = oracle.normalize(oracle.histogram(x), 4096, 9)
    c1, c2 = t[1, 1], t[2, 1]
    assert t[1, 0] == 1 and t[2, 0] == 1
    w = f["data"][: f["bw"][0, 0] & 0xffff].tolist()
    assert w[:2] == [c1, c2]


def test_container_size_formula():
    rng = np.random.default_rng(1)
    for n in [0, 1, 31, 32, 33, 4095, 4096, 4097, 8192 * 3 + 5, 100_000]:
        x = rng.integers(0, 40, n).astype(np.uint8)
        for fmt in (REF, NAT):
            c = oracle.stream_encode(x, 11, fmt)
            f = parse(c, fmt)
            nb = -(-n // fmt[1])
            assert f["nb"] == nb
            padded = [((w & 0xffff) + 7) // 8 * 8 for w in f["bw"][:, 0]]
            assert f["words"] == sum(padded)
            assert f["bw"][:, 1].tolist() == list(np.cumsum([0] + padded)[:-1])
            assert len(c) == oracle.stream_overhead(nb, fmt) + 2 * f["words"]
            assert (oracle.stream_decode(c, 11) == x).all()


@pytest.mark.parametrize("fmt", [REF, NAT])
@pytest.mark.parametrize("p", [9, 10, 11])
def test_round_trip_distributions(fmt, p):
    rng = np.random.default_rng(p)
    cases = [
        rng.integers(0, 256, 50_000).astype(np.uint8),
        np.clip(128 + np.round(rng.laplace(0, 3, 70_001)), 0, 255).astype(np.uint8),
        (rng.random(30_000) < 0.001).astype(np.uint8) * 255,
        rng.integers(0, 2, 12_345).astype(np.uint8),
    ]
    for x in cases:
        c = oracle.stream_encode(x, p, fmt)
        assert (oracle.stream_decode(c, p) == x).all()


def test_worst_case_word_bound():
    # every symbol at pdf 1 (p = 11): 255 symbols at 1/2048 plus one common
    # one; lanes fed only rare symbols emit the most words.  The measured
    # count must stay within lane_max_words, which sizes the GPU LDS staging.
    for p in (9, 10, 11):
        n = 1 << 20
        x = np.zeros(n, dtype=np.uint8)
        x[: 4096 * 4] = np.tile(np.arange(1, 256, dtype=np.uint8), 4096 * 4 // 255 + 1)[: 4096 * 4]
        for fmt in (REF, NAT):
            c = oracle.stream_encode(x, p, fmt)
            f = parse(c, fmt)
            per_block = f["bw"][:, 0] & 0xffff
            bound = fmt[0] * oracle.lane_max_words(fmt[1] // fmt[0], p)
            assert per_block.max() <= bound
            assert per_block.max() > 0.8 * bound * (p - 1) / (p + 0.1)  # the bound is close to tight
            assert (oracle.stream_decode(c, p) == x).all()


def test_native_ratio_within_0p1_percent_of_reference():
    rng = np.random.default_rng(0)
    x = np.clip(32768 + np.round(rng.laplace(0, 4, 1 << 20)), 0, 65535).astype(np.uint16).view(np.uint8)
    ref = len(oracle.stream_encode(x, 10, REF))
    nat = len(oracle.stream_encode(x, 10, NAT))
    assert nat <= ref * 1.001


def test_corruption_is_detected():
    rng = np.random.default_rng(3)
    x = np.clip(128 + np.round(rng.laplace(0, 5, 20_000)), 0, 255).astype(np.uint8)
    c = oracle.stream_encode(x, 10, REF)
    bad = c.copy()
    bad[0] ^= 1  # magic
    with pytest.raises(oracle.OracleError):
        oracle.stream_decode(bad, 10)
    with pytest.raises(oracle.OracleError):
        oracle.stream_decode(c, 11)  # precision mismatch
    bad = c.copy()
    bad[parse(c)["data_off"] + 6] ^= 0x40  # a data word
    with pytest.raises(oracle.OracleError):
        oracle.stream_decode(bad, 10)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32])
@pytest.mark.parametrize("fmt", [REF, NAT])
def test_frame_planes_round_trip(dt, fmt):
    rng = np.random.default_rng(5)
    for n in [0, 1, 100, 4096, 8193, 40_000]:
        x = (rng.zipf(1.3, n) % np.iinfo(dt).max).astype(dt)
        c = oracle.frame_encode(x, oracle.MODE_PLANES, 10, fmt)
        assert (oracle.frame_decode(c) == x).all()


def test_frame_plane_is_the_strided_byte_stream():
    # plane k of the frame holds the same symbols / table as a stream of
    # byte k of each element
    rng = np.random.default_rng(6)
    x = rng.integers(0, 1 << 20, 10_000).astype(np.uint32)
    c = oracle.frame_encode(x, oracle.MODE_PLANES, 10, NAT)
    for k in range(4):
        sec = int(np.frombuffer(bytes(c[64 + 8 * k:72 + 8 * k]), dtype="<u8")[0])
        probs = np.frombuffer(bytes(c[sec + 32:sec + 544]), dtype="<u2")
        h = oracle.histogram(x, stride=4, offset=k)
        assert (probs == oracle.normalize(h, x.size, 10)[:, 0]).all()


def test_frame_adm_round_trip():
    rng = np.random.default_rng(8)
    x = np.clip(30000 + np.round(rng.laplace(0, 20, 100_000)), 0, 65535).astype(np.uint16)
    for fmt in (REF, NAT):
        c = oracle.frame_encode(x, oracle.MODE_ADM, 10, fmt)
        assert (oracle.frame_decode(c) == x).all()
    wide = x.copy()
    wide[::1000] = 65535  # d ~ 35500 -> 283 signal bits in one lane > 256
    with pytest.raises(oracle.OracleError):
        oracle.frame_encode(wide, oracle.MODE_ADM, 10, NAT)
