"""Benchmark-sized inputs through the C ABI.

* the Lorenzo quantisation codes (uint16) at a reduced 64 x 512 x 1024 shape,
  frames and the drop-in byte stream compared byte for byte with the oracle;
* the 1 GiB Lorenzo headline (1024 x 512 x 1024), the 512 MiB Zipf uint32
  input and the 512 MiB detector uint16 input (byte planes, and ADM for the
  detector frames) at full size: bit-exact round trips, and the native
  64-lane format's compressed size at most 0.1 % above the 32-lane
  reference format's on the same planes;
* a 32 MiB prefix of the Zipf and detector inputs compared byte for byte with
  the oracle in every form.
"""
import os

import numpy as np
import pytest

import mans_amd
import oracle
from mans_amd import workloads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NP = {1: np.uint8, 2: np.uint16, 4: np.uint32}
FMTS = {32: oracle.FORMAT_REF, 64: oracle.FORMAT_NATIVE}


def threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host(t) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def frame_vs_oracle(x, ws, mode, lanes, p=10):
    """Frame bytes of the HIP path against the oracle's for the same input."""
    eb = x.element_size()
    prm = mans_amd.params(p, lanes, eb, mode)
    out, size = mans_amd.frame_encode(x, prm, ws)
    got = host(out[:size])
    ref = oracle.frame_encode(host(x).view(NP[eb]), oracle.MODE_ADM if mode == mans_amd.MODE_ADM else oracle.MODE_PLANES,
                              p, FMTS[lanes], threads=threads())
    assert size == ref.size, (size, ref.size)
    assert (got == ref).all(), np.flatnonzero(got != ref)[:10]
    back = mans_amd.frame_decode(out, size, x.dtype, x.numel(), ws)
    assert torch.equal(back, x)
    ws.check()
    return size


def round_trip(x, ws, mode, lanes, p=10):
    """Full-size frame round trip; returns the frame size."""
    eb = x.element_size()
    prm = mans_amd.params(p, lanes, eb, mode)
    out, size = mans_amd.frame_encode(x, prm, ws)
    back = mans_amd.frame_decode(out, size, x.dtype, x.numel(), ws)
    ok = torch.equal(back, x)
    ws.check()
    del out, back
    assert ok, f"round trip differs (mode {mode}, {lanes} lanes)"
    return size


def native_ratio_close(x, ws, mode):
    """Native (64-lane) and reference-format (32-lane) frames of the same
    planes: round trips of both; the native frame is at most 0.1 % larger
    (it carries half as many block-word entries and block paddings, so it is
    usually a little smaller)."""
    s64 = round_trip(x, ws, mode, 64)
    s32 = round_trip(x, ws, mode, 32)
    print(f"  native {s64} B, reference format {s32} B ({100.0 * (s64 - s32) / s32:+.3f} %)", flush=True)
    assert s64 <= s32 * 1.001, (s64, s32)
    return x.numel() * x.element_size() / s64


def test_lorenzo_reduced_matches_oracle(ws):
    print("lorenzo 64x512x1024: frames 64/32 lanes, byte stream 32 lanes", flush=True)
    x = workloads.lorenzo_codes((64, 512, 1024))
    for lanes in (64, 32):
        frame_vs_oracle(x, ws, mans_amd.MODE_PLANES, lanes)
    # the drop-in path: the codes' raw bytes as one reference-format stream
    xb = x.view(torch.uint8)
    prm = mans_amd.params(10, 32)
    out, size = mans_amd.encode(xb, prm, ws)
    ref = oracle.stream_encode(host(xb), 10, oracle.FORMAT_REF, threads=threads())
    assert size == ref.size and (host(out[:size]) == ref).all()
    back = mans_amd.decode(out, xb.numel(), prm, ws, c=size)
    assert torch.equal(back, xb)
    ws.check()


def test_lorenzo_1gib_round_trip(ws):
    print("lorenzo 1024x512x1024 (1 GiB): frames, byte stream", flush=True)
    x = workloads.lorenzo_codes((1024, 512, 1024))
    assert x.numel() == 1 << 29
    ratio = native_ratio_close(x, ws, mans_amd.MODE_PLANES)
    assert 1.8 < ratio < 2.5, ratio
    # the drop-in byte stream at the reference's block limit (262,144 blocks)
    xb = x.view(torch.uint8)
    prm = mans_amd.params(10, 32)
    out, size = mans_amd.encode(xb, prm, ws)
    hdr = host(out[:32]).view(np.uint32)
    assert hdr[1] == 262_144 and hdr[2] == 1 << 30  # numBlocks, totalUncompressedWords
    back = mans_amd.decode(out, xb.numel(), prm, ws, c=size)
    assert torch.equal(back, xb)
    ws.check()


def test_zipf_u32_512mib(ws):
    print("zipf u32 512 MiB: 32 MiB prefix vs oracle, full-size round trips", flush=True)
    x = workloads.zipf_u32(1 << 27)
    pre = x[: (32 << 20) // 4]
    for lanes in (64, 32):
        frame_vs_oracle(pre, ws, mans_amd.MODE_PLANES, lanes)
    ratio = native_ratio_close(x, ws, mans_amd.MODE_PLANES)
    assert 1.2 < ratio < 1.8, ratio


def test_detector_u16_512mib(ws):
    print("detector u16 512 MiB: 32 MiB prefix vs oracle (planes, ADM), full-size round trips", flush=True)
    x = workloads.detector_u16()
    pre = x[: (32 << 20) // 2]
    for lanes in (64, 32):
        frame_vs_oracle(pre, ws, mans_amd.MODE_PLANES, lanes)
        frame_vs_oracle(pre, ws, mans_amd.MODE_ADM, lanes)
    r_planes = native_ratio_close(x, ws, mans_amd.MODE_PLANES)
    r_adm = native_ratio_close(x, ws, mans_amd.MODE_ADM)
    assert 2.0 < r_planes < 3.5 and 2.0 < r_adm < 3.5, (r_planes, r_adm)
