"""HIP kernels (through the C ABI) against the CPU oracle, bit for bit.

Sizes are small enough for the oracle to finish in seconds; the 1 GiB
workloads are checked by round trip in test_gpu_large.py.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import mans_amd
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

REF, NAT = oracle.FORMAT_REF, oracle.FORMAT_NATIVE
FMTS = {32: REF, 64: NAT}


def dev(a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def u16_dev(x: np.ndarray):
    return dev(x.astype(np.uint16).view(np.int16))


def datasets(seed=0):
    rng = np.random.default_rng(seed)
    lap = np.clip(32768 + np.round(rng.laplace(0, 4, 300_000)), 0, 65535).astype(np.uint16)
    return {
        "uniform": rng.integers(0, 256, 200_003).astype(np.uint8),
        "laplace_bytes": lap.view(np.uint8).copy(),
        "skewed": ((rng.random(150_000) < 0.002) * rng.integers(1, 256, 150_000)).astype(np.uint8),
        "single": np.full(70_000, 9, dtype=np.uint8),
        "two": rng.integers(0, 2, 33_333).astype(np.uint8),
    }


# ------------------------------------------------------------------ A1
@pytest.mark.parametrize("stride", [1, 2, 4])
def test_histogram(ws, stride):
    rng = np.random.default_rng(stride)
    # (a workgroup counts 32 KiB rows of 2048 uint4s, taken in chunks of 4
    # rows: sizes on, around and between those boundaries)
    for n in [0, 1, 15, 16, 17, 1000, 8192 // stride, 32_768 // stride, 32_768 // stride + 9, 65_536 + 7,
              131_072 // stride, 131_072 * 3 // stride + 5, 8192 * 43 // stride, 1_000_003]:
        raw = rng.integers(0, 256, n * stride + 64).astype(np.uint8)
        raw[: n * stride: 3] = 7  # a hot bin
        for off in ([0, 4, 8] if stride > 1 else [0, 1, 5, 15]):
            d = dev(raw)[off: off + n * stride]
            print(f"hist stride={stride} n={n} off={off}", flush=True)
            h = host(mans_amd.histogram(d, ws, stride=stride)).astype(np.uint32)
            ref = oracle.histogram(raw[off: off + n * stride], stride=stride)
            assert (h == ref).all(), (n, off)


# ------------------------------------------------------------------ A2
@pytest.mark.parametrize("p", [9, 10, 11])
def test_normalize(ws, p):
    rng = np.random.default_rng(100 + p)
    hists = []
    for trial in range(60):
        k = int(rng.integers(1, 257))
        h = np.zeros(256, dtype=np.uint32)
        h[rng.choice(256, size=k, replace=False)] = rng.integers(1, 10 ** int(rng.integers(1, 7)), size=k)
        hists.append(h)
    h = np.zeros(256, dtype=np.uint32)
    h[[3, 7, 9]] = [600_000, 300_000, 99_000]
    h[200:255] = 1
    hists.append(h)
    h = np.zeros(256, dtype=np.uint32)
    h[[10, 20]] = 500_000
    h[100:199] = 1
    hists.append(h)
    hists.append(np.zeros(256, dtype=np.uint32))
    for h in hists:
        total = int(h.sum())
        t = host(mans_amd.normalize(dev(h.view(np.int32)), total, p, ws)).view(np.uint32)
        r = oracle.normalize(h, total, p)
        assert (t[:, :2] == r[:, :2]).all()
        used = r[:, 0] > 0
        assert (t[used, 2:] == r[used, 2:]).all()


# ------------------------------------------------------------------ A3-A8
@pytest.mark.parametrize("lanes", [32, 64])
@pytest.mark.parametrize("p", [9, 10, 11])
def test_encode_matches_oracle_bytes(ws, lanes, p):
    fmt = FMTS[lanes]
    prm = mans_amd.params(p, lanes)
    for name, x in datasets(p).items():
        for n in sorted({0, 1, 31, 33, 4095, 4096, 4097, 8193, x.size}):
            xs = x[:n]
            print(f"encode {name} n={n} lanes={lanes} p={p}", flush=True)
            out, size = mans_amd.encode(dev(xs) if n else torch.empty(0, dtype=torch.uint8, device="cuda"), prm, ws)
            got = host(out[:size])
            ref = oracle.stream_encode(xs, p, fmt)
            assert size == ref.size, (name, n)
            assert (got == ref).all(), (name, n, np.flatnonzero(got != ref)[:10])
            ws.check()


@pytest.mark.parametrize("lanes", [32, 64])
@pytest.mark.parametrize("p", [9, 10, 11])
def test_decode_oracle_streams(ws, lanes, p):
    fmt = FMTS[lanes]
    prm = mans_amd.params(p, lanes)
    for name, x in datasets(10 + p).items():
        for n in sorted({1, 32, 4097, 8191, x.size}):
            xs = x[:n]
            c = oracle.stream_encode(xs, p, fmt)
            print(f"decode {name} n={n} lanes={lanes} p={p}", flush=True)
            got = host(mans_amd.decode(dev(c), n, prm, ws, c=c.size))
            assert (got == xs).all(), (name, n)
            ws.check()


def test_decode_without_sizes_like_reference(ws):
    # the reference's ansDecodeBatch gets no sizes: c = 0, capacity unchecked
    x = datasets(3)["laplace_bytes"]
    c = oracle.stream_encode(x, 10, REF)
    got = host(mans_amd.decode(dev(c), x.size, mans_amd.params(10, 32), ws, c=0))
    assert (got == x).all()


def test_encode_is_deterministic(ws):
    x = dev(datasets(4)["uniform"])
    prm = mans_amd.params(11, 64)
    a, sa = mans_amd.encode(x, prm, ws)
    a = host(a[:sa]).copy()
    for _ in range(3):
        b, sb = mans_amd.encode(x, prm, ws)
        assert sb == sa and (host(b[:sb]) == a).all()


def test_device_size_output(ws):
    x = dev(datasets(5)["laplace_bytes"])
    out, d_size = mans_amd.encode(x, mans_amd.params(10, 32), ws, sync_size=False)
    n = int(host(d_size)[0])
    ref = oracle.stream_encode(host(x), 10, REF)
    assert n == ref.size


def test_corrupt_streams_are_flagged(ws):
    x = datasets(6)["laplace_bytes"][:50_000]
    c = oracle.stream_encode(x, 10, REF)
    prm = mans_amd.params(10, 32)
    bad = c.copy()
    bad[2] ^= 1  # magic
    mans_amd.decode(dev(bad), x.size, prm, ws, c=bad.size)
    with pytest.raises(mans_amd.MansError):
        ws.check()
    mans_amd.decode(dev(c), x.size, mans_amd.params(11, 32), ws, c=c.size)  # precision
    with pytest.raises(mans_amd.MansError):
        ws.check()
    nb = -(-x.size // 4096)
    data_off = 544 + 128 * nb + 8 * (nb + (nb & 1))
    bad = c.copy()
    bad[data_off + 100] ^= 0x10
    mans_amd.decode(dev(bad), x.size, prm, ws, c=bad.size)
    with pytest.raises(mans_amd.MansError):
        ws.check()
    bad = c.copy()
    bad[544 + 128 * nb + 4: 544 + 128 * nb + 8] = 255  # block 0 offset far out
    mans_amd.decode(dev(bad), x.size, prm, ws, c=bad.size)
    with pytest.raises(mans_amd.MansError):
        ws.check()
    # a good stream after the bad ones still decodes
    got = host(mans_amd.decode(dev(c), x.size, prm, ws, c=c.size))
    assert (got == x).all()
    ws.check()


def test_argument_errors(ws):
    x = dev(np.zeros(100, dtype=np.uint8))
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.encode(x, mans_amd.params(12, 32), ws)
    assert e.value.status == 2
    out = torch.empty(100_000, dtype=torch.uint8, device="cuda")
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.encode(x, mans_amd.params(10, 32), ws, out=out[1:])
    assert e.value.status == 3
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.encode(x, mans_amd.params(10, 32), ws, out=out[:100])
    assert e.value.status == 5


# ------------------------------------------------------------------ M1-M3
def golden():
    import glob
    return sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "adm_v1_*.npz")))


@pytest.mark.parametrize("path", golden(), ids=lambda p: os.path.basename(p)[7:-4])
def test_adm_map_matches_reference_vectors(ws, path):
    z = np.load(path)
    x = z["input"]
    print(f"adm map {os.path.basename(path)}", flush=True)
    ok, offs, cen, codes, bits = mans_amd.adm_map(u16_dev(x), ws)
    g = z["offsets"].size
    assert ok
    assert (host(offs)[:g] == z["offsets"]).all()
    assert (host(cen).view(np.uint16) == z["centers"]).all()
    assert (host(codes) == z["codes"]).all()
    b = host(bits)
    assert b.size == z["bits"].size and (b == z["bits"]).all()
    back = host(mans_amd.adm_unmap(offs, cen, codes, bits, x.size, ws)).view(np.uint16)
    assert (back == x).all()


def test_adm_map_matches_oracle_random(ws):
    rng = np.random.default_rng(21)
    for n in [1, 16, 17, 511, 512, 513, 100_000, 1_000_001]:
        x = np.clip(20000 + np.round(rng.laplace(0, 60, n)), 0, 65535).astype(np.uint16)
        print(f"adm random n={n}", flush=True)
        ok, offs, cen, codes, bits = mans_amd.adm_map(u16_dev(x), ws)
        rok, roffs, rcen, rcodes, rbits = oracle.adm_map(x)
        assert ok == rok
        assert (host(offs) == roffs).all() and (host(cen).view(np.uint16) == rcen).all()
        assert (host(codes) == rcodes).all() and (host(bits) == rbits).all()


def test_adm_map_multiword_lanes_match_oracle(ws):
    # wide spreads: a lane's 16 elements take up to ~200 bits, so its bit
    # string spans several 32-bit words (and pads inside a later word); whole
    # and ragged 16-group tickets
    rng = np.random.default_rng(22)
    # (b = 1200 leaves some lanes beyond 256 bits: only the verdict is
    # defined there)
    for n, b in [(3 * 8192, 300.0), (5 * 8192 + 1000, 600.0), (8192 * 40 + 17, 700.0), (8192 * 8, 1200.0)]:
        x = np.clip(32768 + np.round(rng.laplace(0, b, n)), 0, 65535).astype(np.uint16)
        print(f"adm multiword n={n} b={b}", flush=True)
        ok, offs, cen, codes, bits = mans_amd.adm_map(u16_dev(x), ws)
        rok, roffs, rcen, rcodes, rbits = oracle.adm_map(x)
        assert ok == rok
        if b > 1000:
            assert not ok
            continue
        assert ok
        assert (host(offs) == roffs).all() and (host(cen).view(np.uint16) == rcen).all()
        assert (host(codes) == rcodes).all() and (host(bits) == rbits).all()
        back = host(mans_amd.adm_unmap(offs, cen, codes, bits, n, ws)).view(np.uint16)
        assert (back == x).all()


def test_adm_domain_flag(ws):
    x = np.full(4096, 20000, dtype=np.uint16)
    x[1000] = 65535
    ok = mans_amd.adm_map(u16_dev(x), ws)[0]
    assert ok is False


# ------------------------------------------------------------------ frames
@pytest.mark.parametrize("lanes", [32, 64])
@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32])
def test_frame_planes_match_oracle(ws, lanes, dt):
    rng = np.random.default_rng(31)
    fmt = FMTS[lanes]
    for n in [0, 1, 4097, 100_001]:
        x = (rng.zipf(1.2, n) % np.iinfo(dt).max).astype(dt)
        print(f"frame planes dt={np.dtype(dt).name} n={n} lanes={lanes}", flush=True)
        xt = dev(x.view({1: np.uint8, 2: np.int16, 4: np.int32}[x.itemsize]))
        out, size = mans_amd.frame_encode(xt, mans_amd.params(10, lanes, x.itemsize, mans_amd.MODE_PLANES), ws)
        ref = oracle.frame_encode(x, oracle.MODE_PLANES, 10, fmt)
        got = host(out[:size])
        assert size == ref.size and (got == ref).all(), (n, np.flatnonzero(got != ref)[:8])
        back = host(mans_amd.frame_decode(out, size, xt.dtype, n, ws))
        ws.check()
        assert (back.view(x.dtype) == x).all()


@pytest.mark.parametrize("dt,n", [(np.uint16, 40_000_003), (np.uint32, 12_000_001)])
def test_frame_large_mixed_entropy(ws, dt, n):
    """Many units per wavefront: blocks alternate between low-entropy
    (Laplace) and near-uniform bytes, so the 64-lane coder's LDS rings spill
    and refill on every chain, and look-back spans more than 64 tickets."""
    rng = np.random.default_rng(41)
    bits = 8 * np.dtype(dt).itemsize
    lap = np.round(rng.laplace(0, 3, n)).astype(np.int64) + (1 << (bits - 1))
    uni = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.int64)
    block = (np.arange(n) // 8192) % 3 == 0
    x = np.where(block, uni, lap).astype(dt)
    print(f"frame large dt={np.dtype(dt).name} n={n}", flush=True)
    xt = dev(x.view({2: np.int16, 4: np.int32}[x.itemsize]))
    prm = mans_amd.params(10, 64, x.itemsize, mans_amd.MODE_PLANES)
    out, size = mans_amd.frame_encode(xt, prm, ws)
    ref = oracle.frame_encode(x, oracle.MODE_PLANES, 10, NAT, threads=8)
    got = host(out[:size])
    assert size == ref.size and (got == ref).all(), np.flatnonzero(got != ref)[:8]
    back = host(mans_amd.frame_decode(out, size, xt.dtype, n, ws))
    ws.check()
    assert (back.view(x.dtype) == x).all()


@pytest.mark.parametrize("lanes", [32, 64])
def test_frame_adm_matches_oracle(ws, lanes):
    rng = np.random.default_rng(41)
    fmt = FMTS[lanes]
    for n in [1, 600, 65_536, 300_001]:
        x = np.clip(32768 + np.round(rng.laplace(0, 4, n)), 0, 65535).astype(np.uint16)
        print(f"frame adm n={n} lanes={lanes}", flush=True)
        xt = u16_dev(x)
        out, size = mans_amd.frame_encode(xt, mans_amd.params(10, lanes, 2, mans_amd.MODE_ADM), ws)
        ref = oracle.frame_encode(x, oracle.MODE_ADM, 10, fmt)
        got = host(out[:size])
        assert size == ref.size and (got == ref).all(), (n, np.flatnonzero(got != ref)[:8])
        back = host(mans_amd.frame_decode(out, size, torch.int16, n, ws)).view(np.uint16)
        ws.check()
        assert (back == x).all()


@pytest.mark.parametrize("p", [9, 11])
def test_frame_adm_long_signals(ws, p):
    # wide spread: |x - c| beyond 126 gives unary signals of 2-3 bits, and
    # several 8192-element blocks end in a ragged group
    rng = np.random.default_rng(60 + p)
    for n in [5 * 8192 + 777, 16 * 8192]:
        x = np.clip(32768 + np.round(rng.laplace(0, 90, n)), 0, 65535).astype(np.uint16)
        print(f"frame adm long signals n={n} p={p}", flush=True)
        out, size = mans_amd.frame_encode(u16_dev(x), mans_amd.params(p, 64, 2, mans_amd.MODE_ADM), ws)
        ref = oracle.frame_encode(x, oracle.MODE_ADM, p, FMTS[64])
        got = host(out[:size])
        assert size == ref.size and (got == ref).all()
        back = host(mans_amd.frame_decode(out, size, torch.int16, n, ws)).view(np.uint16)
        ws.check()
        assert (back == x).all()


def test_frame_adm_rejects_bad_offsets(ws):
    rng = np.random.default_rng(62)
    n = 3 * 8192
    x = np.clip(32768 + np.round(rng.laplace(0, 4, n)), 0, 65535).astype(np.uint16)
    out, size = mans_amd.frame_encode(u16_dev(x), mans_amd.params(10, 64, 2, mans_amd.MODE_ADM), ws)
    frame = host(out[:size]).copy()
    meta = int(frame[32:40].view(np.uint64)[0])
    offs = frame[meta:meta + 8].view(np.int32)
    offs[1] = offs[0] + 1000  # a group longer than 32 bytes per lane
    print("frame adm with a corrupt offset table", flush=True)
    bad = torch.from_numpy(frame).cuda()
    mans_amd.frame_decode(bad, size, torch.int16, n, ws)
    with pytest.raises(mans_amd.MansError):
        ws.check()


@pytest.mark.parametrize("p", [9, 10, 11])
def test_frame_adm_fused_decode(ws, p):
    # mans_frame_decode_fused: the restore inside the code plane's decode
    # launch gives the elements the oracle's frame holds, on ragged sizes
    # (a partial last lane, group and block) and on wide signals
    rng = np.random.default_rng(80 + p)
    for n, spread in [(1, 4), (513, 4), (8191, 4), (8192, 4), (8193, 30), (5 * 8192 + 777, 90), (300_001, 4)]:
        x = np.clip(32768 + np.round(rng.laplace(0, spread, n)), 0, 65535).astype(np.uint16)
        print(f"frame adm fused decode n={n} p={p}", flush=True)
        out, size = mans_amd.frame_encode(u16_dev(x), mans_amd.params(p, 64, 2, mans_amd.MODE_ADM), ws)
        ref = oracle.frame_encode(x, oracle.MODE_ADM, p, FMTS[64])
        assert size == ref.size and (host(out[:size]) == ref).all()
        back = host(mans_amd.frame_decode(out, size, torch.int16, n, ws, fused=True)).view(np.uint16)
        ws.check()
        assert (back == x).all(), (n, np.flatnonzero(back != x)[:8])


def test_frame_fused_decode_other_frames(ws):
    # byte-plane frames and 32-lane ADM frames take the ordinary path
    rng = np.random.default_rng(84)
    x = np.clip(32768 + np.round(rng.laplace(0, 4, 70_001)), 0, 65535).astype(np.uint16)
    for lanes, mode in ((64, mans_amd.MODE_PLANES), (32, mans_amd.MODE_ADM), (32, mans_amd.MODE_PLANES)):
        print(f"fused decode entry, lanes={lanes} mode={mode}", flush=True)
        out, size = mans_amd.frame_encode(u16_dev(x), mans_amd.params(10, lanes, 2, mode), ws)
        back = host(mans_amd.frame_decode(out, size, torch.int16, x.size, ws, fused=True)).view(np.uint16)
        ws.check()
        assert (back == x).all()


def test_frame_adm_fused_rejects_bad_offsets(ws):
    rng = np.random.default_rng(63)
    n = 3 * 8192
    x = np.clip(32768 + np.round(rng.laplace(0, 4, n)), 0, 65535).astype(np.uint16)
    out, size = mans_amd.frame_encode(u16_dev(x), mans_amd.params(10, 64, 2, mans_amd.MODE_ADM), ws)
    frame = host(out[:size]).copy()
    meta = int(frame[32:40].view(np.uint64)[0])
    offs = frame[meta:meta + 8].view(np.int32)
    offs[1] = offs[0] + 1000  # a group longer than 32 bytes per lane
    print("fused decode of a frame with a corrupt offset table", flush=True)
    bad = torch.from_numpy(frame).cuda()
    mans_amd.frame_decode(bad, size, torch.int16, n, ws, fused=True)
    with pytest.raises(mans_amd.MansError):
        ws.check()


def test_frame_auto_falls_back_outside_domain(ws):
    rng = np.random.default_rng(51)
    x = np.clip(32768 + np.round(rng.laplace(0, 4, 50_000)), 0, 65535).astype(np.uint16)
    x[::777] = 0  # out of the ADM domain
    out, size = mans_amd.frame_encode(u16_dev(x), mans_amd.params(10, 64, 2, mans_amd.MODE_AUTO), ws)
    got = host(out[:size])
    assert got[7] == oracle.MODE_PLANES
    assert (got == oracle.frame_encode(x, oracle.MODE_PLANES, 10, NAT)).all()
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.frame_encode(u16_dev(x), mans_amd.params(10, 64, 2, mans_amd.MODE_ADM), ws)
    assert e.value.status == 9


# ------------------------------------------------------------------ D1
def test_compress_cli_round_trip():
    rng = np.random.default_rng(61)
    x = np.clip(32768 + np.round(rng.laplace(0, 4, 1 << 20)), 0, 65535).astype(np.uint16)
    with tempfile.TemporaryDirectory() as td:
        src, tmp, dst = (os.path.join(td, f) for f in ("in.u2", "c.ans", "out.u2"))
        x.tofile(src)
        r = subprocess.run([mans_amd.CLI_PATH, src, tmp, dst], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        assert "compress ratio:" in r.stdout and "comp   time" in r.stdout and "decomp time" in r.stdout
        assert (np.fromfile(dst, dtype=np.uint16) == x).all()
        assert (np.fromfile(tmp, dtype=np.uint8) == oracle.stream_encode(x.view(np.uint8), 10, REF)).all()
        r = subprocess.run([mans_amd.CLI_PATH, "--frame", "--dtype", "u16", "--mode", "auto", src, tmp, dst],
                           capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        assert (np.fromfile(dst, dtype=np.uint16) == x).all()
        assert (oracle.frame_decode(np.fromfile(tmp, dtype=np.uint8)) == x).all()


# ------------------------------------------------------------------ multi-input
@pytest.mark.parametrize("lanes", [32, 64])
def test_multi_input_streams_match_single(ws, lanes):
    # DietGPU-style pointer arrays: every container equals mans_encode of
    # that input alone (and, for the reference format, the oracle's)
    rng = np.random.default_rng(50 + lanes)
    xs = [rng.integers(0, 256, n).astype(np.uint8) for n in (0, 1, 4095, 4096, 70_001)]
    xs += [datasets(51)["laplace_bytes"], datasets(52)["skewed"]]
    prm = mans_amd.params(10, lanes)
    ts = [dev(x) if x.size else torch.empty(0, dtype=torch.uint8, device="cuda") for x in xs]
    outs, sizes = mans_amd.encode_multi(ts, prm, ws)
    sizes = host(sizes).tolist()
    for i, (x, t) in enumerate(zip(xs, ts)):
        ref, ref_size = mans_amd.encode(t, prm, ws)
        assert sizes[i] == ref_size and (host(outs[i][:ref_size]) == host(ref[:ref_size])).all(), i
        if lanes == 32:
            o = oracle.stream_encode(x, 10, REF)
            assert o.size == ref_size and (host(outs[i][:ref_size]) == o).all(), i
    backs = mans_amd.decode_multi(outs, [x.size for x in xs], prm, ws)
    for i, x in enumerate(xs):
        assert (host(backs[i])[:x.size] == x).all(), i
    ws.check()
