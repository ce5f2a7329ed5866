"""Batched encode (mans_batch_encode_*): every frame of a batch must be
byte-identical to the single-buffer frame encode of that buffer (and to the
oracle), decode back to the input, and stay identical over repeated and
graph-replayed launches."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import mans_amd  # noqa: E402

NAT = oracle.FORMAT_NATIVE
DT = {1: np.uint8, 2: np.uint16, 4: np.uint32}
VIEW = {1: np.uint8, 2: np.int16, 4: np.int32}


@pytest.fixture(scope="module")
def ws():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    w = mans_amd.Workspace(0)
    yield w
    w.close()


def gen_buffer(rng, n: int, eb: int) -> np.ndarray:
    bits = 8 * eb
    if rng.random() < 0.5:
        b = float(np.exp(rng.uniform(0, np.log(64))))
        v = np.round(rng.laplace(0, b, n)).astype(np.int64) + (1 << (bits - 1))
    else:
        v = rng.zipf(1.1, n).astype(np.int64) * 2654435761
    return (v % (1 << bits)).astype(DT[eb])


def make_batch(seed: int, sizes: list[int]):
    rng = np.random.default_rng(seed)
    xs = []
    for i, n in enumerate(sizes):
        eb = int(rng.choice([1, 2, 4], p=[0.2, 0.6, 0.2]))
        xs.append(gen_buffer(np.random.default_rng(1000 + seed * 10_000 + i), n, eb))
    ts = [torch.from_numpy(x.view(VIEW[x.itemsize])).cuda() if x.size else
          torch.empty(0, dtype=torch.from_numpy(x.view(VIEW[x.itemsize])).dtype, device="cuda") for x in xs]
    return xs, ts


EDGE_SIZES = [0, 1, 3, 15, 16, 17, 63, 64, 65, 8191, 8192, 8193, 16384, 3 * 8192 + 5,
              (1 << 18) - 1, 1 << 18, (1 << 18) + 1, 100_000, 1_000_003]


@pytest.mark.parametrize("p", [10, 11])
def test_batch_frames_match_single_frames(ws, p):
    rng = np.random.default_rng(p)
    sizes = EDGE_SIZES + [int(v) for v in np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 21), 40))]
    xs, ts = make_batch(p, sizes)
    print(f"batch p={p} buffers={len(xs)} bytes={sum(x.nbytes for x in xs)}", flush=True)
    b = mans_amd.EncodeBatch(ts, mans_amd.params(p, 64), ws)
    sizes_d = b.run()
    got_sizes = sizes_d.cpu().numpy()
    ws.check()
    for i, (x, t) in enumerate(zip(xs, ts)):
        q = mans_amd.params(p, 64, x.itemsize, mans_amd.MODE_PLANES)
        ref, ref_size = mans_amd.frame_encode(t, q, ws)
        assert got_sizes[i] == ref_size, (i, x.size, x.itemsize)
        got = b.outs[i][:ref_size].cpu().numpy()
        assert (got == ref[:ref_size].cpu().numpy()).all(), (i, x.size, x.itemsize)
        if i % 7 == 0:
            o = oracle.frame_encode(x, oracle.MODE_PLANES, p, NAT)
            assert o.size == ref_size and (o == got).all(), (i, x.size)
        back = mans_amd.frame_decode(b.outs[i], int(ref_size), t.dtype, x.size, ws)
        assert (back.cpu().numpy().view(x.dtype) == x).all(), i
    ws.check()
    # a second launch of the same plan rewrites identical frames
    for o in b.outs:
        o.zero_()
    b.run()
    torch.cuda.synchronize()
    assert (b.sizes.cpu().numpy() == got_sizes).all()
    for i, x in enumerate(xs):
        q = mans_amd.params(p, 64, x.itemsize, mans_amd.MODE_PLANES)
        back = mans_amd.frame_decode(b.outs[i], int(got_sizes[i]), ts[i].dtype, x.size, ws)
        assert (back.cpu().numpy().view(x.dtype) == x).all(), i
    b.close()


@pytest.mark.parametrize("p", [9, 10])
def test_batch_32_lane_frames_match_single_frames(ws, p):
    """The reference-format (32-lane, 4096-symbol blocks) batched form: every
    frame equals the single-buffer 32-lane frame encode (and the oracle's),
    and the batched decode restores every buffer."""
    rng = np.random.default_rng(70 + p)
    sizes = EDGE_SIZES + [4095, 4096, 4097, 12_289] + [int(v) for v in np.exp(rng.uniform(np.log(1 << 10),
                                                                                          np.log(1 << 21), 30))]
    xs, ts = make_batch(70 + p, sizes)
    prm = mans_amd.params(p, 32)
    print(f"batch 32 lanes p={p} buffers={len(xs)}", flush=True)
    b = mans_amd.EncodeBatch(ts, prm, ws)
    got_sizes = b.run().cpu().tolist()
    ws.check()
    assert b.shape[2] == 32
    for i, (x, t) in enumerate(zip(xs, ts)):
        q = mans_amd.params(p, 32, x.itemsize, mans_amd.MODE_PLANES)
        ref, n = mans_amd.frame_encode(t, q, ws)
        assert got_sizes[i] == n, (i, x.size, x.itemsize)
        assert torch.equal(b.outs[i][:n], ref[:n]), (i, x.size, x.itemsize)
        if i % 5 == 0:
            o = oracle.frame_encode(x, oracle.MODE_PLANES, p, oracle.FORMAT_REF)
            assert o.size == n and (o == b.outs[i][:n].cpu().numpy()).all(), (i, x.size)
    dec = mans_amd.DecodeBatch(b.outs, got_sizes, ts, prm, ws)
    outs = dec.run()
    torch.cuda.synchronize()
    ws.check()
    for i, x in enumerate(xs):
        assert (outs[i].cpu().numpy().view(x.dtype) == x).all(), (i, x.size, x.itemsize)
    # a 64-lane frame handed to a 32-lane decode batch is rejected, alone
    if len(xs) > 3 and xs[3].size > 0:
        q64 = mans_amd.params(p, 64, xs[3].itemsize, mans_amd.MODE_PLANES)
        f64, n64 = mans_amd.frame_encode(ts[3], q64, ws)
        frames = list(b.outs)
        fsz = list(got_sizes)
        frames[3], fsz[3] = f64, int(n64)
        bad = mans_amd.DecodeBatch(frames, fsz, ts, prm, ws)
        bad.run()
        torch.cuda.synchronize()
        with pytest.raises(mans_amd.MansError):
            ws.check()
        flags = bad.flags()
        assert flags[3] != 0 and all(f == 0 for k, f in enumerate(flags) if k != 3)
        bad.close()
    dec.close()
    b.close()


def test_batch_graph_replay(ws):
    rng = np.random.default_rng(7)
    sizes = [int(v) for v in np.exp(rng.uniform(np.log(4096), np.log(1 << 20), 64))]
    xs, ts = make_batch(7, sizes)
    b = mans_amd.EncodeBatch(ts, mans_amd.params(10, 64), ws)
    b.run()
    torch.cuda.synchronize()
    eager = [o.clone() for o in b.outs]
    eager_sizes = b.sizes.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        b.run(stream=s)
    for rep in range(3):
        for o in b.outs:
            o.zero_()
        b.sizes.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"graph replay {rep}", flush=True)
        assert torch.equal(b.sizes, eager_sizes), rep
        for i in range(len(xs)):
            n = int(eager_sizes[i])
            assert torch.equal(b.outs[i][:n], eager[i][:n]), (rep, i)
    ws.check()
    b.close()


def test_batch_decode_round_trip(ws):
    rng = np.random.default_rng(11)
    sizes = EDGE_SIZES + [int(v) for v in np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 21), 30))]
    xs, ts = make_batch(11, sizes)
    p = mans_amd.params(10, 64)
    enc = mans_amd.EncodeBatch(ts, p, ws)
    sizes_h = enc.run().cpu().tolist()
    ws.check()
    print(f"batch decode buffers={len(xs)}", flush=True)
    dec = mans_amd.DecodeBatch(enc.outs, sizes_h, ts, p, ws)
    outs = dec.run()
    torch.cuda.synchronize()
    ws.check()
    for i, x in enumerate(xs):
        assert (outs[i].cpu().numpy().view(x.dtype) == x).all(), (i, x.size, x.itemsize)
    # again, from a captured graph
    for o in dec.outs:
        o.zero_()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        dec.run(stream=s)
    for rep in range(2):
        for o in dec.outs:
            o.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i, x in enumerate(xs):
            got = dec.outs[i][:x.size].cpu().numpy().view(x.dtype)
            assert (got == x).all(), (rep, i)
    ws.check()
    dec.close()
    enc.close()


def test_batch_decode_flags_a_corrupt_frame(ws):
    xs, ts = make_batch(12, [50_000, 70_001, 9_000])
    p = mans_amd.params(10, 64)
    enc = mans_amd.EncodeBatch(ts, p, ws)
    sizes_h = enc.run().cpu().tolist()
    ws.check()
    bad = enc.outs[1].clone()
    bad[0] ^= 0xFF  # frame magic
    frames = [enc.outs[0], bad, enc.outs[2]]
    dec = mans_amd.DecodeBatch(frames, sizes_h, ts, p, ws)
    outs = dec.run()
    torch.cuda.synchronize()
    with pytest.raises(mans_amd.MansError):
        ws.check()
    flags = dec.flags()
    assert flags[0] == 0 and flags[1] != 0 and flags[2] == 0, flags
    for i in (0, 2):
        assert (outs[i].cpu().numpy().view(xs[i].dtype) == xs[i]).all(), i
    # a stream corrupted inside its data (a block's state), not its header
    bad2 = enc.outs[2].clone()
    st0 = int(np.frombuffer(bytes(bad2[64:72].cpu().numpy()), dtype="<u8")[0])  # plane 0 section
    bad2[st0 + 544: st0 + 548] = 0
    dec2 = mans_amd.DecodeBatch([enc.outs[0], enc.outs[1], bad2], sizes_h, ts, p, ws)
    dec2.run()
    flags = dec2.flags()
    assert flags[0] == 0 and flags[1] == 0 and flags[2] != 0, flags
    with pytest.raises(mans_amd.MansError):
        ws.check()
    # the flags are per run: a clean run clears them
    dec.close()
    dec = mans_amd.DecodeBatch(enc.outs, sizes_h, ts, p, ws)
    dec.run()
    assert dec.flags() == [0, 0, 0]
    ws.check()
    dec2.close()
    dec.close()
    enc.close()


def test_workspace_frozen_after_capture():
    # a frame encode captured in a graph keeps the workspace's buffers: a
    # later call that would need larger ones fails instead of freeing them,
    # and the graph still replays correctly
    ws = mans_amd.Workspace(0)
    xs, ts = make_batch(13, [300_000])
    t = ts[0]
    q = mans_amd.params(10, 64, t.element_size(), mans_amd.MODE_PLANES)
    ref, ref_size = mans_amd.frame_encode(t, q, ws)
    ref = ref[:ref_size].clone()
    out = torch.empty(mans_amd.frame_max_size(t.numel(), q), dtype=torch.uint8, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        mans_amd.frame_encode(t, q, ws, out=out, stream=s, sync_size=False, d_size=d_size)
    big = torch.zeros(1 << 26, dtype=torch.int16, device="cuda")
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.frame_encode(big, mans_amd.params(10, 64, 2, mans_amd.MODE_PLANES), ws)
    assert e.value.status == 5  # MANS_ERR_CAPACITY
    with pytest.raises(mans_amd.MansError):
        ws.reserve(1 << 30, mans_amd.params(10, 64, 2, mans_amd.MODE_PLANES))
    for rep in range(3):
        out.zero_()
        d_size.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"frozen workspace replay {rep}", flush=True)
        assert int(d_size.item()) == ref_size and torch.equal(out[:ref_size], ref), rep
    # a call that fits still runs eagerly on the frozen workspace
    again, again_size = mans_amd.frame_encode(t, q, ws)
    assert again_size == ref_size and torch.equal(again[:ref_size], ref)
    ws.check()
    del g
    ws.close()


def test_batch_archive_round_trip(ws):
    rng = np.random.default_rng(15)
    sizes = EDGE_SIZES + [int(v) for v in np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 21), 20))]
    xs, ts = make_batch(15, sizes)
    p = mans_amd.params(10, 64)
    enc = mans_amd.EncodeBatch(ts, p, ws)
    fsz = enc.run().cpu().tolist()
    arch, d_size = enc.pack()
    total = int(d_size.item())
    ws.check()
    print(f"archive of {len(xs)} frames, {total} bytes", flush=True)
    entries = mans_amd.archive_index(arch)
    assert len(entries) == len(xs)
    end = 64 + 32 * len(xs)
    for i, (off, sz, n, eb) in enumerate(entries):
        assert off == (end + 15) // 16 * 16, i
        assert sz == fsz[i] and n == xs[i].size and eb == xs[i].itemsize, i
        assert torch.equal(arch[off:off + sz], enc.outs[i][:sz]), i
        pad = (off + sz + 15) // 16 * 16 - (off + sz)
        assert int(arch[off + sz: off + sz + pad].sum()) == 0, i
        end = off + sz
    assert total == (end + 15) // 16 * 16
    dec = mans_amd.ArchiveDecodeBatch(arch, total, p, ws)
    outs = dec.run()
    torch.cuda.synchronize()
    ws.check()
    for i, x in enumerate(xs):
        assert (outs[i].cpu().numpy().view(x.dtype) == x).all(), i
    # encode + pack captured together, replayed
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    arch2 = torch.zeros_like(arch)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        enc.run(stream=s)
        enc.pack(archive=arch2, d_size=d_size, stream=s)
    for rep in range(2):
        arch2.zero_()
        d_size.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert int(d_size.item()) == total and torch.equal(arch2[:total], arch[:total]), rep
    # a corrupt index is refused
    bad = arch.clone()
    bad[64:72] = 255
    with pytest.raises(mans_amd.MansError):
        mans_amd.ArchiveDecodeBatch(bad, total, p, ws)
    dec.close()
    enc.close()


def test_batch_decode_reads_dims_from_headers(ws):
    xs, ts = make_batch(16, [0, 5, 8192, 70_001, 300_000, 1_000_003])
    p = mans_amd.params(10, 64)
    enc = mans_amd.EncodeBatch(ts, p, ws)
    fsz = enc.run().cpu().tolist()
    dec = mans_amd.DecodeBatch(enc.outs, fsz, ts, p, ws, from_headers=True)
    outs = dec.run()
    torch.cuda.synchronize()
    ws.check()
    for i, x in enumerate(xs):
        assert (outs[i].cpu().numpy().view(x.dtype) == x).all(), i
    # a frame of another precision is refused at create
    with pytest.raises(mans_amd.MansError):
        mans_amd.DecodeBatch(enc.outs, fsz, ts, mans_amd.params(11, 64), ws, from_headers=True)
    dec.close()
    enc.close()


SHAPES = [(64 * nw, kb) for nw in (4, 8, 16) for kb in (4, 8, 16)]
DEC_SHAPES = [(64 * nw, kb) for nw in (4, 8, 16) for kb in (1, 4, 8, 16)]


def lds_fits(p: int, wg: int, kb: int, decode: bool) -> bool:
    nw = wg // 64
    if decode:
        h = max(512, 64 * kb)
        return nw * (2 * (1 << p) * 4 + 4 * (2 * h + 64)) <= 160 << 10
    return nw * (2 * 256 * 4 + 1024 + 64) * 4 <= 160 << 10


@pytest.mark.parametrize("p", [9, 10, 11])
def test_batch_every_shape_is_bit_identical(ws, p):
    rng = np.random.default_rng(30 + p)
    sizes = EDGE_SIZES + [int(v) for v in np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 21), 24))]
    xs, ts = make_batch(30 + p, sizes)
    prm = mans_amd.params(p, 64)
    enc = mans_amd.EncodeBatch(ts, prm, ws)
    with pytest.raises(mans_amd.MansError):
        enc.set_shape(256, 1)  # one symbol per lane is a decode shape only
    enc.set_shape(256, 16)
    ref_sizes = enc.run().cpu().tolist()
    ws.check()
    ref = [o[:n].clone() for o, n in zip(enc.outs, ref_sizes)]
    dec = mans_amd.DecodeBatch(ref, ref_sizes, ts, prm, ws)
    for dirn, b, shapes in (("encode", enc, SHAPES), ("decode", dec, DEC_SHAPES)):
        for wg, kb in shapes:
            fits = lds_fits(p, wg, kb, dirn == "decode")
            if not fits:
                with pytest.raises(mans_amd.MansError):
                    b.set_shape(wg, kb)
                continue
            print(f"p={p} {dirn} shape wg={wg} kb={kb}", flush=True)
            b.set_shape(wg, kb)
            assert b.shape == (wg, kb, 64)
            for o in b.outs:
                o.zero_()
            b.run()
            torch.cuda.synchronize()
            ws.check()
            if dirn == "encode":
                assert b.sizes.cpu().tolist()[:len(xs)] == ref_sizes
                for i in range(len(xs)):
                    assert torch.equal(b.outs[i][:ref_sizes[i]], ref[i]), (wg, kb, i)
            else:
                for i, x in enumerate(xs):
                    got = b.outs[i][:x.size].cpu().numpy().view(x.dtype)
                    assert (got == x).all(), (wg, kb, i, x.size, x.itemsize)
    dec.close()
    enc.close()


GRAINS = [(1 << 16, 1, 1, 64), (1 << 17, 2, 4, 64), (1 << 18, 4, 1, 64), (1 << 19, 4, 4, 64), (1 << 19, 1, 4, 64),
          (1 << 16, 2, 1, 64), (1 << 16, 1, 4, 32), (1 << 19, 4, 1, 32)]


@pytest.mark.parametrize("grain", GRAINS, ids=lambda g: f"chunk{g[0] >> 10}k-r{g[1]}-run{g[2]}-l{g[3]}")
def test_batch_every_grain_is_bit_identical(ws, grain):
    """Every work-item grain (histogram chunk, encode ticket rounds, decode
    run) gives the frames of the single-buffer encode, and decodes back."""
    rng = np.random.default_rng(50)
    sizes = EDGE_SIZES + [int(v) for v in np.exp(rng.uniform(np.log(1 << 10), np.log(3 << 20), 24))]
    xs, ts = make_batch(50, sizes)
    lanes = grain[3]
    grain = grain[:3]
    prm = mans_amd.params(10, lanes)
    print(f"grain {grain} buffers={len(xs)}", flush=True)
    mans_amd.set_batch_grain(*grain)
    try:
        enc = mans_amd.EncodeBatch(ts, prm, ws)
        got_sizes = enc.run().cpu().tolist()
        ws.check()
        for i, (x, t) in enumerate(zip(xs, ts)):
            q = mans_amd.params(10, lanes, x.itemsize, mans_amd.MODE_PLANES)
            ref, n = mans_amd.frame_encode(t, q, ws)
            assert got_sizes[i] == n, (grain, i, x.size, x.itemsize)
            assert torch.equal(enc.outs[i][:n], ref[:n]), (grain, i, x.size, x.itemsize)
        dec = mans_amd.DecodeBatch(enc.outs, got_sizes, ts, prm, ws)
        outs = dec.run()
        torch.cuda.synchronize()
        ws.check()
        for i, x in enumerate(xs):
            assert (outs[i].cpu().numpy().view(x.dtype) == x).all(), (grain, i, x.size, x.itemsize)
        dec.close()
        enc.close()
    finally:
        mans_amd.set_batch_grain(0, 0, 0)


def test_batch_grain_rejects_bad_values():
    for g in [(1 << 15, 0, 0), (3 << 16, 0, 0), (1 << 20, 0, 0), (0, 3, 0), (0, 8, 0), (0, 0, 2)]:
        with pytest.raises(mans_amd.MansError):
            mans_amd.set_batch_grain(*g)
    mans_amd.set_batch_grain(0, 0, 0)


def test_batch_tune_picks_the_fastest_and_caches_it(ws):
    rng = np.random.default_rng(40)
    sizes = [int(v) for v in np.exp(rng.uniform(np.log(4096), np.log(1 << 21), 96))]
    xs, ts = make_batch(40, sizes)
    prm = mans_amd.params(10, 64)
    enc = mans_amd.EncodeBatch(ts, prm, ws)
    rows = enc.tune(reps=2)
    ws.check()
    assert [(r["wg_threads"], r["syms_per_lane"]) for r in rows] == SHAPES
    timed = [r for r in rows if r["ms"] >= 0]
    assert timed and all(lds_fits(10, r["wg_threads"], r["syms_per_lane"], False) for r in timed)
    best = min(timed, key=lambda r: r["ms"])
    assert enc.shape == (best["wg_threads"], best["syms_per_lane"], 64)
    sizes_h = enc.run().cpu().tolist()
    ws.check()
    for i, (x, t) in enumerate(zip(xs, ts)):
        q = mans_amd.params(10, 64, x.itemsize, mans_amd.MODE_PLANES)
        one, n = mans_amd.frame_encode(t, q, ws)
        assert sizes_h[i] == n and torch.equal(enc.outs[i][:n], one[:n]), i
    # a new batch with the same key starts from the tuned shape
    again = mans_amd.EncodeBatch(ts, prm, ws)
    assert again.shape == enc.shape
    again.close()
    dec = mans_amd.DecodeBatch(enc.outs, sizes_h, ts, prm, ws)
    drows = dec.tune(reps=2)
    ws.check()
    assert [(r["wg_threads"], r["syms_per_lane"]) for r in drows] == DEC_SHAPES
    dbest = min((r for r in drows if r["ms"] >= 0), key=lambda r: r["ms"])
    assert dec.shape == (dbest["wg_threads"], dbest["syms_per_lane"], 64)
    for o in dec.outs:
        o.zero_()
    dec.run()
    torch.cuda.synchronize()
    ws.check()
    for i, x in enumerate(xs):
        assert (dec.outs[i][:x.size].cpu().numpy().view(x.dtype) == x).all(), i
    dec.close()
    enc.close()


PROBE = r'''
import sys
sys.path.insert(0, sys.argv[2])
import torch
import mans_amd
g = torch.Generator(device="cuda")
g.manual_seed(5)
ts = [torch.randint(0, 60, (n,), generator=g, device="cuda", dtype=torch.int16) for n in (30000, 70000, 250000)]
ws = mans_amd.Workspace(0)
enc = mans_amd.EncodeBatch(ts, mans_amd.params(10, 64), ws)
if sys.argv[1] == "tune":
    enc.tune(reps=1)
enc.run()
torch.cuda.synchronize()
ws.check()
print("shape", enc.shape, flush=True)
'''


def test_tune_cache_file(ws, tmp_path):
    # a tuned shape goes to the MANS_TUNE_CACHE file; a new process creating
    # a batch with the same key starts from the shape the file holds
    import os
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cache = tmp_path / "tune.txt"
    script = tmp_path / "probe.py"
    script.write_text(PROBE)
    env = dict(os.environ, MANS_TUNE_CACHE=str(cache))
    print("tune cache: tuning process", flush=True)
    r = subprocess.run([sys.executable, str(script), "tune", repo], env=env, capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    fields = cache.read_text().split()
    assert len(fields) == 3 and fields[0].startswith("enc-p10-"), fields
    cache.write_text(f"{fields[0]} 8 4\n")  # 512 threads, 4 symbols per lane: not the default
    print("tune cache: reading process", flush=True)
    r = subprocess.run([sys.executable, str(script), "read", repo], env=env, capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1] == "shape (512, 4, 64)", r.stdout
