"""mans_encode_noalloc (no allocation, library-held device scratch) against
mans_encode, through the C ABI: byte-identical containers on consecutive
calls (the second smaller than the first, so stale scratch from the larger
call would show), round trips, and the entry point's argument errors."""
import numpy as np
import pytest

import mans_amd
import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ws():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    w = mans_amd.Workspace(0)
    yield w
    w.close()


def laplace_bytes(n: int, b: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = np.clip(32768 + np.round(rng.laplace(0, b, (n + 1) // 2)), 0, 65535).astype(np.uint16)
    return v.view(np.uint8)[:n].copy()


@pytest.mark.parametrize("p", [9, 10, 11])
def test_noalloc_matches_encode(ws, p):
    prm = mans_amd.params(p, 32)
    # large then smaller: the second call must not see the first's scratch
    for n, b, seed in ((3_000_017, 6.0, p), (70_001, 2.0, p + 10), (0, 1.0, 0), (4096 * 9, 30.0, p + 20)):
        x = laplace_bytes(n, b, seed)
        t = torch.from_numpy(x).cuda() if n else torch.empty(0, dtype=torch.uint8, device="cuda")
        print(f"noalloc p={p} n={n}", flush=True)
        ref, ref_size = mans_amd.encode(t, prm, ws)
        out = torch.full((mans_amd.max_compressed_size(n, prm),), 0xA5, dtype=torch.uint8, device="cuda")
        d_size = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        mans_amd.encode_noalloc(t, prm, out, d_size)
        torch.cuda.synchronize()
        size = int(d_size.item())
        assert size == ref_size, (size, ref_size)
        assert torch.equal(out[:size], ref[:size])
        if n and n < 200_000:
            o = oracle.stream_encode(x, p, oracle.FORMAT_REF)
            assert o.size == size and (out[:size].cpu().numpy() == o).all()
        if n:
            back = mans_amd.decode(out, n, prm, ws, c=size)
            assert torch.equal(back, t)
    ws.check()


def test_noalloc_repeated_calls_on_a_stream(ws):
    # calls queued back to back on one stream reuse the scratch in order
    prm = mans_amd.params(10, 32)
    xs = [laplace_bytes(n, 3.0 + i, 40 + i) for i, n in enumerate((1 << 20, 12_345, 1 << 20, 5))]
    ts = [torch.from_numpy(x).cuda() for x in xs]
    outs = [torch.empty(mans_amd.max_compressed_size(x.size, prm), dtype=torch.uint8, device="cuda") for x in xs]
    sizes = [torch.zeros(1, dtype=torch.int32, device="cuda") for _ in xs]
    for t, o, s in zip(ts, outs, sizes):
        mans_amd.encode_noalloc(t, prm, o, s)
    torch.cuda.synchronize()
    for t, o, s in zip(ts, outs, sizes):
        ref, ref_size = mans_amd.encode(t, prm, ws)
        assert int(s.item()) == ref_size and torch.equal(o[:ref_size], ref[:ref_size])


def test_noalloc_argument_errors():
    prm = mans_amd.params(10, 32)
    t = torch.zeros(64, dtype=torch.uint8, device="cuda")
    out = torch.empty(mans_amd.max_compressed_size(64, prm), dtype=torch.uint8, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int32, device="cuda")
    L = mans_amd.lib()
    import ctypes as C

    def call(p, n, cap):
        return L.mans_encode_noalloc(C.byref(p), t.data_ptr(), n, out.data_ptr(), cap, d_size.data_ptr(), None)

    assert call(mans_amd.params(10, 64), 64, out.numel()) == 1       # native lanes: not this entry point
    assert call(mans_amd.params(12, 32), 64, out.numel()) == 2       # precision
    assert call(prm, (1 << 30) + 1, 1 << 40) == 4                   # beyond the scratch's limit
    assert call(prm, 64, 16) == 5                                    # output too small
    assert L.mans_encode_noalloc(C.byref(prm), t.data_ptr(), 64, out.data_ptr(), out.numel(), None, None) == 1
