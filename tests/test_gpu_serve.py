"""Request server (mans_server_*): every result must equal the same request
run alone (mans_frame_encode / mans_frame_decode of that buffer), whatever it
was batched with, including when admission waits for arena memory."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import mans_amd  # noqa: E402

VIEW = {1: np.uint8, 2: np.int16, 4: np.int32}


@pytest.fixture(scope="module")
def ws():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    w = mans_amd.Workspace(0)
    yield w
    w.close()


def requests(seed: int, count: int, lo: int, hi: int, ws):
    """(kind, input tensor, frame tensor, frame bytes, elements tensor)"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        eb = int(rng.choice([1, 2, 4], p=[0.2, 0.6, 0.2]))
        n = max(1, int(np.exp(rng.uniform(np.log(lo), np.log(hi)))) // eb)
        bits = 8 * eb
        v = np.round(rng.laplace(0, float(np.exp(rng.uniform(0, np.log(64)))), n)).astype(np.int64) + (1 << (bits - 1))
        x = np.clip(v, 0, (1 << bits) - 1).astype({1: np.uint8, 2: np.uint16, 4: np.uint32}[eb])
        t = torch.from_numpy(x.view(VIEW[eb])).cuda()
        frame, size = mans_amd.frame_encode(t, mans_amd.params(10, 64, eb, mans_amd.MODE_PLANES), ws)
        kind = "c" if i % 2 == 0 else "d"
        out.append((kind, t, frame, size, eb))
    return out


def check_result(srv, res, kind, t, frame, size):
    assert res.status == 0
    if kind == "c":
        got = srv.fetch(res)
        assert got.numel() == size
        assert torch.equal(got, frame[:size])
    else:
        got = srv.fetch(res, t.dtype)
        assert torch.equal(got, t)


def test_server_mix_equals_alone(ws):
    reqs = requests(21, 48, 4 << 10, 4 << 20, ws)
    srv = mans_amd.Server(mans_amd.params(10, 64), 8 << 30)
    print("server: one request alone", flush=True)
    kind, t, frame, size, eb = reqs[0]
    rid = srv.submit_compress(t)
    res = srv.wait(rid, 60_000)
    check_result(srv, res, kind, t, frame, size)
    srv.release(rid)
    print(f"server: {len(reqs)} requests in a mix", flush=True)
    ids = []
    for kind, t, frame, size, eb in reqs:
        if kind == "c":
            ids.append(srv.submit_compress(t))
        else:
            ids.append(srv.submit_decompress(frame, size, t.numel(), eb))
    for rid, (kind, t, frame, size, eb) in zip(ids, reqs):
        res = srv.wait(rid, 60_000)
        assert res is not None, "request did not finish"
        check_result(srv, res, kind, t, frame, size)
        srv.release(rid)
    print(f"server: the same {len(reqs)} requests in one submit_many", flush=True)
    batch = [srv.compress_request(t) if kind == "c" else srv.decompress_request(frame, size, t.numel(), eb)
             for kind, t, frame, size, eb in reqs]
    ids = srv.submit_many(batch)
    assert len(set(ids)) == len(reqs)
    for rid, (kind, t, frame, size, eb) in zip(ids, reqs):
        res = srv.wait(rid, 60_000)
        assert res is not None, "request did not finish"
        check_result(srv, res, kind, t, frame, size)
        srv.release(rid)
    # one bad request in the list: the call fails and queues none of them
    bad = list(batch[:3])
    bad[1] = mans_amd.Request(mans_amd.REQ_COMPRESS, 3, batch[1].d_in, 0, 16)
    with pytest.raises(mans_amd.MansError):
        srv.submit_many(bad)
    assert srv.submit_many([]) == []
    st = srv.stats()
    assert st["arena_used"] == 0 and st["waiting"] == 0
    srv.close()
    ws.check()


def big_requests(count: int, n: int, ws):
    out = []
    g = torch.Generator(device="cuda")
    g.manual_seed(22)
    for i in range(count):
        v = torch.round(torch.randn(n, generator=g, device="cuda") * (4.0 + 4 * i)).to(torch.int32) + 32768
        t = torch.clamp(v, 0, 65535)
        t = torch.where(t >= 32768, t - 65536, t).to(torch.int16)
        frame, size = mans_amd.frame_encode(t, mans_amd.params(10, 64, 2, mans_amd.MODE_PLANES), ws)
        out.append(("c" if i % 2 == 0 else "d", t, frame, size, 2))
    return out


def test_server_admission_waits_for_memory(ws):
    # outputs of all requests together exceed the arena: later requests are
    # admitted only as earlier results are released
    # (the server's own state, two slots of batch workspaces, takes about
    # 2.1 GB of the 4 GiB)
    srv = mans_amd.Server(mans_amd.params(10, 64), 4 << 30)
    arena = srv.stats()["arena_bytes"]
    n = 16 << 20  # 32 MiB of uint16 per request
    count = int(1.5 * arena // (2 * n)) + 4
    reqs = big_requests(count, n, ws)
    print(f"server: arena {arena} bytes, {len(reqs)} requests of {2 * n} bytes", flush=True)
    ids = []
    done = []
    errors = []

    def collect():
        try:
            for rid, (kind, t, frame, size, eb) in zip(ids_ready, reqs):
                res = srv.wait(rid, 120_000)
                assert res is not None, "request did not finish"
                check_result(srv, res, kind, t, frame, size)
                srv.release(rid)
                done.append(rid)
        except Exception as e:  # pragma: no cover - reported below
            errors.append(e)

    for kind, t, frame, size, eb in reqs:
        if kind == "c":
            ids.append(srv.submit_compress(t))
        else:
            ids.append(srv.submit_decompress(frame, size, t.numel(), eb))
    ids_ready = list(ids)
    th = threading.Thread(target=collect)
    th.start()
    th.join(300)
    assert not th.is_alive(), "collector did not finish"
    assert not errors, errors
    assert len(done) == len(reqs)
    srv.close()
    ws.check()


def test_server_results_match_oracle(ws):
    # compress results byte for byte against the oracle's frame of the same
    # input; decompress requests take frames the ORACLE wrote and must give
    # back the oracle's input
    import oracle

    rng = np.random.default_rng(23)
    srv = mans_amd.Server(mans_amd.params(10, 64), 4 << 30)
    cases = []
    for i in range(24):
        eb = [1, 2, 4][i % 3]
        n = max(1, int(np.exp(rng.uniform(np.log(2 << 10), np.log(1 << 20)))) // eb)
        bits = 8 * eb
        v = np.round(rng.laplace(0, 3.0 + i, n)).astype(np.int64) + (1 << (bits - 1))
        x = np.clip(v, 0, (1 << bits) - 1).astype({1: np.uint8, 2: np.uint16, 4: np.uint32}[eb])
        ref = oracle.frame_encode(x, oracle.MODE_PLANES, 10, oracle.FORMAT_NATIVE)
        t = torch.from_numpy(x.view(VIEW[eb])).cuda()
        if i % 2 == 0:
            rid = srv.submit_compress(t)
        else:
            f = torch.from_numpy(ref).cuda()
            rid = srv.submit_decompress(f, ref.size, n, eb)
        cases.append((rid, i % 2 == 0, x, ref, t))
    print(f"server: {len(cases)} requests against the oracle", flush=True)
    for rid, is_c, x, ref, t in cases:
        res = srv.wait(rid, 60_000)
        assert res is not None and res.status == 0
        if is_c:
            got = srv.fetch(res).cpu().numpy()
            assert got.size == ref.size and (got == ref).all()
        else:
            got = srv.fetch(res, t.dtype).cpu().numpy().view(x.dtype)
            assert (got == x).all()
        srv.release(rid)
    srv.close()
    ws.check()


def test_server_corrupt_frame_fails_only_its_request(ws):
    reqs = requests(24, 12, 8 << 10, 1 << 20, ws)
    srv = mans_amd.Server(mans_amd.params(10, 64), 4 << 30)
    # request 3 decompresses a frame whose first block state is zeroed
    kind, t, frame, size, eb = reqs[3]
    assert kind == "d"
    bad = frame.clone()
    st0 = int(np.frombuffer(bytes(bad[64:72].cpu().numpy()), dtype="<u8")[0])
    bad[st0 + 544: st0 + 548] = 0
    batch = []
    for i, (kind, t, frame, size, eb) in enumerate(reqs):
        if kind == "c":
            batch.append(srv.compress_request(t))
        else:
            batch.append(srv.decompress_request(bad if i == 3 else frame, size, t.numel(), eb))
    print(f"server: {len(batch)} requests, one corrupt frame", flush=True)
    ids = srv.submit_many(batch)
    for i, (rid, (kind, t, frame, size, eb)) in enumerate(zip(ids, reqs)):
        res = srv.wait(rid, 60_000)
        assert res is not None
        if i == 3:
            assert res.status == 6, res.status  # MANS_ERR_CORRUPT
        else:
            check_result(srv, res, kind, t, frame, size)
        srv.release(rid)
    srv.close()
