"""The Lorenzo front end's HIP kernels (mans_lorenzo_quantize /
mans_lorenzo_reconstruct, through the C ABI) against the CPU restatement
(oracle/lorenzo.py): codes, the outlier list (indices, residuals, order) and
the reconstructed field bit for bit, on ragged shapes, spikes, the fix-up
path (more outliers than a thread holds), capacity and domain errors, and a
benchmark-shaped field."""
import numpy as np
import pytest

import mans_amd
from oracle import lorenzo as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def field(shape, seed, spikes=0):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.zeros(shape, dtype=np.float64)
    for _ in range(6):
        k = rng.normal(size=3) / rng.uniform(8, 64)
        f += np.sin(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    f += rng.normal(0, 0.01, shape)
    f = f.astype(np.float32)
    for _ in range(spikes):
        f[rng.integers(nz), rng.integers(ny), rng.integers(nx)] += np.float32(rng.uniform(-40, 40))
    return f


def check(f, eb, radius, ws, cap=None):
    print(f"lorenzo shape={f.shape} eb={eb} radius={radius}", flush=True)
    d = torch.from_numpy(f).cuda()
    codes, idx, val = mans_amd.lorenzo_quantize(d, eb, ws, radius=radius, outlier_capacity=cap)
    rc, ri, rv = L.lorenzo_quantize(f, eb, radius)
    got = codes.cpu().numpy().view(np.uint16)
    assert (got == rc).all(), np.flatnonzero(got != rc)[:10]
    assert idx.numel() == ri.size
    assert (idx.cpu().numpy().view(np.uint32) == ri).all()
    assert (val.cpu().numpy() == rv).all()
    rec = mans_amd.lorenzo_reconstruct(codes, idx, val, f.shape, eb, ws).cpu().numpy()
    ref = L.lorenzo_reconstruct(rc, ri, rv, f.shape, eb)
    assert (rec == ref).all()
    # codes 2 and 4 bytes past an 8-byte boundary (2 bytes: the inverse's
    # one-element-a-thread path): same field
    odd = torch.empty(codes.numel() + 2, dtype=codes.dtype, device=codes.device)
    for off in (1, 2):
        odd[off:off + codes.numel()] = codes
        rec2 = mans_amd.lorenzo_reconstruct(odd[off:off + codes.numel()], idx, val, f.shape, eb, ws).cpu().numpy()
        assert (rec2 == ref).all(), off
    # the field written at an address that is 4 but not 8 bytes aligned:
    # same field
    buf = torch.empty(codes.numel() + 1, dtype=torch.float32, device=codes.device)
    rec3 = mans_amd.lorenzo_reconstruct(codes, idx, val, f.shape, eb, ws, out=buf[1:].view(f.shape))
    assert (rec3.cpu().numpy() == ref).all()
    # the error bound, up to the float32 roundings of f / (2 eb) and q * 2 eb
    assert np.abs(rec.astype(np.float64) - f).max() <= eb + 4 * np.finfo(np.float32).eps * np.abs(f).max()
    ws.check()
    return ri.size


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 5, 7), (17, 33, 300), (20, 16, 256), (5, 70, 1000),
                                   (3, 19, 2500), (2, 4, 8192), (2, 9, 2049), (3, 9, 1026)])
def test_lorenzo_matches_oracle(ws, shape):
    check(field(shape, sum(shape), spikes=5), 1e-3, 512, ws)


@pytest.mark.parametrize("shape", [(18, 40, 300), (17, 32, 512)])
def test_lorenzo_fixup_path(ws, shape):
    # radius 1: every nonzero residual is an outlier, so threads overflow the
    # outliers they hold and the fix-up launch writes their tiles (the second
    # shape has whole tiles only: the forward launch without bounds)
    f = field(shape, 7, spikes=3)
    n = check(f, 1e-4, 1, ws, cap=f.size)
    assert n > f.size // 2


def test_lorenzo_fixup_list_over_capacity(ws):
    # radius 1 with a small capacity: more tiles need the fix-up than the
    # capacity can hold outliers for, which is a capacity error; the same
    # field with room for every outlier still matches the oracle afterwards
    f = field((18, 40, 300), 8, spikes=3)
    d = torch.from_numpy(f).cuda()
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.lorenzo_quantize(d, 1e-4, ws, radius=1, outlier_capacity=40)
    assert e.value.status == 5  # MANS_ERR_CAPACITY
    check(f, 1e-4, 1, ws, cap=f.size)


def test_lorenzo_capacity_and_domain(ws):
    f = field((8, 32, 256), 9, spikes=40)
    d = torch.from_numpy(f).cuda()
    _, _, rv = L.lorenzo_quantize(f, 1e-3, 512)
    assert rv.size > 10
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.lorenzo_quantize(d, 1e-3, ws, outlier_capacity=rv.size - 1)
    assert e.value.status == 5  # MANS_ERR_CAPACITY
    codes, idx, val = mans_amd.lorenzo_quantize(d, 1e-3, ws, outlier_capacity=rv.size)
    assert idx.numel() == rv.size
    big = d.clone()
    big[3, 3, 3] = 1e9
    with pytest.raises(mans_amd.MansError) as e:
        mans_amd.lorenzo_quantize(big, 1e-3, ws)
    assert e.value.status == 9  # MANS_ERR_DOMAIN
    ws.check()


def test_lorenzo_benchmark_field(ws):
    # the benchmark generator's field at 64 x 512 x 1024 against the oracle,
    # and the full 512 x 512 x 1024 field by round trip and error bound
    from mans_amd import workloads

    f = workloads.sinusoid_field((64, 512, 1024))
    check(f.cpu().numpy(), 1e-3, 512, ws)
    del f
    f = workloads.sinusoid_field((512, 512, 1024))
    codes, idx, val = mans_amd.lorenzo_quantize(f, 1e-3, ws)
    rec = mans_amd.lorenzo_reconstruct(codes, idx, val, tuple(f.shape), 1e-3, ws)
    err = (rec - f).abs().max().item()
    print(f"512x512x1024: {idx.numel()} outliers, max error {err:.3e}", flush=True)
    assert err <= 1e-3 + 4 * float(np.finfo(np.float32).eps) * f.abs().max().item()
    assert idx.numel() < f.numel() // 100
    ws.check()
