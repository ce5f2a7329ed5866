"""Seeded synthetic inputs of the benchmark shapes.

* laplace_u16      -- clip(32768 + round(Laplace(0, b)), 0, 65535), the small
                      CPU-runnable case (2,097,152 values, b = 4, seed 0).
* sinusoid_field   -- a float32 3-D field: the sum of 64 random-phase plane
                      sinusoids (wavelengths 16-256 cells) plus N(0, 0.01)
                      noise, generated on the GPU.
* lorenzo_codes    -- uint16 error-bounded quantisation codes of that field
                      from the library's Lorenzo front end: q = rint(f / 2e-3)
                      (absolute error bound 1e-3), 3-D Lorenzo residual r of q,
                      code = 32768 + r if |r| < 512 else 0.  The 1 GiB case is
                      the 1024 x 512 x 1024 field (2^29 codes).
* zipf_u32         -- Zipf(s) ranks over 2^20 values mapped through a seeded
                      injective map to distinct 32-bit words with the high
                      byte in [0, 15].  Generated on the GPU.
* detector_u16     -- frames of round(N(1000, 8)) + 40 * Poisson(0.05).
                      Generated on the GPU.
* batch_buffers    -- independent buffers: sizes log-uniform in [4 KiB, 16 MiB],
                      element width uint8 / uint16 / uint32 at 20 / 60 / 20 %,
                      values 50 / 50 from a Laplace source (b log-uniform in
                      [1, 64], centred at half the range, clipped) or the Zipf
                      source (s = 1.1 ranks over 2^20 values through the seeded
                      injective map to words with the high byte in [0, 15], low
                      bytes kept for narrower widths); buffer i from seed
                      1000 + i.  Generated on the GPU.
"""
from __future__ import annotations

import numpy as np


def laplace_u16(n: int = 2_097_152, b: float = 4.0, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.clip(32768 + np.round(rng.laplace(0, b, n)), 0, 65535).astype(np.uint16)


def sinusoid_field(shape=(1024, 512, 1024), seed: int = 1, device: str = "cuda", slab: int = 32):
    """float32 CUDA tensor (nz, ny, nx): the sum of 64 random-phase plane
    sinusoids (unit directions, wavelengths log-uniform in 16-256 cells)
    plus N(0, 0.01) noise, generated in z-slabs."""
    import torch

    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    lam = np.exp(rng.uniform(np.log(16.0), np.log(256.0), 64))
    phi = rng.uniform(0, 2 * np.pi, 64)
    kvec = 2 * np.pi * d / lam[:, None]  # (64, 3) over (x, y, z)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    xs = torch.arange(nx, device=device, dtype=torch.float32)
    ys = torch.arange(ny, device=device, dtype=torch.float32)
    zs = torch.arange(nz, device=device, dtype=torch.float32)
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=device)
    for z0 in range(0, nz, slab):
        zc = zs[z0:z0 + slab]
        f = out[z0:z0 + zc.numel()]
        f.zero_()
        for j in range(64):
            kx, ky, kz = (float(v) for v in kvec[j])
            px = (kx * xs)[None, None, :]
            py = (ky * ys)[None, :, None]
            pz = (kz * zc + float(phi[j]))[:, None, None]
            f += torch.sin(px + py + pz)
        f += 0.01 * torch.randn(f.shape, generator=gen, device=device, dtype=torch.float32)
    return out


def lorenzo_codes(shape=(1024, 512, 1024), seed: int = 1, error_bound: float = 1e-3, radius: int = 512,
                  device: str = "cuda", ws=None):
    """int16-typed CUDA tensor holding the uint16 codes (z-major, then y, x):
    the library's Lorenzo front end (mans_lorenzo_quantize) on
    sinusoid_field(shape, seed)."""
    import mans_amd

    f = sinusoid_field(shape, seed, device)
    own = ws is None
    if own:
        ws = mans_amd.Workspace(0)
    codes, _, _ = mans_amd.lorenzo_quantize(f, error_bound, ws, radius=radius)
    del f
    if own:
        ws.close()
    return codes


def zipf_words(seed: int = 2) -> np.ndarray:
    """The 2^20 distinct 32-bit words (high byte in [0, 15]) that Zipf ranks
    map to: a seeded injective map of the ranks, drawn without replacement."""
    rng = np.random.default_rng(seed)
    return rng.choice(1 << 28, size=1 << 20, replace=False).astype(np.int64)


def _zipf_cdf(s: float, device: str):
    import torch

    w = torch.arange(1, (1 << 20) + 1, device=device, dtype=torch.float64).pow(-s)
    cdf = torch.cumsum(w, 0)
    return cdf / cdf[-1].clone()


def zipf_u32(n: int = 1 << 27, s: float = 1.1, seed: int = 2, device: str = "cuda", chunk: int = 1 << 24):
    """int32-typed CUDA tensor of uint32 values: Zipf(s) ranks over 2^20
    values (inverse CDF of the truncated distribution) mapped to
    zipf_words(seed).  Generated on the GPU in chunks."""
    import torch

    words = torch.from_numpy(zipf_words(seed)).to(device)
    cdf = _zipf_cdf(s, device)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    out = torch.empty(n, dtype=torch.int32, device=device)
    for i in range(0, n, chunk):
        m = min(chunk, n - i)
        u = torch.rand(m, generator=g, device=device, dtype=torch.float64)
        r = torch.clamp(torch.searchsorted(cdf, u), max=(1 << 20) - 1)
        out[i:i + m] = words[r].to(torch.int32)  # words < 2^28: no sign wrap
    return out


def detector_u16(frames: int = 64, side: int = 2048, seed: int = 3, device: str = "cuda"):
    """int16-typed CUDA tensor of uint16 detector counts: round(N(1000, 8))
    + 40 * Poisson(0.05) per pixel, clipped to [0, 65535].  Generated on the
    GPU one frame at a time."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    m = side * side
    out = torch.empty(frames * m, dtype=torch.int16, device=device)
    rate = torch.full((m,), 0.05, device=device, dtype=torch.float32)
    for f in range(frames):
        ped = torch.round(torch.normal(1000.0, 8.0, (m,), generator=g, device=device, dtype=torch.float32))
        ph = torch.poisson(rate, generator=g) * 40.0
        v = torch.clamp(ped + ph, 0, 65535).to(torch.int32)
        out[f * m:(f + 1) * m] = torch.where(v >= 32768, v - 65536, v).to(torch.int16)
    return out


def batch_buffers(count: int = 4096, seed_base: int = 1000, lo: int = 4096, hi: int = 16 << 20,
                  device: str = "cuda"):
    """List of `count` CUDA tensors (int8/int16/int32 views of the unsigned
    values) and their element widths."""
    import torch

    # Zipf source shared by all buffers: inverse-CDF table over 2^20 ranks
    words = torch.from_numpy(zipf_words(2)).to(device)
    cdf = _zipf_cdf(1.1, device)
    out, widths = [], []
    dt = {1: torch.int8, 2: torch.int16, 4: torch.int32}
    for i in range(count):
        rng = np.random.default_rng(seed_base + i)
        eb = int(rng.choice([1, 2, 4], p=[0.2, 0.6, 0.2]))
        nbytes = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        n = max(1, nbytes // eb)
        g = torch.Generator(device=device)
        g.manual_seed(seed_base + i)
        bits = 8 * eb
        if rng.random() < 0.5:
            b = float(np.exp(rng.uniform(0.0, np.log(64.0))))
            u = torch.rand(n, generator=g, device=device, dtype=torch.float64) - 0.5
            lap = -b * torch.sign(u) * torch.log1p(-2.0 * u.abs())
            v = torch.clamp(torch.round(lap).to(torch.int64) + (1 << (bits - 1)), 0, (1 << bits) - 1)
        else:
            u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
            r = torch.clamp(torch.searchsorted(cdf, u), max=(1 << 20) - 1)
            v = words[r] & ((1 << bits) - 1)
        v = torch.where(v >= (1 << (bits - 1)), v - (1 << bits), v).to(dt[eb])
        out.append(v)
        widths.append(eb)
    return out, widths
