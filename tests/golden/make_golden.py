"""Write ADM golden vectors with the reference's own CPU mapping program.

Runs oracle/_ref/mappingcpuV1 (built by ``make -C oracle ref`` from
mapping/cpu/mappingcpuV1.cpp of the reference tree) on small seeded uint16
inputs and stores input + the reference's planes as .npz fixtures:

  offsets  int32[gsize]   exclusive prefix of the per-group byte length
  centers  uint16[gsize]
  codes    uint8[n]
  bits     uint8[32 * total]

The program writes them back to back (mappingcpuV1.cpp:316-331); the bit plane
is the remainder of the file.  Only in-domain inputs are used: outside it the
reference overflows its 32-byte lane buffer (mappingcpuV1.cpp:48, 97).

Usage:  python tests/golden/make_golden.py   (needs the reference mounted)
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def cases() -> dict[str, np.ndarray]:
    out: dict[str, np.ndarray] = {}
    rng = np.random.default_rng(0)
    # Laplace(0, 4) around 32768 (the small CPU config's distribution).
    out["laplace_b4"] = np.clip(32768 + np.round(rng.laplace(0, 4, 65536)), 0, 65535).astype(np.uint16)
    # Wider Laplace so several elements need more than one signal bit.
    rng = np.random.default_rng(1)
    out["laplace_b300"] = np.clip(20000 + np.round(rng.laplace(0, 300, 40000)), 0, 65535).astype(np.uint16)
    # Hand cases: distances 0, 1, 125, 126, 127, 251, 252, 253 either side of
    # an exact centre, and lanes whose signal is exactly 256 bits.
    c = 32768
    g = np.full(512, c, dtype=np.int64)
    ds = [1, 125, 126, 127, 251, 252, 253, 378, 379]
    for k, d in enumerate(ds):
        g[32 + 2 * k] = c + d
        g[33 + 2 * k] = c - d
    g[0] = c + 126 * 240 + 3      # len 241 + 15 * 1 = 256 bits in lane 0
    g[16] = c - (126 * 240 + 3)   # lane 1 the same, mirrored
    hand = [g]
    g2 = np.full(512, 5000, dtype=np.int64)  # lane of 16 elements at len 16 = 256 bits
    for i in range(16):
        g2[64 + i] = 5000 + 126 * 15 + 1
        g2[128 + i] = 5000 - (126 * 15 + 1)
    hand.append(g2)
    out["hand_edges"] = np.concatenate(hand).astype(np.uint16)
    # Ragged sizes: partial last group, partial and empty lanes.
    rng = np.random.default_rng(2)
    for n in (1, 17, 552, 1000, 3 * 512 + 100):
        out[f"ragged_{n}"] = (5000 + rng.integers(-200, 200, n)).astype(np.uint16)
    # Values near 0 / 65535 (uint16 wrap in the unmap arithmetic).
    rng = np.random.default_rng(3)
    lo = rng.integers(0, 300, 2048)
    hi = 65535 - rng.integers(0, 300, 2048)
    out["extremes"] = np.concatenate([lo, hi]).astype(np.uint16)
    return out


def run_reference(x: np.ndarray) -> dict[str, np.ndarray]:
    exe = oracle.reference_mapping_path()
    n = x.size
    gsize = (n + 511) // 512
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.u2")
        fout = os.path.join(td, "out.bin")
        x.astype("<u2").tofile(fin)
        r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120,
                           env=dict(os.environ, OMP_NUM_THREADS="1"))
        if r.returncode != 0 or "Pass error check" not in r.stdout:
            raise RuntimeError(f"reference mapping failed: rc={r.returncode}\n{r.stdout}\n{r.stderr}")
        raw = np.fromfile(fout, dtype=np.uint8)
    o = 0
    offsets = raw[o:o + 4 * gsize].view("<i4").copy(); o += 4 * gsize
    centers = raw[o:o + 2 * gsize].view("<u2").copy(); o += 2 * gsize
    codes = raw[o:o + n].copy(); o += n
    bits = raw[o:].copy()
    return {"offsets": offsets, "centers": centers, "codes": codes, "bits": bits}


def main() -> None:
    if not oracle.reference_mapping_available():
        oracle.build()
    if not oracle.reference_mapping_available():
        raise SystemExit("reference mapping program not built (is the reference mounted?)")
    for name, x in cases().items():
        ref = run_reference(x)
        path = os.path.join(HERE, f"adm_v1_{name}.npz")
        np.savez_compressed(path, input=x, **ref)
        print(f"{name}: n={x.size} gsize={ref['offsets'].size} bits={ref['bits'].size} -> {os.path.basename(path)}")


if __name__ == "__main__":
    main()
