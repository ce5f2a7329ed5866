"""ADM mapping oracle against golden vectors written by the reference's own
CPU mapping program (mapping/cpu/mappingcpuV1.cpp, built from its source by
oracle/Makefile and run by tests/golden/make_golden.py)."""
import glob
import os

import numpy as np
import pytest

import oracle

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "adm_v1_*.npz")))


def test_have_golden_vectors():
    assert len(GOLDEN) >= 9


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[7:-4] for p in GOLDEN])
def test_map_matches_reference(path):
    z = np.load(path)
    x = z["input"]
    ok, offs, cen, codes, bits = oracle.adm_map(x)
    g = z["offsets"].size
    assert ok
    assert (offs[:g] == z["offsets"]).all()
    assert (cen == z["centers"]).all()
    assert (codes == z["codes"]).all()
    assert bits.size == z["bits"].size and (bits == z["bits"]).all()
    assert offs[g] * 32 == bits.size


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[7:-4] for p in GOLDEN])
def test_unmap_reference_planes(path):
    z = np.load(path)
    x = z["input"]
    offs = np.concatenate([z["offsets"], [z["bits"].size // 32]]).astype(np.int32)
    assert (oracle.adm_unmap(offs, z["centers"], z["codes"], z["bits"], x.size) == x).all()


def test_domain_detection():
    x = np.full(512, 32768, dtype=np.uint16)
    x[0] = 32768 + 126 * 240 + 3  # 241 + 15 = 256 bits: the last lane length inside
    x[16] = 32768 - 126 * 240 - 3
    assert oracle.adm_map(x)[0] is True
    x[0] = 32768 + 126 * 241 + 1  # ceil(30367 / 126) = 242; 242 + 15 = 257 bits
    x[16] = 32768 - 126 * 241 - 1
    assert oracle.adm_map(x)[0] is False
    y = np.full(512, 20000, dtype=np.uint16)
    y[0] = 65535  # 45535 / 126 -> 362 bits
    assert oracle.adm_map(y)[0] is False


def test_round_trip_random_in_domain():
    rng = np.random.default_rng(11)
    for n in [1, 511, 512, 513, 10_000, 123_457]:
        x = (40000 + np.round(rng.laplace(0, 50, n))).astype(np.uint16)
        ok, offs, cen, codes, bits = oracle.adm_map(x)
        assert ok
        assert (oracle.adm_unmap(offs, cen, codes, bits, n) == x).all()
