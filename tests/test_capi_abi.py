"""The C ABI library without a GPU: it loads, exports every function that
include/mans.h declares (and the reference-compatible C++ entry points of
include/gpu_ans.h), and its host-only functions agree with the oracle."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import mans_amd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mans.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mans_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 19
    L = mans_amd.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(mans_amd.SIGNATURES), "ctypes signatures out of sync with mans.h"


def test_reference_cpp_api_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", mans_amd.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ("_ZN13multibyte_ans14ansEncodeBatchEiPhjS0_PjP12ihipStream_t",
                "_ZN13multibyte_ans14ansDecodeBatchEiPhS0_P12ihipStream_t",
                "_ZN13multibyte_ans20getMaxCompressedSizeEj"):
        assert sym in out, sym


def test_status_strings_and_version():
    L = mans_amd.lib()
    assert L.mans_abi_version() == 1
    for s in range(12):
        assert L.mans_status_string(s)
    assert b"precision" in L.mans_status_string(2)


@pytest.mark.parametrize("lanes,fmt", [(32, oracle.FORMAT_REF), (64, oracle.FORMAT_NATIVE)])
@pytest.mark.parametrize("p", [9, 10, 11])
def test_max_size_matches_oracle(lanes, fmt, p):
    for n in [0, 1, 4096, 4097, 1 << 20, (1 << 30) + 3]:
        assert mans_amd.max_compressed_size(n, mans_amd.params(p, lanes)) == oracle.stream_max_size(n, p, fmt)


@pytest.mark.parametrize("eb", [1, 2, 4])
def test_frame_max_size_covers_oracle(eb):
    for lanes, fmt in ((32, oracle.FORMAT_REF), (64, oracle.FORMAT_NATIVE)):
        for n in [0, 5, 100_000]:
            p = mans_amd.params(10, lanes, eb, mans_amd.MODE_PLANES)
            assert mans_amd.frame_max_size(n, p) == oracle.frame_max_size(n, eb, 0, 10, fmt)
            if eb == 2:
                p = mans_amd.params(10, lanes, 2, mans_amd.MODE_ADM)
                assert mans_amd.frame_max_size(n, p) >= oracle.frame_max_size(n, 2, 1, 10, fmt)


def test_frame_info_parses_oracle_frames():
    x = np.arange(1000, dtype=np.uint16)
    c = oracle.frame_encode(x, oracle.MODE_PLANES, 10, oracle.FORMAT_NATIVE)
    n, eb, mode, total = C.c_uint64(), C.c_int32(), C.c_int32(), C.c_uint64()
    buf = (C.c_uint8 * 128).from_buffer_copy(bytes(c[:128]))
    rc = mans_amd.lib().mans_frame_info(buf, 128, C.byref(n), C.byref(eb), C.byref(mode), C.byref(total))
    assert rc == 0 and n.value == 1000 and eb.value == 2 and mode.value == 0 and total.value == len(c)
    bad = (C.c_uint8 * 128).from_buffer_copy(b"\0" * 128)
    assert mans_amd.lib().mans_frame_info(bad, 128, None, None, None, None) == 6


def test_compress_cli_built():
    assert os.access(mans_amd.CLI_PATH, os.X_OK)
