"""CPU oracle for the multi-byte ANS codec path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it; the product package
``mans_amd`` never does.  The C restatement lives in ``mans_oracle.c`` (each
function cites the reference file:line it follows) and is wrapped here with
ctypes over numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_MAPPING = os.path.join(_HERE, "_ref", "mappingcpuV1")

FORMAT_REF = (32, 4096, 1)
FORMAT_NATIVE = (64, 8192, 2)
MODE_PLANES = 0
MODE_ADM = 1


class _Fmt(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("block_bytes", C.c_uint32), ("version", C.c_uint32)]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []) + ["oracle"], check=True)
    if os.path.isdir("/root/reference/mapping/cpu"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u8p = C.POINTER(C.c_uint8)
        L.mo_histogram.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p]
        L.mo_normalize.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
        L.mo_normalize.restype = C.c_int
        L.mo_magic_div.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.mo_magic_div.restype = C.c_uint32
        L.mo_stream_overhead.argtypes = [C.c_uint64, _Fmt]
        L.mo_stream_overhead.restype = C.c_uint64
        L.mo_stream_max_size.argtypes = [C.c_uint64, C.c_int, _Fmt]
        L.mo_stream_max_size.restype = C.c_uint64
        L.mo_lane_max_words.argtypes = [C.c_uint32, C.c_int]
        L.mo_lane_max_words.restype = C.c_uint32
        L.mo_stream_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, _Fmt,
                                       C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.c_int]
        L.mo_stream_encode.restype = C.c_int
        L.mo_stream_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_uint32,
                                       C.c_uint64, C.POINTER(C.c_uint64), C.c_int]
        L.mo_stream_decode.restype = C.c_int
        L.mo_adm_groups.argtypes = [C.c_uint64]
        L.mo_adm_groups.restype = C.c_uint64
        L.mo_adm_map.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
        L.mo_adm_map.restype = C.c_int
        L.mo_adm_unmap.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.mo_frame_max_size.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, _Fmt]
        L.mo_frame_max_size.restype = C.c_uint64
        L.mo_frame_encode.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_int, _Fmt,
                                      C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.c_int]
        L.mo_frame_encode.restype = C.c_int
        L.mo_frame_decode.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.c_int]
        L.mo_frame_decode.restype = C.c_int
        del u8p
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _fmt(fmt) -> _Fmt:
    return _Fmt(*fmt)


class OracleError(RuntimeError):
    pass


def histogram(data: np.ndarray, stride: int = 1, offset: int = 0) -> np.ndarray:
    """256-bin byte histogram of bytes data[offset + i*stride]."""
    b = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    n = (b.size - offset + stride - 1) // stride if b.size > offset else 0
    h = np.zeros(256, dtype=np.uint32)
    lib().mo_histogram(_ptr(b) + offset, n, stride, _ptr(h))
    return h


def normalize(hist: np.ndarray, total: int, prob_bits: int) -> np.ndarray:
    """(256, 4) uint32 table {pdf, cdf, magic, shift}."""
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    t = np.zeros((256, 4), dtype=np.uint32)
    rc = lib().mo_normalize(_ptr(h), total, prob_bits, _ptr(t))
    if rc:
        raise OracleError(f"mo_normalize rc={rc}")
    return t


def magic_div(x: int, magic: int, shift: int) -> int:
    return lib().mo_magic_div(x, magic, shift)


def lane_max_words(nsym: int, prob_bits: int) -> int:
    return lib().mo_lane_max_words(nsym, prob_bits)


def stream_overhead(num_blocks: int, fmt=FORMAT_REF) -> int:
    return lib().mo_stream_overhead(num_blocks, _fmt(fmt))


def stream_max_size(n: int, prob_bits: int, fmt=FORMAT_REF) -> int:
    return lib().mo_stream_max_size(n, prob_bits, _fmt(fmt))


def stream_encode(data: np.ndarray, prob_bits: int = 10, fmt=FORMAT_REF, stride: int = 1,
                  offset: int = 0, threads: int = 1) -> np.ndarray:
    """Single GpuANS container for bytes data[offset + i*stride]."""
    b = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
    n = (b.size - offset + stride - 1) // stride if b.size > offset else 0
    cap = stream_max_size(n, prob_bits, fmt)
    out = np.zeros(cap, dtype=np.uint8)
    size = C.c_uint64(0)
    rc = lib().mo_stream_encode(_ptr(b) + offset, n, stride, prob_bits, _fmt(fmt), _ptr(out), cap,
                                C.byref(size), threads)
    if rc:
        raise OracleError(f"mo_stream_encode rc={rc}")
    return out[: size.value].copy()


def stream_decode(comp: np.ndarray, prob_bits: int = 10, n: int | None = None,
                  threads: int = 1) -> np.ndarray:
    c = np.ascontiguousarray(comp, dtype=np.uint8)
    if n is None:
        n = int(c[8:12].view(np.uint32)[0]) if c.size >= 12 else 0
    out = np.zeros(max(n, 1), dtype=np.uint8)
    nout = C.c_uint64(0)
    rc = lib().mo_stream_decode(_ptr(c), c.size, prob_bits, _ptr(out), 1, out.size, C.byref(nout), threads)
    if rc:
        raise OracleError(f"mo_stream_decode rc={rc}")
    return out[: nout.value].copy()


def adm_map(x: np.ndarray):
    """ADM byte mapping of uint16 data: (in_domain, offsets[gsize+1], centers, codes, bits)."""
    x = np.ascontiguousarray(x, dtype=np.uint16)
    n = x.size
    g = lib().mo_adm_groups(n)
    offs = np.zeros(g + 1, dtype=np.int32)
    cen = np.zeros(max(g, 1), dtype=np.uint16)
    codes = np.zeros(max(n, 1), dtype=np.uint8)
    cap = g * 32 * 32
    bits = np.zeros(max(cap, 1), dtype=np.uint8)
    blen = C.c_uint64(0)
    rc = lib().mo_adm_map(_ptr(x), n, _ptr(cen), _ptr(offs), _ptr(codes), _ptr(bits), cap, C.byref(blen))
    if rc < 0:
        raise OracleError(f"mo_adm_map rc={rc}")
    return bool(rc == 1), offs, cen[:g].copy(), codes[:n].copy(), bits[: blen.value].copy()


def adm_unmap(offsets, centers, codes, bits, n: int) -> np.ndarray:
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    centers = np.ascontiguousarray(centers, dtype=np.uint16)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    bits = np.ascontiguousarray(bits, dtype=np.uint8)
    out = np.zeros(max(n, 1), dtype=np.uint16)
    lib().mo_adm_unmap(_ptr(centers) if centers.size else 0, _ptr(offsets), _ptr(codes) if codes.size else 0,
                       _ptr(bits) if bits.size else 0, n, _ptr(out))
    return out[:n].copy()


def frame_max_size(n_elems: int, elem_bytes: int, mode: int, prob_bits: int, fmt) -> int:
    return lib().mo_frame_max_size(n_elems, elem_bytes, mode, prob_bits, _fmt(fmt))


def frame_encode(x: np.ndarray, mode: int = MODE_PLANES, prob_bits: int = 10, fmt=FORMAT_NATIVE,
                 threads: int = 1) -> np.ndarray:
    x = np.ascontiguousarray(x).reshape(-1)
    eb = x.dtype.itemsize
    cap = frame_max_size(x.size, eb, mode, prob_bits, fmt)
    out = np.zeros(cap, dtype=np.uint8)
    size = C.c_uint64(0)
    rc = lib().mo_frame_encode(_ptr(x), x.size, eb, mode, prob_bits, _fmt(fmt), _ptr(out), cap,
                               C.byref(size), threads)
    if rc:
        raise OracleError(f"mo_frame_encode rc={rc}")
    return out[: size.value].copy()


def frame_decode(comp: np.ndarray, threads: int = 1) -> np.ndarray:
    c = np.ascontiguousarray(comp, dtype=np.uint8)
    eb = int(c[6]) if c.size > 6 else 1
    n = int(c[16:24].view(np.uint64)[0]) if c.size >= 24 else 0
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[eb]
    out = np.zeros(max(n, 1), dtype=dt)
    nout = C.c_uint64(0)
    rc = lib().mo_frame_decode(_ptr(c), c.size, _ptr(out), out.nbytes, C.byref(nout), threads)
    if rc:
        raise OracleError(f"mo_frame_decode rc={rc}")
    return out[: nout.value].copy()


def reference_mapping_available() -> bool:
    return os.path.exists(_REF_MAPPING)


def reference_mapping_path() -> str:
    return _REF_MAPPING
