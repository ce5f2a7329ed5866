"""MI355X multi-byte ANS codec: Python host bindings over libmans.so.

The product is the C ABI in ``include/mans.h`` (hand-written HIP kernels for
gfx950 behind it); this module binds it with ctypes for tests, benchmarks and
Python callers.  Device buffers are torch tensors on ``cuda`` (torch is only
used for device memory and streams).  There is no CPU fallback: every entry
point raises if the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmans.so")
CLI_PATH = os.path.join(HERE, "bin", "compress")

MODE_PLANES = 0
MODE_ADM = 1
MODE_AUTO = 2

STATUS = {
    0: "MANS_OK", 1: "MANS_ERR_INVALID_ARG", 2: "MANS_ERR_PRECISION", 3: "MANS_ERR_ALIGNMENT",
    4: "MANS_ERR_TOO_LARGE", 5: "MANS_ERR_CAPACITY", 6: "MANS_ERR_CORRUPT", 7: "MANS_ERR_HIP",
    8: "MANS_ERR_NO_DEVICE", 9: "MANS_ERR_DOMAIN", 10: "MANS_ERR_BUDGET", 11: "MANS_ERR_BUSY",
}

# Every function declared in include/mans.h, with its ctypes signature.
_u64, _u32, _i32, _vp = C.c_uint64, C.c_uint32, C.c_int32, C.c_void_p


class Params(C.Structure):
    _fields_ = [("prob_bits", C.c_int32), ("lanes", C.c_int32), ("elem_bytes", C.c_int32), ("mode", C.c_int32)]


class Buffer(C.Structure):
    _fields_ = [("d_in", C.c_void_p), ("in_bytes", C.c_uint64), ("d_out", C.c_void_p),
                ("out_capacity", C.c_uint64), ("n_elems", C.c_uint64), ("elem_bytes", C.c_int32),
                ("reserved", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("elem_bytes", C.c_int32), ("d_in", C.c_void_p), ("in_bytes", C.c_uint64),
                ("n_elems", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("reserved", C.c_int32), ("d_out", C.c_void_p), ("out_bytes", C.c_uint64),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]


class BatchShape(C.Structure):
    _fields_ = [("wg_threads", C.c_int32), ("syms_per_lane", C.c_int32), ("lanes", C.c_int32),
                ("reserved", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("shape", BatchShape), ("ms", C.c_double)]


REQ_COMPRESS = 0
REQ_DECOMPRESS = 1

_P = C.POINTER(Params)
_B = C.POINTER(Buffer)
SIGNATURES = {
    "mans_status_string": (C.c_char_p, [C.c_int]),
    "mans_abi_version": (_i32, []),
    "mans_workspace_create": (C.c_int, [C.POINTER(_vp), _i32]),
    "mans_workspace_destroy": (C.c_int, [_vp]),
    "mans_workspace_reserve": (C.c_int, [_vp, _u64, _P]),
    "mans_workspace_check": (C.c_int, [_vp, _vp, C.POINTER(_u32)]),
    "mans_max_compressed_size": (_u64, [_u64, _P]),
    "mans_encode": (C.c_int, [_P, _vp, _u64, _vp, _u64, _vp, C.POINTER(_u64), _vp, _vp]),
    "mans_encode_noalloc": (C.c_int, [_P, _vp, _u64, _vp, _u64, _vp, _vp]),
    "mans_decode": (C.c_int, [_P, _vp, _u64, _vp, _u64, _vp, _vp]),
    "mans_encode_multi": (C.c_int, [_P, _i32, C.POINTER(_vp), C.POINTER(_u32), C.POINTER(_vp), C.POINTER(_u64), _vp,
                                    _vp, _vp]),
    "mans_decode_multi": (C.c_int, [_P, _i32, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_u64), _vp, _vp]),
    "mans_histogram": (C.c_int, [_vp, _u64, _u32, _vp, _vp, _vp]),
    "mans_normalize": (C.c_int, [_vp, _u32, _i32, _vp, _vp, _vp]),
    "mans_adm_map": (C.c_int, [_vp, _u64, _vp, _vp, _vp, _vp, _u64, C.POINTER(_u64), C.POINTER(_i32), _vp, _vp]),
    "mans_adm_unmap": (C.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp]),
    "mans_lorenzo_quantize": (C.c_int, [_vp, _u32, _u32, _u32, C.c_double, _u32, _vp, _vp, _vp, _u64, _vp,
                                        C.POINTER(_u64), _vp, _vp]),
    "mans_lorenzo_reconstruct": (C.c_int, [_vp, _vp, _vp, _u64, _u32, _u32, _u32, C.c_double, _vp, _vp, _vp]),
    "mans_frame_max_size": (_u64, [_u64, _P]),
    "mans_frame_encode": (C.c_int, [_P, _vp, _u64, _vp, _u64, _vp, C.POINTER(_u64), _vp, _vp]),
    "mans_frame_decode": (C.c_int, [_vp, _u64, _vp, _vp, _u64, C.POINTER(_u64), _vp, _vp]),
    "mans_frame_decode_fused": (C.c_int, [_vp, _u64, _vp, _vp, _u64, C.POINTER(_u64), _vp, _vp]),
    "mans_frame_info": (C.c_int, [_vp, _u64, C.POINTER(_u64), C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_u64)]),
    "mans_batch_encode_create": (C.c_int, [C.POINTER(_vp), _P, _B, _i32, _vp]),
    "mans_batch_encode_run": (C.c_int, [_vp, _vp, _vp]),
    "mans_batch_decode_create": (C.c_int, [C.POINTER(_vp), _P, _B, _i32, _vp]),
    "mans_batch_decode_run": (C.c_int, [_vp, _vp]),
    "mans_batch_decode_flags": (C.c_int, [_vp, C.POINTER(_u32), _i32, _vp]),
    "mans_batch_archive_max_size": (_u64, [_vp]),
    "mans_batch_pack": (C.c_int, [_vp, _vp, _vp, _u64, _vp, _vp]),
    "mans_archive_info": (C.c_int, [_vp, _u64, C.POINTER(_u32), C.POINTER(_u64)]),
    "mans_batch_decode_create_archive": (C.c_int, [C.POINTER(_vp), _P, _vp, _u64, C.POINTER(_vp), C.POINTER(_u64),
                                                   _vp]),
    "mans_batch_destroy": (C.c_int, [_vp]),
    "mans_batch_get_shape": (C.c_int, [_vp, C.POINTER(BatchShape)]),
    "mans_batch_set_shape": (C.c_int, [_vp, C.POINTER(BatchShape)]),
    "mans_batch_tune": (C.c_int, [_vp, _vp, _i32, C.POINTER(TuneEntry), _i32, C.POINTER(_i32)]),
    "mans_batch_set_grain": (C.c_int, [_u32, _u32, _u32]),
    "mans_encode_batch": (C.c_int, [_P, _B, _i32, _vp, _vp, _vp]),
    "mans_decode_batch": (C.c_int, [_P, _B, _i32, _vp, _vp]),
    "mans_server_create": (C.c_int, [C.POINTER(_vp), _P, _u64, _i32]),
    "mans_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(_u64)]),
    "mans_server_submit_many": (C.c_int, [_vp, C.POINTER(Request), _i32, C.POINTER(_u64)]),
    "mans_server_wait": (C.c_int, [_vp, _u64, _i32, C.POINTER(Result)]),
    "mans_server_release": (C.c_int, [_vp, _u64]),
    "mans_server_stats": (C.c_int, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "mans_server_destroy": (C.c_int, [_vp]),
    "mans_set_profiling": (C.c_int, [_vp, _i32]),
    "mans_last_kernel_times": (_i32, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_char_p), _i32]),
}


class MansError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)}")


_lib = None


def lib() -> C.CDLL:
    """Load libmans.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m mans_amd.build` first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise MansError(rc, what)


def params(prob_bits: int = 10, lanes: int = 32, elem_bytes: int = 1, mode: int = MODE_PLANES) -> Params:
    return Params(prob_bits, lanes, elem_bytes, mode)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("mans_amd needs a GPU (torch.cuda.is_available() is False)")
    return torch


def _stream(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Workspace:
    """Device scratch for one stream (see mans_workspace_create)."""

    def __init__(self, device: int = 0):
        _torch()
        self.device = device
        h = C.c_void_p()
        _check(lib().mans_workspace_create(C.byref(h), device), "mans_workspace_create")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            lib().mans_workspace_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, max_bytes: int, p: Params | None = None) -> None:
        _check(lib().mans_workspace_reserve(self.handle, max_bytes, C.byref(p) if p else None), "reserve")

    def check(self, stream=None) -> int:
        flags = C.c_uint32(0)
        rc = lib().mans_workspace_check(self.handle, _stream(stream), C.byref(flags))
        if rc != 0:
            raise MansError(rc, f"device flags 0x{flags.value:x}")
        return flags.value

    def set_profiling(self, on: bool = True) -> None:
        _check(lib().mans_set_profiling(self.handle, 1 if on else 0), "set_profiling")

    def kernel_times(self, max_records: int = 4096) -> list[tuple[str, float]]:
        """(kernel, ms) for every launch since the previous query."""
        ms = (C.c_double * max_records)()
        names = (C.c_char_p * max_records)()
        n = lib().mans_last_kernel_times(self.handle, ms, names, max_records)
        return [(names[i].decode(), ms[i]) for i in range(n)]


def max_compressed_size(n: int, p: Params) -> int:
    return lib().mans_max_compressed_size(n, C.byref(p))


def frame_max_size(n_elems: int, p: Params) -> int:
    return lib().mans_frame_max_size(n_elems, C.byref(p))


def _ptr(t) -> int:
    return t.data_ptr() if t is not None and t.numel() > 0 else 0


def encode(data, p: Params, ws: Workspace, out=None, stream=None, sync_size: bool = True):
    """Encode a uint8 CUDA tensor into one GpuANS container.

    Returns (out_buffer, size) when sync_size, else (out_buffer, d_size) with
    the size as a device uint32 tensor."""
    torch = _torch()
    n = data.numel() * data.element_size()
    cap = max_compressed_size(n, p)
    if out is None:
        out = torch.empty(cap, dtype=torch.uint8, device=data.device)
    d_size = torch.zeros(1, dtype=torch.int32, device=data.device)
    h_size = C.c_uint64(0)
    rc = lib().mans_encode(C.byref(p), _ptr(data), n, out.data_ptr(), out.numel(), d_size.data_ptr(),
                           C.byref(h_size) if sync_size else None, ws.handle, _stream(stream))
    _check(rc, "mans_encode")
    return (out, h_size.value) if sync_size else (out, d_size)


def encode_noalloc(data, p: Params, out, d_size, stream=None):
    """mans_encode_noalloc: the caller passes the output buffer (at least
    max_compressed_size bytes) and a one-element int32 CUDA tensor for the
    size; the call allocates nothing."""
    n = data.numel() * data.element_size()
    rc = lib().mans_encode_noalloc(C.byref(p), _ptr(data), n, out.data_ptr(), out.numel(), d_size.data_ptr(),
                                   _stream(stream))
    _check(rc, "mans_encode_noalloc")
    return out, d_size


def encode_multi(inputs, p: Params, ws: Workspace, stream=None):
    """mans_encode_multi over uint8 CUDA tensors: (containers, device sizes)."""
    torch = _torch()
    k = len(inputs)
    outs = [torch.empty(max_compressed_size(x.numel(), p), dtype=torch.uint8, device="cuda") for x in inputs]
    sizes = torch.zeros(max(1, k), dtype=torch.int32, device="cuda")
    ins = (C.c_void_p * max(1, k))(*[_ptr(x) for x in inputs])
    ns = (C.c_uint32 * max(1, k))(*[x.numel() for x in inputs])
    ops = (C.c_void_p * max(1, k))(*[o.data_ptr() for o in outs])
    caps = (C.c_uint64 * max(1, k))(*[o.numel() for o in outs])
    _check(lib().mans_encode_multi(C.byref(p), k, ins, ns, ops, caps, sizes.data_ptr(), ws.handle, _stream(stream)),
           "mans_encode_multi")
    return outs, sizes


def decode_multi(comps, sizes_n, p: Params, ws: Workspace, stream=None):
    """mans_decode_multi: containers -> uint8 CUDA tensors of the given lengths."""
    torch = _torch()
    k = len(comps)
    outs = [torch.empty(max(1, n), dtype=torch.uint8, device="cuda") for n in sizes_n]
    ins = (C.c_void_p * max(1, k))(*[c.data_ptr() for c in comps])
    ops = (C.c_void_p * max(1, k))(*[o.data_ptr() for o in outs])
    caps = (C.c_uint64 * max(1, k))(*[o.numel() for o in outs])
    _check(lib().mans_decode_multi(C.byref(p), k, ins, ops, caps, ws.handle, _stream(stream)), "mans_decode_multi")
    return [o[:n] for o, n in zip(outs, sizes_n)]


def decode(comp, n: int, p: Params, ws: Workspace, out=None, stream=None, c: int | None = None):
    torch = _torch()
    if out is None:
        out = torch.empty(max(n, 1), dtype=torch.uint8, device=comp.device)
    rc = lib().mans_decode(C.byref(p), comp.data_ptr(), comp.numel() if c is None else c, out.data_ptr(),
                           out.numel(), ws.handle, _stream(stream))
    _check(rc, "mans_decode")
    return out[:n]


def histogram(data, ws: Workspace, stride: int = 1, stream=None):
    torch = _torch()
    nbytes = data.numel() * data.element_size()
    h = torch.empty(256, dtype=torch.int32, device=data.device)
    _check(lib().mans_histogram(_ptr(data), nbytes // stride, stride, h.data_ptr(), ws.handle, _stream(stream)),
           "mans_histogram")
    return h


def normalize(hist, total: int, prob_bits: int, ws: Workspace, stream=None):
    torch = _torch()
    t = torch.empty((256, 4), dtype=torch.int32, device=hist.device)
    _check(lib().mans_normalize(hist.data_ptr(), total, prob_bits, t.data_ptr(), ws.handle, _stream(stream)),
           "mans_normalize")
    return t


def adm_map(x, ws: Workspace, stream=None):
    """ADM mapping of a uint16 (int16-typed) CUDA tensor.

    Returns (in_domain, offsets[gsize+1], centers[gsize], codes[n], bits)."""
    torch = _torch()
    n = x.numel()
    g = (n + 511) // 512
    offs = torch.empty(g + 1, dtype=torch.int32, device=x.device)
    cen = torch.empty(max(g, 1), dtype=torch.int16, device=x.device)
    codes = torch.empty(max(n, 16), dtype=torch.uint8, device=x.device)
    cap = g * 1024
    bits = torch.empty(max(cap, 16), dtype=torch.uint8, device=x.device)
    blen = C.c_uint64(0)
    dom = C.c_int32(0)
    _check(lib().mans_adm_map(_ptr(x), n, offs.data_ptr(), cen.data_ptr(), codes.data_ptr(), bits.data_ptr(), cap,
                              C.byref(blen), C.byref(dom), ws.handle, _stream(stream)), "mans_adm_map")
    return bool(dom.value), offs, cen[:g], codes[:n], bits[: blen.value]


def adm_unmap(offsets, centers, codes, bits, n: int, ws: Workspace, stream=None):
    torch = _torch()
    out = torch.empty(max(n, 8), dtype=torch.int16, device=offsets.device)
    _check(lib().mans_adm_unmap(offsets.data_ptr(), _ptr(centers), _ptr(codes), _ptr(bits), n, out.data_ptr(),
                                ws.handle, _stream(stream)), "mans_adm_unmap")
    return out[:n]


def lorenzo_quantize(field, error_bound: float, ws: Workspace, radius: int = 512, outlier_capacity: int | None = None,
                     codes=None, stream=None):
    """Lorenzo codes of a float32 CUDA tensor of shape (nz, ny, nx)
    (mans_lorenzo_quantize).  Returns (codes as an int16-typed tensor of the
    uint16 codes, outlier indices, outlier residuals), synchronising."""
    torch = _torch()
    nz, ny, nx = field.shape
    n = field.numel()
    cap = outlier_capacity if outlier_capacity is not None else max(1024, n // 64)
    if codes is None:
        codes = torch.empty(n, dtype=torch.int16, device=field.device)
    idx = torch.empty(max(1, cap), dtype=torch.int32, device=field.device)
    val = torch.empty(max(1, cap), dtype=torch.int32, device=field.device)
    d_cnt = torch.zeros(1, dtype=torch.int64, device=field.device)
    h_cnt = C.c_uint64(0)
    rc = lib().mans_lorenzo_quantize(_ptr(field), nz, ny, nx, error_bound, radius, _ptr(codes), idx.data_ptr(),
                                     val.data_ptr(), cap, d_cnt.data_ptr(), C.byref(h_cnt), ws.handle, _stream(stream))
    _check(rc, "mans_lorenzo_quantize")
    k = h_cnt.value
    return codes, idx[:k], val[:k]


def lorenzo_reconstruct(codes, idx, val, shape, error_bound: float, ws: Workspace, out=None, stream=None):
    """float32 field of `shape` from Lorenzo codes and outliers
    (mans_lorenzo_reconstruct)."""
    torch = _torch()
    nz, ny, nx = shape
    if out is None:
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device=codes.device)
    rc = lib().mans_lorenzo_reconstruct(_ptr(codes), _ptr(idx), _ptr(val), idx.numel(), nz, ny, nx, error_bound,
                                        _ptr(out), ws.handle, _stream(stream))
    _check(rc, "mans_lorenzo_reconstruct")
    return out


def frame_encode(x, p: Params, ws: Workspace, out=None, stream=None, sync_size: bool = True, d_size=None):
    """Encode a CUDA tensor of 1/2/4-byte elements into a MANS frame.  The
    size is returned on the host (sync_size) or left in `d_size` (a one-element
    int64 CUDA tensor, allocated when not given)."""
    torch = _torch()
    n = x.numel()
    p = Params(p.prob_bits, p.lanes, x.element_size(), p.mode)
    cap = frame_max_size(n, p)
    if out is None:
        out = torch.empty(cap, dtype=torch.uint8, device=x.device)
    if d_size is None:
        d_size = torch.zeros(1, dtype=torch.int64, device=x.device)
    h_size = C.c_uint64(0)
    rc = lib().mans_frame_encode(C.byref(p), _ptr(x), n, out.data_ptr(), out.numel(), d_size.data_ptr(),
                                 C.byref(h_size) if sync_size else None, ws.handle, _stream(stream))
    _check(rc, "mans_frame_encode")
    return (out, h_size.value) if sync_size else (out, d_size)


def frame_decode(comp, size: int, dtype, n: int, ws: Workspace, out=None, stream=None,
                 header: bytes | None = None, fused: bool = False):
    """Decode a MANS frame.  With `header` (the frame's first 128 bytes on the
    host) the call does not synchronise.  `fused`: mans_frame_decode_fused
    (64-lane ADM frames restore their elements inside the code plane's decode
    launch; other frames decode as usual)."""
    torch = _torch()
    if out is None:
        out = torch.empty(max(n, 8), dtype=dtype, device=comp.device)
    nout = C.c_uint64(0)
    hbuf = (C.c_uint8 * 128).from_buffer_copy(header[:128]) if header is not None else None
    fn = lib().mans_frame_decode_fused if fused else lib().mans_frame_decode
    rc = fn(comp.data_ptr(), size, hbuf, out.data_ptr(), out.numel() * out.element_size(), C.byref(nout), ws.handle,
            _stream(stream))
    _check(rc, "mans_frame_decode_fused" if fused else "mans_frame_decode")
    return out[:n]


class _Tunable:
    """Launch shape and autotuner of a planned batch (mans_batch_*_shape,
    mans_batch_tune)."""

    @property
    def shape(self) -> tuple[int, int, int]:
        """(threads per workgroup, symbols per lane, lanes)."""
        sh = BatchShape()
        _check(lib().mans_batch_get_shape(self.handle, C.byref(sh)), "mans_batch_get_shape")
        return sh.wg_threads, sh.syms_per_lane, sh.lanes

    def set_shape(self, wg_threads: int, syms_per_lane: int, lanes: int = 64) -> None:
        sh = BatchShape(wg_threads, syms_per_lane, lanes, 0)
        _check(lib().mans_batch_set_shape(self.handle, C.byref(sh)), "mans_batch_set_shape")

    def tune(self, reps: int = 3, stream=None) -> list[dict]:
        """Times every candidate shape on this batch and keeps the fastest;
        returns one row per candidate (ms < 0: does not fit the LDS)."""
        ent = (TuneEntry * 16)()
        n = C.c_int32(0)
        _check(lib().mans_batch_tune(self.handle, _stream(stream), reps, ent, 16, C.byref(n)), "mans_batch_tune")
        return [{"wg_threads": e.shape.wg_threads, "syms_per_lane": e.shape.syms_per_lane,
                 "lanes": e.shape.lanes, "ms": e.ms} for e in ent[:n.value]]


class EncodeBatch(_Tunable):
    """A planned batch of independent buffers, each coded as a byte-plane
    frame in p.lanes' format, 64 (native) or 32 (reference-format streams)
    (mans_batch_encode_create).  `inputs`: CUDA tensors of 1/2/4-byte
    elements (16-byte aligned storage); frames go to `outs` (allocated at
    frame_max_size when not given).  run() only launches, so it can be
    captured in a CUDA/HIP graph; sizes are left on the device."""

    def __init__(self, inputs, p: Params, ws: Workspace, outs=None):
        torch = _torch()
        self.inputs = list(inputs)
        self.ws = ws
        if outs is None:
            outs = []
            for x in self.inputs:
                q = Params(p.prob_bits, p.lanes, x.element_size(), MODE_PLANES)
                outs.append(torch.empty(frame_max_size(x.numel(), q), dtype=torch.uint8, device=x.device))
        self.outs = list(outs)
        arr = (Buffer * max(1, len(self.inputs)))()
        for i, (x, o) in enumerate(zip(self.inputs, self.outs)):
            arr[i] = Buffer(_ptr(x), 0, o.data_ptr(), o.numel(), x.numel(), x.element_size(), 0)
        self.sizes = torch.zeros(max(1, len(self.inputs)), dtype=torch.int64,
                                 device=self.inputs[0].device if self.inputs else "cuda")
        h = C.c_void_p()
        _check(lib().mans_batch_encode_create(C.byref(h), C.byref(p), arr, len(self.inputs), ws.handle),
               "mans_batch_encode_create")
        self.handle = h

    def run(self, stream=None):
        _check(lib().mans_batch_encode_run(self.handle, self.sizes.data_ptr(), _stream(stream)),
               "mans_batch_encode_run")
        return self.sizes

    def archive_max_size(self) -> int:
        return lib().mans_batch_archive_max_size(self.handle)

    def pack(self, archive=None, d_size=None, stream=None):
        """Frames of the last run in one archive (mans_batch_pack); returns
        (archive, d_size) with the size left on the device."""
        torch = _torch()
        if archive is None:
            archive = torch.empty(self.archive_max_size(), dtype=torch.uint8, device="cuda")
        if d_size is None:
            d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
        _check(lib().mans_batch_pack(self.handle, self.sizes.data_ptr(), archive.data_ptr(), archive.numel(),
                                     d_size.data_ptr(), _stream(stream)), "mans_batch_pack")
        return archive, d_size

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().mans_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DecodeBatch(_Tunable):
    """A planned batch decode (mans_batch_decode_create) of byte-plane frames
    of p.lanes' format: `frames` are uint8 CUDA tensors holding each frame (at least
    its size in bytes), `like` gives each buffer's element count and dtype
    (tensors, only their numel / dtype are used; with from_headers the
    library reads them from the frames and `like` only sizes the outputs).
    Outputs are allocated in `outs`.  run() only launches."""

    def __init__(self, frames, sizes, like, p: Params, ws: Workspace, from_headers: bool = False):
        torch = _torch()
        self.frames = list(frames)
        self.ws = ws
        self.outs = [torch.empty(max(1, t.numel()), dtype=t.dtype, device="cuda") for t in like]
        arr = (Buffer * max(1, len(self.frames)))()
        for i, (f, sz, t, o) in enumerate(zip(self.frames, sizes, like, self.outs)):
            # from_headers: element count and width come from each frame's header
            arr[i] = Buffer(f.data_ptr(), int(sz), o.data_ptr(), o.numel() * o.element_size(),
                            0 if from_headers else t.numel(), 0 if from_headers else t.element_size(), 0)
        h = C.c_void_p()
        _check(lib().mans_batch_decode_create(C.byref(h), C.byref(p), arr, len(self.frames), ws.handle),
               "mans_batch_decode_create")
        self.handle = h
        self.like_n = [t.numel() for t in like]

    def run(self, stream=None):
        _check(lib().mans_batch_decode_run(self.handle, _stream(stream)), "mans_batch_decode_run")
        return [o[:n] for o, n in zip(self.outs, self.like_n)]

    def flags(self, stream=None) -> list[int]:
        """Per-buffer error flags of the last run (0: decoded)."""
        n = len(self.frames)
        arr = (C.c_uint32 * max(1, n))()
        _check(lib().mans_batch_decode_flags(self.handle, arr, n, _stream(stream)), "mans_batch_decode_flags")
        return [arr[i] for i in range(n)]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().mans_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_NP_DT = {1: "int8", 2: "int16", 4: "int32"}


def set_batch_grain(hist_chunk: int = 0, rounds: int = 0, run: int = 0) -> None:
    """Force the work-item grain of batches planned from now on (process-wide;
    0 = the planner's pick): histogram item bytes (power of two, 64 KiB ..
    512 KiB), encode ticket rounds (1, 2, 4), decode items per grab (1, 4).
    Output bytes do not depend on it."""
    _check(lib().mans_batch_set_grain(hist_chunk, rounds, run), "mans_batch_set_grain")


def archive_index(archive) -> list[tuple[int, int, int, int]]:
    """(offset, frame bytes, elements, element bytes) of every frame of an
    archive (mans_archive_info + its index, read from the device)."""
    import numpy as np

    head = archive[:64].cpu().numpy().tobytes()
    cnt, total = C.c_uint32(0), C.c_uint64(0)
    _check(lib().mans_archive_info(head, 64, C.byref(cnt), C.byref(total)), "mans_archive_info")
    n = cnt.value
    idx = archive[64:64 + 32 * n].cpu().numpy().view(np.uint64).reshape(n, 4) if n else np.zeros((0, 4), np.uint64)
    return [(int(r[0]), int(r[1]), int(r[2]), int(r[3]) & 0xffffffff) for r in idx]


class ArchiveDecodeBatch(DecodeBatch):
    """A decode batch over every frame of an archive
    (mans_batch_decode_create_archive); outputs are allocated from the
    archive's index.  run() only launches."""

    def __init__(self, archive, archive_bytes: int, p: Params, ws: Workspace):
        torch = _torch()
        self.ws = ws
        entries = archive_index(archive)
        dt = {1: torch.int8, 2: torch.int16, 4: torch.int32}
        self.outs = [torch.empty(max(1, n), dtype=dt[eb], device="cuda") for _, _, n, eb in entries]
        self.like_n = [n for _, _, n, _ in entries]
        self.frames = [None] * len(entries)
        k = max(1, len(entries))
        ptrs = (C.c_void_p * k)(*[o.data_ptr() for o in self.outs])
        caps = (C.c_uint64 * k)(*[o.numel() * o.element_size() for o in self.outs])
        h = C.c_void_p()
        _check(lib().mans_batch_decode_create_archive(C.byref(h), C.byref(p), archive.data_ptr(), archive_bytes,
                                                      ptrs, caps, ws.handle), "mans_batch_decode_create_archive")
        self.handle = h


class Server:
    """Request server (mans_server_*): submit compress / decompress requests
    of CUDA tensors; results stay in the server's device arena until
    released."""

    def __init__(self, p: Params, memory_budget: int, device: int = 0):
        _torch()
        h = C.c_void_p()
        _check(lib().mans_server_create(C.byref(h), C.byref(p), memory_budget, device), "mans_server_create")
        self.handle = h

    @staticmethod
    def compress_request(x) -> Request:
        return Request(REQ_COMPRESS, x.element_size(), _ptr(x), 0, x.numel())

    @staticmethod
    def decompress_request(frame, frame_bytes: int, n: int, elem_bytes: int) -> Request:
        return Request(REQ_DECOMPRESS, elem_bytes, frame.data_ptr(), frame_bytes, n)

    def submit(self, r: Request) -> int:
        rid = C.c_uint64(0)
        _check(lib().mans_server_submit(self.handle, C.byref(r), C.byref(rid)), "mans_server_submit")
        return rid.value

    def submit_compress(self, x) -> int:
        return self.submit(self.compress_request(x))

    def submit_decompress(self, frame, frame_bytes: int, n: int, elem_bytes: int) -> int:
        return self.submit(self.decompress_request(frame, frame_bytes, n, elem_bytes))

    def submit_many(self, reqs) -> list[int]:
        """Queue a list of Requests at once (one call, one worker wake-up);
        returns their ids in order.  On an error none is queued."""
        n = len(reqs)
        arr = (Request * max(1, n))(*reqs)
        ids = (C.c_uint64 * max(1, n))()
        _check(lib().mans_server_submit_many(self.handle, arr, n, ids), "mans_server_submit_many")
        return [ids[i] for i in range(n)]

    def wait(self, rid: int, timeout_ms: int = -1) -> Result | None:
        res = Result()
        rc = lib().mans_server_wait(self.handle, rid, timeout_ms, C.byref(res))
        if rc == 11:  # MANS_ERR_BUSY
            return None
        _check(rc, "mans_server_wait")
        return res

    def fetch(self, res: Result, dtype=None):
        """Copy a result out of the arena into a new CUDA tensor (uint8, or
        viewed as dtype)."""
        torch = _torch()
        out = torch.empty(max(1, res.out_bytes), dtype=torch.uint8, device="cuda")
        if res.out_bytes:
            t = _torch_from_ptr(res.d_out, res.out_bytes)
            out[:res.out_bytes].copy_(t)
        out = out[:res.out_bytes]
        return out.view(dtype) if dtype is not None else out

    def release(self, rid: int) -> None:
        _check(lib().mans_server_release(self.handle, rid), "mans_server_release")

    def stats(self) -> dict:
        v = [C.c_uint64(0) for _ in range(4)]
        _check(lib().mans_server_stats(self.handle, *[C.byref(x) for x in v]), "mans_server_stats")
        return {"arena_bytes": v[0].value, "arena_used": v[1].value, "batches": v[2].value, "waiting": v[3].value}

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().mans_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _torch_from_ptr(ptr: int, nbytes: int):
    """A uint8 CUDA tensor aliasing device memory owned by the library."""
    torch = _torch()

    class _Arr:
        __cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False), "version": 3}

    return torch.as_tensor(_Arr(), device="cuda")
