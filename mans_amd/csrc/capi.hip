// capi.hip -- the C ABI of libmans.so (include/mans.h): argument checks,
// container layout, workspace management and kernel launches.  Host code
// only; the kernels live in hist.hip, encode.hip, decode.hip and adm.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mans.h"
#include "mans_internal.hpp"

#include <atomic>

using namespace mans;

struct mans_workspace {
  int device = 0;
  int cu = 0;
  uint32_t* hist = nullptr;      // [4][256] self-cleaning accumulators
  uint32_t* counters = nullptr;  // 0 done, 1 encode ticket, 2 adm ticket, 3 err, 4-5 scratch scalars,
                                 // 8-9 histogram row grabs
  uint4* table = nullptr;        // [4][256]
  uint32_t* hist_out = nullptr;  // [4][256]
  uint64_t* status = nullptr;
  uint64_t status_cap = 0;
  uint8_t* scratch = nullptr;    // ADM planes of frames
  uint64_t scratch_cap = 0;
  uint16_t* spill = nullptr;     // encoder spill areas (one per resident wave)
  uint64_t spill_cap = 0;
  uint64_t* d_u64 = nullptr;     // [4]
  uint64_t* h_pinned = nullptr;  // [32] host pinned
  // set once a call ran on a capturing stream: the buffers above may then
  // be referenced by a graph, so they are never freed or replaced again
  bool frozen = false;
  bool profiling = false;
  std::vector<hipEvent_t> ev;
  std::vector<const char*> ev_names;
  int ev_used = 0;
};

namespace {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

#define HIP_TRY(x)                        \
  do {                                    \
    if ((x) != hipSuccess) return MANS_ERR_HIP; \
  } while (0)

bool aligned(const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

mans_status check_params(const mans_params* p) {
  if (!p) return MANS_ERR_INVALID_ARG;
  if (p->prob_bits < 9 || p->prob_bits > 11) return MANS_ERR_PRECISION;
  if (p->lanes != 32 && p->lanes != 64) return MANS_ERR_INVALID_ARG;
  return MANS_OK;
}

// A call on a stream under capture freezes the workspace (see
// mans_workspace::frozen): later calls that would need larger buffers fail
// with MANS_ERR_CAPACITY instead of freeing memory a captured graph uses.
void note_capture(mans_workspace* ws, hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) ws->frozen = true;
}

mans_status ensure_status(mans_workspace* ws, uint64_t n) {
  if (n <= ws->status_cap) return MANS_OK;
  if (ws->frozen) return MANS_ERR_CAPACITY;
  uint64_t cap = std::max<uint64_t>(n, ws->status_cap * 2);
  if (ws->status) (void)hipFree(ws->status);
  ws->status = nullptr;
  ws->status_cap = 0;
  HIP_TRY(hipMalloc(&ws->status, cap * sizeof(uint64_t)));
  HIP_TRY(hipMemset(ws->status, 0, cap * sizeof(uint64_t)));
  ws->status_cap = cap;
  return MANS_OK;
}

mans_status ensure_scratch(mans_workspace* ws, uint64_t n) {
  if (n <= ws->scratch_cap) return MANS_OK;
  if (ws->frozen) return MANS_ERR_CAPACITY;
  if (ws->scratch) (void)hipFree(ws->scratch);
  ws->scratch = nullptr;
  ws->scratch_cap = 0;
  HIP_TRY(hipMalloc(&ws->scratch, n));
  ws->scratch_cap = n;
  return MANS_OK;
}

mans_status ensure_spill(mans_workspace* ws, uint64_t bytes) {
  if (bytes <= ws->spill_cap) return MANS_OK;
  if (ws->frozen) return MANS_ERR_CAPACITY;
  if (ws->spill) (void)hipFree(ws->spill);
  ws->spill = nullptr;
  ws->spill_cap = 0;
  HIP_TRY(hipMalloc(&ws->spill, bytes));
  ws->spill_cap = bytes;
  return MANS_OK;
}

// Spill area the encoder needs on a large input of P planes of stride S.
uint64_t spill_for(mans_workspace* ws, int lanes, int pb, int P, uint32_t S, int order) {
  EncodeArgs e;
  memset(&e, 0, sizeof(e));
  e.d.lanes = lanes;
  e.d.prob_bits = pb;
  e.d.num_planes = P;
  e.d.order = order;
  for (int k = 0; k < P; ++k) e.d.plane[k].stride = S;
  e.d.total_units = 1ull << 30;
  uint64_t bytes = 0;
  if (encode_spill_bytes(e, ws->cu, &bytes) != hipSuccess) return 0;
  return bytes;
}

// Profiling: events recorded around launches when enabled.
void prof_mark(mans_workspace* ws, hipStream_t s, const char* name) {
  if (!ws->profiling) return;
  if ((size_t)ws->ev_used >= ws->ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ws->ev.push_back(e);
  }
  (void)hipEventRecord(ws->ev[ws->ev_used++], s);
  ws->ev_names.push_back(name);
}

uint64_t chain_len(uint64_t units, int lanes) { return div_up(units, lanes == 32 ? 2 : 1); }

uint64_t adm_bits_cap(uint64_t n) { return div_up(n, kAdmGroup) * kAdmLanes * kAdmLaneBytes; }
uint64_t adm_meta_bytes(uint64_t gsize) { return round_up(4ull * (gsize + 1) + 2ull * gsize, 16); }

uint64_t planes_frame_max(uint64_t n, uint32_t eb, int pb, Format f) {
  const uint64_t nb = div_up(n, f.block_bytes);
  const uint64_t cap = round_up(unit_max_words(f.lanes, f.block_bytes, pb), 8);
  return kFrameHeaderBytes + eb * (section_bytes(nb, f) + 2ull * nb * cap);
}
uint64_t adm_frame_max(uint64_t n, int pb, Format f) {
  const uint64_t g = div_up(n, kAdmGroup);
  const uint64_t nb0 = div_up(n, f.block_bytes), nb1 = div_up(adm_bits_cap(n), f.block_bytes);
  const uint64_t cap = round_up(unit_max_words(f.lanes, f.block_bytes, pb), 8);
  return kFrameHeaderBytes + adm_meta_bytes(g) + section_bytes(nb0, f) + section_bytes(nb1, f) +
         2ull * (nb0 + nb1) * cap;
}

void put64(uint32_t* w, int byte_off, uint64_t v) {
  w[byte_off / 4] = (uint32_t)v;
  w[byte_off / 4 + 1] = (uint32_t)(v >> 32);
}

// The 128-byte frame header of n elements of eb bytes coded as `d`.
void frame_header(uint32_t (&hdr)[32], uint64_t n, uint32_t eb, int mode, Format f, int prob_bits,
                  const CodecDesc& d, uint64_t gsize, uint64_t meta_off) {
  memset(hdr, 0, sizeof(hdr));
  uint8_t* hb = (uint8_t*)hdr;
  hdr[0] = kFrameMagic;
  hb[4] = (uint8_t)kFrameVersion;
  hb[5] = 0;
  hb[6] = (uint8_t)eb;
  hb[7] = (uint8_t)mode;
  hb[8] = (uint8_t)f.lanes;
  hb[9] = (uint8_t)prob_bits;
  hb[10] = (uint8_t)d.num_planes;
  hb[11] = (uint8_t)d.order;
  hdr[3] = f.block_bytes;
  put64(hdr, 16, n);
  put64(hdr, 24, gsize);
  put64(hdr, 32, meta_off);
  put64(hdr, 40, d.data_off);
  for (int k = 0; k < d.num_planes; ++k) {
    put64(hdr, 64 + 8 * k, d.plane[k].section);
    put64(hdr, 96 + 8 * k, d.plane[k].nsym);
  }
}

// Device scratch of one encode: histogram accumulators, counters (0 done,
// 1 ticket, 3 error flags), tables, look-back words and encoder spill areas.
struct EncodeScratch {
  uint32_t* hist;
  uint32_t* counters;
  uint4* table;
  uint64_t* status;
  uint16_t* spill;
  uint64_t spill_bytes;
};

// One hist+normalise launch and one encode launch over the planes of `d`,
// with the scratch in `sc` (large enough: checked by the callers).
mans_status launch_encode_pair(const CodecDesc& d, const HistSource* src, int nsrc, uint8_t* out,
                               uint32_t* d_out_size32, uint64_t* d_out_size64, const uint32_t* frame_hdr,
                               const EncodeScratch& sc, int cu, mans_workspace* prof, hipStream_t s) {
  const Format f = format_for_lanes(d.lanes);
  const uint64_t chain = chain_len(d.total_units, d.lanes);
  HistArgs h;
  memset(&h, 0, sizeof(h));
  for (int i = 0; i < nsrc; ++i) h.src[i] = src[i];
  h.nsrc = nsrc;
  h.num_planes = d.num_planes;
  h.prob_bits = d.prob_bits;
  h.write_headers = 1;
  for (int k = 0; k < d.num_planes; ++k) {
    h.plane_nsym[k] = d.plane[k].nsym;
    h.plane_nb[k] = d.plane[k].nb;
    h.plane_section[k] = d.plane[k].section;
  }
  h.hist = sc.hist;
  h.done = sc.counters + 0;
  h.grab = sc.counters + 8;
  h.table = sc.table;
  h.out = out;
  h.version = f.version;
  h.status = sc.status;
  h.status_count = chain;
  h.ticket = sc.counters + 1;
  h.d_out_size32 = d_out_size32;
  h.d_out_size64 = d_out_size64;
  h.empty_size = d.data_off;
  h.total_units = d.total_units;
  if (frame_hdr) {
    h.has_frame = 1;
    memcpy(h.frame_hdr, frame_hdr, sizeof(h.frame_hdr));
    if (d.total_units == 0) put64(h.frame_hdr, 56, d.data_off);
  }
  if (prof) prof_mark(prof, s, "hist_normalize");
  HIP_TRY(launch_hist(h, cu, s));
  EncodeArgs e;
  memset(&e, 0, sizeof(e));
  e.d = d;
  e.table = sc.table;
  e.out = out;
  e.status = sc.status;
  e.ticket = sc.counters + 1;
  e.err = sc.counters + 3;
  e.d_out_size32 = d_out_size32;
  e.d_out_size64 = d_out_size64;
  e.frame = frame_hdr ? 1 : 0;
  e.spill = sc.spill;
  e.spill_bytes = sc.spill_bytes;
  if (prof) prof_mark(prof, s, "encode");
  HIP_TRY(launch_encode(e, cu, s));
  if (prof) prof_mark(prof, s, nullptr);
  return MANS_OK;
}

// Shared by mans_encode and mans_frame_encode: sizes the workspace, then
// launches with its buffers.
mans_status run_encode(mans_workspace* ws, const CodecDesc& d, const HistSource* src, int nsrc,
                       uint8_t* out, uint32_t* d_out_size32, uint64_t* d_out_size64,
                       const uint32_t* frame_hdr, hipStream_t s) {
  const uint64_t chain = chain_len(d.total_units, d.lanes);
  mans_status st = ensure_status(ws, std::max<uint64_t>(chain, 1));
  if (st != MANS_OK) return st;
  EncodeArgs probe;
  memset(&probe, 0, sizeof(probe));
  probe.d = d;
  uint64_t spill = 0;
  HIP_TRY(encode_spill_bytes(probe, ws->cu, &spill));
  st = ensure_spill(ws, spill);
  if (st != MANS_OK) return st;
  const EncodeScratch sc{ws->hist, ws->counters, ws->table, ws->status, ws->spill, ws->spill_cap};
  return launch_encode_pair(d, src, nsrc, out, d_out_size32, d_out_size64, frame_hdr, sc, ws->cu, ws, s);
}

mans_status read_u64(mans_workspace* ws, const uint64_t* d, uint64_t* h, hipStream_t s) {
  HIP_TRY(hipMemcpyAsync(ws->h_pinned, d, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *h = ws->h_pinned[0];
  return MANS_OK;
}

// ---- allocation-free encode: device scratch held by the library -----------
// mans_encode_noalloc keeps its scratch in the file-scope array `scratch`
// below, one copy per GPU for the life of the process, instead of a
// workspace.  Its largest input is 2^30 bytes: 262,144 blocks of the 32-lane
// reference format, the reference's own limit (its block-count prefix scan
// asserts nb <= 512 x 512, BatchPrefixSum.h).
constexpr uint64_t kNoAllocMaxBytes = 1ull << 30;
constexpr uint64_t kNoAllocStatus = kNoAllocMaxBytes / 4096 / 2;  // one look-back word per two blocks
struct NoAllocLayout {
  uint32_t hist[4 * 256];    // histogram accumulators (cleared by each call)
  uint32_t counters[64];     // 0 done, 1 ticket, 3 error flags, 8-9 row grabs (cleared by each call)
  uint4 table[4 * 256];      // normalised tables
  uint64_t status[kNoAllocStatus];  // look-back words (cleared by the histogram launch)
};
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
NoAllocLayout* g_scratch[kMaxDevices];  // device address of `scratch` per device
int g_scratch_cu[kMaxDevices];

}  // namespace

// 1.02 MiB of device memory per GPU the library runs mans_encode_noalloc on.
// In host code this name is only a stand-in for the symbol: kernels and
// copies use the address hipGetSymbolAddress returns for it (g_scratch).
static __device__ __align__(256) uint8_t scratch[sizeof(NoAllocLayout)];

namespace {

// The device address of `scratch` on device dev (current), looked up once
// per device with that device current, and the device's compute units.
mans_status noalloc_scratch(int dev, NoAllocLayout** sc, int* cu) {
  if (dev < 0 || dev >= kMaxDevices) return MANS_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return MANS_ERR_HIP;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
      return MANS_ERR_NO_DEVICE;
    g_scratch_cu[dev] = n;
    g_scratch[dev] = (NoAllocLayout*)ptr;
  }
  *sc = g_scratch[dev];
  *cu = g_scratch_cu[dev];
  return MANS_OK;
}

}  // namespace

extern "C" {

const char* mans_status_string(mans_status s) {
  switch (s) {
    case MANS_OK: return "ok";
    case MANS_ERR_INVALID_ARG: return "invalid argument";
    case MANS_ERR_PRECISION: return "unsupported probability precision (use 9, 10 or 11)";
    case MANS_ERR_ALIGNMENT: return "pointer not aligned as required";
    case MANS_ERR_TOO_LARGE: return "size exceeds the format's 32-bit fields";
    case MANS_ERR_CAPACITY: return "output buffer too small";
    case MANS_ERR_CORRUPT: return "malformed or inconsistent compressed stream";
    case MANS_ERR_HIP: return "HIP runtime error";
    case MANS_ERR_NO_DEVICE: return "no usable GPU";
    case MANS_ERR_DOMAIN: return "data outside the ADM mapping's lossless domain";
    case MANS_ERR_BUDGET: return "request exceeds the device-memory budget";
    case MANS_ERR_BUSY: return "queue full";
  }
  return "unknown status";
}

int32_t mans_abi_version(void) { return MANS_ABI_VERSION; }

mans_status mans_workspace_create(mans_workspace** out, int32_t device) {
  if (!out) return MANS_ERR_INVALID_ARG;
  *out = nullptr;
  int dev = device;
  if (dev < 0 && hipGetDevice(&dev) != hipSuccess) return MANS_ERR_NO_DEVICE;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || dev >= count) return MANS_ERR_NO_DEVICE;
  DeviceGuard g(dev);
  mans_workspace* ws = new (std::nothrow) mans_workspace();
  if (!ws) return MANS_ERR_INVALID_ARG;
  ws->device = dev;
  if (hipDeviceGetAttribute(&ws->cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    delete ws;
    return MANS_ERR_NO_DEVICE;
  }
  bool ok = hipMalloc(&ws->hist, 4 * 256 * 4) == hipSuccess && hipMalloc(&ws->counters, 64 * 4) == hipSuccess &&
            hipMalloc(&ws->table, 4 * 256 * 16) == hipSuccess && hipMalloc(&ws->hist_out, 4 * 256 * 4) == hipSuccess &&
            hipMalloc(&ws->d_u64, 4 * 8) == hipSuccess &&
            hipHostMalloc((void**)&ws->h_pinned, 32 * 8, hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMemset(ws->hist, 0, 4 * 256 * 4) == hipSuccess && hipMemset(ws->counters, 0, 64 * 4) == hipSuccess &&
       hipMemset(ws->table, 0, 4 * 256 * 16) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    mans_workspace_destroy(ws);
    return MANS_ERR_HIP;
  }
  *out = ws;
  return MANS_OK;
}

mans_status mans_workspace_destroy(mans_workspace* ws) {
  if (!ws) return MANS_OK;
  DeviceGuard g(ws->device);
  (void)hipDeviceSynchronize();
  if (ws->hist) (void)hipFree(ws->hist);
  if (ws->counters) (void)hipFree(ws->counters);
  if (ws->table) (void)hipFree(ws->table);
  if (ws->hist_out) (void)hipFree(ws->hist_out);
  if (ws->status) (void)hipFree(ws->status);
  if (ws->scratch) (void)hipFree(ws->scratch);
  if (ws->spill) (void)hipFree(ws->spill);
  if (ws->d_u64) (void)hipFree(ws->d_u64);
  if (ws->h_pinned) (void)hipHostFree(ws->h_pinned);
  for (hipEvent_t e : ws->ev) (void)hipEventDestroy(e);
  delete ws;
  return MANS_OK;
}

mans_status mans_workspace_reserve(mans_workspace* ws, uint64_t max_bytes, const mans_params* p) {
  if (!ws) return MANS_ERR_INVALID_ARG;
  DeviceGuard g(ws->device);
  const int lanes = (p && p->lanes == 64) ? 64 : 32;
  const Format f = format_for_lanes(lanes);
  const uint64_t units = 4 * div_up(max_bytes, f.block_bytes) + 8;
  mans_status st = ensure_status(ws, std::max(chain_len(units, 32), div_up(max_bytes, kAdmGroup) + 8));
  if (st != MANS_OK) return st;
  if (p && p->mode != MANS_MODE_PLANES && p->elem_bytes == 2) {
    const uint64_t n = max_bytes / 2;
    st = ensure_scratch(ws, round_up(n, 16) + adm_bits_cap(n) + 16);
    if (st != MANS_OK) return st;
  }
  if (p && check_params(p) == MANS_OK) {
    // the encoder's spill areas for this element width (byte planes) and for
    // single byte streams (single streams, ADM planes)
    const int eb = p->elem_bytes == 4 ? 4 : (p->elem_bytes == 2 ? 2 : 1);
    const uint64_t need = std::max({spill_for(ws, lanes, p->prob_bits, eb, (uint32_t)eb, 0),
                                    spill_for(ws, lanes, p->prob_bits, 1, 1u, 0),
                                    spill_for(ws, lanes, p->prob_bits, 1, 1u, 1)});
    st = ensure_spill(ws, need);
  }
  return st;
}

mans_status mans_workspace_check(mans_workspace* ws, void* stream, uint32_t* h_flags) {
  if (!ws) return MANS_ERR_INVALID_ARG;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  uint32_t* hf = (uint32_t*)ws->h_pinned;
  HIP_TRY(hipMemcpyAsync(hf, ws->counters + 3, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  uint32_t flags = hf[0];
  if (h_flags) *h_flags = flags;
  if (flags) {
    HIP_TRY(hipMemsetAsync(ws->counters + 3, 0, 4, s));
    HIP_TRY(hipStreamSynchronize(s));
    return MANS_ERR_CORRUPT;
  }
  return MANS_OK;
}

uint64_t mans_max_compressed_size(uint64_t n, const mans_params* p) {
  const int lanes = (p && p->lanes == 64) ? 64 : 32;
  const int pb = (p && p->prob_bits >= 9 && p->prob_bits <= 11) ? p->prob_bits : 11;
  const Format f = format_for_lanes(lanes);
  const uint64_t nb = div_up(n, f.block_bytes);
  return section_bytes(nb, f) + 2ull * nb * round_up(unit_max_words(f.lanes, f.block_bytes, pb), 8);
}

mans_status mans_encode(const mans_params* p, const void* d_in, uint64_t n, void* d_out, uint64_t out_capacity,
                        uint32_t* d_out_size, uint64_t* h_out_size, mans_workspace* ws, void* stream) {
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (!ws || !d_out || (n > 0 && !d_in)) return MANS_ERR_INVALID_ARG;
  // input 4-byte aligned as the reference requires (kANSRequiredAlignment,
  // GpuANSCodec.h:216); the encoder reads whole dwords
  if (!aligned(d_out, 16) || !aligned(d_in, 4)) return MANS_ERR_ALIGNMENT;
  if (n > 0xffffffffull) return MANS_ERR_TOO_LARGE;
  if (out_capacity < mans_max_compressed_size(n, p)) return MANS_ERR_CAPACITY;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  const Format f = format_for_lanes(p->lanes);
  CodecDesc d;
  memset(&d, 0, sizeof(d));
  d.num_planes = 1;
  d.order = 0;
  d.prob_bits = p->prob_bits;
  d.lanes = p->lanes;
  d.plane[0].src = (const uint8_t*)d_in;
  d.plane[0].nsym = n;
  d.plane[0].nb = div_up(n, f.block_bytes);
  d.plane[0].section = 0;
  d.plane[0].unit_base = 0;
  d.plane[0].stride = 1;
  d.total_units = d.plane[0].nb;
  d.data_off = section_bytes(d.plane[0].nb, f);
  HistSource src = {(const uint8_t*)d_in, n, 1u, 0u};
  st = run_encode(ws, d, &src, n ? 1 : 0, (uint8_t*)d_out, d_out_size, h_out_size ? ws->d_u64 : nullptr, nullptr, s);
  if (st != MANS_OK) return st;
  if (h_out_size) return read_u64(ws, ws->d_u64, h_out_size, s);
  return MANS_OK;
}

mans_status mans_encode_noalloc(const mans_params* p, const void* d_in, uint64_t n, void* d_out,
                                uint64_t out_capacity, uint32_t* d_out_size, void* stream) {
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (p->lanes != 32 || !d_out || !d_out_size || (n > 0 && !d_in)) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_out, 16) || !aligned(d_in, 4) || !aligned(d_out_size, 4)) return MANS_ERR_ALIGNMENT;
  if (n > kNoAllocMaxBytes) return MANS_ERR_TOO_LARGE;
  if (out_capacity < mans_max_compressed_size(n, p)) return MANS_ERR_CAPACITY;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MANS_ERR_NO_DEVICE;
  NoAllocLayout* sc = nullptr;
  int cu = 0;
  st = noalloc_scratch(dev, &sc, &cu);
  if (st != MANS_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  // the accumulators and counters start at zero in every call, whatever an
  // earlier call left (hist and counters are adjacent: one fill kernel)
  static_assert(offsetof(NoAllocLayout, counters) == sizeof(uint32_t) * 4 * 256, "adjacent");
  HIP_TRY(launch_fill_u32(sc->hist, 4 * 256 + 64, 0, s));
  const Format f = format_for_lanes(32);
  CodecDesc d;
  memset(&d, 0, sizeof(d));
  d.num_planes = 1;
  d.order = 0;
  d.prob_bits = p->prob_bits;
  d.lanes = 32;
  d.plane[0] = PlaneDesc{(const uint8_t*)d_in, nullptr, n, div_up(n, f.block_bytes), 0, 0, 1, 0};
  d.total_units = d.plane[0].nb;
  d.data_off = section_bytes(d.plane[0].nb, f);
  if (chain_len(d.total_units, 32) > kNoAllocStatus) return MANS_ERR_TOO_LARGE;
  const HistSource src = {(const uint8_t*)d_in, n, 1u, 0u};
  // the 32-lane encoder stages whole blocks in LDS: no spill areas
  const EncodeScratch es{sc->hist, sc->counters, sc->table, sc->status, nullptr, 0};
  return launch_encode_pair(d, &src, n ? 1 : 0, (uint8_t*)d_out, d_out_size, nullptr, nullptr, es, cu, nullptr, s);
}

mans_status mans_decode(const mans_params* p, const void* d_in, uint64_t c, void* d_out, uint64_t out_capacity,
                        mans_workspace* ws, void* stream) {
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (!ws || !d_in || !d_out) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_in, 16)) return MANS_ERR_ALIGNMENT;
  if (c && c < kSectionFixed) return MANS_ERR_CORRUPT;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  DecodeArgs a;
  memset(&a, 0, sizeof(a));
  a.d.num_planes = 1;
  a.d.order = 0;
  a.d.prob_bits = p->prob_bits;
  a.d.lanes = p->lanes;
  a.d.plane[0].dst = (uint8_t*)d_out;
  a.d.plane[0].stride = 1;
  a.in = (const uint8_t*)d_in;
  a.in_bytes = c;
  a.out_capacity = out_capacity;
  a.err = ws->counters + 3;
  a.from_header = 1;
  prof_mark(ws, s, "decode");
  HIP_TRY(launch_decode(a, ws->cu, s));
  prof_mark(ws, s, nullptr);
  return MANS_OK;
}

mans_status mans_encode_multi(const mans_params* p, int32_t count, const void* const* h_in,
                              const uint32_t* h_in_size, void* const* h_out, const uint64_t* h_out_capacity,
                              uint32_t* d_out_sizes, mans_workspace* ws, void* stream) {
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (!ws || count < 0 || (count > 0 && (!h_in || !h_in_size || !h_out || !h_out_capacity || !d_out_sizes)))
    return MANS_ERR_INVALID_ARG;
  // every argument first: on an error nothing is launched
  for (int32_t i = 0; i < count; ++i) {
    if (!h_out[i] || (h_in_size[i] > 0 && !h_in[i])) return MANS_ERR_INVALID_ARG;
    if (!aligned(h_out[i], 16) || !aligned(h_in[i], 4)) return MANS_ERR_ALIGNMENT;
    if (h_out_capacity[i] < mans_max_compressed_size(h_in_size[i], p)) return MANS_ERR_CAPACITY;
  }
  for (int32_t i = 0; i < count; ++i) {
    st = mans_encode(p, h_in[i], h_in_size[i], h_out[i], h_out_capacity[i], d_out_sizes + i, nullptr, ws, stream);
    if (st != MANS_OK) return st;
  }
  return MANS_OK;
}

mans_status mans_decode_multi(const mans_params* p, int32_t count, const void* const* h_in, void* const* h_out,
                              const uint64_t* h_out_capacity, mans_workspace* ws, void* stream) {
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (!ws || count < 0 || (count > 0 && (!h_in || !h_out || !h_out_capacity))) return MANS_ERR_INVALID_ARG;
  for (int32_t i = 0; i < count; ++i) {
    if (!h_in[i] || !h_out[i]) return MANS_ERR_INVALID_ARG;
    if (!aligned(h_in[i], 16)) return MANS_ERR_ALIGNMENT;
  }
  for (int32_t i = 0; i < count; ++i) {
    st = mans_decode(p, h_in[i], 0, h_out[i], h_out_capacity[i], ws, stream);
    if (st != MANS_OK) return st;
  }
  return MANS_OK;
}

mans_status mans_histogram(const void* d_in, uint64_t n, uint32_t stride, uint32_t* d_hist, mans_workspace* ws,
                           void* stream) {
  if (!ws || !d_hist || (n > 0 && !d_in)) return MANS_ERR_INVALID_ARG;
  if (stride != 1 && stride != 2 && stride != 4) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_in, stride)) return MANS_ERR_ALIGNMENT;
  if (n > 0xffffffffull) return MANS_ERR_TOO_LARGE;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  HistArgs h;
  memset(&h, 0, sizeof(h));
  h.src[0] = HistSource{(const uint8_t*)d_in, n * stride, stride, 0u};
  h.nsrc = n ? 1 : 0;
  h.num_planes = (int)stride;
  h.prob_bits = 10;
  for (uint32_t k = 0; k < stride; ++k) h.plane_nsym[k] = n;
  h.hist = ws->hist;
  h.done = ws->counters + 0;
  h.grab = ws->counters + 8;
  h.table = ws->table;
  h.hist_out = ws->hist_out;
  h.ticket = ws->counters + 1;
  prof_mark(ws, s, "hist_normalize");
  HIP_TRY(launch_hist(h, ws->cu, s));
  prof_mark(ws, s, nullptr);
  HIP_TRY(hipMemcpyAsync(d_hist, ws->hist_out, 256 * 4, hipMemcpyDeviceToDevice, s));
  return MANS_OK;
}

mans_status mans_normalize(const uint32_t* d_hist, uint32_t total, int32_t prob_bits, uint32_t* d_table,
                           mans_workspace* ws, void* stream) {
  if (!ws || !d_hist || !d_table) return MANS_ERR_INVALID_ARG;
  if (prob_bits < 9 || prob_bits > 11) return MANS_ERR_PRECISION;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  HIP_TRY(hipMemcpyAsync(ws->hist, d_hist, 256 * 4, hipMemcpyDeviceToDevice, s));
  HistArgs h;
  memset(&h, 0, sizeof(h));
  h.nsrc = 0;
  h.num_planes = 1;
  h.prob_bits = prob_bits;
  h.plane_nsym[0] = total;
  h.hist = ws->hist;
  h.done = ws->counters + 0;
  h.grab = ws->counters + 8;
  h.table = ws->table;
  h.ticket = ws->counters + 1;
  prof_mark(ws, s, "normalize");
  HIP_TRY(launch_hist(h, ws->cu, s));
  prof_mark(ws, s, nullptr);
  HIP_TRY(hipMemcpyAsync(d_table, ws->table, 256 * 16, hipMemcpyDeviceToDevice, s));
  return MANS_OK;
}

static mans_status adm_map_impl(const uint16_t* d_in, uint64_t n, int32_t* d_offsets, uint16_t* d_centers,
                                uint8_t* d_codes, uint8_t* d_bits, uint64_t bits_capacity, uint64_t* h_bits_len,
                                int32_t* h_in_domain, mans_workspace* ws, hipStream_t s) {
  const uint64_t gsize = div_up(n, kAdmGroup);
  const uint64_t chain = div_up(gsize, kAdmTicketGroups);
  mans_status st = ensure_status(ws, std::max<uint64_t>(chain, 1));
  if (st != MANS_OK) return st;
  HIP_TRY(launch_fill_u64(ws->status, chain, 0, s));
  // counters[2] = ticket, counters[4] = this call's domain / capacity flags
  // (kept apart from the sticky stream-error word counters[3])
  HIP_TRY(launch_fill_u32(ws->counters + 2, 1, 0, s));
  HIP_TRY(launch_fill_u32(ws->counters + 4, 1, 0, s));
  if (gsize == 0) HIP_TRY(launch_fill_u32((uint32_t*)d_offsets, 1, 0, s));
  AdmMapArgs a;
  memset(&a, 0, sizeof(a));
  a.in = d_in;
  a.n = n;
  a.gsize = gsize;
  a.offsets = d_offsets;
  a.centers = d_centers;
  a.codes = d_codes;
  a.bits = d_bits;
  a.bits_capacity = bits_capacity;
  a.status = ws->status;
  a.ticket = ws->counters + 2;
  a.err = ws->counters + 4;
  prof_mark(ws, s, "adm_map");
  HIP_TRY(launch_adm_map(a, ws->cu, s));
  if (h_bits_len || h_in_domain) {
    uint32_t* hp = (uint32_t*)ws->h_pinned;
    HIP_TRY(hipMemcpyAsync(hp, d_offsets + gsize, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hp + 1, ws->counters + 4, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h_bits_len) *h_bits_len = 32ull * hp[0];
    if (h_in_domain) *h_in_domain = (hp[1] & kErrAdmDomain) ? 0 : 1;
    if (hp[1] & kErrCapacity) return MANS_ERR_CAPACITY;
  }
  return MANS_OK;
}

mans_status mans_adm_map(const uint16_t* d_in, uint64_t n, int32_t* d_offsets, uint16_t* d_centers, uint8_t* d_codes,
                         uint8_t* d_bits, uint64_t bits_capacity, uint64_t* h_bits_len, int32_t* h_in_domain,
                         mans_workspace* ws, void* stream) {
  if (!ws || !d_offsets || (n > 0 && (!d_in || !d_centers || !d_codes || !d_bits))) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_in, 16) || !aligned(d_codes, 16)) return MANS_ERR_ALIGNMENT;
  if (n > 0xffffffffull) return MANS_ERR_TOO_LARGE;
  DeviceGuard g(ws->device);
  note_capture(ws, (hipStream_t)stream);
  mans_status st = adm_map_impl(d_in, n, d_offsets, d_centers, d_codes, d_bits, bits_capacity, h_bits_len,
                                h_in_domain, ws, (hipStream_t)stream);
  prof_mark(ws, (hipStream_t)stream, nullptr);
  return st;
}

mans_status mans_adm_unmap(const int32_t* d_offsets, const uint16_t* d_centers, const uint8_t* d_codes,
                           const uint8_t* d_bits, uint64_t n, uint16_t* d_out, mans_workspace* ws, void* stream) {
  if (!ws || (n > 0 && (!d_offsets || !d_centers || !d_codes || !d_bits || !d_out))) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_out, 16) || !aligned(d_codes, 16)) return MANS_ERR_ALIGNMENT;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  AdmUnmapArgs a;
  memset(&a, 0, sizeof(a));
  a.offsets = d_offsets;
  a.centers = d_centers;
  a.codes = d_codes;
  a.bits = d_bits;
  a.n = n;
  a.gsize = div_up(n, kAdmGroup);
  a.bits_len = ~0ull;
  a.out = d_out;
  a.err = ws->counters + 3;
  prof_mark(ws, s, "adm_unmap");
  HIP_TRY(launch_adm_unmap(a, s));
  prof_mark(ws, s, nullptr);
  return MANS_OK;
}

mans_status mans_lorenzo_quantize(const float* d_field, uint32_t nz, uint32_t ny, uint32_t nx, double error_bound,
                                  uint32_t radius, uint16_t* d_codes, uint32_t* d_outlier_idx, int32_t* d_outlier_val,
                                  uint64_t outlier_capacity, uint64_t* d_outlier_count, uint64_t* h_outlier_count,
                                  mans_workspace* ws, void* stream) {
  const uint64_t n = (uint64_t)nz * ny * nx;
  if (!ws || !d_outlier_count || (n > 0 && (!d_field || !d_codes))) return MANS_ERR_INVALID_ARG;
  if (outlier_capacity > 0 && (!d_outlier_idx || !d_outlier_val)) return MANS_ERR_INVALID_ARG;
  if (!(error_bound > 0.0) || radius < 1 || radius > 32768 || nx > 8192) return MANS_ERR_INVALID_ARG;
  if (n >= (1ull << 32)) return MANS_ERR_TOO_LARGE;
  if (!aligned(d_field, 4) || !aligned(d_codes, 2) || !aligned(d_outlier_count, 8)) return MANS_ERR_ALIGNMENT;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  const uint32_t ntiles = n ? lorenzo_tiles(nz, ny, nx) : 0;
  // scratch: tile offsets (8 B a tile), the fix-up list (4 B a tile), its
  // thread offsets and per-plane-group counts (2 KiB a listed tile; a tile
  // is listed only with more than kLorenzoHold outliers, so more than
  // out_cap / (kLorenzoHold + 1) + 1 listed tiles means more than out_cap
  // outliers, which is a capacity error anyway)
  const uint64_t nt = std::max<uint32_t>(ntiles, 1);
  const uint32_t fix_cap = (uint32_t)std::min<uint64_t>(nt, outlier_capacity / (kLorenzoHold + 1) + 1);
  const uint64_t list_off = 8 * nt, pos_off = list_off + 4 * (nt + 1), grp_off = pos_off + 1024ull * fix_cap;
  mans_status st = ensure_status(ws, nt);
  if (st == MANS_OK) st = ensure_scratch(ws, grp_off + 1024ull * fix_cap);
  if (st != MANS_OK) return st;
  // counters[5]: tile ticket; counters[4]: this call's flags
  HIP_TRY(launch_fill_u64(ws->status, ntiles, 0, s));
  HIP_TRY(launch_fill_u32(ws->counters + 4, 2, 0, s));
  HIP_TRY(launch_fill_u32((uint32_t*)(ws->scratch + list_off), 1, 0, s));
  if (ntiles == 0) HIP_TRY(launch_fill_u64(d_outlier_count, 1, 0, s));
  LorenzoArgs a;
  memset(&a, 0, sizeof(a));
  a.field = d_field;
  a.codes = d_codes;
  a.nz = nz;
  a.ny = ny;
  a.nx = nx;
  a.radius = radius;
  a.inv2eb = (float)(1.0 / (2.0 * error_bound));
  a.twoeb = (float)(2.0 * error_bound);
  a.out_idx = d_outlier_idx;
  a.out_val = d_outlier_val;
  a.out_cap = outlier_capacity;
  a.count = d_outlier_count;
  a.status = ws->status;
  a.ticket = ws->counters + 5;
  a.tile_off = (uint64_t*)ws->scratch;
  a.fix_list = (uint32_t*)(ws->scratch + list_off);
  a.fix_pos = (uint32_t*)(ws->scratch + pos_off);
  a.fix_grp = (uint32_t*)(ws->scratch + grp_off);
  a.fix_cap = fix_cap;
  a.ntiles = ntiles;
  a.err = ws->counters + 4;
  prof_mark(ws, s, "lorenzo");
  HIP_TRY(launch_lorenzo(a, s));
  prof_mark(ws, s, nullptr);
  if (h_outlier_count) {
    HIP_TRY(hipMemcpyAsync(ws->h_pinned, d_outlier_count, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ws->h_pinned + 1, ws->counters + 4, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *h_outlier_count = ws->h_pinned[0];
    const uint32_t flags = (uint32_t)ws->h_pinned[1];
    if (flags & kErrAdmDomain) return MANS_ERR_DOMAIN;
    if (ws->h_pinned[0] > outlier_capacity) return MANS_ERR_CAPACITY;
  }
  return MANS_OK;
}

mans_status mans_lorenzo_reconstruct(const uint16_t* d_codes, const uint32_t* d_outlier_idx,
                                     const int32_t* d_outlier_val, uint64_t outlier_count, uint32_t nz, uint32_t ny,
                                     uint32_t nx, double error_bound, float* d_field, mans_workspace* ws,
                                     void* stream) {
  const uint64_t n = (uint64_t)nz * ny * nx;
  if (!ws || (n > 0 && (!d_codes || !d_field)) || (outlier_count > 0 && (!d_outlier_idx || !d_outlier_val)))
    return MANS_ERR_INVALID_ARG;
  if (!(error_bound > 0.0) || nx > 8192) return MANS_ERR_INVALID_ARG;
  if (n >= (1ull << 32)) return MANS_ERR_TOO_LARGE;
  if (!aligned(d_field, 4) || !aligned(d_codes, 2)) return MANS_ERR_ALIGNMENT;
  if (n == 0) return MANS_OK;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  mans_status st = ensure_scratch(ws, 4 * n);
  if (st != MANS_OK) return st;
  LorenzoArgs a;
  memset(&a, 0, sizeof(a));
  a.codes = (uint16_t*)d_codes;
  a.nz = nz;
  a.ny = ny;
  a.nx = nx;
  a.twoeb = (float)(2.0 * error_bound);
  a.in_idx = d_outlier_idx;
  a.in_val = d_outlier_val;
  a.n_out = outlier_count;
  a.q = (int32_t*)ws->scratch;
  a.recon = d_field;
  a.err = ws->counters + 3;
  prof_mark(ws, s, "lorenzo_inverse");
  HIP_TRY(launch_lorenzo_inverse(a, ws->cu, s));
  prof_mark(ws, s, nullptr);
  return MANS_OK;
}

uint64_t mans_frame_max_size(uint64_t n, const mans_params* p) {
  if (!p) return 0;
  const int pb = (p->prob_bits >= 9 && p->prob_bits <= 11) ? p->prob_bits : 11;
  const Format f = format_for_lanes(p->lanes == 64 ? 64 : 32);
  const uint32_t eb = (uint32_t)(p->elem_bytes == 1 || p->elem_bytes == 2 || p->elem_bytes == 4 ? p->elem_bytes : 4);
  uint64_t m = planes_frame_max(n, eb, pb, f);
  if (eb == 2 && p->mode != MANS_MODE_PLANES) m = std::max(m, adm_frame_max(n, pb, f));
  return m;
}

mans_status mans_frame_encode(const mans_params* p, const void* d_in, uint64_t n, void* d_out, uint64_t out_capacity,
                              uint64_t* d_size, uint64_t* h_size, mans_workspace* ws, void* stream) {
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (!ws || !d_out || (n > 0 && !d_in)) return MANS_ERR_INVALID_ARG;
  const uint32_t eb = (uint32_t)p->elem_bytes;
  if (eb != 1 && eb != 2 && eb != 4) return MANS_ERR_INVALID_ARG;
  if (p->mode < MANS_MODE_PLANES || p->mode > MANS_MODE_AUTO) return MANS_ERR_INVALID_ARG;
  if (p->mode == MANS_MODE_ADM && eb != 2) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_out, 16) || !aligned(d_in, 4)) return MANS_ERR_ALIGNMENT;
  if (n > 0xffffffffull) return MANS_ERR_TOO_LARGE;
  if (out_capacity < mans_frame_max_size(n, p)) return MANS_ERR_CAPACITY;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  const Format f = format_for_lanes(p->lanes);
  uint8_t* out = (uint8_t*)d_out;
  int mode = p->mode == MANS_MODE_AUTO ? (eb == 2 && aligned(d_in, 16) ? MANS_MODE_ADM : MANS_MODE_PLANES) : p->mode;
  if (mode == MANS_MODE_ADM && !aligned(d_in, 16)) return MANS_ERR_ALIGNMENT;

  CodecDesc d;
  memset(&d, 0, sizeof(d));
  d.prob_bits = p->prob_bits;
  d.lanes = p->lanes;
  HistSource src[2];
  int nsrc = 0;
  uint64_t gsize = 0, meta_off = 0, off = kFrameHeaderBytes;
  if (mode == MANS_MODE_ADM) {
    gsize = div_up(n, kAdmGroup);
    meta_off = off;
    off += adm_meta_bytes(gsize);
    const uint64_t codes_bytes = round_up(n, 16);
    st = ensure_scratch(ws, codes_bytes + adm_bits_cap(n) + 16);
    if (st != MANS_OK) return st;
    uint8_t* codes = ws->scratch;
    uint8_t* bits = ws->scratch + codes_bytes;
    int32_t* offs = (int32_t*)(out + meta_off);
    uint16_t* cen = (uint16_t*)(out + meta_off + 4ull * (gsize + 1));
    uint64_t bits_len = 0;
    int32_t in_domain = 0;
    st = adm_map_impl((const uint16_t*)d_in, n, offs, cen, codes, bits, adm_bits_cap(n), &bits_len, &in_domain, ws, s);
    if (st != MANS_OK) return st;
    if (!in_domain) {
      if (p->mode == MANS_MODE_ADM) return MANS_ERR_DOMAIN;
      mode = MANS_MODE_PLANES;
      off = kFrameHeaderBytes;
      gsize = 0;
      meta_off = 0;
    } else {
      // zero the meta padding (the map kernel writes offsets and centers)
      const uint64_t used = 4ull * (gsize + 1) + 2ull * gsize;
      const uint64_t pad = adm_meta_bytes(gsize) - used;
      if (pad) HIP_TRY(hipMemsetAsync(out + meta_off + used, 0, pad, s));
      d.num_planes = 2;
      d.order = 1;
      d.plane[0] = PlaneDesc{codes, nullptr, n, div_up(n, f.block_bytes), 0, 0, 1, 0};
      d.plane[1] = PlaneDesc{bits, nullptr, bits_len, div_up(bits_len, f.block_bytes), 0, d.plane[0].nb, 1, 0};
      src[0] = HistSource{codes, n, 1u, 0u};
      src[1] = HistSource{bits, bits_len, 1u, 1u};
      nsrc = 2;
    }
  }
  if (mode == MANS_MODE_PLANES) {
    d.num_planes = (int)eb;
    d.order = 0;
    for (uint32_t k = 0; k < eb; ++k)
      d.plane[k] = PlaneDesc{(const uint8_t*)d_in + k, nullptr, n, div_up(n, f.block_bytes), 0, 0, eb, 0};
    src[0] = HistSource{(const uint8_t*)d_in, n * eb, eb, 0u};
    nsrc = n ? 1 : 0;
  }
  d.total_units = 0;
  for (int k = 0; k < d.num_planes; ++k) {
    d.plane[k].section = off;
    off += section_bytes(d.plane[k].nb, f);
    d.total_units += d.plane[k].nb;
  }
  d.data_off = off;
  uint32_t hdr[32];
  frame_header(hdr, n, eb, mode, f, p->prob_bits, d, gsize, meta_off);
  uint64_t* dsz = d_size ? d_size : (h_size ? ws->d_u64 : nullptr);
  st = run_encode(ws, d, src, nsrc, out, nullptr, dsz, hdr, s);
  if (st != MANS_OK) return st;
  if (h_size) return read_u64(ws, dsz, h_size, s);
  return MANS_OK;
}

mans_status mans_frame_info(const void* h_header, uint64_t header_bytes, uint64_t* n_elems, int32_t* elem_bytes,
                            int32_t* mode, uint64_t* total_size) {
  if (!h_header || header_bytes < kFrameHeaderBytes) return MANS_ERR_INVALID_ARG;
  const uint8_t* h = (const uint8_t*)h_header;
  uint32_t magic;
  uint64_t n, total;
  memcpy(&magic, h, 4);
  memcpy(&n, h + 16, 8);
  memcpy(&total, h + 56, 8);
  if (magic != kFrameMagic || h[4] != kFrameVersion || h[5] != 0) return MANS_ERR_CORRUPT;
  if (n_elems) *n_elems = n;
  if (elem_bytes) *elem_bytes = h[6];
  if (mode) *mode = h[7];
  if (total_size) *total_size = total;
  return MANS_OK;
}

static mans_status frame_decode_impl(const void* d_in, uint64_t c, const void* h_header, void* d_out,
                                     uint64_t out_capacity, uint64_t* h_n_elems, mans_workspace* ws, void* stream,
                                     bool fuse) {
  if (!ws || !d_in || c < kFrameHeaderBytes) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_in, 16)) return MANS_ERR_ALIGNMENT;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  note_capture(ws, s);
  uint8_t hcopy[kFrameHeaderBytes];
  const uint8_t* hb = hcopy;
  if (h_header) {
    memcpy(hcopy, h_header, kFrameHeaderBytes);
  } else {
    HIP_TRY(hipMemcpyAsync(ws->h_pinned, d_in, kFrameHeaderBytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    memcpy(hcopy, ws->h_pinned, kFrameHeaderBytes);
  }
  uint32_t magic, block_bytes;
  uint64_t n, gsize, meta, data_off, data_words, total, sect[4], nsym[4];
  memcpy(&magic, hb, 4);
  memcpy(&block_bytes, hb + 12, 4);
  memcpy(&n, hb + 16, 8);
  memcpy(&gsize, hb + 24, 8);
  memcpy(&meta, hb + 32, 8);
  memcpy(&data_off, hb + 40, 8);
  memcpy(&data_words, hb + 48, 8);
  memcpy(&total, hb + 56, 8);
  memcpy(sect, hb + 64, 32);
  memcpy(nsym, hb + 96, 32);
  const uint32_t eb = hb[6], mode = hb[7], lanes = hb[8], pb = hb[9], P = hb[10], order = hb[11];
  if (magic != kFrameMagic || hb[4] != kFrameVersion || hb[5] != 0) return MANS_ERR_CORRUPT;
  if ((lanes != 32 && lanes != 64) || pb < 9 || pb > 11 || P < 1 || P > 4 || order > 1) return MANS_ERR_CORRUPT;
  const Format f = format_for_lanes((int)lanes);
  if (block_bytes != f.block_bytes || (eb != 1 && eb != 2 && eb != 4)) return MANS_ERR_CORRUPT;
  if (total > c || data_off + 2ull * data_words > c || (data_off & 15)) return MANS_ERR_CORRUPT;
  if (mode == MANS_MODE_PLANES && (P != eb || order != 0)) return MANS_ERR_CORRUPT;
  if (mode == MANS_MODE_ADM && (eb != 2 || P != 2 || order != 1 || gsize != div_up(n, kAdmGroup))) return MANS_ERR_CORRUPT;
  if (mode > MANS_MODE_ADM) return MANS_ERR_CORRUPT;
  if (h_n_elems) *h_n_elems = n;
  if (n * eb > out_capacity) return MANS_ERR_CAPACITY;
  if (n > 0 && !d_out) return MANS_ERR_INVALID_ARG;
  if (mode == MANS_MODE_ADM && !aligned(d_out, 16)) return MANS_ERR_ALIGNMENT;

  DecodeArgs a;
  memset(&a, 0, sizeof(a));
  a.d.num_planes = (int)P;
  a.d.order = (int)order;
  a.d.prob_bits = (int)pb;
  a.d.lanes = (int)lanes;
  uint64_t units = 0;
  uint8_t* codes = nullptr;
  uint8_t* bits = nullptr;
  if (mode == MANS_MODE_ADM && fuse && lanes == 64) {
    // the bit plane alone, then the code plane with the restore fused in
    // (no code scratch); the checks are those of the two-launch path
    if (nsym[0] != n || nsym[1] > adm_bits_cap(n)) return MANS_ERR_CORRUPT;
    if (meta + adm_meta_bytes(gsize) > c) return MANS_ERR_CORRUPT;
    for (uint32_t k = 0; k < 2; ++k) {
      if (sect[k] + kSectionFixed > c || nsym[k] > 0xffffffffull) return MANS_ERR_CORRUPT;
      if (sect[k] + section_bytes(div_up(nsym[k], f.block_bytes), f) > data_off) return MANS_ERR_CORRUPT;
    }
    mans_status st = ensure_scratch(ws, round_up(nsym[1], 16) + 16);
    if (st != MANS_OK) return st;
    bits = ws->scratch;
    a.d.num_planes = 1;
    a.d.order = 0;
    a.d.data_off = data_off;
    a.in = (const uint8_t*)d_in;
    a.in_bytes = c;
    a.data_words = data_words;
    a.err = ws->counters + 3;
    a.from_header = 0;
    const uint64_t nb1 = div_up(nsym[1], f.block_bytes), nb0 = div_up(nsym[0], f.block_bytes);
    a.d.plane[0] = PlaneDesc{nullptr, bits, nsym[1], nb1, sect[1], 0, 1, 0};
    a.d.total_units = nb1;
    prof_mark(ws, s, "decode");
    HIP_TRY(launch_decode(a, ws->cu, s));
    a.d.plane[0] = PlaneDesc{nullptr, nullptr, nsym[0], nb0, sect[0], 0, 1, 0};
    a.d.total_units = nb0;
    a.adm.offsets = (const int32_t*)((const uint8_t*)d_in + meta);
    a.adm.centers = (const uint16_t*)((const uint8_t*)d_in + meta + 4ull * (gsize + 1));
    a.adm.codes = nullptr;
    a.adm.bits = bits;
    a.adm.n = n;
    a.adm.gsize = gsize;
    a.adm.bits_len = nsym[1];
    a.adm.out = (uint16_t*)d_out;
    a.adm.err = ws->counters + 3;
    prof_mark(ws, s, "decode_restore");
    HIP_TRY(launch_decode_fused(a, ws->cu, s));
    prof_mark(ws, s, nullptr);
    return MANS_OK;
  }
  if (mode == MANS_MODE_ADM) {
    if (nsym[0] != n || nsym[1] > adm_bits_cap(n)) return MANS_ERR_CORRUPT;
    const uint64_t codes_bytes = round_up(n, 16);
    mans_status st = ensure_scratch(ws, codes_bytes + round_up(nsym[1], 16) + 16);
    if (st != MANS_OK) return st;
    codes = ws->scratch;
    bits = ws->scratch + codes_bytes;
  }
  for (uint32_t k = 0; k < P; ++k) {
    if (sect[k] + kSectionFixed > c || nsym[k] > 0xffffffffull) return MANS_ERR_CORRUPT;
    uint8_t* dst = mode == MANS_MODE_ADM ? (k == 0 ? codes : bits) : (uint8_t*)d_out + k;
    const uint32_t stride = mode == MANS_MODE_ADM ? 1u : eb;
    if (mode == MANS_MODE_PLANES && nsym[k] != n) return MANS_ERR_CORRUPT;
    const uint64_t nb = div_up(nsym[k], f.block_bytes);
    if (sect[k] + section_bytes(nb, f) > data_off) return MANS_ERR_CORRUPT;
    a.d.plane[k] = PlaneDesc{nullptr, dst, nsym[k], nb, sect[k], units, stride, 0};
    units += nb;
  }
  a.d.total_units = units;
  a.d.data_off = data_off;
  a.in = (const uint8_t*)d_in;
  a.in_bytes = c;
  a.data_words = data_words;
  a.err = ws->counters + 3;
  a.from_header = 0;
  prof_mark(ws, s, "decode");
  HIP_TRY(launch_decode(a, ws->cu, s));
  if (mode == MANS_MODE_ADM) {
    if (meta + adm_meta_bytes(gsize) > c) return MANS_ERR_CORRUPT;
    AdmUnmapArgs u;
    memset(&u, 0, sizeof(u));
    u.offsets = (const int32_t*)((const uint8_t*)d_in + meta);
    u.centers = (const uint16_t*)((const uint8_t*)d_in + meta + 4ull * (gsize + 1));
    u.codes = codes;
    u.bits = bits;
    u.n = n;
    u.gsize = gsize;
    u.bits_len = nsym[1];
    u.out = (uint16_t*)d_out;
    u.err = ws->counters + 3;
    prof_mark(ws, s, "adm_unmap");
    HIP_TRY(launch_adm_unmap(u, s));
  }
  prof_mark(ws, s, nullptr);
  return MANS_OK;
}

mans_status mans_frame_decode(const void* d_in, uint64_t c, const void* h_header, void* d_out, uint64_t out_capacity,
                              uint64_t* h_n_elems, mans_workspace* ws, void* stream) {
  return frame_decode_impl(d_in, c, h_header, d_out, out_capacity, h_n_elems, ws, stream, false);
}

mans_status mans_frame_decode_fused(const void* d_in, uint64_t c, const void* h_header, void* d_out,
                                    uint64_t out_capacity, uint64_t* h_n_elems, mans_workspace* ws, void* stream) {
  return frame_decode_impl(d_in, c, h_header, d_out, out_capacity, h_n_elems, ws, stream, true);
}

// ---- batches -----------------------------------------------------------------

}  // extern "C"

struct mans_batch {
  mans_workspace* ws = nullptr;
  int pb = 10;
  bool decode = false;
  // capacity
  uint32_t cap_bufs = 0;
  uint64_t cap_items = 0, cap_ticks = 0;
  // current plan
  uint32_t nbufs = 0, nhist = 0, nitems = 0, nticks = 0;
  BatchBuf* bufs = nullptr;   // device [cap_bufs]
  uint2* item_buf = nullptr;  // device [cap_items] queue items
  uint32_t* hist = nullptr;   // [cap_bufs][4][256]
  uint32_t* done = nullptr;   // [cap_bufs]
  uint32_t* ready = nullptr;  // [cap_bufs]
  uint4* tables = nullptr;    // [cap_bufs][4][256]
  uint64_t* status = nullptr; // [cap_ticks]
  uint32_t* ptot = nullptr;   // [cap_bufs][4]
  uint32_t* fin = nullptr;    // [cap_bufs]
  uint32_t* queue = nullptr;  // [4] next item, exited waves, generation, next chunk
  uint32_t* luts = nullptr;   // decode: [cap_bufs][4][2^p]
  BatchBuf* h_bufs = nullptr; // pinned staging for asynchronous plans (or null):
                              // descriptors then items, one upload
  const uint2* items_dev = nullptr;  // the plan's items (item_buf, or after the
                                     // descriptors in `bufs` for pinned plans)
  // the batch's own encoder spill areas, sized at create for every launch
  // shape, so neither a shape change nor the workspace can move them under
  // a captured run
  uint16_t* spill_buf = nullptr;
  uint64_t spill_cap = 0;
  uint32_t* bad = nullptr;    // decode: [cap_bufs] error flags of the last run
  uint64_t* arch_off = nullptr;  // encode: [cap_bufs] frame offsets of the last pack
  uint64_t max_frame = 0;     // encode: largest frame bound of the plan
  uint64_t archive_max = 0;   // encode: archive bound of the plan
  // launch shape, and what the tuning cache keys it by
  BatchShape shape = kBatchEncodeDefault;
  uint64_t total_bytes = 0;
  uint32_t width_mask = 0;
  BatchGrain grain = {kBatchHistChunkMax, 4, 4};  // work-item grain of the plan
  uint32_t lanes = 64;       // interleave width of every frame of the batch (plan-time)
};

namespace {

void batch_free(mans_batch* b) {
  void* dev[] = {b->bufs, b->item_buf, b->hist, b->done, b->ready, b->tables,
                 b->status, b->ptot, b->fin, b->queue, b->luts, b->spill_buf, b->bad, b->arch_off};
  for (void* q : dev)
    if (q) (void)hipFree(q);
  if (b->h_bufs) (void)hipHostFree(b->h_bufs);
  delete b;
}

// A grain forced by mans_batch_set_grain (0: picked per plan).
std::atomic<uint32_t> g_force_chunk{0}, g_force_rounds{0}, g_force_run{0};

// The work-item grain of a batch of `bytes` uncompressed bytes in `units`
// byte-plane blocks on a chip of `cu` compute units: the largest items that
// still leave about 32 histogram chunks, 32 encode tickets and 128 decode
// runs per CU (a launch keeps 8-16 wavefronts per CU busy and the queue
// balances their tails), within the bounds of each.  4096 buffers of 8 GB
// take the largest grain; one request of a few tens of MB the smallest.
BatchGrain batch_grain(uint64_t bytes, uint64_t units, int cu) {
  const uint64_t c = (uint64_t)std::max(cu, 1);
  BatchGrain g{kBatchHistChunkMin, 1, 1};
  while (g.hist_chunk < kBatchHistChunkMax && bytes / (2ull * g.hist_chunk) >= 32 * c) g.hist_chunk *= 2;
  while (g.rounds < 4 && units / (4ull * 2 * g.rounds) >= 32 * c) g.rounds *= 2;
  if (units / (2ull * 4) >= 128 * c) g.run = 4;
  const uint32_t fc = g_force_chunk.load(), fr = g_force_rounds.load(), fu = g_force_run.load();
  if (fc) g.hist_chunk = fc;
  if (fr) g.rounds = fr;
  if (fu) g.run = fu;
  return g;
}

// The grain of a batch of these buffers (elem_bytes 0 counts as 1).
BatchGrain grain_of(const mans_buffer* h_bufs, int32_t count, int cu, int lanes) {
  uint64_t bytes = 0, units = 0;
  for (int32_t i = 0; i < count; ++i) {
    const uint64_t eb = (uint64_t)std::max<int32_t>(h_bufs[i].elem_bytes, 1);
    bytes += h_bufs[i].n_elems * eb;
    units += eb * div_up(h_bufs[i].n_elems, (uint64_t)format_for_lanes(64).block_bytes);
  }
  BatchGrain g = batch_grain(bytes, units, cu);
  if (lanes == 32) g.rounds = 1;  // 32-lane tickets are one pair of units
  return g;
}

// Lays out every buffer of a batch as the byte-plane frame of
// mans_frame_encode in p->lanes' format and counts its histogram items
// (g.hist_chunk bytes) and tickets (decode: two units; 32-lane encode:
// four units, two for uint32; 64-lane encode: g.rounds rounds of
// enc_chains units).
mans_status plan_batch(const mans_params* p, const mans_buffer* h_bufs, int32_t count, bool decode,
                       const BatchGrain& g, std::vector<BatchBuf>& bufs, uint64_t& nhist, uint64_t& nticks) {
  const Format f = format_for_lanes(p->lanes);
  bufs.assign((size_t)count, BatchBuf{});
  nhist = nticks = 0;
  for (int32_t i = 0; i < count; ++i) {
    const mans_buffer& m = h_bufs[i];
    const uint32_t eb = (uint32_t)m.elem_bytes;
    if (eb != 1 && eb != 2 && eb != 4) return MANS_ERR_INVALID_ARG;
    if (m.reserved != 0) return MANS_ERR_INVALID_ARG;
    if ((m.n_elems > 0 && !m.d_out) || (!decode && !m.d_out) || (m.n_elems > 0 && !m.d_in) || (decode && !m.d_in))
      return MANS_ERR_INVALID_ARG;
    if (!aligned(m.d_in, 16) || !aligned(m.d_out, 16)) return MANS_ERR_ALIGNMENT;
    if (m.n_elems > 0xffffffffull) return MANS_ERR_TOO_LARGE;
    mans_params q = *p;
    q.elem_bytes = (int32_t)eb;
    q.mode = MANS_MODE_PLANES;
    if (!decode && m.out_capacity < mans_frame_max_size(m.n_elems, &q)) return MANS_ERR_CAPACITY;
    if (decode && m.out_capacity < m.n_elems * eb) return MANS_ERR_CAPACITY;
    CodecDesc d;
    memset(&d, 0, sizeof(d));
    d.prob_bits = p->prob_bits;
    d.lanes = (int)f.lanes;
    d.num_planes = (int)eb;
    d.order = 0;
    const uint64_t nb = div_up(m.n_elems, f.block_bytes);
    uint64_t off = kFrameHeaderBytes;
    for (uint32_t k = 0; k < eb; ++k) {
      d.plane[k] = PlaneDesc{(const uint8_t*)m.d_in + k, nullptr, m.n_elems, nb, off, 0, eb, 0};
      off += section_bytes(nb, f);
    }
    d.data_off = off;
    BatchBuf& B = bufs[(size_t)i];
    memset(&B, 0, sizeof(B));
    B.in = (const uint8_t*)m.d_in;
    B.out = (uint8_t*)m.d_out;
    B.n = m.n_elems;
    B.nb = nb;
    for (uint32_t k = 0; k < eb; ++k) B.section[k] = d.plane[k].section;
    B.data_off = off;
    B.eb = eb;
    B.hist_first = (uint32_t)nhist;
    B.nhist = (uint32_t)std::max<uint64_t>(1, div_up(m.n_elems * eb, g.hist_chunk));
    B.tick_first = (uint32_t)nticks;
    // decode tickets: two units; encode tickets: g.rounds rounds of four
    // chains (two for uint32), as enc_ticket_units in batch_kernels.hpp
    // units per ticket: decode 2; 32-lane encode 4 (uint32: 2); 64-lane
    // encode rounds x enc_chains
    const uint32_t per = decode ? 2u : (f.lanes == 32 ? (eb == 4 ? 2u : 4u) : g.rounds * (eb == 4 ? 2u : 4u));
    B.nticks = (uint32_t)div_up(eb * nb, per);
    B.lanes = f.lanes;
    B.in_bytes = decode ? m.in_bytes : 0;
    frame_header(B.frame_hdr, m.n_elems, eb, MANS_MODE_PLANES, f, p->prob_bits, d, 0, 0);
    put64(B.frame_hdr, 56, off);  // total size of an empty frame
    nhist += B.nhist;
    nticks += B.nticks;
    if (nhist + nticks + (uint64_t)count > 0x7fffffffull) return MANS_ERR_TOO_LARGE;
  }
  return MANS_OK;
}

// Allocates a batch able to hold cap_bufs buffers, cap_items work items and
// cap_ticks tickets; every self-cleaning array starts at zero and the
// generation at 1.  pinned: keep host staging for asynchronous plans.
mans_status batch_alloc(mans_batch* b, uint32_t cap_bufs, uint64_t cap_items, uint64_t cap_ticks, bool pinned) {
  const size_t nb1 = std::max<size_t>(1, cap_bufs);
  b->cap_bufs = cap_bufs;
  b->cap_items = cap_items;
  b->cap_ticks = cap_ticks;
  // pinned plans upload descriptors and items in one copy: the items follow
  // the plan's descriptors in the same allocation
  const size_t blob = nb1 * sizeof(BatchBuf) + 16 + std::max<size_t>(1, cap_items) * sizeof(uint2);
  if (pinned) {
    HIP_TRY(hipMalloc(&b->bufs, blob));
  } else {
    HIP_TRY(hipMalloc(&b->bufs, nb1 * sizeof(BatchBuf)));
    HIP_TRY(hipMalloc(&b->item_buf, std::max<size_t>(1, cap_items) * sizeof(uint2)));
    b->items_dev = b->item_buf;
  }
  HIP_TRY(hipMalloc(&b->ready, nb1 * 4));
  HIP_TRY(hipMalloc(&b->queue, 16));
  HIP_TRY(hipMemset(b->ready, 0, nb1 * 4));
  const uint32_t q0[4] = {0, 0, 1, 0};  // next item, exited waves, first generation, next chunk
  HIP_TRY(hipMemcpy(b->queue, q0, 16, hipMemcpyHostToDevice));
  if (b->decode) {
    HIP_TRY(hipMalloc(&b->luts, nb1 * 4 * (4ull << b->pb)));
    HIP_TRY(hipMalloc(&b->bad, nb1 * 4));
    HIP_TRY(hipMemset(b->bad, 0, nb1 * 4));
  } else {
    HIP_TRY(hipMalloc(&b->hist, nb1 * 4096));
    HIP_TRY(hipMalloc(&b->done, nb1 * 4));
    HIP_TRY(hipMalloc(&b->tables, nb1 * 4 * 256 * sizeof(uint4)));
    HIP_TRY(hipMalloc(&b->status, std::max<uint64_t>(1, cap_ticks) * 8));
    HIP_TRY(hipMalloc(&b->ptot, nb1 * 16));
    HIP_TRY(hipMalloc(&b->fin, nb1 * 4));
    HIP_TRY(hipMalloc(&b->arch_off, nb1 * 8));
    HIP_TRY(hipMemset(b->hist, 0, nb1 * 4096));
    HIP_TRY(hipMemset(b->done, 0, nb1 * 4));
    HIP_TRY(hipMemset(b->status, 0, std::max<uint64_t>(1, cap_ticks) * 8));
    HIP_TRY(hipMemset(b->ptot, 0, nb1 * 16));
    HIP_TRY(hipMemset(b->fin, 0, nb1 * 4));
  }
  if (pinned) {
    HIP_TRY(hipHostMalloc((void**)&b->h_bufs, blob, hipHostMallocDefault));
  }
  return MANS_OK;
}

// Plans `count` buffers into b (which must hold them) and uploads the
// descriptors: on stream s from the pinned staging when b has one (the
// staging must not be in use by an earlier upload), else synchronously.
mans_status batch_plan(mans_batch* b, const mans_params* p, const mans_buffer* h_bufs, int32_t count, hipStream_t s) {
  std::vector<BatchBuf> bufs;
  uint64_t nhist = 0, nticks = 0;
  const BatchGrain grain = grain_of(h_bufs, count, b->ws->cu, p->lanes);
  mans_status st = plan_batch(p, h_bufs, count, b->decode, grain, bufs, nhist, nticks);
  if (st != MANS_OK) return st;
  const uint64_t nitems = (b->decode ? (uint64_t)count : nhist) + nticks;
  if ((uint32_t)count > b->cap_bufs || nitems > b->cap_items || nticks > b->cap_ticks) return MANS_ERR_TOO_LARGE;
  std::vector<uint2> items;
  const size_t items_off = round_up((uint64_t)count * sizeof(BatchBuf), 16);
  uint2* dst = b->h_bufs ? (uint2*)((char*)b->h_bufs + items_off) : nullptr;
  if (!dst) {
    items.resize(std::max<size_t>(1, nitems));
    dst = items.data();
  }
  uint64_t k = 0;
  if (b->decode) {
    // every buffer's LUT item, then the tickets in buffer order
    for (int32_t i = 0; i < count; ++i) dst[k++] = make_uint2((uint32_t)i, 0u);
    for (int32_t i = 0; i < count; ++i)
      for (uint32_t r = 0; r < bufs[(size_t)i].nticks; ++r) dst[k++] = make_uint2((uint32_t)i, 0u);
  } else {
    // histogram chunks in buffer order, then the tickets in buffer order
    // (the kernel takes each kind from its own counter)
    for (int32_t i = 0; i < count; ++i)
      for (uint32_t r = 0; r < bufs[(size_t)i].nhist; ++r) dst[k++] = make_uint2((uint32_t)i, r);
    for (int32_t i = 0; i < count; ++i)
      for (uint32_t r = 0; r < bufs[(size_t)i].nticks; ++r) dst[k++] = make_uint2((uint32_t)i, kItemTicket | r);
  }
  if (b->h_bufs) {
    if (count > 0) memcpy(b->h_bufs, bufs.data(), bufs.size() * sizeof(BatchBuf));
    const size_t up = items_off + nitems * sizeof(uint2);
    if (up > 0) HIP_TRY(hipMemcpyAsync(b->bufs, b->h_bufs, up, hipMemcpyHostToDevice, s));
    b->items_dev = (const uint2*)((const char*)b->bufs + items_off);
  } else {
    if (count > 0) HIP_TRY(hipMemcpy(b->bufs, bufs.data(), bufs.size() * sizeof(BatchBuf), hipMemcpyHostToDevice));
    if (nitems > 0) HIP_TRY(hipMemcpy(b->item_buf, dst, nitems * sizeof(uint2), hipMemcpyHostToDevice));
  }
  b->pb = p->prob_bits;
  b->total_bytes = 0;
  b->width_mask = 0;
  b->max_frame = 0;
  b->archive_max = kArchiveHeaderBytes + (uint64_t)kArchiveEntryBytes * (uint64_t)count;
  for (int32_t i = 0; i < count; ++i) {
    b->total_bytes += h_bufs[i].n_elems * (uint64_t)h_bufs[i].elem_bytes;
    b->width_mask |= (uint32_t)h_bufs[i].elem_bytes;
    mans_params q = *p;
    q.lanes = p->lanes;
    q.elem_bytes = h_bufs[i].elem_bytes;
    q.mode = MANS_MODE_PLANES;
    const uint64_t fm = mans_frame_max_size(h_bufs[i].n_elems, &q);
    b->max_frame = std::max(b->max_frame, fm);
    b->archive_max += round_up(fm, 16);
  }
  b->nbufs = (uint32_t)count;
  b->grain = grain;
  b->lanes = (uint32_t)p->lanes;
  b->nhist = b->decode ? 0u : (uint32_t)nhist;
  b->nticks = (uint32_t)nticks;
  b->nitems = (uint32_t)nitems;
  return MANS_OK;
}

BatchEncodeArgs batch_args(const mans_batch* b, uint64_t* d_sizes) {
  BatchEncodeArgs a;
  memset(&a, 0, sizeof(a));
  a.bufs = b->bufs;
  a.items = b->items_dev;
  a.nbufs = b->nbufs;
  a.nhist_items = b->nhist;
  a.nitems = b->nitems;
  a.hist = b->hist;
  a.done = b->done;
  a.ready = b->ready;
  a.tables = b->tables;
  a.status = b->status;
  a.ptot = b->ptot;
  a.fin = b->fin;
  a.queue = b->queue;
  a.err = b->ws->counters + 3;
  a.sizes = d_sizes;
  a.spill = b->spill_buf;
  a.spill_bytes = b->spill_cap;
  a.hist_chunk = b->grain.hist_chunk;
  a.rounds = b->grain.rounds;
  a.lanes = b->lanes;
  return a;
}

// ---- launch-shape tuning cache ---------------------------------------------
// Process-wide map from a batch's key to its tuned shape, seeded once from
// the file named by MANS_TUNE_CACHE (lines "<key> <nw> <kb>"; later lines
// win) and appended to by mans_batch_tune.
std::mutex g_tune_mu;
std::map<std::string, BatchShape>* g_tune = nullptr;

uint32_t log2_bucket(uint64_t v) {
  uint32_t k = 0;
  while (v > 1) {
    v >>= 1;
    ++k;
  }
  return k;
}

std::string tune_key(const mans_batch* b) {
  char buf[96];
  snprintf(buf, sizeof(buf), "%s-p%d-l%u-n%u-b%u-w%u-cu%d", b->decode ? "dec" : "enc", b->pb, b->lanes,
           log2_bucket(b->nbufs), log2_bucket(b->total_bytes), b->width_mask, b->ws->cu);
  return buf;
}

std::map<std::string, BatchShape>& tune_cache() {  // g_tune_mu held
  if (!g_tune) {
    g_tune = new std::map<std::string, BatchShape>();
    const char* path = getenv("MANS_TUNE_CACHE");
    FILE* f = path && *path ? fopen(path, "r") : nullptr;
    if (f) {
      char key[96];
      int nw = 0, kb = 0;
      while (fscanf(f, "%95s %d %d", key, &nw, &kb) == 3) (*g_tune)[key] = BatchShape{nw, kb};
      fclose(f);
    }
  }
  return *g_tune;
}

BatchShape lookup_shape(const mans_batch* b) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto& m = tune_cache();
  auto it = m.find(tune_key(b));
  if (it != m.end() && batch_shape_ok(it->second, b->pb, b->decode)) return it->second;
  return b->decode ? kBatchDecodeDefault : kBatchEncodeDefault;
}

void store_shape(const mans_batch* b, BatchShape sh) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  const std::string key = tune_key(b);
  tune_cache()[key] = sh;
  const char* path = getenv("MANS_TUNE_CACHE");
  FILE* f = path && *path ? fopen(path, "a") : nullptr;
  if (f) {
    fprintf(f, "%s %d %d\n", key.c_str(), sh.nw, sh.kb);
    fclose(f);
  }
}

const int kTuneNws[3] = {4, 8, 16}, kTuneEncKbs[3] = {4, 8, 16}, kTuneDecKbs[4] = {1, 4, 8, 16};

// Allocates an encode batch's spill areas once: the most any launch shape
// needs for its plan (nitems_cap work items), or only `only` when given.
mans_status batch_alloc_spill(mans_batch* b, uint32_t nitems_cap, const BatchShape* only) {
  if (b->decode) return MANS_OK;
  BatchEncodeArgs a = batch_args(b, nullptr);
  a.nitems = nitems_cap;
  uint64_t need = 0;
  for (int i = 0; i < 9; ++i) {
    const BatchShape c = only ? *only : BatchShape{kTuneNws[i / 3], kTuneEncKbs[i % 3]};
    if (!batch_shape_ok(c, b->pb, false)) continue;
    uint64_t v = 0;
    if (launch_batch_encode(a, b->pb, c, b->ws->cu, nullptr, &v) != hipSuccess) return MANS_ERR_HIP;
    need = std::max(need, v);
    if (only) break;
  }
  if (need == 0) return MANS_OK;
  HIP_TRY(hipMalloc(&b->spill_buf, need));
  b->spill_cap = need;
  return MANS_OK;
}

// An encode batch's spill areas cover shape sh (always true for a shape
// that fits: they were sized for every shape at create).
mans_status batch_reserve_spill(mans_batch* b, BatchShape sh) {
  if (b->decode || b->nitems == 0) return MANS_OK;
  uint64_t need = 0;
  if (launch_batch_encode(batch_args(b, nullptr), b->pb, sh, b->ws->cu, nullptr, &need) != hipSuccess)
    return MANS_ERR_HIP;
  return need <= b->spill_cap ? MANS_OK : MANS_ERR_INVALID_ARG;
}

// Decode buffers whose elem_bytes is 0 take their element count and width
// from their frame's header: one gather launch and one copy for all of them
// (the frames must be byte-plane frames of the batch's precision).
mans_status read_frame_dims(const mans_params* p, std::vector<mans_buffer>& v, mans_workspace* ws) {
  std::vector<uint32_t> idx;
  std::vector<const uint8_t*> src;
  for (size_t i = 0; i < v.size(); ++i) {
    if (v[i].elem_bytes != 0) continue;
    if (!v[i].d_in || v[i].in_bytes < kFrameHeaderBytes) return MANS_ERR_INVALID_ARG;
    if (!aligned(v[i].d_in, 16)) return MANS_ERR_ALIGNMENT;
    idx.push_back((uint32_t)i);
    src.push_back((const uint8_t*)v[i].d_in);
  }
  if (idx.empty()) return MANS_OK;
  DeviceGuard g(ws->device);
  const size_t n = idx.size();
  const uint8_t** d_src = nullptr;
  uint4* d_hdr = nullptr;
  std::vector<uint8_t> h((size_t)n * kFrameHeaderBytes);
  mans_status st = MANS_OK;
  if (hipMalloc(&d_src, n * sizeof(void*)) != hipSuccess || hipMalloc(&d_hdr, n * kFrameHeaderBytes) != hipSuccess ||
      hipMemcpy(d_src, src.data(), n * sizeof(void*), hipMemcpyHostToDevice) != hipSuccess ||
      launch_gather_headers(d_src, (uint32_t)n, d_hdr, nullptr) != hipSuccess ||
      hipMemcpy(h.data(), d_hdr, n * kFrameHeaderBytes, hipMemcpyDeviceToHost) != hipSuccess)
    st = MANS_ERR_HIP;
  if (d_src) (void)hipFree(d_src);
  if (d_hdr) (void)hipFree(d_hdr);
  if (st != MANS_OK) return st;
  for (size_t k = 0; k < n; ++k) {
    const uint8_t* hb = h.data() + k * kFrameHeaderBytes;
    uint32_t magic;
    uint64_t ne, total;
    memcpy(&magic, hb, 4);
    memcpy(&ne, hb + 16, 8);
    memcpy(&total, hb + 56, 8);
    const uint32_t eb = hb[6], mode = hb[7], lanes = hb[8], pb = hb[9], P = hb[10], order = hb[11];
    mans_buffer& m = v[idx[k]];
    if (magic != kFrameMagic || hb[4] != kFrameVersion || mode != MANS_MODE_PLANES || (int)lanes != p->lanes ||
        (int)pb != p->prob_bits || P != eb || order != 0 || (eb != 1 && eb != 2 && eb != 4) || total > m.in_bytes)
      return MANS_ERR_CORRUPT;
    m.n_elems = ne;
    m.elem_bytes = (int32_t)eb;
  }
  return MANS_OK;
}

mans_status batch_create(mans_batch** out, const mans_params* p, const mans_buffer* h_bufs_in, int32_t count,
                         mans_workspace* ws, bool decode) {
  if (!out) return MANS_ERR_INVALID_ARG;
  *out = nullptr;
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (!ws || count < 0 || (count > 0 && !h_bufs_in) || (p->lanes != 64 && p->lanes != 32))
    return MANS_ERR_INVALID_ARG;
  std::vector<mans_buffer> given(h_bufs_in, h_bufs_in + count);
  if (decode) {
    st = read_frame_dims(p, given, ws);
    if (st != MANS_OK) return st;
  }
  const mans_buffer* h_bufs = given.data();
  std::vector<BatchBuf> bufs;
  uint64_t nhist = 0, nticks = 0;
  st = plan_batch(p, h_bufs, count, decode, grain_of(h_bufs, count, ws->cu, p->lanes), bufs, nhist, nticks);
  if (st != MANS_OK) return st;
  DeviceGuard g(ws->device);
  mans_batch* b = new mans_batch();
  b->ws = ws;
  b->pb = p->prob_bits;
  b->decode = decode;
  st = batch_alloc(b, (uint32_t)count, (decode ? (uint64_t)count : nhist) + nticks, nticks, false);
  if (st == MANS_OK) st = batch_plan(b, p, h_bufs, count, nullptr);
  if (st == MANS_OK && b->nitems > 0) st = batch_alloc_spill(b, b->nitems, nullptr);
  if (st == MANS_OK) {
    b->shape = lookup_shape(b);
    st = batch_reserve_spill(b, b->shape);
  }
  if (st != MANS_OK) {
    batch_free(b);
    return st;
  }
  *out = b;
  return MANS_OK;
}

}  // namespace

// ---- batches for the request server (mans_amd/csrc/serve.cpp) ---------------
mans_status batch_slot_create(mans_batch** out, const mans_params* p, bool decode, uint32_t cap_bufs,
                              uint64_t cap_bytes, mans_workspace* ws) {
  if (!out || !ws) return MANS_ERR_INVALID_ARG;
  *out = nullptr;
  mans_status st = check_params(p);
  if (st != MANS_OK) return st;
  if (p->lanes != 64) return MANS_ERR_INVALID_ARG;
  DeviceGuard g(ws->device);
  mans_batch* b = new mans_batch();
  b->ws = ws;
  b->pb = p->prob_bits;
  b->decode = decode;
  b->shape = decode ? kBatchDecodeDefault : kBatchEncodeDefault;
  // units of cap_bytes in 8 KiB blocks, plus a partial block per plane and
  // buffer; tickets hold two units or more.  Histogram chunks: batch_grain
  // keeps chunks below kBatchHistChunkMax only while there are fewer than
  // 64 per CU, plus a partial chunk per buffer.
  const uint64_t units = cap_bytes / 8192 + 4ull * cap_bufs;
  const uint64_t ticks = units / 2 + cap_bufs;
  const uint64_t hist = cap_bytes / kBatchHistChunkMax + 64ull * (uint64_t)std::max(ws->cu, 1) + cap_bufs;
  st = batch_alloc(b, cap_bufs, (decode ? cap_bufs : hist) + ticks, ticks, true);
  if (st == MANS_OK && !decode) st = batch_alloc_spill(b, 0x7fffffffu, &b->shape);
  if (st != MANS_OK) {
    batch_free(b);
    return st;
  }
  *out = b;
  return MANS_OK;
}

mans_status batch_slot_plan(mans_batch* b, const mans_params* p, const mans_buffer* bufs, int32_t count,
                            hipStream_t s) {
  if (!b) return MANS_ERR_INVALID_ARG;
  DeviceGuard g(b->ws->device);
  return batch_plan(b, p, bufs, count, s);
}

// Device array of a decode batch's per-buffer error flags, and the bits of
// the workspace's sticky error word that only decode launches raise.
const uint32_t* batch_slot_bad(const mans_batch* b) { return b ? b->bad : nullptr; }
uint32_t batch_decode_error_bits() { return kErrMagic | kErrBlockWords | kErrFinalState; }
// The workspace's sticky error word, and its reset enqueued on stream s.
const uint32_t* workspace_error_word(const mans_workspace* ws) { return ws->counters + 3; }
mans_status workspace_clear_error(mans_workspace* ws, hipStream_t s) {
  DeviceGuard g(ws->device);
  HIP_TRY(launch_fill_u32(ws->counters + 3, 1, 0, s));
  return MANS_OK;
}

extern "C" {

mans_status mans_batch_encode_create(mans_batch** out, const mans_params* p, const mans_buffer* h_bufs,
                                     int32_t count, mans_workspace* ws) {
  return batch_create(out, p, h_bufs, count, ws, false);
}

mans_status mans_batch_encode_run(mans_batch* b, uint64_t* d_sizes, void* stream) {
  if (!b || b->decode || (b->nbufs > 0 && !d_sizes)) return MANS_ERR_INVALID_ARG;
  if (b->nitems == 0) return MANS_OK;
  mans_workspace* ws = b->ws;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  prof_mark(ws, s, "batch_encode");
  HIP_TRY(launch_batch_encode(batch_args(b, d_sizes), b->pb, b->shape, ws->cu, s, nullptr));
  prof_mark(ws, s, nullptr);
  return MANS_OK;
}

mans_status mans_batch_decode_create(mans_batch** out, const mans_params* p, const mans_buffer* h_bufs,
                                     int32_t count, mans_workspace* ws) {
  return batch_create(out, p, h_bufs, count, ws, true);
}

mans_status mans_batch_decode_run(mans_batch* b, void* stream) {
  if (!b || !b->decode) return MANS_ERR_INVALID_ARG;
  if (b->nitems == 0) return MANS_OK;
  mans_workspace* ws = b->ws;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  BatchDecodeArgs a;
  memset(&a, 0, sizeof(a));
  a.bufs = b->bufs;
  a.items = b->items_dev;
  a.nbufs = b->nbufs;
  a.nitems = b->nitems;
  a.ready = b->ready;
  a.luts = b->luts;
  a.queue = b->queue;
  a.err = ws->counters + 3;
  a.bad = b->bad;
  a.run = b->grain.run;
  a.lanes = b->lanes;
  // per-buffer flags start at zero in every run (a kernel, so the run can
  // be captured)
  HIP_TRY(launch_fill_u32(b->bad, b->nbufs, 0, s));
  prof_mark(ws, s, "batch_decode");
  HIP_TRY(launch_batch_decode(a, b->pb, b->shape, ws->cu, s));
  prof_mark(ws, s, nullptr);
  return MANS_OK;
}

mans_status mans_batch_set_grain(uint32_t hist_chunk, uint32_t rounds, uint32_t run) {
  const bool chunk_ok = hist_chunk == 0 || (hist_chunk >= kBatchHistChunkMin && hist_chunk <= kBatchHistChunkMax &&
                                           (hist_chunk & (hist_chunk - 1)) == 0);
  if (!chunk_ok || (rounds != 0 && rounds != 1 && rounds != 2 && rounds != 4) || (run != 0 && run != 1 && run != 4))
    return MANS_ERR_INVALID_ARG;
  g_force_chunk.store(hist_chunk);
  g_force_rounds.store(rounds);
  g_force_run.store(run);
  return MANS_OK;
}

uint64_t mans_batch_archive_max_size(const mans_batch* b) { return (b && !b->decode) ? b->archive_max : 0; }

mans_status mans_batch_pack(const mans_batch* b, const uint64_t* d_sizes, void* d_archive, uint64_t capacity,
                            uint64_t* d_archive_size, void* stream) {
  if (!b || b->decode || !d_archive || (b->nbufs > 0 && !d_sizes)) return MANS_ERR_INVALID_ARG;
  if (!aligned(d_archive, 16)) return MANS_ERR_ALIGNMENT;
  if (capacity < b->archive_max) return MANS_ERR_CAPACITY;
  DeviceGuard g(b->ws->device);
  ArchiveArgs a;
  memset(&a, 0, sizeof(a));
  a.bufs = b->bufs;
  a.sizes = d_sizes;
  a.offsets = b->arch_off;
  a.out = (uint8_t*)d_archive;
  a.d_size = d_archive_size;
  a.data_off = kArchiveHeaderBytes + (uint64_t)kArchiveEntryBytes * b->nbufs;
  a.count = b->nbufs;
  HIP_TRY(launch_archive_pack(a, b->max_frame, (hipStream_t)stream));
  return MANS_OK;
}

mans_status mans_archive_info(const void* h_head, uint64_t head_bytes, uint32_t* count, uint64_t* total_size) {
  if (!h_head || head_bytes < kArchiveHeaderBytes) return MANS_ERR_INVALID_ARG;
  const uint8_t* h = (const uint8_t*)h_head;
  uint32_t magic, n;
  uint64_t total, index_off, data_off;
  memcpy(&magic, h, 4);
  memcpy(&n, h + 8, 4);
  memcpy(&total, h + 16, 8);
  memcpy(&index_off, h + 24, 8);
  memcpy(&data_off, h + 32, 8);
  if (magic != kArchiveMagic || h[4] != 1 || index_off != kArchiveHeaderBytes ||
      data_off != kArchiveHeaderBytes + (uint64_t)kArchiveEntryBytes * n || total < data_off)
    return MANS_ERR_CORRUPT;
  if (count) *count = n;
  if (total_size) *total_size = total;
  return MANS_OK;
}

mans_status mans_batch_decode_create_archive(mans_batch** out, const mans_params* p, const void* d_archive,
                                             uint64_t archive_bytes, void* const* h_outs, const uint64_t* h_out_caps,
                                             mans_workspace* ws) {
  if (!out || !p || !d_archive || !ws) return MANS_ERR_INVALID_ARG;
  *out = nullptr;
  if (!aligned(d_archive, 16)) return MANS_ERR_ALIGNMENT;
  if (archive_bytes < kArchiveHeaderBytes) return MANS_ERR_CORRUPT;
  DeviceGuard g(ws->device);
  uint8_t head[kArchiveHeaderBytes];
  HIP_TRY(hipMemcpy(head, d_archive, kArchiveHeaderBytes, hipMemcpyDeviceToHost));
  uint32_t n = 0;
  uint64_t total = 0;
  mans_status st = mans_archive_info(head, kArchiveHeaderBytes, &n, &total);
  if (st != MANS_OK) return st;
  if (total > archive_bytes || n > 0x7fffffffu) return MANS_ERR_CORRUPT;
  if (n > 0 && (!h_outs || !h_out_caps)) return MANS_ERR_INVALID_ARG;
  std::vector<uint64_t> idx((size_t)n * 4);
  if (n) HIP_TRY(hipMemcpy(idx.data(), (const uint8_t*)d_archive + kArchiveHeaderBytes,
                           (size_t)n * kArchiveEntryBytes, hipMemcpyDeviceToHost));
  std::vector<mans_buffer> bufs(n);
  const uint64_t data_off = kArchiveHeaderBytes + (uint64_t)kArchiveEntryBytes * n;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t off = idx[4 * i], size = idx[4 * i + 1], ne = idx[4 * i + 2];
    const uint32_t eb = (uint32_t)idx[4 * i + 3];
    if ((off & 15) || off < data_off || size > total || off > total - size || (idx[4 * i + 3] >> 32) != 0)
      return MANS_ERR_CORRUPT;
    if (eb != 1 && eb != 2 && eb != 4) return MANS_ERR_CORRUPT;
    mans_buffer& m = bufs[i];
    m.d_in = (const uint8_t*)d_archive + off;
    m.in_bytes = size;
    m.d_out = h_outs[i];
    m.out_capacity = h_out_caps[i];
    m.n_elems = ne;
    m.elem_bytes = (int32_t)eb;
    m.reserved = 0;
  }
  return batch_create(out, p, bufs.data(), (int32_t)n, ws, true);
}

mans_status mans_batch_decode_flags(const mans_batch* b, uint32_t* h_flags, int32_t max, void* stream) {
  if (!b || !b->decode || max < 0 || (max > 0 && !h_flags)) return MANS_ERR_INVALID_ARG;
  DeviceGuard g(b->ws->device);
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = std::min<uint32_t>(b->nbufs, (uint32_t)max);
  HIP_TRY(hipStreamSynchronize(s));
  if (n > 0) HIP_TRY(hipMemcpy(h_flags, b->bad, n * 4ull, hipMemcpyDeviceToHost));
  return MANS_OK;
}

mans_status mans_batch_get_shape(const mans_batch* b, mans_batch_shape* out) {
  if (!b || !out) return MANS_ERR_INVALID_ARG;
  *out = mans_batch_shape{64 * b->shape.nw, b->shape.kb, (int32_t)b->lanes, 0};
  return MANS_OK;
}

mans_status mans_batch_set_shape(mans_batch* b, const mans_batch_shape* sh) {
  if (!b || !sh || sh->lanes != (int32_t)b->lanes || sh->reserved != 0 || sh->wg_threads % 64 != 0)
    return MANS_ERR_INVALID_ARG;
  const BatchShape c{sh->wg_threads / 64, sh->syms_per_lane};
  if (!batch_shape_ok(c, b->pb, b->decode)) return MANS_ERR_INVALID_ARG;
  DeviceGuard g(b->ws->device);
  const mans_status st = batch_reserve_spill(b, c);
  if (st == MANS_OK) b->shape = c;
  return st;
}

mans_status mans_batch_tune(mans_batch* b, void* stream, int32_t reps, mans_tune_entry* h_entries,
                            int32_t max_entries, int32_t* n_entries) {
  if (!b || reps < 1 || max_entries < 0 || (max_entries > 0 && !h_entries)) return MANS_ERR_INVALID_ARG;
  // candidates: workgroup size x symbols per lane (decode adds 1: the ring
  // looked after every step)
  const int* kNws = kTuneNws;
  const int nkb = b->decode ? 4 : 3;
  const int* kbs = b->decode ? kTuneDecKbs : kTuneEncKbs;
  const int ncand = 3 * nkb;
  if (n_entries) *n_entries = ncand;
  mans_workspace* ws = b->ws;
  DeviceGuard g(ws->device);
  hipStream_t s = (hipStream_t)stream;
  const BatchShape keep = b->shape;
  mans_status st = MANS_OK;
  uint64_t* d_sizes = nullptr;
  if (!b->decode) HIP_TRY(hipMalloc(&d_sizes, std::max<uint32_t>(1, b->nbufs) * 8ull));
  // every candidate's spill first, so no run sees the workspace change
  for (int i = 0; i < ncand && st == MANS_OK; ++i) {
    const BatchShape c{kNws[i / nkb], kbs[i % nkb]};
    if (batch_shape_ok(c, b->pb, b->decode)) st = batch_reserve_spill(b, c);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == MANS_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = MANS_ERR_HIP;
  double best = -1.0;
  BatchShape pick = keep;
  for (int i = 0; i < ncand && st == MANS_OK; ++i) {
    const BatchShape c{kNws[i / nkb], kbs[i % nkb]};
    const bool fits = batch_shape_ok(c, b->pb, b->decode);
    double ms = fits && b->nitems == 0 ? 0.0 : -1.0;
    if (fits && b->nitems > 0) {
      b->shape = c;
      auto run = [&]() { return b->decode ? mans_batch_decode_run(b, s) : mans_batch_encode_run(b, d_sizes, s); };
      st = run();
      if (st == MANS_OK && hipEventRecord(e0, s) != hipSuccess) st = MANS_ERR_HIP;
      for (int32_t r = 0; r < reps && st == MANS_OK; ++r) st = run();
      if (st == MANS_OK && hipEventRecord(e1, s) != hipSuccess) st = MANS_ERR_HIP;
      float t = 0;
      if (st == MANS_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        st = MANS_ERR_HIP;
      ms = (double)t / reps;
      if (st == MANS_OK && (best < 0 || ms < best)) {
        best = ms;
        pick = c;
      }
    }
    if (i < max_entries) h_entries[i] = mans_tune_entry{mans_batch_shape{64 * c.nw, c.kb, (int32_t)b->lanes, 0}, ms};
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d_sizes) (void)hipFree(d_sizes);
  b->shape = st == MANS_OK ? pick : keep;
  if (st == MANS_OK && best >= 0) store_shape(b, pick);
  return st;
}

mans_status mans_batch_destroy(mans_batch* b) {
  if (!b) return MANS_OK;
  DeviceGuard g(b->ws->device);
  (void)hipDeviceSynchronize();
  batch_free(b);
  return MANS_OK;
}

mans_status mans_encode_batch(const mans_params* p, const mans_buffer* h_bufs, int32_t count, uint64_t* d_sizes,
                              mans_workspace* ws, void* stream) {
  mans_batch* b = nullptr;
  mans_status st = mans_batch_encode_create(&b, p, h_bufs, count, ws);
  if (st != MANS_OK) return st;
  st = mans_batch_encode_run(b, d_sizes, stream);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess && st == MANS_OK) st = MANS_ERR_HIP;
  mans_batch_destroy(b);
  return st;
}

mans_status mans_decode_batch(const mans_params* p, const mans_buffer* h_bufs, int32_t count, mans_workspace* ws,
                              void* stream) {
  mans_batch* b = nullptr;
  mans_status st = mans_batch_decode_create(&b, p, h_bufs, count, ws);
  if (st != MANS_OK) return st;
  st = mans_batch_decode_run(b, stream);
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess && st == MANS_OK) st = MANS_ERR_HIP;
  mans_batch_destroy(b);
  return st;
}

mans_status mans_set_profiling(mans_workspace* ws, int32_t enable) {
  if (!ws) return MANS_ERR_INVALID_ARG;
  ws->profiling = enable != 0;
  return MANS_OK;
}

int32_t mans_last_kernel_times(mans_workspace* ws, double* h_ms, const char** h_names, int32_t max) {
  if (!ws || !ws->profiling) return 0;
  DeviceGuard g(ws->device);
  int32_t k = 0;
  for (int i = 0; i + 1 < ws->ev_used && k < max; ++i) {
    if (!ws->ev_names[i]) continue;
    float ms = 0;
    if (hipEventSynchronize(ws->ev[i + 1]) != hipSuccess) break;
    if (hipEventElapsedTime(&ms, ws->ev[i], ws->ev[i + 1]) != hipSuccess) break;
    if (h_ms) h_ms[k] = ms;
    if (h_names) h_names[k] = ws->ev_names[i];
    ++k;
  }
  ws->ev_used = 0;
  ws->ev_names.clear();
  return k;
}

}  // extern "C"
