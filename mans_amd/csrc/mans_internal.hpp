// mans_internal.hpp -- constants, container layout and kernel launch
// descriptors shared by the HIP translation units of libmans.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mans {

// ---- rANS constants (reference include/ans/GpuANSCodec.h:18-61) ----------
constexpr uint32_t kNumSymbols = 256;
constexpr uint32_t kStateBits = 31;                 // state < 2^31
constexpr uint32_t kMinState = 1u << 15;            // start state == lower bound
constexpr uint32_t kAnsMagic = 0xd00du;
constexpr uint32_t kAnsHeaderBytes = 32;            // ANSCoalescedHeader
constexpr uint32_t kProbsBytes = 2 * kNumSymbols;   // u16 pdf[256]
constexpr uint32_t kSectionFixed = kAnsHeaderBytes + kProbsBytes;  // 544

// ---- frame container (this project's multi-plane wrapper) -----------------
constexpr uint32_t kFrameMagic = 0x534E414Du;  // "MANS"
constexpr uint32_t kFrameVersion = 1;
constexpr uint32_t kFrameHeaderBytes = 128;
constexpr int kMaxPlanes = 4;

// ---- ADM constants (mapping/cpu/mappingcpuV1.cpp:12-16) --------------------
constexpr uint32_t kAdmLanes = 32;
constexpr uint32_t kAdmChunk = 16;
constexpr uint32_t kAdmGroup = kAdmLanes * kAdmChunk;  // 512 elements
constexpr uint32_t kAdmTicketGroups = 32;              // groups per map ticket (one look-back entry; on 512 MiB 16: 0.55 ms, 32 and 64: 0.46 ms)
constexpr uint32_t kAdmMod = 126;
constexpr uint32_t kAdmLaneBytes = 32;                 // 256-bit lane buffer

__host__ __device__ constexpr uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
__host__ __device__ constexpr uint64_t round_up(uint64_t a, uint64_t b) { return div_up(a, b) * b; }

struct Format {
  uint32_t lanes;
  uint32_t block_bytes;
  uint32_t version;
};
__host__ __device__ inline Format format_for_lanes(int lanes) {
  return lanes == 64 ? Format{64u, 8192u, 2u} : Format{32u, 4096u, 1u};
}

// Bound on the 16-bit words one interleaved lane emits for nsym symbols:
// each step grows the virtual state x*2^(16w) by at most
// (2^p/pdf)(1 + 2^(p-15)) <= 2^(p+0.1), it starts at 2^15 and ends >= 2^15,
// so 16w <= nsym (p + 0.1).  One word of slack.
__host__ __device__ constexpr uint32_t lane_max_words(uint32_t nsym, int p) {
  return (uint32_t)(((uint64_t)nsym * (uint64_t)(10 * p + 1) + 159u) / 160u) + 1u;
}
__host__ __device__ constexpr uint32_t unit_max_words(uint32_t lanes, uint32_t block_bytes, int p) {
  return lanes * lane_max_words(block_bytes / lanes, p);
}
// LDS staging words for one wavefront: 64 lanes in either format.
__host__ __device__ constexpr uint32_t wave_stage_words(int p) {
  return (uint32_t)round_up(64u * lane_max_words(128u, p), 8u);
}

__host__ __device__ inline uint64_t section_bytes(uint64_t nb, Format f) {
  return kSectionFixed + 4ull * f.lanes * nb + 8ull * round_up(nb, 2);
}

// One byte plane coded as one ANS "plane section" (header, probs, states,
// block words) whose data lives in the shared data region.
struct PlaneDesc {
  const uint8_t* src;   // encode: symbol i = src[i * stride]
  uint8_t* dst;         // decode: symbol i -> dst[i * stride]
  uint64_t nsym;        // symbols in the plane
  uint64_t nb;          // blocks = ceil(nsym / block_bytes)
  uint64_t section;     // byte offset of the plane's ANS header in the container
  uint64_t unit_base;   // first unit index of the plane (concatenated order)
  uint32_t stride;
  uint32_t pad_;
};

// Unit u -> (plane, block): order 0 interleaves planes (u % P, u / P) so one
// wavefront's units read the same input bytes; order 1 concatenates them.
struct CodecDesc {
  PlaneDesc plane[kMaxPlanes];
  int32_t num_planes;
  int32_t order;
  int32_t prob_bits;
  int32_t lanes;
  uint64_t total_units;
  uint64_t data_off;    // byte offset of the shared data region
};

__device__ __forceinline__ void unit_to_plane(const CodecDesc& d, uint64_t u, int& k, uint64_t& b) {
  if (d.order == 0) {
    uint32_t P = (uint32_t)d.num_planes;
    k = (int)(u % P);
    b = u / P;
  } else {
    k = 0;
    while (k + 1 < d.num_planes && u >= d.plane[k + 1].unit_base) ++k;
    b = u - d.plane[k].unit_base;
  }
}

// Device-side error flags (workspace).
enum : uint32_t {
  kErrMagic = 1u << 0,
  kErrPrecision = 1u << 1,
  kErrTable = 1u << 2,
  kErrBlockWords = 1u << 3,
  kErrUnderflow = 1u << 4,
  kErrFinalState = 1u << 5,
  kErrCapacity = 1u << 6,
  kErrStageOverflow = 1u << 7,
  kErrAdmDomain = 1u << 8,
};

// Histogram sources: bytes of src are grouped in elements of `stride` bytes;
// byte j of each element feeds plane plane0 + j (stride 1: plane plane0).
struct HistSource {
  const uint8_t* ptr;
  uint64_t nbytes;
  uint32_t stride;
  uint32_t plane0;
};

struct HistArgs {
  HistSource src[2];
  int32_t nsrc;
  int32_t num_planes;
  int32_t prob_bits;
  int32_t write_headers;         // 0: only tables (stage API)
  uint64_t plane_nsym[kMaxPlanes];
  uint64_t plane_nb[kMaxPlanes];
  uint64_t plane_section[kMaxPlanes];
  uint32_t* hist;                // [P][256] self-cleaning accumulators
  uint32_t* done;                // self-cleaning WG counter
  uint32_t* grab;                // [2] self-cleaning row-chunk counters, one per source
  uint4* table;                  // [P][256] {pdf, cdf, magic, shift}
  uint32_t* hist_out;            // optional copy of the final histogram
  uint8_t* out;                  // container base (headers), may be null
  uint32_t version;
  uint32_t pad_;
  // finisher also prepares the encode launch:
  uint64_t* status;              // look-back status to clear
  uint64_t status_count;
  uint32_t* ticket;              // encode ticket counter to clear
  uint32_t* d_out_size32;        // written when there is nothing to encode
  uint64_t* d_out_size64;
  uint64_t empty_size;           // container size when total units == 0
  uint64_t total_units;
  int32_t has_frame;             // copy frame_hdr (128 bytes) to out[0..128)
  uint32_t frame_hdr[32];
};

struct EncodeArgs {
  CodecDesc d;
  const uint4* table;            // [P][256]
  uint8_t* out;
  uint64_t* status;              // [chain] look-back words
  uint32_t* ticket;
  uint32_t* err;
  uint32_t* d_out_size32;
  uint64_t* d_out_size64;
  int32_t frame;                 // 1: patch data_words / total_size in the frame header
  uint16_t* spill;               // per-wave spill areas of the 64-lane encoder
  uint64_t spill_bytes;          // their capacity
};

struct AdmUnmapArgs {
  const int32_t* offsets;
  const uint16_t* centers;
  const uint8_t* codes;
  const uint8_t* bits;
  uint64_t n;
  uint64_t gsize;
  uint64_t bits_len;     // bytes in the bit plane
  uint16_t* out;
  uint32_t* err;         // kErrBlockWords on inconsistent offsets
};

struct DecodeArgs {
  CodecDesc d;                   // when from_header: plane 0 section/dst/stride only
  const uint8_t* in;
  uint64_t in_bytes;             // 0 = unknown
  uint64_t out_capacity;         // 0 = unchecked
  uint64_t data_words;           // frames: from the header; single stream: 0 -> header
  uint32_t* err;
  int32_t from_header;           // single stream: nb / nsym / version read on device
  AdmUnmapArgs adm;              // fused ADM restore (launch_decode_fused): plane 0 is the code plane
};

struct AdmMapArgs {
  const uint16_t* in;
  uint64_t n;
  uint64_t gsize;
  int32_t* offsets;      // [gsize + 1]
  uint16_t* centers;     // [gsize]
  uint8_t* codes;        // [n]
  uint8_t* bits;         // bit plane
  uint64_t bits_capacity;
  uint64_t* status;      // look-back words [gsize]
  uint32_t* ticket;
  uint32_t* err;         // kErrAdmDomain / kErrCapacity
};


// Launchers (one per translation unit).  Each returns hipSuccess or the
// launch error.  `cu` is the device's compute-unit count.
// ---- batched byte-plane frames --------------------------------------------
// One buffer of a batch: its frame layout is that of mans_frame_encode in
// MANS_MODE_PLANES with the 64-lane format.
struct BatchBuf {
  const uint8_t* in;             // elements (16-byte aligned)
  uint8_t* out;                  // frame (16-byte aligned)
  uint64_t n;                    // elements
  uint64_t nb;                   // blocks per plane
  uint64_t section[kMaxPlanes];  // plane section offsets in the frame
  uint64_t data_off;             // data region offset in the frame
  uint32_t eb;                   // element bytes = byte planes (1, 2, 4)
  uint32_t hist_first;           // first histogram item of the buffer
  uint32_t nhist;                // histogram items (at least 1)
  uint32_t tick_first;           // first encode ticket (ticket = 2 units)
  uint32_t nticks;               // encode (decode) tickets
  uint32_t lanes;                // 64 (native frames) or 32 (reference-format streams)
  uint64_t in_bytes;             // decode: bytes readable at `in` (the frame)
  uint32_t frame_hdr[32];        // frame header (data_words / total_size set by the kernel)
};

// A queue item: {buffer, kItemTicket | ticket, or histogram chunk}.
constexpr uint32_t kItemTicket = 1u << 31;
struct BatchEncodeArgs {
  const BatchBuf* bufs;
  const uint2* items;            // [nitems] histogram chunks and encode tickets
  uint32_t nbufs;
  uint32_t nhist_items;          // items [0, nhist_items): histogram chunks in buffer order
  uint32_t nitems;               // then the encode tickets in buffer order
  uint32_t* hist;                // [nbufs][4][256] accumulators (left zeroed)
  uint32_t* done;                // [nbufs] histogram items finished (left zeroed)
  uint32_t* ready;               // [nbufs] generation whose tables are ready
  uint4* tables;                 // [nbufs][4][256] encoder table entries
  uint64_t* status;              // [tickets] look-back words tagged with gen
  uint32_t* ptot;                // [nbufs][4] words per plane (left zeroed)
  uint32_t* fin;                 // [nbufs] encode tickets finished (left zeroed)
  uint32_t* queue;               // [4] next ticket, exited wavefronts (both left
                                 // zeroed), launch generation 1 .. kBatchGenMask
                                 // (advanced by the last wavefront out), next
                                 // histogram chunk (left zeroed)
  uint32_t* err;
  uint64_t* sizes;               // [nbufs] frame sizes out
  uint16_t* spill;               // per-wave spill areas
  uint64_t spill_bytes;
  uint32_t hist_chunk;           // bytes per histogram item (BatchGrain)
  uint32_t rounds;               // rounds per encode ticket (BatchGrain; 64 lanes)
  uint32_t lanes;                // every buffer's frame: 64 or 32 lanes
  uint32_t hist_wgs;             // workgroups that start on histogram chunks (set at launch)
};
// Batched decode: BatchBuf.in is the frame, .out the elements; items
// [0, nbufs) build buffer b's LUTs, then come the decode tickets.
struct BatchDecodeArgs {
  const BatchBuf* bufs;
  const uint2* items;            // [nitems] {buffer, 0}: LUT items, then tickets
  uint32_t nbufs;
  uint32_t nitems;
  uint32_t* ready;               // [nbufs] generation whose LUTs are built (| kBatchBad: frame rejected)
  uint32_t* luts;                // [nbufs][4][2^p] packed LUT entries
  uint32_t* queue;               // [4] as BatchEncodeArgs::queue (next item, exited, generation)
  uint32_t* err;
  uint32_t* bad;                 // [nbufs] the error flags of each buffer (zeroed before the launch)
  uint32_t run;                  // items a wavefront takes per queue grab (BatchGrain)
  uint32_t lanes;                // every buffer's frame: 64 or 32 lanes
};
constexpr uint32_t kBatchBad = 1u << 31;
constexpr uint32_t kBatchGenMask = 0x3fffu;     // 14-bit generations
// Work-item grain of a batched launch, picked per plan from the batch's
// size against the chip (batch_grain in capi.hip): a large batch takes big
// items (fewer queue atomics, table loads and look-backs per byte), a small
// one small items, so that a few requests still spread over every CU.  The
// grain changes no output byte.
constexpr uint32_t kBatchHistChunkMin = 1u << 16;  // bytes per histogram item
constexpr uint32_t kBatchHistChunkMax = 1u << 19;
struct BatchGrain {
  uint32_t hist_chunk;  // power of two in [kBatchHistChunkMin, kBatchHistChunkMax]
  uint32_t rounds;      // encode ticket rounds: 1, 2 or 4
  uint32_t run;         // decode items per grab: 1 or 4
};

// ---- archive: the frames of a batch in one container (archive.hip) --------
constexpr uint32_t kArchiveMagic = 0x414E414Du;  // "MANA"
constexpr uint32_t kArchiveHeaderBytes = 64;
constexpr uint32_t kArchiveEntryBytes = 32;
struct ArchiveArgs {
  const BatchBuf* bufs;          // frame i at bufs[i].out, its elements / widths
  const uint64_t* sizes;         // [count] frame bytes
  uint64_t* offsets;             // [count] scratch: frame offsets in the archive
  uint8_t* out;                  // the archive
  uint64_t* d_size;              // archive bytes out (optional)
  uint64_t data_off;             // 64 + 32 * count
  uint32_t count;
  uint32_t pad_;
};
hipError_t launch_archive_pack(const ArchiveArgs& a, uint64_t max_frame, hipStream_t s);
// 128-byte headers of the frames at src[i] (a device array) into out[8 * i ..]
hipError_t launch_gather_headers(const uint8_t* const* src, uint32_t count, uint4* out, hipStream_t s);

// ---- dual-quantisation Lorenzo front end (lorenzo.hip) ----------------------
// outliers a thread of the Lorenzo forward launch holds in registers; a tile
// with a thread over it goes to the fix-up launch
constexpr int kLorenzoHold = 5;
struct LorenzoArgs {
  const float* field;            // forward: f[nz][ny][nx]
  uint16_t* codes;               // forward: out; inverse: in
  uint32_t nz, ny, nx;
  uint32_t radius;
  float inv2eb;                  // float32(1 / (2 eb))
  float twoeb;                   // float32(2 eb)
  // forward: outliers in tile order
  uint32_t* out_idx;
  int32_t* out_val;
  uint64_t out_cap;
  uint64_t* count;               // outliers found (device)
  uint64_t* status;              // [ntiles] look-back words (zeroed before)
  uint32_t* ticket;              // tile ticket (zeroed before)
  uint64_t* tile_off;            // [ntiles] first outlier of each tile
  uint32_t* fix_list;            // [1 + fix_cap]: count (zeroed before), then tiles for the fix-up launch
  uint32_t* fix_pos;             // [fix_cap][256]: each thread's first outlier inside a listed tile
  uint32_t* fix_grp;             // [fix_cap][256]: its outliers before planes 4, 8, 12 (a byte each)
  uint32_t ntiles;
  uint32_t fix_cap;              // >= 1
  uint32_t* err;                 // kErrCapacity: outliers beyond out_cap; kErrAdmDomain: |q| >= 2^28
  // inverse
  const uint32_t* in_idx;
  const int32_t* in_val;
  uint64_t n_out;
  int32_t* q;                    // [n] scratch
  float* recon;                  // [n] out
};
hipError_t launch_lorenzo(const LorenzoArgs& a, hipStream_t s);
hipError_t launch_lorenzo_inverse(const LorenzoArgs& a, int cu, hipStream_t s);
uint32_t lorenzo_tiles(uint32_t nz, uint32_t ny, uint32_t nx);

hipError_t launch_hist(const HistArgs& a, int cu, hipStream_t s);

// Launch shape of the batched kernels; no choice changes an output bit.
//   nw: wavefronts per workgroup (4, 8, 16: 256 / 512 / 1024 threads)
//   kb: symbols per lane per batch.  Encode (4, 8, 16): the elements each
//       lane loads ahead in registers; decode (1, 4, 8, 16): the steps
//       between two upkeeps of the word ring (which holds
//       2 x max(512, 64 kb) words per chain)
struct BatchShape {
  int nw;
  int kb;
};
constexpr BatchShape kBatchEncodeDefault = {4, 16};
constexpr BatchShape kBatchDecodeDefault = {4, 8};
// LDS bytes per workgroup of a shape (encode: table | ring | dummies per
// wave; decode: two LUTs | two rings per wave)
constexpr uint64_t batch_encode_lds(BatchShape sh) { return (uint64_t)sh.nw * (2u * 256u * 4u + 1024u + 64u) * 4u; }
constexpr uint64_t batch_decode_lds(BatchShape sh, int pb) {
  return (uint64_t)sh.nw * (2ull * (1u << pb) * 4u +
                            2ull * 2u * ((2u * (64u * sh.kb > 512u ? 64u * sh.kb : 512u)) + 64u));
}
constexpr uint64_t kLdsBytes = 160u << 10;
bool batch_shape_ok(BatchShape sh, int pb, bool decode);

// batched encode; with need_spill set, only reports the spill bytes needed
hipError_t launch_batch_encode(const BatchEncodeArgs& a, int pb, BatchShape sh, int cu, hipStream_t s,
                               uint64_t* need_spill);
hipError_t launch_batch_decode(const BatchDecodeArgs& a, int pb, BatchShape sh, int cu, hipStream_t s);
// one translation unit per precision
hipError_t launch_batch_encode_p9(const BatchEncodeArgs&, BatchShape, int, hipStream_t, uint64_t*);
hipError_t launch_batch_encode_p10(const BatchEncodeArgs&, BatchShape, int, hipStream_t, uint64_t*);
hipError_t launch_batch_encode_p11(const BatchEncodeArgs&, BatchShape, int, hipStream_t, uint64_t*);
hipError_t launch_batch_decode_p9(const BatchDecodeArgs&, BatchShape, int, hipStream_t);
hipError_t launch_batch_decode_p10(const BatchDecodeArgs&, BatchShape, int, hipStream_t);
hipError_t launch_batch_decode_p11(const BatchDecodeArgs&, BatchShape, int, hipStream_t);
hipError_t launch_encode(const EncodeArgs& a, int cu, hipStream_t s);
// bytes of spill area launch_encode needs for a (0 for the 32-lane format)
hipError_t encode_spill_bytes(const EncodeArgs& a, int cu, uint64_t* bytes);
hipError_t launch_decode(const DecodeArgs& a, int cu, hipStream_t s);
// 64-lane ADM frames: decodes plane 0 (the code plane) and restores each
// block's elements from a.adm (bit plane already decoded) in the same wave.
hipError_t launch_decode_fused(const DecodeArgs& a, int cu, hipStream_t s);
hipError_t launch_adm_map(const AdmMapArgs& a, int cu, hipStream_t s);
hipError_t launch_adm_unmap(const AdmUnmapArgs& a, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

}  // namespace mans
