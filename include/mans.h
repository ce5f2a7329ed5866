/*
 * mans.h -- C ABI of the MI355X multi-byte ANS codec (libmans.so).
 *
 * Plain C, plain pointers and sizes.  Every data pointer is a *device*
 * pointer unless its name starts with h_.  Streams are hipStream_t passed as
 * void* (NULL = the legacy default stream).  All calls are stream-ordered and
 * return a status; none allocates device memory except the workspace calls
 * (mans_workspace_create / mans_workspace_reserve, and the first call that
 * needs more workspace than reserved).
 *
 * Two wire formats share one container layout (reference
 * include/ans/GpuANSCodec.h:68-209):
 *   lanes = 32 -> the reference format: 32 interleaved rANS states per
 *                 4 KiB block, version 1; byte-identical to the reference's
 *                 ansEncode output (its uninitialised bytes defined as 0).
 *   lanes = 64 -> the native format: 64 states per 8 KiB block (one gfx950
 *                 wavefront per block), version 2.
 */
#ifndef MANS_H
#define MANS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MANS_ABI_VERSION 1

typedef enum {
  MANS_OK = 0,
  MANS_ERR_INVALID_ARG = 1,  /* null pointer, bad enum value */
  MANS_ERR_PRECISION = 2,    /* prob_bits outside {9, 10, 11} */
  MANS_ERR_ALIGNMENT = 3,    /* pointer not aligned as required */
  MANS_ERR_TOO_LARGE = 4,    /* size beyond the 32-bit fields of the format */
  MANS_ERR_CAPACITY = 5,     /* output buffer smaller than required */
  MANS_ERR_CORRUPT = 6,      /* malformed / truncated / inconsistent stream */
  MANS_ERR_HIP = 7,          /* HIP runtime error */
  MANS_ERR_NO_DEVICE = 8,    /* no usable gfx950 device */
  MANS_ERR_DOMAIN = 9,       /* ADM mapping requested for data outside its domain */
  MANS_ERR_BUDGET = 10,      /* request exceeds the configured device-memory budget */
  MANS_ERR_BUSY = 11         /* serving queue full */
} mans_status;

typedef enum {
  MANS_MODE_PLANES = 0,      /* element byte k -> plane k (k < elem_bytes) */
  MANS_MODE_ADM = 1,         /* uint16 ADM mapping: code plane + bit-signal plane */
  MANS_MODE_AUTO = 2         /* ADM when in its lossless domain, else planes */
} mans_mode;

typedef struct {
  int32_t prob_bits;   /* 9, 10 or 11 (reference GpuANSEncode.h:348-360) */
  int32_t lanes;       /* 32 = reference format, 64 = native format */
  int32_t elem_bytes;  /* frame API: 1, 2 or 4 */
  int32_t mode;        /* frame API: mans_mode */
} mans_params;

typedef struct mans_workspace mans_workspace;

/* ---- library / workspace ---------------------------------------------- */
const char* mans_status_string(mans_status s);
int32_t mans_abi_version(void);

/* Creates a workspace on `device` (-1 = current device).  A workspace holds
 * the histogram / table / look-back scratch of one call at a time: calls
 * that overlap on the same workspace race; use one workspace per stream.
 * Once a call has run on a stream under HIP graph capture, the workspace
 * never frees or replaces its buffers (a graph may hold them): a later call
 * or mans_workspace_reserve that would need larger ones returns
 * MANS_ERR_CAPACITY, so reserve before capturing. */
mans_status mans_workspace_create(mans_workspace** ws, int32_t device);
mans_status mans_workspace_destroy(mans_workspace* ws);
/* Pre-sizes the workspace for inputs of up to max_bytes (no allocation in
 * later calls up to that size). */
mans_status mans_workspace_reserve(mans_workspace* ws, uint64_t max_bytes, const mans_params* p);
/* Waits for the workspace's last call and returns MANS_ERR_CORRUPT if a
 * kernel flagged a malformed stream (and clears the flags). */
mans_status mans_workspace_check(mans_workspace* ws, void* stream, uint32_t* h_flags);

/* ---- single stream: drop-in for the reference GpuANS API -------------- */
/* Upper bound of the container size for n input bytes (replaces
 * getMaxCompressedSize, GpuANSCodec.h:237-246; the reference's formula
 * passes 4096 as the block count and can undershoot, this one cannot). */
uint64_t mans_max_compressed_size(uint64_t n, const mans_params* p);

/* Encode n bytes at d_in into the container at d_out (16-byte aligned).
 * Replaces multibyte_ans::ansEncodeBatch (src/GpuANS.cu:16-31) /
 * ansEncode (include/ans/GpuANSEncode.h:286-401).  The container size is
 * written by the device to *d_out_size (uint32, as the reference's outSize)
 * when non-NULL, and to *h_out_size (host; the call then synchronises the
 * stream) when non-NULL. */
mans_status mans_encode(const mans_params* p, const void* d_in, uint64_t n, void* d_out,
                        uint64_t out_capacity, uint32_t* d_out_size, uint64_t* h_out_size,
                        mans_workspace* ws, void* stream);

/* mans_encode for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Writes the same bytes as mans_encode
 * with p->lanes = 32 (the reference format; 64 returns
 * MANS_ERR_INVALID_ARG) and the container size to *d_out_size (device
 * uint32).  It allocates nothing and takes no workspace: its scratch is a
 * fixed 1.02 MiB device array inside the library, one per GPU, held until
 * the process ends and shared by every call on that GPU, so calls that
 * overlap on one GPU race.  n is at most 2^30 bytes (262,144 blocks,
 * MANS_ERR_TOO_LARGE beyond).  Runs on the current device. */
mans_status mans_encode_noalloc(const mans_params* p, const void* d_in, uint64_t n, void* d_out,
                                uint64_t out_capacity, uint32_t* d_out_size, void* stream);

/* Decode the container at d_in (c bytes; 0 = unknown, as the reference)
 * into d_out.  Replaces multibyte_ans::ansDecodeBatch (src/GpuANS.cu:33-44)
 * / ansDecode (include/ans/GpuANSDecode.h:379-451).  No host synchronisation:
 * the block count is read on the device.  out_capacity bounds the bytes the
 * device may write (0 = unchecked, as the reference). */
mans_status mans_decode(const mans_params* p, const void* d_in, uint64_t c, void* d_out,
                        uint64_t out_capacity, mans_workspace* ws, void* stream);

/* ---- many single streams, DietGPU-style pointer arrays ----------------
 * The reference keeps DietGPU's multi-input form only as commented-out
 * vestiges (InProvider / OutProvider, a batch on blockIdx.y:
 * GpuANSStatistics.h:138-151, GpuANSDecode.h:202-214,304-317).  These take
 * host arrays of device pointers: input i (h_in_size[i] bytes) is coded
 * exactly as mans_encode codes it alone (p->lanes 32: the reference
 * container, byte-identical to ansEncodeBatch of that input), its size
 * written to d_out_sizes[i] (device uint32).  The inputs are coded one after
 * another on the stream with the workspace's buffers; the calls allocate
 * nothing beyond what the largest input needs (mans_workspace_reserve).
 * Decode: each container at h_in[i] into h_out[i]; a malformed one raises
 * the sticky error read by mans_workspace_check. */
mans_status mans_encode_multi(const mans_params* p, int32_t count, const void* const* h_in,
                              const uint32_t* h_in_size, void* const* h_out, const uint64_t* h_out_capacity,
                              uint32_t* d_out_sizes, mans_workspace* ws, void* stream);
mans_status mans_decode_multi(const mans_params* p, int32_t count, const void* const* h_in, void* const* h_out,
                              const uint64_t* h_out_capacity, mans_workspace* ws, void* stream);

/* ---- stage entry points (each reference kernel on its own) ------------- */
/* 256-bin histogram of the n bytes d_in[i*stride] (ansHistogramBatch,
 * GpuANSStatistics.h:409-446).  d_hist is overwritten. */
mans_status mans_histogram(const void* d_in, uint64_t n, uint32_t stride, uint32_t* d_hist,
                           mans_workspace* ws, void* stream);
/* Normalised table {pdf, cdf, magic, shift} x 256 from a histogram
 * (ansCalcWeights / quantizeWeights, GpuANSStatistics.h:186-406,449-468). */
mans_status mans_normalize(const uint32_t* d_hist, uint32_t total, int32_t prob_bits,
                           uint32_t* d_table, mans_workspace* ws, void* stream);
/* ADM mapping of n uint16 values (mapping/cpu/mappingcpuV1.cpp:33-193,
 * mapping/nv/mappingV15_fused_uint16.cu:400-730).  Writes offsets[gsize+1]
 * (exclusive prefix of group byte lengths, last = total), centers[gsize],
 * codes[n] and the bit plane (32*offsets[gsize] bytes, bits_capacity bound).
 * gsize = ceil(n/512).  *h_bits_len / *h_in_domain (host, synchronising)
 * when non-NULL. */
mans_status mans_adm_map(const uint16_t* d_in, uint64_t n, int32_t* d_offsets, uint16_t* d_centers,
                         uint8_t* d_codes, uint8_t* d_bits, uint64_t bits_capacity,
                         uint64_t* h_bits_len, int32_t* h_in_domain, mans_workspace* ws, void* stream);
/* Inverse mapping (mappingcpuV1.cpp:195-276, mappingV15_fused_uint16.cu:24-121). */
mans_status mans_adm_unmap(const int32_t* d_offsets, const uint16_t* d_centers, const uint8_t* d_codes,
                           const uint8_t* d_bits, uint64_t n, uint16_t* d_out, mans_workspace* ws,
                           void* stream);

/* ---- dual-quantisation Lorenzo front end -------------------------------
 * No counterpart in the reference (its mapping programs take the uint16
 * codes as input, mapping/cpu/mappingcpuV1.cpp).  The producer of those
 * codes for a float32 field f[nz][ny][nx] and absolute error bound eb:
 * q = rint(f * float32(1 / (2 eb))), r = the 3-D Lorenzo residual of q (q
 * is 0 outside the field), code = 32768 + r when |r| < radius, else 0 with
 * (index, r) appended to the outlier list, in the kernel's tile order
 * (deterministic).  radius in [1, 32768]; |f / (2 eb)| < 2^28 (else
 * MANS_ERR_DOMAIN); nx <= 8192 and nz * ny * nx < 2^32.
 * *d_outlier_count (device) gets the number of outliers, which may exceed
 * outlier_capacity (then only the first outlier_capacity are written).
 * With h_outlier_count non-NULL the call synchronises, writes the count
 * there and returns MANS_ERR_CAPACITY / MANS_ERR_DOMAIN when they apply.
 * Workspace: 12 bytes per tile of 256 x 16 x 16 cells, plus 2 KiB for each
 * of min(tiles, outlier_capacity / 6 + 1) tiles (the tiles where a thread
 * finds more than 5 outliers are coded again by a second launch). */
mans_status mans_lorenzo_quantize(const float* d_field, uint32_t nz, uint32_t ny, uint32_t nx, double error_bound,
                                  uint32_t radius, uint16_t* d_codes, uint32_t* d_outlier_idx, int32_t* d_outlier_val,
                                  uint64_t outlier_capacity, uint64_t* d_outlier_count, uint64_t* h_outlier_count,
                                  mans_workspace* ws, void* stream);
/* The inverse: q as the 3-D prefix sum of the residuals (codes and
 * outliers), d_field = q * float32(2 eb).  Needs 4 bytes of workspace per
 * element. */
mans_status mans_lorenzo_reconstruct(const uint16_t* d_codes, const uint32_t* d_outlier_idx,
                                     const int32_t* d_outlier_val, uint64_t outlier_count, uint32_t nz, uint32_t ny,
                                     uint32_t nx, double error_bound, float* d_field, mans_workspace* ws,
                                     void* stream);

/* ---- multi-byte frames (byte planes or ADM planes, one container) ------ */
uint64_t mans_frame_max_size(uint64_t n_elems, const mans_params* p);
/* Encode n_elems elements of p->elem_bytes bytes.  Frame size to *d_size
 * (device, uint64) and/or *h_size (host, synchronising). */
mans_status mans_frame_encode(const mans_params* p, const void* d_in, uint64_t n_elems, void* d_out,
                              uint64_t out_capacity, uint64_t* d_size, uint64_t* h_size,
                              mans_workspace* ws, void* stream);
/* Decodes the frame at d_in (c bytes) into d_out.  The layout comes from the
 * 128-byte frame header: pass a host copy in h_header for a fully
 * asynchronous call, or NULL to have it read from d_in (synchronising). */
mans_status mans_frame_decode(const void* d_in, uint64_t c, const void* h_header, void* d_out,
                              uint64_t out_capacity, uint64_t* h_n_elems, mans_workspace* ws, void* stream);
/* mans_frame_decode with the ADM restore fused into the decode: for a 64-lane
 * ADM frame the bit plane is decoded first, then one launch decodes the code
 * plane and restores each block's elements from codes kept in LDS, so the
 * codes never reach memory and no unmap launch runs: the reference's
 * restore kernel (`decompress`, mapping/nv/mappingV15_fused_uint16.cu:24-121)
 * moved inside the ANS decode.  Same output and errors as
 * mans_frame_decode; other frames decode exactly as there.  On MI355X the
 * two-launch path of mans_frame_decode measures faster (the restore work
 * inside the latency-bound decode waves costs more than the code bytes it
 * saves), so this entry point is for callers that want one launch per plane
 * and no code scratch. */
mans_status mans_frame_decode_fused(const void* d_in, uint64_t c, const void* h_header, void* d_out,
                                    uint64_t out_capacity, uint64_t* h_n_elems, mans_workspace* ws, void* stream);
/* Parses a frame header already on the host. */
mans_status mans_frame_info(const void* h_header, uint64_t header_bytes, uint64_t* n_elems,
                            int32_t* elem_bytes, int32_t* mode, uint64_t* total_size);

/* ---- batches of independent buffers ------------------------------------ */
/* Many buffers of uneven size and element width, each coded as its own
 * byte-plane frame in one persistent launch.  Frame i is byte-identical to
 * mans_frame_encode of buffer i alone with the same prob_bits, 64 lanes,
 * elem_bytes = bufs[i].elem_bytes and mode MANS_MODE_PLANES.  The reference
 * has no batch entry point (its DietGPU-style multi-input vestiges,
 * GpuANSStatistics.h:138-151 and GpuANSDecode.h:202-214, are fixed at one
 * input); this replaces calling ansEncodeBatch (src/GpuANS.cu:16) once per
 * buffer.  A batch is planned once (create: validates, lays out every frame,
 * uploads descriptors) and run any number of times; run only launches, so it
 * can be captured in a HIP graph.  Runs sharing a workspace must be
 * stream-ordered. */
typedef struct mans_buffer {
  const void* d_in;      /* encode: elements; decode: frame (16-byte aligned) */
  uint64_t in_bytes;     /* decode: bytes readable at d_in (the frame); encode: ignored */
  void* d_out;           /* encode: frame; decode: elements (16-byte aligned) */
  uint64_t out_capacity; /* encode: >= mans_frame_max_size(n_elems, {p, lanes, elem_bytes});
                            decode: >= n_elems * elem_bytes */
  uint64_t n_elems;
  int32_t elem_bytes;    /* 1, 2 or 4 */
  int32_t reserved;      /* 0 */
} mans_buffer;

typedef struct mans_batch mans_batch;

/* p->lanes is the interleave width of every frame of the batch: 64 (the
 * native format) or 32 (reference-format streams, byte-identical to
 * mans_frame_encode with lanes 32); it is fixed for the batch's plan.  The
 * request server takes 64.  p->elem_bytes and p->mode are ignored. */
mans_status mans_batch_encode_create(mans_batch** b, const mans_params* p, const mans_buffer* h_bufs,
                                     int32_t count, mans_workspace* ws);
/* d_sizes: device array of count uint64 frame sizes, written by the launch. */
mans_status mans_batch_encode_run(mans_batch* b, uint64_t* d_sizes, void* stream);
/* Decodes frames written by the batched or the single-buffer frame encode
 * (byte planes, 64 lanes, p->prob_bits).  A buffer given with elem_bytes 0
 * (and in_bytes >= 128) takes its element count and width from its frame's
 * header (read once, at create: the call synchronises); otherwise they are
 * the caller's.  A frame that does not match them, or whose streams are
 * inconsistent, raises the sticky error read by mans_workspace_check
 * (MANS_ERR_CORRUPT) and its buffer's flag (mans_batch_decode_flags); its
 * output is then undefined. */
mans_status mans_batch_decode_create(mans_batch** b, const mans_params* p, const mans_buffer* h_bufs,
                                     int32_t count, mans_workspace* ws);
mans_status mans_batch_decode_run(mans_batch* b, void* stream);
/* ---- archives: a batch's frames in one container ----------------------
 * No counterpart in the reference (its container holds one stream,
 * GpuANSCodec.h:68-209).  Layout: a 64-byte header (u32 "MANA", u8 version
 * 1, u32 count at 8, u64 total size at 16, index offset 64 at 24, data
 * offset at 32), then count index entries of 32 bytes {u64 offset, u64
 * frame bytes, u64 elements, u32 element bytes, u32 0}, then the frames,
 * each 16-byte aligned in index order with zero padding.
 * mans_batch_pack copies the frames of an encode batch's last run (d_sizes:
 * that run's sizes) into the archive at d_archive (capacity at least
 * mans_batch_archive_max_size) and writes its size to *d_archive_size
 * (device, may be NULL).  Stream-ordered, only launches: it can be captured
 * with the run. */
uint64_t mans_batch_archive_max_size(const mans_batch* b);
mans_status mans_batch_pack(const mans_batch* b, const uint64_t* d_sizes, void* d_archive, uint64_t capacity,
                            uint64_t* d_archive_size, void* stream);
/* Parses an archive header already on the host. */
mans_status mans_archive_info(const void* h_head, uint64_t head_bytes, uint32_t* count, uint64_t* total_size);
/* A decode batch over every frame of the archive at d_archive (its header
 * and index are read once, synchronously): frame i decodes into h_outs[i]
 * (device, h_out_caps[i] bytes, at least its elements x element bytes). */
mans_status mans_batch_decode_create_archive(mans_batch** b, const mans_params* p, const void* d_archive,
                                             uint64_t archive_bytes, void* const* h_outs, const uint64_t* h_out_caps,
                                             mans_workspace* ws);
/* Error flags of each buffer of the batch's last decode run (0: decoded;
 * else the buffer's frame was malformed and its output is undefined), for
 * up to max buffers.  Synchronises the stream.  The reference reports no
 * per-input errors (its decode asserts compile out, DeviceUtils.h:23-30). */
mans_status mans_batch_decode_flags(const mans_batch* b, uint32_t* h_flags, int32_t max, void* stream);
mans_status mans_batch_destroy(mans_batch* b);
/* Launch shape of a batch's kernels.  No shape changes an output bit; the
 * reference's counterparts are compile-time constants (warp 32,
 * DeviceUtils.h:49; unroll 8, GpuANSEncode.h:126; 256 threads per block of
 * the encode launch, GpuANSEncode.h:329-338). */
typedef struct mans_batch_shape {
  int32_t wg_threads;    /* threads per workgroup: 256, 512 or 1024 */
  int32_t syms_per_lane; /* symbols each lane codes per batch.  Encode: 4, 8 or 16
                            elements loaded ahead; decode: 1, 4, 8 or 16 steps
                            between two upkeeps of the word ring */
  int32_t lanes;         /* interleave width: the batch's frame format, 64 (native) or 32
                            (reference-format streams), fixed when the batch is planned */
  int32_t reserved;      /* 0 */
} mans_batch_shape;

typedef struct mans_tune_entry {
  mans_batch_shape shape;
  double ms;             /* mean time of one run; < 0 if the shape does not fit the LDS */
} mans_tune_entry;

/* The shape the batch runs with: the tuned one for its key (a previous
 * mans_batch_tune in this process, or the file named by the environment
 * variable MANS_TUNE_CACHE) or else the default. */
mans_status mans_batch_get_shape(const mans_batch* b, mans_batch_shape* shape);
/* MANS_ERR_INVALID_ARG if the shape is not a candidate or does not fit. */
mans_status mans_batch_set_shape(mans_batch* b, const mans_batch_shape* shape);
/* Autotuner: runs the batch `reps` times (after one warm-up run) with every
 * candidate shape (wg_threads x syms_per_lane, lanes 64: 9 for encode, 12
 * for decode; at most 16) on `stream`, keeps
 * the fastest in b and under b's key in the process-wide cache (appended to
 * MANS_TUNE_CACHE when set), and writes up to max_entries timings to
 * h_entries (*n_entries = candidates).  Synchronises; an encode batch's
 * frames are written by the runs (identically for every shape).  The key is
 * (direction, prob_bits, buffer count and total bytes rounded to powers of
 * two, element widths present, compute units). */
mans_status mans_batch_tune(mans_batch* b, void* stream, int32_t reps, mans_tune_entry* h_entries,
                            int32_t max_entries, int32_t* n_entries);
/* Work-item grain of the batches planned after this call, process-wide: the
 * bytes of a histogram item (a power of two in [64 KiB, 512 KiB]), the
 * rounds of an encode ticket (1, 2 or 4; a ticket codes rounds x 4 byte-plane
 * blocks, rounds x 2 for uint32) and the decode items a wavefront takes per
 * queue grab (1 or 4).  0 leaves a value to the planner, which picks the
 * largest grain that still gives every compute unit tens of items (one
 * request of tens of MB: 64 KiB, 1, 1; thousands of buffers of several GB:
 * 512 KiB, 4, 4).  The grain changes no output byte; it is a tuning and test
 * knob.  MANS_ERR_INVALID_ARG for other values.  No counterpart in the
 * reference (its kernels take one buffer per launch). */
mans_status mans_batch_set_grain(uint32_t hist_chunk, uint32_t rounds, uint32_t run);
/* create + run + destroy (synchronises the stream before freeing). */
mans_status mans_encode_batch(const mans_params* p, const mans_buffer* h_bufs, int32_t count, uint64_t* d_sizes,
                              mans_workspace* ws, void* stream);
mans_status mans_decode_batch(const mans_params* p, const mans_buffer* h_bufs, int32_t count,
                              mans_workspace* ws, void* stream);

/* ---- request server (continuous batching) ------------------------------ */
/* No counterpart in the reference (its compress driver, examples/compress.cu,
 * codes one file per process).  A server owns a device-memory arena of
 * `memory_budget` bytes (less its own batch state) and a worker thread.
 * Requests (device-resident input) are admitted in arrival order as arena
 * memory frees up and batched continuously into the batched launches; a
 * request's output stays in the arena until mans_server_release.  A
 * compress request yields the frame mans_frame_encode writes for its input
 * (byte planes, 64 lanes, the server's prob_bits), a decompress request the
 * elements of such a frame.  All entry points are thread-safe. */
enum { MANS_REQ_COMPRESS = 0, MANS_REQ_DECOMPRESS = 1 };

typedef struct mans_request {
  int32_t kind;       /* MANS_REQ_COMPRESS or MANS_REQ_DECOMPRESS */
  int32_t elem_bytes; /* 1, 2 or 4 */
  const void* d_in;   /* compress: elements; decompress: frame (device, 16-byte aligned) */
  uint64_t in_bytes;  /* decompress: frame bytes */
  uint64_t n_elems;   /* elements (both kinds) */
} mans_request;

typedef struct mans_result {
  int32_t status;     /* MANS_OK, or the request's error */
  int32_t reserved;
  const void* d_out;  /* frame or elements, in the server's arena */
  uint64_t out_bytes;
  double queue_ms;    /* submit -> admitted into a launch */
  double latency_ms;  /* submit -> output ready */
} mans_result;

typedef struct mans_server mans_server;

/* p->lanes must be 64.  MANS_ERR_BUDGET when the budget does not cover the
 * server's own state. */
mans_status mans_server_create(mans_server** s, const mans_params* p, uint64_t memory_budget, int32_t device);
/* MANS_ERR_BUDGET if the request's output can never fit the arena. */
mans_status mans_server_submit(mans_server* s, const mans_request* r, uint64_t* id);
/* count requests queued together, in order (one worker wake-up); ids[i] is
 * request i's id.  All are checked first: on any error none is queued. */
mans_status mans_server_submit_many(mans_server* s, const mans_request* r, int32_t count, uint64_t* ids);
/* Waits up to timeout_ms (< 0: no limit); MANS_ERR_BUSY if not done yet. */
mans_status mans_server_wait(mans_server* s, uint64_t id, int32_t timeout_ms, mans_result* out);
/* Frees a finished request's output. */
mans_status mans_server_release(mans_server* s, uint64_t id);
mans_status mans_server_stats(mans_server* s, uint64_t* arena_bytes, uint64_t* arena_used, uint64_t* batches,
                              uint64_t* waiting);
/* Finishes every submitted request, then frees everything. */
mans_status mans_server_destroy(mans_server* s);

/* ---- profiling --------------------------------------------------------- */
/* When enabled, every call records HIP events on its stream around each of
 * its kernels.  mans_last_kernel_times waits for them and returns the
 * durations (ms) of all kernels launched since the previous query, in launch
 * order with their names (static strings), then starts a new record. */
mans_status mans_set_profiling(mans_workspace* ws, int32_t enable);
int32_t mans_last_kernel_times(mans_workspace* ws, double* h_ms, const char** h_names, int32_t max);

#ifdef __cplusplus
}
#endif
#endif /* MANS_H */
