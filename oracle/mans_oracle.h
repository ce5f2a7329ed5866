/*
 * mans_oracle.h -- CPU restatement of the multi-byte ANS codec path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * Every function restates the reference algorithm from its source text and
 * cites the file:line it follows (paths relative to the reference root).
 * The restatement is written independently of the HIP product code: no
 * header or helper is shared with mans_amd/csrc.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   mo_histogram       brute-force count + sum invariant
 *   mo_normalize       hand-computed known answers for both fix-up branches,
 *                      closed-form diff>0 rule, sum invariant, IEEE fp32 step
 *   mo_magic_*         brute force q == x / pdf over sampled x, all pdf
 *   mo_stream_encode   decode(encode(x)) == x with the independently written
 *                      decoder, KATs (single symbol -> no words), container
 *                      size formula, hand-traced 1-lane / 2-lane examples
 *   mo_adm_map/unmap   golden vectors written by the reference's own
 *                      mapping/cpu/mappingcpuV1.cpp (built under oracle/_ref)
 *   mo_planes_*        closed form + round trip
 */
#ifndef MANS_ORACLE_H
#define MANS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- rANS constants (include/ans/GpuANSCodec.h:18-61) ---- */
#define MO_NUM_SYMBOLS 256
#define MO_STATE_BITS 31          /* GpuANSCodec.h:42 */
#define MO_ENCODED_BITS 16        /* GpuANSCodec.h:43 */
#define MO_MIN_STATE (1u << 15)   /* GpuANSCodec.h:47-50 */
#define MO_MAGIC 0xd00du          /* GpuANSCodec.h:53 */
#define MO_HEADER_BYTES 32u       /* GpuANSCodec.h:211 */

/* Stream formats.  Format 0 is the reference layout (32 interleaved lanes,
 * 4 KiB blocks, version 1).  Format 1 is this project's native layout
 * (64 interleaved lanes, 8 KiB blocks, version 2); identical container
 * structure, only lanes / block size / version differ. */
typedef struct {
  uint32_t lanes;
  uint32_t block_bytes;
  uint32_t version;
} mo_format;

mo_format mo_format_ref(void);
mo_format mo_format_native(void);

/* ---- A1: histogram (GpuANSStatistics.h:23-136) ---- */
void mo_histogram(const uint8_t* in, uint64_t n, uint32_t stride,
                  uint32_t* hist /* [256] */);

/* ---- A2: normalisation (GpuANSStatistics.h:186-385) ----
 * table[4*s + {0,1,2,3}] = {pdf, cdf, magic, shift}.  total == 0 leaves all
 * entries zero (the reference leaves them unwritten).  Returns 0 on success,
 * -1 on an impossible fix-up state. */
int mo_normalize(const uint32_t* hist, uint32_t total, int prob_bits,
                 uint32_t* table /* [256*4] */);

/* Reference integer division by magic multiply (GpuANSEncode.h:54-65). */
uint32_t mo_magic_div(uint32_t x, uint32_t magic, uint32_t shift);

/* ---- A3-A6: single stream container (GpuANSEncode.h:24-284,
 * GpuANSCodec.h:68-246) ---- */
uint64_t mo_stream_overhead(uint64_t num_blocks, mo_format f);
/* Exact upper bound of mo_stream_encode's output for n symbols. */
uint64_t mo_stream_max_size(uint64_t n, int prob_bits, mo_format f);
/* Upper bound of the 16-bit words one lane can emit for `nsym` symbols. */
uint32_t mo_lane_max_words(uint32_t nsym, int prob_bits);

/* Encode n symbols taken from in[i*stride] into the container at out.
 * Returns 0 and sets *out_size, or a negative error. */
int mo_stream_encode(const uint8_t* in, uint64_t n, uint32_t stride,
                     int prob_bits, mo_format f, uint8_t* out,
                     uint64_t out_cap, uint64_t* out_size, int threads);

/* Decode the container at in (size c) into out[i*stride].  Returns 0 and sets
 * *n_out, or a negative error (bad magic/version, precision mismatch,
 * truncated stream, capacity). */
int mo_stream_decode(const uint8_t* in, uint64_t c, int prob_bits,
                     uint8_t* out, uint32_t stride, uint64_t out_cap,
                     uint64_t* n_out, int threads);

/* ---- M1-M3: ADM byte mapping for uint16 (mapping/cpu/mappingcpuV1.cpp) ----
 * gsize = ceil(n/512).  offsets[gsize+1] = exclusive prefix of the per-group
 * byte length M_g, with offsets[gsize] = total.  The bit plane is
 * 32*offsets[gsize] bytes.  Returns 1 if every 16-element lane stays inside
 * the reference's 256-bit lane buffer (lossless domain), 0 otherwise (the
 * planes are then not a valid mapping), -1 if bits_cap is too small. */
int mo_adm_map(const uint16_t* x, uint64_t n, uint16_t* centers,
               int32_t* offsets, uint8_t* codes, uint8_t* bits,
               uint64_t bits_cap, uint64_t* bits_len);
uint64_t mo_adm_groups(uint64_t n);
void mo_adm_unmap(const uint16_t* centers, const int32_t* offsets,
                  const uint8_t* codes, const uint8_t* bits, uint64_t n,
                  uint16_t* out);

/* ---- Multi-plane frame (this project's container around the streams) ----
 * mode 0: byte planes (plane k = byte k of each element, k < elem_bytes)
 * mode 1: ADM (uint16 only; plane 0 = codes, plane 1 = bit plane) */
uint64_t mo_frame_max_size(uint64_t n_elems, uint32_t elem_bytes, int mode,
                           int prob_bits, mo_format f);
int mo_frame_encode(const void* in, uint64_t n_elems, uint32_t elem_bytes,
                    int mode, int prob_bits, mo_format f, uint8_t* out,
                    uint64_t out_cap, uint64_t* out_size, int threads);
int mo_frame_decode(const uint8_t* in, uint64_t c, void* out,
                    uint64_t out_cap_bytes, uint64_t* n_elems, int threads);

#ifdef __cplusplus
}
#endif
#endif
