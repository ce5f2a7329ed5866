// codec_device.hpp -- device-side pieces shared by the single-buffer and
// batched rANS kernels: the per-symbol encode and decode steps, the
// decoupled look-back, and the staging bound.
#pragma once
#include "mans_internal.hpp"

namespace mans {

template <int S> struct ElemOf;
template <> struct ElemOf<1> { using T = uint8_t; };
template <> struct ElemOf<2> { using T = uint16_t; };
template <> struct ElemOf<4> { using T = uint32_t; };

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;

// Words per chain per wavefront of one unit at most: 64 lanes x 128 symbols
// (one native unit, or two 32-lane reference units, one per half), rounded
// to 16 words so every area is 32-byte aligned.
template <int PB>
constexpr uint32_t chain_stage_words() {
  return (uint32_t)round_up(64ull * lane_max_words(128u, PB), 16u);
}

// One symbol of one chain with its table entry e already in registers.  The
// emit is branch-free: every lane stores, lanes that emit nothing store into
// their own dummy slot, so the steps of several chains can be interleaved by
// the scheduler (the table reads of the next step are issued before these
// stores; no wait on the LDS counter sits between them).  Slots are u16
// indices from the wave's staging base wbase: an emitted word goes to soff
// (the chain's region) + ((cnt + rank) & MASK), where cnt is the unit's words
// so far (wave-uniform in the 64-lane format) and rank the number of
// emitting lanes of its unit below it; a lane that emits nothing writes its
// dummy slot doff.
template <int L, uint32_t MASK>
__device__ __forceinline__ void enc_step(uint32_t& x, uint32_t& cnt, uint16_t* wbase, uint32_t soff, uint32_t doff,
                                         const uint4& e, bool v, uint32_t half) {
  const bool w = v && (x >= e.x);
  const uint64_t m = __ballot(w);
  const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)(m >> 32);
  uint32_t add, below;
  if (L == 64) {
    add = (uint32_t)__popcll(m);
    below = 0;
  } else {
    add = half ? (uint32_t)__popc(mhi) : (uint32_t)__popc(mlo);
    below = half ? (uint32_t)__popc(mlo) : 0u;
  }
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u)) - below;
  // the slot goes through an opaque register so the compiler keeps one
  // unconditional store instead of splitting the select into two branches
  uint32_t slot = w ? soff + ((cnt + rank) & MASK) : doff;
  __asm__ volatile("" : "+v"(slot));
  wbase[slot] = (uint16_t)x;
  x = w ? x >> 16 : x;
  cnt += add;
  const uint32_t q = (__umulhi(x, e.y) + x) >> (e.z & 31u);
  const uint32_t xn = __umul24(q, e.w) + x + (e.z >> 8);
  x = v ? xn : x;
}

typedef __attribute__((address_space(3))) uint16_t lds_u16;

// LDS byte address of a generic pointer into shared memory.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// Encoder table entry of the 64-lane coder, for probability precision p:
// {pdf << (31 - p) | cdf, 2^p - pdf, lo, hi of rcp}, rcp = 1 / pdf rounded
// up to a double.  (p <= 15, so the threshold pdf << (31 - p) has no bits
// below bit 16 and cdf fits under it.)  For 0 <= x < 2^31 and pdf <= 2^11,
// floor(RN(x * rcp)) = floor(x / pdf): x * rcp >= x / pdf >= q, and
// x * rcp < x / pdf + 2^-21 <= q + 1 - 2^-11 + 2^-21, so neither the
// product's rounding nor the truncating conversion can leave [q, q + 1).
// This replaces the reference's magic-number division
// (GpuANSCodec.h:36-61: q = (umulhi(x, magic) + x) >> shift) by two full-rate
// conversions and one full-rate f64 multiply on gfx950; q is the same.
__device__ __forceinline__ uint4 enc64_entry(uint32_t pdf, uint32_t cdf, int p) {
  double r = 0.0;
  if (pdf > 0) {
    r = 1.0 / (double)pdf;
    // round up: the exact residual r * pdf - 1 (one fma) below zero means
    // the correctly rounded quotient is below 1 / pdf
    if (__fma_rn(r, (double)pdf, -1.0) < 0.0) r = __longlong_as_double(__double_as_longlong(r) + 1);
  }
  const uint64_t rb = (uint64_t)__double_as_longlong(r);
  return make_uint4((pdf << (31 - p)) | cdf, (1u << p) - pdf, (uint32_t)rb, (uint32_t)(rb >> 32));
}

// x / pdf for the entry e (see enc64_entry).
__device__ __forceinline__ uint32_t enc64_quot(uint32_t x, const uint4& e) {
  const double r = __hiloint2double((int)e.w, (int)e.z);
  return (uint32_t)((double)x * r);
}

// sbase += 2 * popcount(m) in two scalar instructions
__device__ __forceinline__ void advance2(uint32_t& sbase, uint64_t m) {
  uint32_t c;
  __asm__("s_bcnt1_i32_b64 %0, %1" : "=s"(c) : "s"(m) : "scc");
  __asm__("s_lshl1_add_u32 %0, %1, %0" : "+s"(sbase) : "s"(c) : "scc");
}

// One symbol of one chain in the 64-lane format with a linear stage, table
// entry e (enc64_entry).  An emitted word goes to LDS byte address
// sbase + 2 * rank, where sbase (wave uniform) is where the chain's next word
// goes and rank the number of emitting lanes below this one; a lane that
// emits nothing writes its own dummy slot at byte address dummy.  sbase
// advances by two bytes per emitted word.  Lanes with v false keep their
// state and emit nothing.
__device__ __forceinline__ void enc_step64(uint32_t& x, uint32_t& sbase, uint32_t dummy, const uint4& e,
                                           bool v = true) {
  const bool w = v && (x >> 16) >= (e.x >> 16);
  const uint64_t m = __ballot(w);
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  // sbase as one scalar and the emit address computed by every lane: one
  // v_lshl_add and one v_cndmask, no exec-mask branch around them
  __asm__ volatile("" : "+s"(sbase));
  uint32_t emit = sbase + 2u * rank;
  __asm__ volatile("" : "+v"(emit));
  uint32_t addr = w ? emit : dummy;
  // an opaque address keeps one unconditional store (no branch)
  __asm__ volatile("" : "+v"(addr));
  *(lds_u16*)(uintptr_t)addr = (uint16_t)x;
  x = w ? x >> 16 : x;
  advance2(sbase, m);
  const uint32_t xn = __umul24(enc64_quot(x, e), e.y) + x + (e.x & 0xffffu);
  x = v ? xn : x;
}

// enc_step64 for a step where every lane codes a symbol, written with
// sub-dword operand selects: the threshold compare reads the high halves of
// x and e.x, the renormalising shift is a select of x's high half, and cdf
// is added from e.x's low half.
__device__ __forceinline__ void enc_step64f(uint32_t& x, uint32_t& sbase, uint32_t dummy, const uint4& e) {
  uint64_t m;
  uint32_t xr;
  __asm__(
      "v_cmp_ge_u32_sdwa vcc, %2, %3 src0_sel:WORD_1 src1_sel:WORD_1\n\t"
      "v_cndmask_b32_sdwa %1, %2, %2, vcc dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1\n\t"
      "s_mov_b64 %0, vcc"
      : "=s"(m), "=&v"(xr)
      : "v"(x), "v"(e.x)
      : "vcc");
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  __asm__ volatile("" : "+s"(sbase));
  const uint32_t emit = sbase + 2u * rank;
  // only the emitting lanes store: exec is narrowed to m for the store and
  // restored inside the same block (s_and_saveexec also writes SCC)
  uint64_t saved;
  __asm__ volatile(
      "s_and_saveexec_b64 %0, %1\n\t"
      "ds_write_b16 %2, %3\n\t"
      "s_mov_b64 exec, %0"
      : "=&s"(saved)
      : "s"(m), "v"(emit), "v"(x)
      : "memory", "scc");
  (void)dummy;
  advance2(sbase, m);
  const uint32_t xn = __umul24(enc64_quot(xr, e), e.y) + xr;
  __asm__("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
          : "=v"(x)
          : "v"(xn), "v"(e.x));
}

// The 32-lane format on a 64-lane wavefront: lanes 0-31 code one unit and
// lanes 32-63 another (halfv = 0 / 1 per lane), each unit with its own stage.
// An emitted word of the lower half goes to sbA + 2 * rank, one of the upper
// half to sbB + 2 * (rank - emitters of the lower half); the two bases are
// scalars, the per-lane choice one 24-bit multiply-add:
// emit = (sbA + 2 rank) + halfv * (sbB - sbA - 2 lo).
__device__ __forceinline__ uint32_t emit32(uint64_t m, uint32_t& sbA, uint32_t& sbB, uint32_t halfv) {
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  uint32_t lo, hi;
  __asm__("s_bcnt1_i32_b32 %0, %1" : "=s"(lo) : "s"((uint32_t)m) : "scc");
  __asm__("s_bcnt1_i32_b32 %0, %1" : "=s"(hi) : "s"((uint32_t)(m >> 32)) : "scc");
  __asm__ volatile("" : "+s"(sbA), "+s"(sbB));
  uint32_t a2 = sbA;
  __asm__("s_lshl1_add_u32 %0, %1, %0" : "+s"(a2) : "s"(lo) : "scc");
  const int32_t delta = (int32_t)sbB - (int32_t)a2;
  const uint32_t emit = (uint32_t)((int32_t)(sbA + 2u * rank) + __mul24((int32_t)halfv, delta));
  sbA = a2;
  __asm__("s_lshl1_add_u32 %0, %1, %0" : "+s"(sbB) : "s"(hi) : "scc");
  return emit;
}

// enc_step64f for the 32-lane format (every lane codes a symbol).
__device__ __forceinline__ void enc_step32f(uint32_t& x, uint32_t& sbA, uint32_t& sbB, uint32_t halfv,
                                            const uint4& e) {
  uint64_t m;
  uint32_t xr;
  __asm__(
      "v_cmp_ge_u32_sdwa vcc, %2, %3 src0_sel:WORD_1 src1_sel:WORD_1\n\t"
      "v_cndmask_b32_sdwa %1, %2, %2, vcc dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1\n\t"
      "s_mov_b64 %0, vcc"
      : "=s"(m), "=&v"(xr)
      : "v"(x), "v"(e.x)
      : "vcc");
  const uint32_t emit = emit32(m, sbA, sbB, halfv);
  uint64_t saved;
  __asm__ volatile(
      "s_and_saveexec_b64 %0, %1\n\t"
      "ds_write_b16 %2, %3\n\t"
      "s_mov_b64 exec, %0"
      : "=&s"(saved)
      : "s"(m), "v"(emit), "v"(x)
      : "memory", "scc");
  const uint32_t xn = __umul24(enc64_quot(xr, e), e.y) + xr;
  __asm__("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
          : "=v"(x)
          : "v"(xn), "v"(e.x));
}

// enc_step64 (lanes with v false keep their state) for the 32-lane format.
__device__ __forceinline__ void enc_step32(uint32_t& x, uint32_t& sbA, uint32_t& sbB, uint32_t halfv, uint32_t dummy,
                                           const uint4& e, bool v) {
  const bool w = v && (x >> 16) >= (e.x >> 16);
  const uint64_t m = __ballot(w);
  uint32_t emit = emit32(m, sbA, sbB, halfv);
  __asm__ volatile("" : "+v"(emit));
  uint32_t addr = w ? emit : dummy;
  __asm__ volatile("" : "+v"(addr));
  *(lds_u16*)(uintptr_t)addr = (uint16_t)x;
  x = w ? x >> 16 : x;
  const uint32_t xn = __umul24(enc64_quot(x, e), e.y) + x + (e.x & 0xffffu);
  x = v ? xn : x;
}

// 16 * byte B of w: a byte offset into a table of 16-byte entries.
template <int B>
__device__ __forceinline__ uint32_t byte_x16(uint32_t w) {
  uint32_t r;
  if constexpr (B == 0)
    __asm__("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
            : "=v"(r) : "v"(w));
  else if constexpr (B == 1)
    __asm__("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1"
            : "=v"(r) : "v"(w));
  else if constexpr (B == 2)
    __asm__("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2"
            : "=v"(r) : "v"(w));
  else
    __asm__("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3"
            : "=v"(r) : "v"(w));
  return r;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum of the aggregates of status entries j, j - 1, ... back to and including
// the nearest inclusive prefix (entries below `first` read as a zero prefix).
// Lane l reads entries j - l - 64 k, k < K: 64 K entries per round trip.
// Groups of 64 are taken in order; a group with an unpublished entry before
// its nearest prefix ends the round (the groups before it are kept) and the
// wave sleeps briefly before reading again; with TRY it returns ~0 instead
// (an entry still unpublished: the caller resolves later).  dec(raw, flag,
// value) splits a status word (flag 0: not yet published, 1: aggregate,
// 2: inclusive prefix).
template <int K, bool TRY = false, typename Dec>
__device__ __forceinline__ uint64_t look_back_window(const uint64_t* status, int64_t j, int64_t first,
                                                     uint64_t zero_prefix, int lane, Dec dec) {
  uint64_t excl = 0;
  for (;;) {
    uint64_t raw[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t i = j - lane - 64 * k;
      raw[k] = i >= first ? __hip_atomic_load(status + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : zero_prefix;
    }
    uint64_t part = 0;
    int groups = 0;
    bool found = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      uint32_t f;
      uint64_t val;
      dec(raw[k], f, val);
      const uint64_t pm = __ballot(f == 2), nm = __ballot(f == 0);
      // lanes up to and including the nearest prefix (all 64 if none)
      const uint64_t upto = pm ? (pm & (0 - pm)) * 2u - 1u : ~0ull;
      if (nm & upto) {
        if constexpr (TRY) return ~0ull;
        break;
      }
      part += ((upto >> lane) & 1u) ? val : 0u;
      ++groups;
      if (pm) {
        found = true;
        break;
      }
    }
    excl += wave_sum(part);
    if (found) return excl;
    j -= 64 * groups;
    if (groups < K) __builtin_amdgcn_s_sleep(1);
  }
}

// First half of a look-back: publish entry `chain`'s own count (the first
// entry's is already its inclusive prefix).
__device__ __forceinline__ void publish_aggregate(uint64_t* status, uint64_t chain, uint64_t agg, int lane) {
  if (lane == 0)
    __hip_atomic_store(status + chain, (chain == 0 ? kFlagPrefix : kFlagAgg) | agg, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
// Second half: sum the predecessors back to the nearest published prefix,
// publish this entry's inclusive prefix and return the exclusive one.  Any
// time may pass between the two halves; the predecessors only need to have
// published their aggregates.  K groups of 64 entries per round trip.
template <int K = 8>
__device__ __forceinline__ uint64_t resolve_prefix(uint64_t* status, uint64_t chain, uint64_t agg, int lane) {
  if (chain == 0) return 0;
  const uint64_t excl = look_back_window<K>(status, (int64_t)chain - 1, 0, kFlagPrefix, lane,
                                            [](uint64_t v, uint32_t& f, uint64_t& val) {
                                              f = (uint32_t)(v >> 62);
                                              val = v & kValueMask;
                                            });
  if (lane == 0)
    __hip_atomic_store(status + chain, kFlagPrefix | (excl + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}
// Decoupled look-back for ticket group `chain` with aggregate agg (all
// lanes): publishes the aggregate, sums predecessors back to the nearest
// inclusive prefix (512 per round trip), publishes this group's inclusive
// prefix and returns the exclusive one.  Each status word holds its flag and
// value together and carries nothing else, so relaxed device-scope atomics
// suffice (release / acquire at agent scope would write back / invalidate
// the XCD's whole L2 on every publish and poll).
__device__ __forceinline__ uint64_t look_back(uint64_t* status, uint64_t chain, uint64_t agg, int lane) {
  publish_aggregate(status, chain, agg, lane);
  return resolve_prefix(status, chain, agg, lane);
}

// One symbol of each of the CH chains; returns the CH decoded bytes packed
// into an element.  Every lane reads a word (no branch, so the chains'
// steps interleave): lanes that take none read some in-range slot and keep
// their state.  ptr (words of the unit not yet taken) never drops below a
// step's count on a valid stream; on a corrupt one it is clamped at 0 so
// every read stays inside the stage, and `taken` (the words taken so far)
// then ends above the unit's word count.  Word i of a unit sits at
// stage[i & MASK] (MASK = R - 1 for a ring, all ones for a whole-unit
// stage); a step reads stage[(base & MASK) + rank] with rank <= 64, which the
// ring's mirror covers.  The three phases (LUT reads; state update and word
// reads; selects) each run over all chains, so one chain's LDS latency
// overlaps the others' work.
// Decoder LUT entries.  uint2 {pdf | sym << 24, slot - cdf}: the 24-bit
// multiply ignores the symbol byte, so the update is one shift and one
// multiply-add.  uint32_t (slot - cdf) << 20 | pdf << 8 | sym: half the LDS,
// two more operations per symbol.
template <int PB>
__device__ __forceinline__ uint32_t lut_next(uint2 e, uint32_t x) { return __umul24(e.x, x >> PB) + e.y; }
template <int PB>
__device__ __forceinline__ uint32_t lut_next(uint32_t e, uint32_t x) {
  return __umul24((e >> 8) & 0xfffu, x >> PB) + (e >> 20);
}
__device__ __forceinline__ uint32_t lut_sym(uint2 e) { return e.x >> 24; }
__device__ __forceinline__ uint32_t lut_sym(uint32_t e) { return e & 0xffu; }
// The CH decoded symbols packed into an element (chain c's byte at bits
// 8c..8c+7) with byte permutes: a uint2 entry's symbol is byte 3 of e.x, a
// uint32_t entry's byte 0.  v_perm selector bytes 0-3 pick the second
// operand's bytes, 4-7 the first's, 0x0c gives zero.
__device__ __forceinline__ uint32_t sym_word(uint2 e) { return e.x; }
__device__ __forceinline__ uint32_t sym_word(uint32_t e) { return e; }
template <typename LE>
constexpr uint32_t sym_byte() { return sizeof(LE) == 8 ? 3u : 0u; }
template <int CH, typename LE>
__device__ __forceinline__ uint32_t pack_syms(const LE (&e)[CH]) {
  constexpr uint32_t b = sym_byte<LE>();
  if constexpr (CH == 1) {
    return lut_sym(e[0]);
  } else if constexpr (CH == 2) {
    return __builtin_amdgcn_perm(sym_word(e[1]), sym_word(e[0]), 0x0c0c0000u | ((4u + b) << 8) | b);
  } else if constexpr (CH == 4) {
    const uint32_t lo = __builtin_amdgcn_perm(sym_word(e[1]), sym_word(e[0]), 0x0c0c0000u | ((4u + b) << 8) | b);
    const uint32_t hi = __builtin_amdgcn_perm(sym_word(e[3]), sym_word(e[2]), ((4u + b) << 24) | (b << 16) | 0x0c0cu);
    return lo | hi;
  } else {
    uint32_t el = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) el |= lut_sym(e[c]) << (8 * c);
    return el;
  }
}

// CLAMP: ptr stops at 0 (whole-unit stages, where a read past the unit's
// words must stay inside the stage); without it (rings, every read masked
// into the ring) ptr goes below 0 on a corrupt stream and stays there, and
// `taken` is not kept.
template <int L, int PB, int CH, uint32_t MASK, typename LE = uint2, bool CLAMP = true>
__device__ __forceinline__ uint32_t dec_steps(uint32_t (&x)[CH], uint32_t (&ptr)[CH], uint32_t (&taken)[CH],
                                              const LE* const (&lut)[CH], const uint16_t* const (&stage)[CH],
                                              const bool (&v)[CH], uint32_t half) {
  constexpr uint32_t W = 1u << PB;
  LE e[CH];
  uint32_t word[CH];
  bool r[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) e[c] = lut[c][x[c] & (W - 1u)];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const uint32_t xn = lut_next<PB>(e[c], x[c]);
    x[c] = v[c] ? xn : x[c];
    r[c] = v[c] && (x[c] < kMinState);
    const uint64_t m = __ballot(r[c]);
    const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)(m >> 32);
    uint32_t cnt, below;
    if (L == 64) {
      cnt = (uint32_t)__popcll(m);
      below = 0;
    } else {
      cnt = half ? (uint32_t)__popc(mhi) : (uint32_t)__popc(mlo);
      below = half ? (uint32_t)__popc(mlo) : 0u;
    }
    uint32_t base;
    if constexpr (CLAMP) {
      taken[c] += cnt;
      // ptr < 2^16 and cnt <= 64: a signed max keeps this scalar (s_max_i32)
      const int32_t bd = (int32_t)ptr[c] - (int32_t)cnt;
      base = (uint32_t)(bd > 0 ? bd : 0);
    } else {
      base = ptr[c] - cnt;
    }
    // slot of this lane's word: (base & MASK) + rank, the base taken in as
    // mbcnt's accumulator (no scalar shift / add for the address)
    const uint32_t slot = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, base & MASK)) - below;
    word[c] = stage[c][slot];
    ptr[c] = base;
  }
  // keep the reads unconditional: without this the compiler moves each under
  // its `r` and puts a branch between the chains
#pragma unroll
  for (int c = 0; c < CH; ++c) __asm__ volatile("" : "+v"(word[c]));
#pragma unroll
  for (int c = 0; c < CH; ++c) x[c] = r[c] ? (x[c] << 16) | word[c] : x[c];
  return pack_syms<CH>(e);
}

// Sum of a 32-bit value over the 64 lanes (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Normalises one plane with a single wavefront: lane l owns symbols
// 4l .. 4l + 3 (count[j] is symbol 4l + j).  The same restatement of
// GpuANSStatistics.h:186-385 as the 256-thread finisher of the histogram
// launch (fp32 quantisation, rank by (q << 16 | s) descending, the +1 rule
// for symbol ids below diff % 256, repeated decrements at ranks
// [G - it, G)), with wave-wide sums and scans.  s_keys: 256 words of this
// wave's LDS.  Returns {pdf, cdf, magic, shift} per owned symbol.
__device__ inline void normalize_wave(const uint32_t (&count)[4], uint32_t total, int p, uint32_t* s_keys,
                                      uint4 (&out)[4], bool& bad) {
  const uint32_t lane = threadIdx.x & 63u;
  if (total == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = make_uint4(0, 0, 0, 0);
    return;
  }
  const uint32_t W = 1u << p;
  uint32_t q[4], key[4], qs = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float ratio = __fdiv_rn((float)count[j], (float)total);
    const float prod = __fmul_rn((float)W, ratio);
    q[j] = (uint32_t)prod;
    if (count[j] > 0 && q[j] == 0) q[j] = 1;
    qs += q[j];
    key[j] = (q[j] << 16) | (4u * lane + (uint32_t)j);
    s_keys[4u * lane + (uint32_t)j] = key[j];
  }
  const uint32_t qsum = wave_sum32(qs);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  int rank[4] = {0, 0, 0, 0};
  for (int i = 0; i < (int)kNumSymbols; ++i) {
    const uint32_t k = s_keys[i];
#pragma unroll
    for (int j = 0; j < 4; ++j) rank[j] += k > key[j];
  }
  const int diff = (int)W - (int)qsum;
  if (diff > 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) q[j] += (uint32_t)(diff >> 8) + ((4u * lane + (uint32_t)j < (uint32_t)(diff & 255)) ? 1u : 0u);
  } else if (diff < 0) {
    int d = -diff;
    while (d > 0) {
      uint32_t gl = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) gl += q[j] > 1 ? 1u : 0u;
      const int g = (int)wave_sum32(gl);
      const int it = d < g ? d : g;
      if (it <= 0) {
        bad = true;
        break;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (rank[j] >= g - it && rank[j] < g) q[j] -= 1;
      d -= it;
    }
  }
  // exclusive scan in symbol order: lane sums, then a wave scan
  const uint32_t ls = q[0] + q[1] + q[2] + q[3];
  uint32_t inc = ls;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += u;
  }
  uint32_t cdf = inc - ls;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t pdf = q[j];
    uint32_t shift = 0, magic = 0;
    if (pdf > 0) {
      shift = 32u - (uint32_t)__clz((int)(pdf - 1u));
      const uint64_t m = ((1ull << 32) * ((1ull << shift) - pdf)) / pdf + 1ull;
      magic = (uint32_t)m;
    }
    out[j] = make_uint4(pdf, cdf, magic, shift);
    cdf += pdf;
  }
}

}  // namespace mans
