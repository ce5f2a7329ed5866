// batch_kernels.hpp -- many independent buffers, each coded as a byte-plane
// frame, in one persistent launch.  Included by one translation unit per
// probability precision (batch_p9.hip, batch_p10.hip, batch_p11.hip), each
// instantiating the kernels of every launch shape (BatchShape) for it.
//
// Every buffer's frame is byte-identical to what mans_frame_encode writes for
// it alone (MANS_MODE_PLANES, in the batch's 64- or 32-lane format): the same
// histogram and normalisation (GpuANSStatistics.h:23-151, :186-406), the
// same rANS arithmetic and container (GpuANSEncode.h:24-164, :180-284).  What
// changes is the schedule.  One queue of work items (an atomic counter) holds
// the histogram chunks of every buffer and its encode tickets, the tickets of
// a buffer queued a few hundred items after its last chunk; every wavefront
// takes items one at a time:
//  * a histogram item counts one chunk of one buffer (all its byte planes)
//    into LDS and adds the counts to the buffer's accumulators; the wavefront
//    that finishes a buffer's last chunk normalises its planes (one wavefront
//    per buffer), writes the frame and plane headers and publishes the
//    encoder tables, then marks the buffer ready;
//  * a 64-lane encode ticket is 1, 2 or 4 rounds of four chains per lane (two
//    for uint32) with the single-buffer coder's step; a 32-lane ticket is two
//    units, one per half-wave.  A ticket waits until its buffer is ready,
//    codes its units and resolves its output offset by decoupled look-back
//    within its buffer.
// A ticket only waits on items taken before it (its buffer's histogram
// chunks and earlier tickets of its buffer), all held by running
// wavefronts, so the launch always drains.  Data published inside the launch
// (accumulators, tables, ready flags, look-back words) goes through
// device-scope atomics; look-back words and ready flags carry a 14-bit launch
// generation, so nothing needs clearing between launches.
#pragma once
#include "codec_device.hpp"

namespace mans {
namespace {

constexpr uint32_t kL = 64;            // lanes per unit
constexpr uint32_t kBlock = 8192;      // symbols per unit
constexpr uint32_t kRB = 10;           // ring: 1024 words per chain
constexpr uint32_t kR = 1u << kRB, kH = kR / 2, kMask = kR - 1;
constexpr uint32_t kGenShift = 48;
constexpr uint64_t kValue48 = (1ull << kGenShift) - 1;
// per-wave LDS in 32-bit words: two encoder tables (or the histogram
// counters, or the normalisation keys) | two rings | dummy slots
constexpr uint32_t kTabWords = 2 * 256 * 4;
constexpr uint32_t kRingWords = 2 * kR / 2;
constexpr uint32_t kWaveWords = kTabWords + kRingWords + 64;
// Queue items a decode wavefront takes at once (BatchDecodeArgs::run, 1 or
// 4): consecutive tickets of a buffer then run on one wavefront, which keeps
// the buffer's LUTs in its LDS.  (Encode takes one at a time: a ticket's
// look-back waits for its predecessor's count, and with runs that
// predecessor would be the last of another wavefront's run, so the runs of a
// buffer would follow one another.)  An item waits only on lower items, each
// held by a running wavefront that reaches it, so the launch drains.

template <typename T>
using gptr = const __attribute__((address_space(1))) T*;

__device__ __forceinline__ uint32_t wave_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t atomic_ld64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atomic_ld32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// orders this wavefront's LDS accesses (a workgroup-scope fence emits no
// wait on global memory operations)
__device__ __forceinline__ void drain() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); }
// Precedes each flag or counter that tells another wavefront it may read
// data this wavefront published.  All such data is written with
// device-scope atomics (stores carry the device scope bit, adds are
// performed where every XCD sees them) and read with device-scope atomic
// loads, so it is enough that those writes have completed: wait for every
// outstanding vector memory operation of the wavefront.  (A device-scope
// release fence would also write back the XCD's L2, which holds none of
// this data and costs a writeback per item.)
__device__ __forceinline__ void publish() { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Look-back within one buffer's tickets [first, ...), each status word
// tagged by the launch generation (an entry of another generation reads as
// not yet published).  A ticket publishes its aggregate as soon as it is
// coded (publish_agg_gen: the buffer's first ticket publishes it as its
// inclusive prefix) and resolves its offset when its predecessors have
// published theirs (resolve_gen; with TRY, ~0 if one of them has not yet).
__device__ __forceinline__ void publish_agg_gen(uint64_t* status, uint64_t idx, uint64_t first, uint64_t agg,
                                                uint32_t gen, int lane) {
  const uint64_t G = (uint64_t)gen << kGenShift;
  if (lane == 0) atomic_st64(status + idx, (idx == first ? kFlagPrefix : kFlagAgg) | G | agg);
}
template <bool TRY>
__device__ uint64_t resolve_gen(uint64_t* status, uint64_t idx, uint64_t first, uint64_t agg, uint32_t gen,
                                int lane) {
  if (idx == first) return 0;
  const uint64_t G = (uint64_t)gen << kGenShift;
  const uint64_t excl = look_back_window<8, TRY>(status, (int64_t)idx - 1, (int64_t)first, kFlagPrefix | G, lane,
                                            [gen](uint64_t v, uint32_t& f, uint64_t& val) {
                                              f = ((v >> kGenShift) & kBatchGenMask) == gen ? (uint32_t)(v >> 62) : 0u;
                                              val = v & kValue48;
                                            });
  if (TRY && excl == ~0ull) return excl;
  if (lane == 0) atomic_st64(status + idx, kFlagPrefix | G | (excl + agg));
  return excl;
}

// ---- histogram chunks -----------------------------------------------------
// A histogram item (bytes [c0, c1) of a buffer; plane = byte offset % P) is
// counted by a whole workgroup into its LDS counters [P][256 bins][R]
// (32-bit words, R = 32 / P replicas, thread t adding into replica t % R:
// the 32 lanes of an LDS pass hit at most two words of one bank, and for
// uint8 buffers none), which it then adds to the buffer's accumulators.
// Loads: kHistLoads 16-byte loads per thread in flight while the previous
// ones are counted.
constexpr uint32_t kHistCntWords = 8192;  // P x 256 x 32 / P
constexpr int kHistLoads = 4;

template <uint32_t EB, int NW>
__device__ __forceinline__ void wg_hist_chunk(const uint8_t* in, uint64_t c0, uint64_t c1, uint32_t* cnt) {
  constexpr uint32_t P = EB, R = 32u / EB, T = 64u * NW;
  const uint32_t t = threadIdx.x;
  uint32_t* const c = cnt + t % R;
  const uint4* p4 = (const uint4*)(in + c0);
  const uint32_t n16 = (uint32_t)((c1 - c0) / 16u);  // (a chunk is at most 512 KiB)
  auto count16 = [&](const uint4& v) {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int m = 0; m < 4; ++m)
        atomicAdd(c + ((uint32_t)((4 * q + m) % (int)P) * 256u + __builtin_amdgcn_ubfe(w[q], 8u * (uint32_t)m, 8u)) * R,
                  1u);
  };
  // rows of kHistLoads uint4 per thread, two buffers; loads are
  // unconditional (indices past the end re-read uint4 0 and are not counted)
  constexpr uint32_t ROW = T * kHistLoads;
  auto load = [&](uint4 (&buf)[kHistLoads], uint32_t base) {
#pragma unroll
    for (int k = 0; k < kHistLoads; ++k) {
      const uint32_t i = base + T * (uint32_t)k + t;
      buf[k] = p4[i < n16 ? i : 0u];
    }
  };
  auto count = [&](const uint4 (&buf)[kHistLoads], uint32_t base) {
#pragma unroll
    for (int k = 0; k < kHistLoads; ++k)
      if (base + T * (uint32_t)k + t < n16) count16(buf[k]);
  };
  if (n16 > 0) {
    uint4 bufA[kHistLoads], bufB[kHistLoads];
    uint32_t r = 0;
    load(bufA, 0);
    for (;;) {
      load(bufB, r + ROW);
      count(bufA, r);
      r += ROW;
      if (r >= n16) break;
      load(bufA, r + ROW);
      count(bufB, r);
      r += ROW;
      if (r >= n16) break;
    }
  }
  const uint64_t tail = c0 + 16ull * n16;  // (a multiple of P from c0: plane t % P)
  if (tail + t < c1) atomicAdd(c + ((t % P) * 256u + in[tail + t]) * R, 1u);
}

// Adds the workgroup's counters to buffer b's accumulators and zeroes them:
// thread t sums bin t (t + T, ...) of the P x 256, its replicas read from a
// rotated start so the lanes of a pass spread over the banks.
template <uint32_t EB, int NW>
__device__ __forceinline__ void wg_hist_merge(uint32_t b, uint32_t* cnt, uint32_t* ghist) {
  constexpr uint32_t P = EB, R = 32u / EB, T = 64u * NW;
  for (uint32_t i = threadIdx.x; i < P * 256u; i += T) {
    uint32_t* const row = cnt + i * R;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < R; ++q) {
      const uint32_t qq = (q + i) % R;
      sum += row[qq];
      row[qq] = 0;
    }
    if (sum) atomicAdd(ghist + (uint64_t)b * 1024u + i, sum);
  }
}

// The histogram phase of a workgroup that starts on histograms: it takes
// chunk items from their own counter (queue[3]) until none is left, one
// workgroup barrier round per chunk (the next grab is taken while the
// current chunk is counted); the workgroup that counts a buffer's last chunk
// runs its finisher on wavefront 0.  LDS: the counters, the finisher's 256
// keys, the two grabbed items and the finish flag.
template <int PB>
__device__ void finish_buffer(const BatchBuf& B, uint32_t b, const BatchEncodeArgs& a, uint32_t gen, uint32_t* keys);

template <int PB, int NW>
__device__ void hist_phase(const BatchEncodeArgs& a, uint32_t gen, uint32_t* lds) {
  static_assert(kHistCntWords + 256u + 4u <= NW * kWaveWords, "the histogram phase fits the workgroup's LDS");
  uint32_t* const cnt = lds;
  uint32_t* const keys = lds + kHistCntWords;
  uint32_t* const s_it = keys + 256;  // [0..1] grabbed items, [2] finish flag
  for (uint32_t i = threadIdx.x; i < kHistCntWords; i += 64u * NW) cnt[i] = 0;
  if (threadIdx.x == 0) s_it[0] = atomicAdd(a.queue + 3, 1u);
  __syncthreads();
  for (uint32_t k = 0;; ++k) {
    const uint32_t it = wave_u32(s_it[k & 1u]);
    if (it >= a.nhist_items) break;
    if (threadIdx.x == 0) s_it[(k + 1) & 1u] = atomicAdd(a.queue + 3, 1u);
    const uint2 itm = a.items[it];
    const uint32_t b = wave_u32(itm.x), li = wave_u32(itm.y);
    const BatchBuf& B = a.bufs[b];
    const uint64_t bytes = B.n * B.eb;
    const uint64_t c0 = (uint64_t)li * a.hist_chunk;
    const uint64_t c1 = c0 + a.hist_chunk < bytes ? c0 + a.hist_chunk : bytes;
    if (c0 < c1) {
      if (B.eb == 1) wg_hist_chunk<1, NW>(B.in, c0, c1, cnt);
      else if (B.eb == 2) wg_hist_chunk<2, NW>(B.in, c0, c1, cnt);
      else wg_hist_chunk<4, NW>(B.in, c0, c1, cnt);
    }
    __syncthreads();
    if (B.eb == 1) wg_hist_merge<1, NW>(b, cnt, a.hist);
    else if (B.eb == 2) wg_hist_merge<2, NW>(b, cnt, a.hist);
    else wg_hist_merge<4, NW>(b, cnt, a.hist);
    publish();
    __syncthreads();
    if (threadIdx.x == 0) s_it[2] = atomicAdd(a.done + b, 1u) + 1u == B.nhist ? 1u : 0u;
    __syncthreads();
    if (s_it[2] && threadIdx.x < 64) finish_buffer<PB>(B, b, a, gen, keys);
  }
  __syncthreads();  // (the LDS becomes the wavefronts' ticket areas)
}

// ---- finisher: one wavefront normalises every plane of buffer b -----------
template <int PB>
__device__ void finish_buffer(const BatchBuf& B, uint32_t b, const BatchEncodeArgs& a, uint32_t gen, uint32_t* keys) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t P = B.eb;
  bool bad = false;
  for (uint32_t k = 0; k < P; ++k) {
    uint32_t* h = a.hist + (uint64_t)b * 1024u + k * 256u;
    uint32_t c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      c[j] = atomic_ld32(h + 4 * lane + j);
      atomic_st32(h + 4 * lane + j, 0u);
    }
    uint4 e[4];
    normalize_wave(c, (uint32_t)B.n, PB, keys, e, bad);
    uint64_t* t2 = (uint64_t*)(a.tables + ((uint64_t)b * 4u + k) * 256u);
    uint16_t pr[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // the encoder's form (enc64_entry)
      const uint4 f = enc64_entry(e[j].x, e[j].y, PB);
      atomic_st64(t2 + 2 * (4 * lane + j), ((uint64_t)f.y << 32) | f.x);
      atomic_st64(t2 + 2 * (4 * lane + j) + 1, ((uint64_t)f.w << 32) | f.z);
      pr[j] = (uint16_t)e[j].x;
    }
    uint8_t* sec = B.out + B.section[k];
    ((uint2*)(sec + kAnsHeaderBytes))[lane] =
        make_uint2((uint32_t)pr[0] | ((uint32_t)pr[1] << 16), (uint32_t)pr[2] | ((uint32_t)pr[3] << 16));
    if (lane < 8) {
      uint32_t f = 0;
      if (lane == 0) f = (kAnsMagic << 16) | (B.lanes == 32 ? 1u : 2u);  // reference / native format version
      if (lane == 1) f = (uint32_t)B.nb;
      if (lane == 2) f = (uint32_t)B.n;
      if (lane == 4) f = (uint32_t)PB;
      // totalCompressedWords (field 3) is written by the buffer's last
      // ticket to finish; no other store may touch its bytes before that
      if (lane != 3 || B.nticks == 0) ((uint32_t*)sec)[lane] = f;
    }
    if ((B.nb & 1) && lane < 2) ((uint32_t*)(sec + kSectionFixed + 4ull * B.lanes * B.nb + 8ull * B.nb))[lane] = 0;
  }
  // frame header; data_words / total_size (bytes 48..63) belong to the
  // buffer's last ticket unless there is none
  if (lane < 32 && (lane < 12 || lane >= 16 || B.nticks == 0)) ((uint32_t*)B.out)[lane] = B.frame_hdr[lane];
  if (B.nticks == 0 && lane == 0) a.sizes[b] = B.data_off;
  if (bad && lane == 0) atomicOr(a.err, kErrTable);
  publish();
  if (lane == 0) {
    atomic_st32(a.done + b, 0u);
    atomic_st32(a.ready + b, gen);
  }
}

// ---- encode ticket --------------------------------------------------------
// An encode ticket of a buffer of EB-byte elements is `rounds` (1, 2 or 4,
// BatchEncodeArgs::rounds) rounds of enc_chains(EB) units: four chains per
// lane (four blocks of a uint8 buffer, or both planes of two uint16 blocks)
// or two (two planes of one uint32 block), so a lane runs four (two)
// interleaved chains per round.  One queue item, one table load and one
// look-back serve the whole ticket.
constexpr uint32_t kMaxRounds = 4;
constexpr uint32_t kMaxTicketUnits = 16;  // kMaxRounds x four chains
// A wavefront's spill slot: an area of chain_stage_words per unit of a
// ticket, then the units' padded word counts (kMaxTicketUnits uint32, in a
// 128-byte tail).
template <int PB>
__host__ __device__ constexpr uint64_t batch_slot_words() {
  return (uint64_t)kMaxTicketUnits * chain_stage_words<PB>() + 64u;
}
__host__ __device__ constexpr uint32_t enc_chains(uint32_t eb) { return eb == 4 ? 2u : 4u; }
__host__ __device__ constexpr uint32_t enc_ticket_units(uint32_t eb, uint32_t rounds) {
  return rounds * enc_chains(eb);
}

// Ticket lt of buffer B: units u0 .. u0 + UT - 1, u0 = UT lt, UT = rounds x
// chains, coded round by round with the single-buffer launch's multi-chain
// step (enc_step64f, a linear LDS stage of H + 256 words per chain, spilled
// to the wave's slot in chunks of H; unit j of the ticket has slot area j).
// uint32 tickets of four rounds code planes 0-1 of both blocks, then planes
// 2-3, so the tables change once; of two rounds, planes 0-1 then 2-3 of one
// block; of one round, two planes of a block.
// Once all its rounds are coded the ticket publishes its aggregate and
// leaves its units' padded sizes at the tail of its slot; its offset is
// resolved by the caller (resolve_gen, place_ticket).  Returns the aggregate.  tkey: which
// tables the wave's LDS holds ((buffer << 4) | first plane, ~0 for none).
// KBB: elements loaded ahead per block.
template <int PB, int EB, int KBB>
__device__ uint64_t encode_ticket(const BatchBuf& B, uint32_t b, uint32_t lt, uint64_t tix, const BatchEncodeArgs& a,
                              uint32_t gen, uint32_t* wl, uint16_t* spill, uint64_t& tkey) {
  constexpr int CH = (int)enc_chains(EB);                    // chains per round
  constexpr int UT = (int)enc_ticket_units(EB, kMaxRounds);  // units per ticket, at most
  const uint32_t R = a.rounds, UTR = R * (uint32_t)CH;      // rounds, units of this ticket
  constexpr int SPB = EB == 4 ? 2 : EB;          // chains per element block (planes coded together)
  constexpr int BPT = CH / SPB;                  // element blocks per round
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t H = 256, H4 = H / 8, REG = H + 256;
  static_assert(kBlock / kL % KBB == 0 && KBB % 4 == 0, "KBB steps per prefetch batch, checks every 4");
  static_assert((uint32_t)CH * REG <= 2 * kRingWords, "stages fit the wave's stage area");
  using T = typename ElemOf<EB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t P = EB;
  const uint64_t U = (uint64_t)P * B.nb;
  const uint64_t u0 = (uint64_t)UTR * lt;
  uint4* s_tab = (uint4*)wl;  // [SPB <= 2][256]
  uint16_t* wstage = (uint16_t*)(wl + kTabWords);
  const uint32_t dummy_addr = lds_addr(wstage + CH * REG + 2u * lane);
  static_assert(UT <= kMaxTicketUnits, "a ticket's units fit its slot");
  uint32_t padj[UT];  // per unit of the ticket (wave-uniform)
#pragma unroll
  for (int j = 0; j < UT; ++j) padj[j] = 0;

#pragma unroll 1
  for (uint32_t r = 0; r < R; ++r) {
    // the round's units: position j0 + c in the ticket (uint32 of four
    // rounds: block r % 2, planes 2 (r / 2) + c)
    const uint32_t j0 = EB == 4 && R == 4 ? 4u * (r % 2u) + 2u * (r / 2u) : (uint32_t)CH * r;
    if (u0 + j0 >= U) continue;  // (uint32: the second block's units may be missing)
    const uint32_t p0 = EB == 4 ? (uint32_t)((u0 + j0) % 4u) : 0u;  // plane of chain 0
    uint32_t x[CH], n[CH], fl[CH], cb[CH], sb[CH];
    uint64_t blk[CH];
    bool act[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint64_t u = u0 + j0 + c;
      act[c] = u < U;
      blk[c] = act[c] ? u / P : 0u;
      const uint64_t start = blk[c] * kBlock;
      n[c] = act[c] ? (uint32_t)((B.n - start) < kBlock ? (B.n - start) : kBlock) : 0u;
      x[c] = kMinState;
      fl[c] = 0;
      cb[c] = lds_addr(wstage + c * REG);
      sb[c] = cb[c];
    }
    // tables of the round's planes, published by the finisher
    const uint64_t key = ((uint64_t)b << 4) | p0;
    if (key != tkey) {
      drain();  // the previous round's table reads are done
#pragma unroll
      for (int q = 0; q < SPB; ++q) {
        const uint64_t* g = (const uint64_t*)(a.tables + ((uint64_t)b * 4u + p0 + q) * 256u);
        uint64_t* d = (uint64_t*)(s_tab + q * 256);
#pragma unroll
        for (int k = 0; k < 8; ++k) d[lane + 64 * k] = atomic_ld64(g + lane + 64 * k);
      }
      tkey = key;
    }
    uint32_t n_min = n[0], n_max = n[0];
#pragma unroll
    for (int c = 1; c < CH; ++c) {
      if (act[c]) n_min = n[c] < n_min ? n[c] : n_min;
      n_max = n[c] > n_max ? n[c] : n_max;
    }
    n_min = wave_u32(n_min);
    n_max = wave_u32(n_max);
    const bool all_act = act[CH - 1];
    const uint32_t all = (uint32_t)div_up(n_max, kL);
    drain();
    uint16_t* const rspill = spill + (uint64_t)j0 * CSTAGE;  // chain c: slot area j0 + c

    auto spill_full = [&]() {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const uint32_t un = (sb[c] - cb[c]) >> 1;  // unspilled words
        if (un >= H) {
          uint16_t* st = wstage + c * REG;
          if (lane < H4) ((uint4*)(rspill + c * CSTAGE + fl[c]))[lane] = ((const uint4*)st)[lane];
          if (4u * lane < un - H) ((uint2*)st)[lane] = ((const uint2*)(st + H))[lane];
          fl[c] += H;
          sb[c] -= 2u * H;
        }
      }
    };

    // block d of the round: its elements; chain d * SPB + q codes byte
    // p0 + q of each (uint32 planes 2-3: the element is shifted down 16 bits)
    const T* eb[BPT];
#pragma unroll
    for (int d = 0; d < BPT; ++d) eb[d] = (const T*)B.in + blk[d * SPB] * kBlock;
    uint32_t s = 0;
    if (n_min == kBlock && all_act) {
      auto load = [&](uint32_t (&buf)[BPT][KBB], uint32_t s0) {
#pragma unroll
        for (int d = 0; d < BPT; ++d) {
          const gptr<T> bp = (gptr<T>)eb[d] + lane + (uint64_t)s0 * kL;
#pragma unroll
          for (int k = 0; k < KBB; ++k) buf[d][k] = bp[k * kL];
        }
      };
      auto lookup = [&](uint4 (&e)[CH], const uint32_t (&buf)[BPT][KBB], int k) {
#pragma unroll
        for (int d = 0; d < BPT; ++d) {
          const uint32_t w = EB == 4 ? buf[d][k] >> (8u * p0) : buf[d][k];
          e[d * SPB] = *(const uint4*)((const char*)s_tab + byte_x16<0>(w));
          if constexpr (SPB > 1) e[d * SPB + 1] = *(const uint4*)((const char*)(s_tab + 256) + byte_x16<1>(w));
        }
      };
      auto code = [&](const uint32_t (&buf)[BPT][KBB]) {
        uint4 e[CH];
        lookup(e, buf, 0);
#pragma unroll
        for (int k = 0; k < KBB; ++k) {
          uint4 en[CH];
          if (k + 1 < KBB) lookup(en, buf, k + 1);
#pragma unroll
          for (int c = 0; c < CH; ++c) enc_step64f(x[c], sb[c], dummy_addr, e[c]);
          if ((k & 3) == 3) spill_full();
          if (k + 1 < KBB) {
#pragma unroll
            for (int c = 0; c < CH; ++c) e[c] = en[c];
          }
        }
      };
      const uint32_t last = kBlock / kL - KBB;
      uint32_t bA[BPT][KBB], bB[BPT][KBB];
      load(bA, 0);
      for (;;) {
        load(bB, s + KBB <= last ? s + KBB : last);
        code(bA);
        s += KBB;
        if (s > last) break;
        load(bA, s + KBB <= last ? s + KBB : last);
        code(bB);
        s += KBB;
        if (s > last) break;
      }
    }
    for (; s < all; ++s) {
      const uint32_t i = s * kL + lane;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const bool v = i < n[c];
        const uint32_t w = v ? (uint32_t)eb[c / SPB][i] : 0u;
        const uint32_t sy = (w >> (8u * (p0 + (uint32_t)(c % SPB)))) & 0xffu;
        enc_step64(x[c], sb[c], dummy_addr, s_tab[(c % SPB) * 256 + sy], v);
      }
      spill_full();
    }

    // pad, stage remainder to the slot, states, block sizes
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint32_t rem = (sb[c] - cb[c]) >> 1;
      const uint32_t cnt = fl[c] + rem;
      const uint32_t padded = act[c] ? (uint32_t)round_up(cnt, 8) : 0u;
      if (cnt > CSTAGE && lane == 0) atomicOr(a.err, kErrStageOverflow);
      uint16_t* st = wstage + c * REG;
      if (lane < padded - cnt) st[rem + lane] = 0;
      const uint32_t nr = padded > fl[c] ? (padded - fl[c]) / 8 : 0u;
      for (uint32_t k = lane; k < nr; k += kL) ((uint4*)(rspill + c * CSTAGE + fl[c]))[k] = ((const uint4*)st)[k];
      const uint32_t kp = p0 + (uint32_t)(c % SPB);
      if (act[c]) {
        ((uint32_t*)(B.out + B.section[kp] + kSectionFixed))[blk[c] * kL + lane] = x[c];
        if (lane == 0) {
          ((uint32_t*)(B.out + B.section[kp] + kSectionFixed + 4ull * kL * B.nb + 8ull * blk[c]))[0] =
              (n[c] << 16) | cnt;
          atomicAdd(a.ptot + 4ull * b + kp, padded);
        }
      }
#pragma unroll
      for (int j = 0; j < UT; ++j)
        if ((uint32_t)j == j0 + (uint32_t)c) padj[j] = padded;
    }
  }

  // the units' padded sizes to the slot's tail, the aggregate to the
  // ticket's status word
  uint64_t agg = 0;
  uint32_t pv = 0;
#pragma unroll
  for (int j = 0; j < UT; ++j) {
    agg += padj[j];
    pv = lane == (uint32_t)j ? padj[j] : pv;
  }
  if (lane < kMaxTicketUnits) ((uint32_t*)(spill + kMaxTicketUnits * CSTAGE))[lane] = pv;
  publish_agg_gen(a.status, tix, tix - lt, agg, gen, (int)lane);
  return agg;
}

// ---- 32-lane (reference-format) encode ticket ------------------------------
// The 32-lane format (4096-symbol blocks): a ticket is enc32_units(EB) =
// 2 NC consecutive units u0 = 2 NC lt, ..., lanes 0-31 coding units u0 ..
// u0 + NC - 1 and lanes 32-63 the next NC, NC chains per lane (unit
// u0 + NC h + c is chain c of half h), with the single-buffer 32-lane step
// (enc_step32f: one ballot per chain, words compacted per half).  NC = 2
// for uint8 (two blocks per half) and uint16 (both planes of one block per
// half: one element load feeds both chains); uint32 tickets are the two
// units u0, u0 + 1 (planes p0, p0 + 1 of one block), one chain per lane,
// since four consecutive uint32 units would need four planes' tables.  Each
// unit has a linear LDS stage of H + 256 words (checked every 4 steps, at
// most 32 words per step) and a slot area of CSTAGE / 2 words.  Published
// and resolved as the 64-lane tickets are.
__host__ __device__ constexpr uint32_t enc32_chains(uint32_t eb) { return eb == 4 ? 1u : 2u; }
__host__ __device__ constexpr uint32_t enc32_units(uint32_t eb) { return 2u * enc32_chains(eb); }

template <int PB, int EB, int KBB>
__device__ uint64_t encode_ticket32(const BatchBuf& B, uint32_t b, uint32_t lt, uint64_t tix, const BatchEncodeArgs& a,
                                    uint32_t gen, uint32_t* wl, uint16_t* spill, uint64_t& tkey) {
  constexpr uint32_t L = 32, BLOCK = 4096;
  constexpr int NC = (int)enc32_chains(EB);    // chains per lane
  constexpr int NU = 2 * NC;                   // units per ticket
  constexpr int NBK = EB == 1 ? NC : 1;        // element blocks per lane
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t SSTAGE = CSTAGE / 2;  // a 32-lane unit's words at most, rounded
  constexpr uint32_t H = 256, H4 = H / 8, REG = H + 256;
  constexpr uint32_t NT = EB == 1 ? 1u : 2u;  // tables of the ticket
  static_assert(BLOCK / L % KBB == 0 && KBB % 4 == 0, "KBB steps per prefetch batch, checks every 4");
  static_assert(NU * REG + 128 <= 2 * (kRingWords + 64), "stages and dummy slots fit the wave's stage area");
  static_assert(SSTAGE % 8 == 0 && NU * SSTAGE <= kMaxTicketUnits * CSTAGE && NU <= (int)kMaxTicketUnits, "slot areas");
  using T = typename ElemOf<EB>::T;
  const uint32_t lane = threadIdx.x & 63u, half = lane >> 5, lin = lane & 31u;
  const uint32_t P = EB;
  const uint64_t U = (uint64_t)P * B.nb;
  const uint64_t u0 = (uint64_t)NU * lt;
  const uint32_t p0 = (uint32_t)(u0 % P);
  // this lane's chains: unit, plane, block, symbols
  bool act[NC];
  uint32_t kp[NC], n[NC];
  uint64_t blk[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const uint64_t u = u0 + (uint64_t)NC * half + c;
    act[c] = u < U;
    kp[c] = act[c] ? (uint32_t)(u % P) : p0;
    blk[c] = act[c] ? u / P : 0u;
    n[c] = act[c] ? (uint32_t)((B.n - blk[c] * BLOCK) < BLOCK ? (B.n - blk[c] * BLOCK) : BLOCK) : 0u;
  }
  uint4* s_tab = (uint4*)wl;  // [NT][256]
  uint16_t* wstage = (uint16_t*)(wl + kTabWords);
  const uint32_t dummy_addr = lds_addr(wstage + NU * REG + 2u * lane);
  const uint64_t key = ((uint64_t)b << 4) | p0;
  if (key != tkey) {
    drain();  // the previous ticket's table reads are done
#pragma unroll
    for (uint32_t q = 0; q < NT; ++q) {
      const uint64_t* g = (const uint64_t*)(a.tables + ((uint64_t)b * 4u + p0 + q) * 256u);
      uint64_t* d = (uint64_t*)(s_tab + q * 256);
#pragma unroll
      for (int k = 0; k < 8; ++k) d[lane + 64 * k] = atomic_ld64(g + lane + 64 * k);
    }
    tkey = key;
  }
  const uint4* tab[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) tab[c] = s_tab + (kp[c] - p0) * 256u;
  uint32_t x[NC];
  // per unit j = NC h + c: LDS byte address of the stage and of the next
  // word, words moved to the slot area
  uint32_t cb[NU], sb[NU], fl[NU];
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    cb[j] = wave_u32(lds_addr(wstage + j * REG));
    sb[j] = cb[j];
    fl[j] = 0;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) x[c] = kMinState;
  uint32_t n_min = BLOCK, n_max = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const uint32_t a0 = __builtin_amdgcn_readlane(n[c], 0), a1 = __builtin_amdgcn_readlane(n[c], 32);
    n_min = a0 < n_min ? a0 : n_min;
    n_min = a1 < n_min ? a1 : n_min;
    n_max = a0 > n_max ? a0 : n_max;
    n_max = a1 > n_max ? a1 : n_max;
  }
  const uint32_t all = (uint32_t)div_up(n_max, L);
  drain();

  // a stage with H or more unspilled words moves its first H to its slot
  // area and the rest (fewer than 256) to the front
  auto spill_full = [&]() {
#pragma unroll
    for (int j = 0; j < NU; ++j) {
      const uint32_t un = wave_u32((sb[j] - cb[j]) >> 1);  // unspilled words
      if (un >= H) {
        uint16_t* st = wstage + j * REG;
        if (lane < H4) ((uint4*)(spill + j * SSTAGE + fl[j]))[lane] = ((const uint4*)st)[lane];
        if (4u * lane < un - H) ((uint2*)st)[lane] = ((const uint2*)(st + H))[lane];
        fl[j] += H;
        sb[j] = wave_u32(sb[j] - 2u * H);
      }
    }
  };

  // block d of this lane (uint8: chain d's block; otherwise the half's
  // block, chain c taking byte kp[c] of each element)
  const T* ep[NBK];
#pragma unroll
  for (int d = 0; d < NBK; ++d) ep[d] = (const T*)B.in + blk[d] * BLOCK;
  auto shift_of = [&](int c) -> uint32_t { return EB == 1 ? 0u : 8u * kp[c]; };
  uint32_t s = 0;
  if (n_min == BLOCK) {
    auto load = [&](uint32_t (&buf)[NBK][KBB], uint32_t s0) {
#pragma unroll
      for (int d = 0; d < NBK; ++d) {
        const gptr<T> bp = (gptr<T>)ep[d] + lin + (uint64_t)s0 * L;
#pragma unroll
        for (int k = 0; k < KBB; ++k) buf[d][k] = bp[k * L];
      }
    };
    auto code = [&](const uint32_t (&buf)[NBK][KBB]) {
#pragma unroll
      for (int k = 0; k < KBB; ++k) {
        uint4 e[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) e[c] = tab[c][(buf[EB == 1 ? c : 0][k] >> shift_of(c)) & 0xffu];
#pragma unroll
        for (int c = 0; c < NC; ++c) enc_step32f(x[c], sb[c], sb[NC + c], half, e[c]);
        if ((k & 3) == 3) spill_full();
      }
    };
    const uint32_t last = BLOCK / L - KBB;
    uint32_t bA[NBK][KBB], bB[NBK][KBB];
    load(bA, 0);
    for (;;) {
      load(bB, s + KBB <= last ? s + KBB : last);
      code(bA);
      s += KBB;
      if (s > last) break;
      load(bA, s + KBB <= last ? s + KBB : last);
      code(bB);
      s += KBB;
      if (s > last) break;
    }
  }
  for (; s < all; ++s) {
    const uint32_t i = s * L + lin;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const bool v = i < n[c];
      const uint32_t w = v ? (uint32_t)ep[EB == 1 ? c : 0][i] : 0u;
      enc_step32(x[c], sb[c], sb[NC + c], half, dummy_addr, tab[c][(w >> shift_of(c)) & 0xffu], v);
    }
    spill_full();
  }

  // pad each unit, its stage remainder to its slot area, states, block sizes
  uint32_t pad[NU], cnt[NU];
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    const uint32_t rem = (sb[j] - cb[j]) >> 1;
    const bool actj = u0 + (uint64_t)j < U;
    cnt[j] = fl[j] + rem;
    pad[j] = actj ? (uint32_t)round_up(cnt[j], 8) : 0u;
    if (cnt[j] > SSTAGE && lane == 0) atomicOr(a.err, kErrStageOverflow);
    uint16_t* st = wstage + j * REG;
    if (lane < pad[j] - cnt[j]) st[rem + lane] = 0;
    const uint32_t nr = pad[j] > fl[j] ? (pad[j] - fl[j]) / 8 : 0u;
    for (uint32_t k = lane; k < nr; k += 64) ((uint4*)(spill + j * SSTAGE + fl[j]))[k] = ((const uint4*)st)[k];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (act[c]) {
      uint8_t* sec = B.out + B.section[kp[c]] + kSectionFixed;
      ((uint32_t*)sec)[blk[c] * L + lin] = x[c];
      if (lin == 0) {
        const uint32_t cu = half ? cnt[NC + c] : cnt[c];
        const uint32_t pu = half ? pad[NC + c] : pad[c];
        ((uint32_t*)(sec + 4ull * L * B.nb + 8ull * blk[c]))[0] = (n[c] << 16) | cu;
        atomicAdd(a.ptot + 4ull * b + kp[c], pu);
      }
    }
  }

  // the units' padded sizes to the slot's tail, the aggregate to the
  // ticket's status word
  uint64_t agg = 0;
  uint32_t pv = 0;
#pragma unroll
  for (int j = 0; j < NU; ++j) {
    agg += pad[j];
    pv = lane == (uint32_t)j ? pad[j] : pv;
  }
  if (lane < kMaxTicketUnits) ((uint32_t*)(spill + kMaxTicketUnits * CSTAGE))[lane] = pv;
  publish_agg_gen(a.status, tix, tix - lt, agg, gen, (int)lane);
  return agg;
}

// ---- placing a resolved ticket ---------------------------------------------
// Ticket lt of buffer B, coded by this wavefront into `slot` with aggregate
// agg and exclusive offset excl (in words): each unit's block word offset,
// its words from the slot areas (unit j at j x AREA words) to their place in
// unit order; the buffer's last ticket writes the frame's sizes, the last to
// finish each plane's totalCompressedWords (section header field 3).
template <int PB, int L>
__device__ void place_ticket(const BatchBuf& B, uint32_t b, uint32_t lt, uint64_t agg, uint64_t excl,
                             const BatchEncodeArgs& a, const uint16_t* slot) {
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t AREA = L == 32 ? CSTAGE / 2 : CSTAGE;  // a unit's slot area
  constexpr int UT = L == 32 ? 4 : (int)kMaxTicketUnits;     // units per ticket, at most
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t P = B.eb;
  const uint64_t U = (uint64_t)P * B.nb;
  const uint32_t ut = L == 32 ? enc32_units(P) : enc_ticket_units(P, a.rounds);
  const uint64_t u0 = (uint64_t)ut * lt;
  // lane j < UT: unit j's padded words and the first uint4 of its words in
  // the ticket's output (exclusive scan over the units)
  publish();  // the slot's stores (by every lane of this wave) are done
  const uint32_t pq = lane < (uint32_t)UT ? ((const uint32_t*)(slot + kMaxTicketUnits * CSTAGE))[lane] / 8u : 0u;
  uint32_t inc = pq;
#pragma unroll
  for (int o = 1; o < UT; o <<= 1) {
    const uint32_t v = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += v;
  }
  const uint32_t exq = inc - pq;
  const uint32_t nq = __builtin_amdgcn_readlane(inc, UT - 1);  // the ticket's uint4s
  uint32_t qb[UT];  // first uint4 of each unit (wave-uniform)
#pragma unroll
  for (int k = 1; k < UT; ++k) qb[k] = __builtin_amdgcn_readlane(exq, k);
  {
    const uint64_t u = u0 + lane;
    if (lane < ut && u < U) {
      const uint32_t kp = (uint32_t)(u % P);
      ((uint32_t*)(B.out + B.section[kp] + kSectionFixed + 4ull * L * B.nb + 8ull * (u / P)))[1] =
          (uint32_t)(excl + 8ull * exq);
    }
  }
  {
    const uint4* s4 = (const uint4*)slot;
    uint4* dst = (uint4*)(B.out + B.data_off + 2ull * excl);
    constexpr int CBQ = 8;
    for (uint32_t j0 = 0; j0 < nq; j0 += CBQ * 64) {
      uint4 q[CBQ];
#pragma unroll
      for (int u = 0; u < CBQ; ++u) {
        const uint32_t j = j0 + u * 64 + lane;
        const uint32_t jj = j < nq ? j : 0u;
        uint32_t dd = 0;
#pragma unroll
        for (int k = 1; k < UT; ++k) dd = jj >= qb[k] ? k * (AREA / 8) - qb[k] : dd;
        q[u] = s4[jj + dd];
      }
#pragma unroll
      for (int u = 0; u < CBQ; ++u) {
        const uint32_t j = j0 + u * 64 + lane;
        if (j < nq) dst[j] = q[u];
      }
    }
  }
  if (lt + 1 == B.nticks && lane == 0) {
    const uint64_t total_words = excl + agg;
    const uint64_t size = B.data_off + 2ull * total_words;
    a.sizes[b] = size;
    ((uint64_t*)B.out)[6] = total_words;  // data_words at byte 48
    ((uint64_t*)B.out)[7] = size;         // total_size at byte 56
  }
  publish();
  uint32_t f = 0;
  if (lane == 0) f = atomicAdd(a.fin + b, 1u);
  if (wave_u32(f) + 1 == B.nticks && lane < B.eb) {
    const uint32_t v = atomic_ld32(a.ptot + 4ull * b + lane);
    atomic_st32(a.ptot + 4ull * b + lane, 0u);
    ((uint32_t*)(B.out + B.section[lane]))[3] = v;
    if (lane == 0) atomic_st32(a.fin + b, 0u);
  }
}

// L: every buffer's frame format, 64 or 32 lanes (a kernel of its own for
// each, so neither ticket coder shares a kernel body with the other).
// (At most 168 VGPRs: the wave's LDS allows three four-wave workgroups per
// CU, three wavefronts per SIMD, and a larger register budget would halve
// that to two.)
template <int PB, int NW, int KB, int L>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(3))) void batch_encode_kernel(
    BatchEncodeArgs a) {
  extern __shared__ uint32_t lds[];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = wave_u32(threadIdx.x >> 6);
  uint32_t* wl = lds + wave * kWaveWords;
  // two slots, used in turn: the ticket being coded and the one whose offset
  // is still to be resolved
  uint16_t* const spill = a.spill + ((uint64_t)blockIdx.x * NW + wave) * (2ull * batch_slot_words<PB>());
  // this launch's generation (every wavefront reads it before the last one
  // out advances it)
  const uint32_t gen = wave_u32(lane == 0 ? atomic_ld32(a.queue + 2) : 0u);
  // the first hist_wgs workgroups count histogram chunks until none is left
  // (the rest start on tickets, whose buffers become ready as their chunks
  // are counted)
  if (blockIdx.x < a.hist_wgs) hist_phase<PB, NW>(a, gen, lds);
  uint64_t tkey = ~0ull;  // tables held in this wave's LDS
  // the ticket coded last if not resolved yet (wave-uniform), in the slot
  // other than `slot`
  bool pend = false;
  uint32_t p_b = 0, p_lt = 0, slot = 0;
  uint64_t p_agg = 0;
  auto resolve_pending = [&]() {
    const BatchBuf& PBuf = a.bufs[p_b];
    const uint64_t ptix = PBuf.tick_first + p_lt;
    const uint64_t excl = resolve_gen<false>(a.status, ptix, ptix - p_lt, p_agg, gen, (int)lane);
    place_ticket<PB, L>(PBuf, p_b, p_lt, p_agg, excl, a, spill + (slot ^ 1u) * batch_slot_words<PB>());
  };

  // tickets, a wavefront at a time, from their own counter (queue[0])
  const uint32_t nt = a.nitems - a.nhist_items;
  uint32_t it = 0;
  if (lane == 0) it = atomicAdd(a.queue, 1u);
  it = wave_u32(it);
  while (it < nt) {
    const uint2 itm = a.items[a.nhist_items + it];
    const uint32_t b = wave_u32(itm.x), li = wave_u32(itm.y);
    const BatchBuf& B = a.bufs[b];
    const uint32_t lt = li & ~kItemTicket;
    const uint32_t tix = B.tick_first + lt;
    while (wave_u32(lane == 0 ? atomic_ld32(a.ready + b) : 0u) != gen) __builtin_amdgcn_s_sleep(2);
    uint16_t* const sl = spill + slot * batch_slot_words<PB>();
    uint64_t agg;
    if constexpr (L == 32) {
      if (B.eb == 1) agg = encode_ticket32<PB, 1, KB>(B, b, lt, tix, a, gen, wl, sl, tkey);
      else if (B.eb == 2) agg = encode_ticket32<PB, 2, KB>(B, b, lt, tix, a, gen, wl, sl, tkey);
      else agg = encode_ticket32<PB, 4, KB>(B, b, lt, tix, a, gen, wl, sl, tkey);
    } else {
      if (B.eb == 1) agg = encode_ticket<PB, 1, (KB / 4 < 4 ? 4 : KB / 4)>(B, b, lt, tix, a, gen, wl, sl, tkey);
      else if (B.eb == 2) agg = encode_ticket<PB, 2, (KB / 2 < 4 ? 4 : KB / 2)>(B, b, lt, tix, a, gen, wl, sl, tkey);
      else agg = encode_ticket<PB, 4, KB>(B, b, lt, tix, a, gen, wl, sl, tkey);
    }
    // a ticket left unresolved last time (its predecessors have had this
    // ticket's coding time to publish), then this one if its predecessors
    // have all published, else it waits in its slot for the next ticket
    if (pend) {
      resolve_pending();
      pend = false;
    }
    const uint64_t excl = resolve_gen<true>(a.status, tix, tix - lt, agg, gen, (int)lane);
    if (excl != ~0ull) {
      place_ticket<PB, L>(B, b, lt, agg, excl, a, sl);
    } else {
      pend = true;
      p_b = b;
      p_lt = lt;
      p_agg = agg;
      slot ^= 1u;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    // the next item only once this one is done (items start in order)
    uint32_t nxt = 0;
    if (lane == 0) nxt = atomicAdd(a.queue, 1u);
    it = wave_u32(nxt);
  }
  if (pend) resolve_pending();
  // the last wavefront out resets the queue and advances the generation for
  // the next launch (a replayed graph passes the same arguments every time)
  if (lane == 0) {
    const uint32_t total = gridDim.x * NW;
    if (atomicAdd(a.queue + 1, 1u) + 1 == total) {
      atomic_st32(a.queue, 0u);
      atomic_st32(a.queue + 1, 0u);
      atomic_st32(a.queue + 3, 0u);
      atomic_st32(a.queue + 2, gen % kBatchGenMask + 1u);
    }
  }
}

template <int PB, int NW, int KB>
hipError_t launch_enc_shape(const BatchEncodeArgs& a, int cu, hipStream_t s, uint64_t* need_spill) {
  constexpr size_t lds = (size_t)NW * kWaveWords * 4u;
  static_assert(lds == batch_encode_lds(BatchShape{NW, KB}), "encode LDS layout");
  if constexpr (lds > kLdsBytes) {
    return hipErrorInvalidConfiguration;
  } else {
    int per_cu = 0;
    hipError_t e = a.lanes == 32
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_encode_kernel<PB, NW, KB, 32>, 64 * NW, lds)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_encode_kernel<PB, NW, KB, 64>, 64 * NW, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    uint64_t grid = (uint64_t)cu * (uint64_t)per_cu;
    const uint64_t want = div_up(a.nitems, NW);
    if (want < grid) grid = want;
    if (grid < 1) grid = 1;
    const uint64_t spill_bytes = grid * NW * 2ull * batch_slot_words<PB>() * 2ull;  // two slots per wave
    if (need_spill) {
      *need_spill = spill_bytes;
      return hipSuccess;
    }
    if (spill_bytes > a.spill_bytes) return hipErrorInvalidValue;
    // a quarter of the workgroups start on histogram chunks (at least one
    // when there are any: the tickets wait for them).  (4096 buffers of
    // 8.5 GB: 9.53 ms with a quarter, 9.65 with half or all of them.)
    BatchEncodeArgs g = a;
    g.hist_wgs = a.nhist_items == 0 ? 0u : (uint32_t)(grid > 3 ? grid / 4 : 1);
    if (a.lanes == 32) hipLaunchKernelGGL((batch_encode_kernel<PB, NW, KB, 32>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, g);
    else hipLaunchKernelGGL((batch_encode_kernel<PB, NW, KB, 64>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, g);
    return hipGetLastError();
  }
}

template <int PB, int NW>
hipError_t launch_enc_nw(const BatchEncodeArgs& a, int kb, int cu, hipStream_t s, uint64_t* need_spill) {
  switch (kb) {
    case 4: return launch_enc_shape<PB, NW, 4>(a, cu, s, need_spill);
    case 8: return launch_enc_shape<PB, NW, 8>(a, cu, s, need_spill);
    case 16: return launch_enc_shape<PB, NW, 16>(a, cu, s, need_spill);
  }
  return hipErrorInvalidValue;
}

template <int PB>
hipError_t launch_enc(const BatchEncodeArgs& a, BatchShape sh, int cu, hipStream_t s, uint64_t* need_spill) {
  switch (sh.nw) {
    case 4: return launch_enc_nw<PB, 4>(a, sh.kb, cu, s, need_spill);
    case 8: return launch_enc_nw<PB, 8>(a, sh.kb, cu, s, need_spill);
    case 16: return launch_enc_nw<PB, 16>(a, sh.kb, cu, s, need_spill);
  }
  return hipErrorInvalidValue;
}

// ---- batched decode --------------------------------------------------------
// Decode ring of a chain for KB steps per ring upkeep: two halves of H
// words plus a mirror of its first 64 words.
//  * KB = 1: the half below the ring is fetched into registers once the
//    upper half is being read (ptr <= lo + H) and written when the reads
//    come within one step of the ring's bottom (ptr < lo + 64).
//  * KB > 1: it is fetched as soon as the previous one is written, and
//    written once the half it replaces is consumed (ptr <= lo + H); until
//    then ptr > lo + H >= lo + 64 KB, so the KB steps to the next upkeep
//    (at most 64 words each) read inside the ring.
template <int KB>
struct DecRing {
  static constexpr uint32_t H = 64u * KB > 512u ? 64u * KB : 512u;
  static constexpr uint32_t R = 2 * H, MASK = R - 1, WORDS = R + 64;
  static_assert(64u * KB <= H, "KB steps must fit in a half");
};

// LUT item: check buffer b's frame against its plan and build the packed LUT
// (slot - cdf) << 20 | pdf << 8 | sym of every plane into global memory.
template <int PB>
__device__ void build_luts(const BatchBuf& B, uint32_t b, const BatchDecodeArgs& a, uint32_t gen, uint32_t* wl) {
  constexpr uint32_t W = 1u << PB;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t P = B.eb;
  bool bad = B.in_bytes < B.data_off;
  if (!bad) {
    const uint32_t* fh = (const uint32_t*)B.in;
    const uint8_t* fb = B.in;
    // magic, version, element width, mode 0, lanes, precision, planes, order,
    // block bytes, n, data offset and section offsets must match the plan
    bad = fh[0] != kFrameMagic || fb[4] != kFrameVersion || fb[5] != 0 || fb[6] != P || fb[7] != 0 ||
          fb[8] != B.lanes || fb[9] != PB || fb[10] != P || fb[11] != 0 || fh[3] != 128u * B.lanes ||
          ((const uint64_t*)B.in)[2] != B.n || ((const uint64_t*)B.in)[5] != B.data_off;
    for (uint32_t k = 0; k < P && !bad; ++k) bad = ((const uint64_t*)B.in)[8 + k] != B.section[k];
    const uint64_t data_words = ((const uint64_t*)B.in)[6];
    bad = bad || B.data_off + 2ull * data_words > B.in_bytes;
  }
  uint32_t* cdf = wl;  // 257 words
  for (uint32_t k = 0; k < P && !bad; ++k) {
    const uint8_t* sec = B.in + B.section[k];
    const uint32_t* h = (const uint32_t*)sec;
    if ((h[0] >> 16) != kAnsMagic || (h[0] & 0xffffu) != (B.lanes == 32 ? 1u : 2u) || (h[4] & 0xfu) != (uint32_t)PB ||
        h[1] != B.nb ||
        h[2] != (uint32_t)B.n) {
      bad = true;
      break;
    }
    const uint16_t* probs = (const uint16_t*)(sec + kAnsHeaderBytes);
    uint32_t v[4], ls = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = probs[4 * lane + j];
      ls += v[j];
    }
    uint32_t inc = ls;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(inc, o, 64);
      if (lane >= (uint32_t)o) inc += u;
    }
    uint32_t run = inc - ls;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cdf[4 * lane + j] = run;
      run += v[j];
    }
    if (lane == 63) cdf[256] = run;
    drain();
    if (B.n > 0 && cdf[256] != W) {
      bad = true;
      break;
    }
    uint32_t* g = a.luts + ((uint64_t)b * 4u + k) * W;
    for (uint32_t slot = 2 * lane; slot < W; slot += 128) {
      uint32_t ent[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const uint32_t sl = slot + q;
        uint32_t lo = 0, hi = kNumSymbols;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (cdf[mid] <= sl) lo = mid; else hi = mid;
        }
        const uint32_t c = cdf[lo], pdf = cdf[lo + 1] - c;
        ent[q] = ((sl - c) << 20) | (pdf << 8) | lo;
      }
      atomic_st64((uint64_t*)(g + slot), ((uint64_t)ent[1] << 32) | ent[0]);
    }
    drain();
  }
  bad = __ballot(bad) != 0;
  if (bad && lane == 0) {
    atomicOr(a.err, kErrMagic);
    atomicOr(a.bad + b, kErrMagic);
  }
  publish();
  if (lane == 0) atomic_st32(a.ready + b, gen | (bad ? kBatchBad : 0u));
}

// Decode ticket: units u0, u0 + 1 of buffer B with the 64-lane ring decoder.
// SAME: two planes of the same elements (stored together as 16 bits at byte
// kp[0] of each element); otherwise two blocks of a uint8 buffer.
// tkey: which LUTs the wave's LDS holds, as for encode_ticket.
template <int PB, bool SAME, int KB>
__device__ void decode_ticket(const BatchBuf& B, uint32_t b, uint32_t lt, const BatchDecodeArgs& a, uint32_t* wl,
                              uint64_t& tkey) {
  constexpr uint32_t W = 1u << PB;
  constexpr uint32_t kH = DecRing<KB>::H, kR = DecRing<KB>::R, kMask = DecRing<KB>::MASK;
  constexpr uint32_t kDecRing = DecRing<KB>::WORDS;
  constexpr uint32_t H4 = kH / 8;    // uint4 per half
  constexpr int QH = (int)(H4 / 64);  // uint4 per lane per half
  constexpr uint32_t UNIT_CAP = kL * lane_max_words(kBlock / kL, PB);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t P = B.eb;
  const uint64_t U = (uint64_t)P * B.nb;
  const uint64_t u0 = 2ull * lt;
  const uint64_t data_words = ((const uint64_t*)B.in)[6];
  uint32_t* s_lut = wl;                                   // [2][W]
  uint16_t* wstage = (uint16_t*)(wl + 2 * W);             // [2][R + 64]
  uint32_t x[2], ptr[2], n[2], wds[2], taken[2], lo[2], kp[2];
  bool act[2], pend[2];
  uint64_t blk[2];
  uint4 rv[2][QH];
  const uint4* src4[2];
  const uint32_t* lut[2];
  const uint16_t* stage[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const uint64_t u = u0 + c;
    act[c] = u < U;
    kp[c] = act[c] ? (uint32_t)(u % P) : 0u;
    blk[c] = act[c] ? u / P : 0u;
    const uint8_t* sec = B.in + B.section[kp[c]];
    uint32_t nn = 0, words = 0, y = 0;
    if (act[c]) {
      const uint2 bw = *(const uint2*)(sec + kSectionFixed + 4ull * kL * B.nb + 8ull * blk[c]);
      nn = bw.x >> 16;
      words = bw.x & 0xffffu;
      y = bw.y;
      const uint64_t start = blk[c] * kBlock;
      const uint64_t expect = (B.n - start) < kBlock ? (B.n - start) : kBlock;
      if (nn != expect || words > UNIT_CAP || (uint64_t)y + round_up(words, 8) > data_words ||
          B.data_off + 2ull * ((uint64_t)y + round_up(words, 8)) > B.in_bytes) {
        if (lane == 0) {
          atomicOr(a.err, kErrBlockWords);
          atomicOr(a.bad + b, kErrBlockWords);
        }
        act[c] = false;
        nn = 0;
        words = 0;
      }
    }
    n[c] = nn;
    words = wave_u32(words);
    y = wave_u32(y);
    lut[c] = s_lut + c * W;
    src4[c] = (const uint4*)(B.in + B.data_off + 2ull * y);
    uint16_t* reg = wstage + c * kDecRing;
    stage[c] = reg;
    lo[c] = 0;
    pend[c] = false;
#pragma unroll
    for (int q = 0; q < QH; ++q) rv[c][q] = make_uint4(0, 0, 0, 0);
    if (words > 0) {
      const uint32_t jt = (words - 1) / kH;
      const uint32_t j0 = jt > 0 ? jt - 1 : 0;
      lo[c] = j0 * kH;
      for (uint32_t j = j0; j <= jt; ++j) {
        const uint32_t left = words - j * kH;
        const uint32_t n4 = (uint32_t)div_up(left < kH ? left : kH, 8);
        for (uint32_t i = lane; i < n4; i += 64) {
          const uint4 q = src4[c][j * H4 + i];
          ((uint4*)(reg + (j & 1) * kH))[i] = q;
          if ((j & 1) == 0 && i < 8) ((uint4*)(reg + kR))[i] = q;
        }
      }
    }
    x[c] = act[c] ? ((const uint32_t*)(sec + kSectionFixed))[blk[c] * kL + lane] : kMinState;
    ptr[c] = words;
    wds[c] = words;
    taken[c] = 0;
  }
  // the chains' LUTs, built by the buffer's LUT item
  const uint64_t key = ((uint64_t)b << 4) | (kp[0] << 2) | kp[1];
  if (key != tkey) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const uint64_t* g = (const uint64_t*)(a.luts + ((uint64_t)b * 4u + kp[c]) * W);
      uint64_t* d = (uint64_t*)(s_lut + c * W);
      for (uint32_t j = lane; j < W / 2; j += 64) d[j] = atomic_ld64(g + j);
    }
    tkey = key;
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) __asm__ volatile("" : "+v"(x[c]));
  drain();

  auto fetch = [&](int c) {
    const uint32_t j = lo[c] / kH - 1;
#pragma unroll
    for (int q = 0; q < QH; ++q) rv[c][q] = src4[c][j * H4 + 64 * q + lane];
    pend[c] = true;
  };
  auto commit = [&](int c) {
    const uint32_t j = lo[c] / kH - 1;
    uint16_t* reg = wstage + c * kDecRing;
#pragma unroll
    for (int q = 0; q < QH; ++q) ((uint4*)(reg + (j & 1) * kH))[64 * q + lane] = rv[c][q];
    if ((j & 1) == 0 && lane < 8) ((uint4*)(reg + kR))[lane] = rv[c][0];
    lo[c] -= kH;
    pend[c] = false;
  };
  auto refill = [&]() {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if constexpr (KB == 1) {
        if (!pend[c] && lo[c] > 0 && ptr[c] <= lo[c] + kH) fetch(c);
        if (pend[c] && ptr[c] < lo[c] + 64u) commit(c);
      } else {
        if (!pend[c] && lo[c] > 0) fetch(c);
        if (pend[c] && ptr[c] <= lo[c] + kH) {
          commit(c);
          if (lo[c] > 0) fetch(c);
        }
      }
    }
  };

  const uint32_t n_min = wave_u32(n[0] < n[1] || !act[1] ? n[0] : n[1]);
  const uint32_t n_max = wave_u32(n[0] > n[1] ? n[0] : n[1]);
  const uint32_t full = act[1] ? n_min / kL : 0u;
  const uint32_t all = (uint32_t)div_up(n_max, kL);
  // element outputs: SAME stores 16 bits at byte kp[0] of each element
  uint8_t* o0 = B.out + blk[0] * kBlock * P + kp[0];
  uint8_t* o1 = B.out + blk[1] * kBlock * P + kp[1];
  auto store = [&](uint32_t s, uint32_t el, bool v0, bool v1) {
    const uint64_t i = (uint64_t)s * kL + lane;
    if constexpr (SAME) {
      if (v0) *(__attribute__((address_space(1))) uint16_t*)(o0 + i * P) = (uint16_t)el;
    } else {
      if (v0) *(__attribute__((address_space(1))) uint8_t*)(o0 + i) = (uint8_t)el;
      if (v1) *(__attribute__((address_space(1))) uint8_t*)(o1 + i) = (uint8_t)(el >> 8);
    }
  };
  const bool vall[2] = {true, true};
  if constexpr (KB == 1) {
    for (uint32_t s = all; s > full;) {
      --s;
      const uint32_t i = s * kL + lane;
      bool v[2] = {i < n[0], i < n[1]};
      const uint32_t el = dec_steps<kL, PB, 2, kMask, uint32_t>(x, ptr, taken, lut, stage, v, 0u);
      store(s, el, v[0], v[1]);
      refill();
    }
    for (uint32_t s = full; s > 0;) {
      --s;
      const uint32_t el = dec_steps<kL, PB, 2, kMask, uint32_t>(x, ptr, taken, lut, stage, vall, 0u);
      store(s, el, true, true);
      refill();
    }
  } else {
    // steps above `full` and full steps down to a multiple of KB one at a
    // time (upkeep after each step s with s % KB == 0), then KB full steps
    // per upkeep
    uint32_t s = all;
    while (s > full || s % KB != 0) {
      --s;
      const uint32_t i = s * kL + lane;
      bool v[2] = {i < n[0], i < n[1]};
      const uint32_t el = dec_steps<kL, PB, 2, kMask, uint32_t>(x, ptr, taken, lut, stage, v, 0u);
      store(s, el, v[0], v[1]);
      if (s % KB == 0) refill();
    }
    while (s > 0) {
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        --s;
        const uint32_t el = dec_steps<kL, PB, 2, kMask, uint32_t>(x, ptr, taken, lut, stage, vall, 0u);
        store(s, el, true, true);
      }
      refill();
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const bool bad = act[c] && (taken[c] != wds[c] || x[c] != kMinState);
    if (__ballot(bad) && lane == 0) {
      atomicOr(a.err, kErrFinalState);
      atomicOr(a.bad + b, kErrFinalState);
    }
  }
}

// Decode ticket of the 32-lane format: units u0 = 2 lt and u0 + 1 of buffer
// B, lanes 0-31 decoding the first and lanes 32-63 the second (one chain per
// lane, each unit with its own ring and LUT slot; a unit consumes at most 32
// words per step, inside the ring schedule's 64).  Each half stores its
// plane's bytes.
template <int PB, int KB>
__device__ void decode_ticket32(const BatchBuf& B, uint32_t b, uint32_t lt, const BatchDecodeArgs& a, uint32_t* wl,
                                uint64_t& tkey) {
  constexpr uint32_t L = 32, BLOCK = 4096;
  constexpr uint32_t W = 1u << PB;
  constexpr uint32_t kH = DecRing<KB>::H, kR = DecRing<KB>::R, kMask = DecRing<KB>::MASK;
  constexpr uint32_t kDecRing = DecRing<KB>::WORDS;
  constexpr uint32_t H4 = kH / 8;    // uint4 per half ring
  constexpr int QH = (int)(H4 / 64);  // uint4 per lane per half ring
  constexpr uint32_t UNIT_CAP = L * lane_max_words(BLOCK / L, PB);
  const uint32_t lane = threadIdx.x & 63u, half = lane >> 5, lin = lane & 31u;
  const uint32_t P = B.eb;
  const uint64_t U = (uint64_t)P * B.nb;
  const uint64_t u0 = 2ull * lt;
  const uint64_t data_words = ((const uint64_t*)B.in)[6];
  uint32_t* s_lut = wl;                        // [2][W]
  uint16_t* wstage = (uint16_t*)(wl + 2 * W);  // [2][R + 64]
  // per unit (stream h, wave-uniform): validity, plane, block, symbols,
  // words, where its words are, ring state
  bool acth[2], pend[2];
  uint32_t kph[2], nh[2], wordsh[2], lo[2];
  uint64_t blkh[2];
  const uint4* ssrc4[2];
  uint4 rv[2][QH];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint64_t u = u0 + h;
    acth[h] = u < U;
    kph[h] = acth[h] ? (uint32_t)(u % P) : 0u;
    blkh[h] = acth[h] ? u / P : 0u;
    const uint8_t* sec = B.in + B.section[kph[h]];
    uint32_t nn = 0, words = 0, y = 0;
    if (acth[h]) {
      const uint2 bw = *(const uint2*)(sec + kSectionFixed + 4ull * L * B.nb + 8ull * blkh[h]);
      nn = bw.x >> 16;
      words = bw.x & 0xffffu;
      y = bw.y;
      const uint64_t start = blkh[h] * BLOCK;
      const uint64_t expect = (B.n - start) < BLOCK ? (B.n - start) : BLOCK;
      if (nn != expect || words > UNIT_CAP || (uint64_t)y + round_up(words, 8) > data_words ||
          B.data_off + 2ull * ((uint64_t)y + round_up(words, 8)) > B.in_bytes) {
        if (lane == 0) {
          atomicOr(a.err, kErrBlockWords);
          atomicOr(a.bad + b, kErrBlockWords);
        }
        acth[h] = false;
        nn = 0;
        words = 0;
      }
    }
    nh[h] = wave_u32(nn);
    wordsh[h] = wave_u32(words);
    y = wave_u32(y);
    ssrc4[h] = (const uint4*)(B.in + B.data_off + 2ull * y);
    uint16_t* reg = wstage + h * kDecRing;
    lo[h] = 0;
    pend[h] = false;
#pragma unroll
    for (int q = 0; q < QH; ++q) rv[h][q] = make_uint4(0, 0, 0, 0);
    if (wordsh[h] > 0) {
      const uint32_t jt = (wordsh[h] - 1) / kH;
      const uint32_t j0 = jt > 0 ? jt - 1 : 0;
      lo[h] = j0 * kH;
      for (uint32_t j = j0; j <= jt; ++j) {
        const uint32_t left = wordsh[h] - j * kH;
        const uint32_t n4 = (uint32_t)div_up(left < kH ? left : kH, 8);
        for (uint32_t i = lane; i < n4; i += 64) {
          const uint4 q = ssrc4[h][j * H4 + i];
          ((uint4*)(reg + (j & 1) * kH))[i] = q;
          if ((j & 1) == 0 && i < 8) ((uint4*)(reg + kR))[i] = q;
        }
      }
    }
  }
  // this lane's unit
  const bool act = half ? acth[1] : acth[0];
  const uint32_t kp = half ? kph[1] : kph[0];
  const uint64_t blk = half ? blkh[1] : blkh[0];
  const uint32_t n = half ? nh[1] : nh[0];
  uint32_t x[1], ptr[1], taken[1];
  x[0] = act ? ((const uint32_t*)(B.in + B.section[kp] + kSectionFixed))[blk * L + lin] : kMinState;
  ptr[0] = half ? wordsh[1] : wordsh[0];
  taken[0] = 0;
  const uint32_t* lut[1] = {s_lut + half * W};
  const uint16_t* stage[1] = {wstage + half * kDecRing};
  // the units' LUTs, built by the buffer's LUT item
  const uint64_t key = ((uint64_t)b << 4) | (kph[0] << 2) | kph[1];
  if (key != tkey) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const uint64_t* g = (const uint64_t*)(a.luts + ((uint64_t)b * 4u + kph[c]) * W);
      uint64_t* d = (uint64_t*)(s_lut + c * W);
      for (uint32_t j = lane; j < W / 2; j += 64) d[j] = atomic_ld64(g + j);
    }
    tkey = key;
  }
  __asm__ volatile("" : "+v"(x[0]));
  drain();

  auto fetch = [&](int h) {
    const uint32_t j = lo[h] / kH - 1;
#pragma unroll
    for (int q = 0; q < QH; ++q) rv[h][q] = ssrc4[h][j * H4 + 64 * q + lane];
    pend[h] = true;
  };
  auto commit = [&](int h) {
    const uint32_t j = lo[h] / kH - 1;
    uint16_t* reg = wstage + h * kDecRing;
#pragma unroll
    for (int q = 0; q < QH; ++q) ((uint4*)(reg + (j & 1) * kH))[64 * q + lane] = rv[h][q];
    if ((j & 1) == 0 && lane < 8) ((uint4*)(reg + kR))[lane] = rv[h][0];
    lo[h] -= kH;
    pend[h] = false;
  };
  auto refill = [&]() {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t p = __builtin_amdgcn_readlane(ptr[0], (uint32_t)h * 32u);
      if constexpr (KB == 1) {
        if (!pend[h] && lo[h] > 0 && p <= lo[h] + kH) fetch(h);
        if (pend[h] && p < lo[h] + 64u) commit(h);
      } else {
        if (!pend[h] && lo[h] > 0) fetch(h);
        if (pend[h] && p <= lo[h] + kH) {
          commit(h);
          if (lo[h] > 0) fetch(h);
        }
      }
    }
  };

  const bool act1 = acth[1];
  const uint32_t n_min = act1 ? (nh[0] < nh[1] ? nh[0] : nh[1]) : 0u;
  const uint32_t n_max = nh[0] > nh[1] ? nh[0] : nh[1];
  const uint32_t full = n_min / L;
  const uint32_t all = (uint32_t)div_up(n_max, L);
  uint8_t* o = B.out + blk * BLOCK * P + kp;
  auto store = [&](uint32_t s, uint32_t el, bool v) {
    const uint64_t i = (uint64_t)s * L + lin;
    if (v) *(__attribute__((address_space(1))) uint8_t*)(o + i * P) = (uint8_t)el;
  };
  const bool vall[1] = {true};
  if constexpr (KB == 1) {
    for (uint32_t s = all; s > full;) {
      --s;
      const bool v[1] = {s * L + lin < n};
      store(s, dec_steps<L, PB, 1, kMask, uint32_t>(x, ptr, taken, lut, stage, v, half), v[0]);
      refill();
    }
    for (uint32_t s = full; s > 0;) {
      --s;
      store(s, dec_steps<L, PB, 1, kMask, uint32_t>(x, ptr, taken, lut, stage, vall, half), true);
      refill();
    }
  } else {
    uint32_t s = all;
    while (s > full || s % KB != 0) {
      --s;
      const bool v[1] = {s * L + lin < n};
      store(s, dec_steps<L, PB, 1, kMask, uint32_t>(x, ptr, taken, lut, stage, v, half), v[0]);
      if (s % KB == 0) refill();
    }
    while (s > 0) {
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        --s;
        store(s, dec_steps<L, PB, 1, kMask, uint32_t>(x, ptr, taken, lut, stage, vall, half), true);
      }
      refill();
    }
  }
  const bool bad = act && (taken[0] != (half ? wordsh[1] : wordsh[0]) || x[0] != kMinState);
  if (__ballot(bad) && lane == 0) {
    atomicOr(a.err, kErrFinalState);
    atomicOr(a.bad + b, kErrFinalState);
  }
}

template <int PB, int NW, int KB, int L>
__global__ __launch_bounds__(64 * NW) void batch_decode_kernel(BatchDecodeArgs a) {
  constexpr uint32_t W = 1u << PB;
  constexpr uint32_t WAVE_WORDS = 2 * W + DecRing<KB>::WORDS;  // two LUTs | two rings (u16)
  extern __shared__ uint32_t lds[];
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = wave_u32(threadIdx.x >> 6);
  uint32_t* wl = lds + wave * WAVE_WORDS;
  const uint32_t gen = wave_u32(lane == 0 ? atomic_ld32(a.queue + 2) : 0u);
  uint64_t tkey = ~0ull;  // LUTs held in this wave's LDS
  uint32_t it = 0;
  const uint32_t run = a.run;
  if (lane == 0) it = atomicAdd(a.queue, run);
  it = wave_u32(it);
  uint32_t run_end = it + run;
  while (it < a.nitems) {
    const uint32_t b = wave_u32(a.items[it].x);
    const BatchBuf& B = a.bufs[b];
    if (it < a.nbufs) {
      build_luts<PB>(B, b, a, gen, wl);  // uses the LUT area as scratch
      tkey = ~0ull;
    } else {
      uint32_t r = 0;
      while (((r = wave_u32(lane == 0 ? atomic_ld32(a.ready + b) : 0u)) & ~kBatchBad) != gen)
        __builtin_amdgcn_s_sleep(2);
      if (!(r & kBatchBad)) {
        const uint32_t lt = it - a.nbufs - B.tick_first;
        if constexpr (L == 32) decode_ticket32<PB, KB>(B, b, lt, a, wl, tkey);
        else if (B.eb == 1) decode_ticket<PB, false, KB>(B, b, lt, a, wl, tkey);
        else decode_ticket<PB, true, KB>(B, b, lt, a, wl, tkey);
      }
    }
    drain();
    if (++it == run_end) {
      uint32_t nxt = 0;
      if (lane == 0) nxt = atomicAdd(a.queue, run);
      it = wave_u32(nxt);
      run_end = it + run;
    }
  }
  if (lane == 0) {
    const uint32_t total = gridDim.x * NW;
    if (atomicAdd(a.queue + 1, 1u) + 1 == total) {
      atomic_st32(a.queue, 0u);
      atomic_st32(a.queue + 1, 0u);
      atomic_st32(a.queue + 2, gen % kBatchGenMask + 1u);
    }
  }
}

template <int PB, int NW, int KB>
hipError_t launch_dec_shape(const BatchDecodeArgs& a, int cu, hipStream_t s) {
  constexpr size_t lds = (size_t)NW * (2 * (1u << PB) + DecRing<KB>::WORDS) * 4u;
  static_assert(lds == batch_decode_lds(BatchShape{NW, KB}, PB), "decode LDS layout");
  if constexpr (lds > kLdsBytes) {
    return hipErrorInvalidConfiguration;
  } else {
    int per_cu = 0;
    hipError_t e = a.lanes == 32
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_decode_kernel<PB, NW, KB, 32>, 64 * NW, lds)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_decode_kernel<PB, NW, KB, 64>, 64 * NW, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    uint64_t grid = (uint64_t)cu * (uint64_t)per_cu;
    const uint64_t want = div_up(a.nitems, (uint64_t)NW * a.run);
    if (want < grid) grid = want;
    if (grid < 1) grid = 1;
    if (a.lanes == 32) hipLaunchKernelGGL((batch_decode_kernel<PB, NW, KB, 32>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, a);
    else hipLaunchKernelGGL((batch_decode_kernel<PB, NW, KB, 64>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, a);
    return hipGetLastError();
  }
}

template <int PB, int NW>
hipError_t launch_dec_nw(const BatchDecodeArgs& a, int kb, int cu, hipStream_t s) {
  switch (kb) {
    case 1: return launch_dec_shape<PB, NW, 1>(a, cu, s);
    case 4: return launch_dec_shape<PB, NW, 4>(a, cu, s);
    case 8: return launch_dec_shape<PB, NW, 8>(a, cu, s);
    case 16: return launch_dec_shape<PB, NW, 16>(a, cu, s);
  }
  return hipErrorInvalidValue;
}

template <int PB>
hipError_t launch_dec(const BatchDecodeArgs& a, BatchShape sh, int cu, hipStream_t s) {
  switch (sh.nw) {
    case 4: return launch_dec_nw<PB, 4>(a, sh.kb, cu, s);
    case 8: return launch_dec_nw<PB, 8>(a, sh.kb, cu, s);
    case 16: return launch_dec_nw<PB, 16>(a, sh.kb, cu, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace
}  // namespace mans
