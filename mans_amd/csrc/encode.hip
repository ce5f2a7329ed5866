// encode.hip -- interleaved rANS encoder on gfx950 wavefronts.
//
// Reference: include/ans/GpuANSEncode.h:24-164 (encodeOne / ansEncodeBatch),
// :166-178 + BatchPrefixSum.h:29-187 (prefix of padded word counts) and
// :180-284 (ansEncodeCoalesceBatch).  The reference encodes each 4 KiB block
// into an uncoalesced staging buffer in HBM, scans the word counts with up to
// three launches (limit 512^2 blocks) and copies every block again into the
// final container.  Here one persistent launch does all of it:
//  * a wavefront takes a ticket (atomic counter) for a group of units: the
//    CH planes of one element block (two blocks, one per half-wave, in the
//    32-lane reference format), taking the next one when it is done.  Each
//    lane runs CH independent rANS chains, one per byte plane, fed
//    by a single element load per step, so the serial per-chain dependency is
//    interleaved CH ways;
//  * element loads are software-pipelined: the next 16 steps are in flight
//    while the current 16 are coded, and the table entries of a step are read
//    one step ahead;
//  * emitted 16-bit words are compacted in lane order within each step
//    (64-bit ballot + mbcnt) into LDS.  In the 64-lane format a chain keeps
//    a linear stage of H + 256 words: at most every 4 steps, once H words are
//    waiting, they go to one of the wave's two spill slots in HBM and the
//    rest moves to the front, so a wave needs a few KiB of LDS and many waves
//    stay resident.  The 32-lane format stages units whole (bounded by
//    lane_max_words);
//  * the wave publishes its padded word count and resolves its output offset
//    by decoupled look-back over earlier tickets, 512 of them per round trip
//    (tickets are taken by running waves, so every wave waited on is resident
//    and finishes).  In the 64-lane format the look-back of a ticket runs
//    only after the wave has coded its next ticket, so it rarely waits;
//  * it then writes states, block words and the unit's words straight to
//    their final place, 16 bytes per lane per store.
// The arithmetic per symbol is the reference's: emit x & 0xffff when
// x >= pdf * 2^(31-p); q = (umulhi(x, magic) + x) >> shift = x / pdf;
// x = q * 2^p + (x - q * pdf) + cdf, computed as x + q * (2^p - pdf) + cdf.
#include "codec_device.hpp"

namespace mans {
namespace {

constexpr int kBatch = 16;  // steps per prefetch batch

template <typename T>
using gptr = const __attribute__((address_space(1))) T*;
// 16 bytes as a native vector: arrays of it stay in registers
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct LanePlane {  // per-plane fields copied to LDS (dynamic index by lane)
  const uint8_t* src;
  uint64_t nsym;
  uint64_t section;
  uint64_t nb;
};

// The 32-lane reference format: L lanes per unit, PB probability bits, S
// element bytes, CH chains per lane, NW wavefronts per group; units are
// staged whole in LDS.  CH == S: the CH chains of a lane code the CH byte
// planes of the same elements (unit order 0 with P == S).  CH == 1, S == 1:
// one chain per lane on any unit (single streams, concatenated ADM planes).
template <int L, int PB, int S, int CH, int NW>
__global__ __launch_bounds__(64 * NW) void encode_kernel(EncodeArgs a) {
  constexpr uint32_t BLOCK = L == 32 ? 4096u : 8192u;
  constexpr uint32_t HALVES = L == 32 ? 2u : 1u;
  constexpr uint32_t UPW = CH * HALVES;  // units per ticket
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t REG = CSTAGE;  // LDS words per chain
  static_assert(CH == S, "chain c codes byte c of each element");
  extern __shared__ uint4 smem[];
  __shared__ LanePlane s_pl[kMaxPlanes];
  __shared__ uint64_t s_ubase[kMaxPlanes + 1];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int P = a.d.num_planes;

  uint4* s_tab = smem;  // [P][256] {thresh, magic, shift | cdf << 8, 2^PB - pdf}
  for (int i = t; i < P * (int)kNumSymbols; i += 64 * NW) {
    uint4 e = a.table[i];
    s_tab[i] = make_uint4(e.x << (31 - PB), e.z, e.w | (e.y << 8), (1u << PB) - e.x);
  }
  if (t < P) {
    s_pl[t].src = a.d.plane[t].src;
    s_pl[t].nsym = a.d.plane[t].nsym;
    s_pl[t].section = a.d.plane[t].section;
    s_pl[t].nb = a.d.plane[t].nb;
    s_ubase[t] = a.d.plane[t].unit_base;
  }
  if (t == 0) s_ubase[P] = a.d.total_units;
  __syncthreads();

  const uint32_t half = L == 32 ? (uint32_t)(lane >> 5) : 0u;
  const uint32_t lin = (uint32_t)lane & (L - 1);
  // staging: per wave, CH regions of REG words ([half] whole units), then
  // one dword-spaced dummy slot per lane
  const uint32_t wave_u = __builtin_amdgcn_readfirstlane((uint32_t)wave);
  uint16_t* wstage = (uint16_t*)(smem + P * kNumSymbols) + wave_u * (CH * REG + 128);
  const uint32_t doff = CH * REG + 2u * (uint32_t)lane;  // dummy slot, from wstage
  // words coded per plane by this wave (added to the section headers once,
  // at the end)
  uint32_t ptot[kMaxPlanes];
#pragma unroll
  for (int k = 0; k < kMaxPlanes; ++k) ptot[k] = 0;

  uint32_t tk = 0;
  if (lane == 0) tk = atomicAdd(a.ticket, UPW);
  tk = __builtin_amdgcn_readfirstlane(tk);
  while ((uint64_t)tk < a.d.total_units) {

    uint32_t x[CH], cnt[CH], n[CH];
    uint32_t soff[CH];  // chain region, from wstage
    const uint4* tab[CH];
    uint64_t blk[CH];
    int kp[CH];
    bool act[CH];
    const uint8_t* src[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint64_t unit = (uint64_t)tk + half * CH + c;
      act[c] = unit < a.d.total_units;
      int k = 0;
      uint64_t b = 0;
      if (act[c]) {
        if (a.d.order == 0) {  // planes of one element array: P is 1, 2 or 4
          k = (int)((uint32_t)unit & (uint32_t)(P - 1));
          b = unit >> (P == 4 ? 2 : (P == 2 ? 1 : 0));
        } else {
          while (k + 1 < P && unit >= s_ubase[k + 1]) ++k;
          b = unit - s_ubase[k];
        }
      }
      kp[c] = k;
      blk[c] = b;
      const LanePlane pl = s_pl[k];
      const uint64_t start = b * BLOCK;
      n[c] = act[c] ? (uint32_t)((pl.nsym - start) < BLOCK ? (pl.nsym - start) : BLOCK) : 0u;
      src[c] = pl.src;
      x[c] = kMinState;
      cnt[c] = 0;
      soff[c] = c * REG + half * (CSTAGE / HALVES);
      tab[c] = s_tab + k * kNumSymbols;
    }
    // wave-uniform step counts over all chains and halves
    uint32_t n_min = n[0], n_max = n[0];
#pragma unroll
    for (int c = 1; c < CH; ++c) {
      n_min = n[c] < n_min ? n[c] : n_min;
      n_max = n[c] > n_max ? n[c] : n_max;
    }
    if (L == 32) {
      const uint32_t a0 = __builtin_amdgcn_readlane(n_min, 0), a1 = __builtin_amdgcn_readlane(n_min, 32);
      const uint32_t b0 = __builtin_amdgcn_readlane(n_max, 0), b1 = __builtin_amdgcn_readlane(n_max, 32);
      n_min = a0 < a1 ? a0 : a1;
      n_max = b0 > b1 ? b0 : b1;
    } else {
      n_min = __builtin_amdgcn_readfirstlane(n_min);
      n_max = __builtin_amdgcn_readfirstlane(n_max);
    }
    const uint32_t all = (uint32_t)div_up(n_max, L);

    auto step = [&](int c, const uint4& e, bool v) {
      enc_step<L, 0xffffffffu>(x[c], cnt[c], wstage, soff[c], doff, e, v, half);
    };

    {
      // One element per lane per step feeds all CH chains (byte c -> chain
      // c).  Chain 0 of a ticket is plane 0 whenever CH > 1, so its plane
      // pointer is the element array; with CH == 1 it is the chain's own
      // byte stream (single streams, concatenated planes).
      const uint8_t* eb = src[0] + blk[0] * BLOCK * S;  // element block start (4-byte aligned)
      uint32_t s = 0;
      // Full blocks: each lane loads the aligned dword holding its element
      // and takes its bytes at a per-lane bit offset.  Two register buffers:
      // while one batch is coded the next is loading.  (Whole dwords keep the
      // compiler from narrowing the buffers and waiting on the next batch.)
      const gptr<uint32_t> dp = (gptr<uint32_t>)eb + ((lin * S) >> 2);
      const uint32_t sh = 8u * ((lin * S) & 3u);
      auto load = [&](uint32_t (&buf)[kBatch], uint32_t s0) {
        const gptr<uint32_t> bp = dp + (uint64_t)s0 * (L * S / 4);
#pragma unroll
        for (int j = 0; j < kBatch; ++j) buf[j] = bp[j * (L * S / 4)];
      };
      // The table entries of step j + 1 are read before step j is coded:
      // they depend on the symbols only, not on the states.
      auto code = [&](const uint32_t (&buf)[kBatch]) {
        uint4 e[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) e[c] = tab[c][__builtin_amdgcn_ubfe(buf[0], sh + 8u * c, 8u)];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
          uint4 en[CH];
          if (j + 1 < kBatch) {
#pragma unroll
            for (int c = 0; c < CH; ++c) en[c] = tab[c][__builtin_amdgcn_ubfe(buf[j + 1], sh + 8u * c, 8u)];
          }
#pragma unroll
          for (int c = 0; c < CH; ++c) step(c, e[c], true);
          if (j + 1 < kBatch) {
#pragma unroll
            for (int c = 0; c < CH; ++c) e[c] = en[c];
          }
        }
      };
      // The prefetch is unconditional (the last batch reloads itself) so the
      // wait before each batch only covers that batch's loads.  Partial blocks
      // (one per plane) take the byte path below, which never reads past the
      // end of the input.
      const uint32_t nbatch = n_min == BLOCK ? BLOCK / L / kBatch : 0u;
      if (nbatch > 0) {
        const uint32_t last = (nbatch - 1) * kBatch;
        uint32_t bufA[kBatch], bufB[kBatch];
        load(bufA, 0);
        for (;;) {
          load(bufB, s + kBatch <= last ? s + kBatch : last);
          code(bufA);
          s += kBatch;
          if (s > last) break;
          load(bufA, s + kBatch <= last ? s + kBatch : last);
          code(bufB);
          s += kBatch;
          if (s > last) break;
        }
      }
      for (; s < all; ++s) {
        const uint32_t i = s * L + lin;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const bool v = i < n[c];
          const uint32_t sym = v ? (uint32_t)eb[(uint64_t)i * S + c] : 0u;
          step(c, tab[c][sym], v);
        }
      }
    }

    // ---- per chain: pad, states; wave aggregate over the ticket's units ---
    uint32_t padded[CH];
    uint64_t agg = 0;
    uint32_t a_lo[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      padded[c] = (uint32_t)round_up(cnt[c], 8);
      if (cnt[c] > CSTAGE / HALVES && lin == 0) atomicOr(a.err, kErrStageOverflow);
      if (lin < padded[c] - cnt[c]) wstage[soff[c] + cnt[c] + lin] = 0;
      if (act[c]) ((uint32_t*)(a.out + s_pl[kp[c]].section + kSectionFixed))[blk[c] * L + lin] = x[c];
      a_lo[c] = __builtin_amdgcn_readlane(padded[c], 0);
      agg += a_lo[c];
      if (L == 32) agg += __builtin_amdgcn_readlane(padded[c], 32);
    }

    const uint64_t excl = look_back(a.status, tk / UPW, agg, lane);

    // unit order inside the ticket: (half 0: chains 0..CH-1), (half 1: ...)
    uint64_t y = excl;
    if (half) {
#pragma unroll
      for (int c = 0; c < CH; ++c) y += a_lo[c];
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      if (act[c] && lin == 0) {
        uint8_t* sec = a.out + s_pl[kp[c]].section;
        uint32_t* bw = (uint32_t*)(sec + kSectionFixed + 4ull * L * s_pl[kp[c]].nb + 8ull * blk[c]);
        bw[0] = (n[c] << 16) | cnt[c];
        bw[1] = (uint32_t)y;
      }
      {
        // this chain's unit per half-wave (two units in the 32-lane format,
        // possibly of different planes when planes are concatenated)
        const uint32_t pa = act[c] ? padded[c] : 0u;
        const uint32_t k0 = __builtin_amdgcn_readlane((uint32_t)kp[c], 0), a0 = __builtin_amdgcn_readlane(pa, 0);
#pragma unroll
        for (int k = 0; k < kMaxPlanes; ++k) ptot[k] += k0 == (uint32_t)k ? a0 : 0u;
        if (L == 32) {
          const uint32_t k1 = __builtin_amdgcn_readlane((uint32_t)kp[c], 32), a1 = __builtin_amdgcn_readlane(pa, 32);
#pragma unroll
          for (int k = 0; k < kMaxPlanes; ++k) ptot[k] += k1 == (uint32_t)k ? a1 : 0u;
        }
      }
      uint4* dst = (uint4*)(a.out + a.d.data_off + 2ull * y);
      const uint4* st4 = (const uint4*)(wstage + soff[c]);
      for (uint32_t j = lin; j < padded[c] / 8; j += L) dst[j] = st4[j];
      y += padded[c];
    }

    if ((uint64_t)tk + UPW >= a.d.total_units && lane == 0) {
      const uint64_t total_words = excl + agg;
      const uint64_t size = a.d.data_off + 2ull * total_words;
      if (a.d_out_size32) *a.d_out_size32 = (uint32_t)size;
      if (a.d_out_size64) *a.d_out_size64 = size;
      if (a.frame) {
        ((uint64_t*)a.out)[6] = total_words;  // data_words at byte 48
        ((uint64_t*)a.out)[7] = size;         // total_size at byte 56
      }
    }
    __threadfence_block();
    // the next ticket only once this one is done: tickets start in ticket
    // order, so the tickets a look-back waits on are already being coded
    if (lane == 0) tk = atomicAdd(a.ticket, UPW);
    tk = __builtin_amdgcn_readfirstlane(tk);
  }
  // total coded words of each plane (the section headers start at 0)
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kMaxPlanes; ++k)
      if (k < P && ptot[k]) atomicAdd((uint32_t*)(a.out + s_pl[k].section + 12), ptot[k]);
  }
}

// The multi-chain coder, for both formats.  A ticket is BPT consecutive
// element blocks per unit half (CH = S * BPT consecutive units: the S planes
// of each block), one chain per unit per lane, so a lane runs several
// interleaved chains whatever the element width.  In the 64-lane format a
// wavefront codes CH units; in the 32-lane format it codes 2 CH, lanes 0-31
// the first CH and lanes 32-63 the next CH (the units of one chain are
// "streams" h * CH + c, h the half).  A stream's words go to a linear LDS
// stage of H + 256 words; once H are waiting (checked at most 4 steps apart)
// they go to the wave's spill slot in HBM and the rest moves to the front.
// When the ticket is coded, the stage's remainder follows them, so the slot
// holds all of the ticket's words, padded, and the wave publishes its
// aggregate.  The offset is resolved only after the NEXT ticket has been
// coded (two slots per wave, used in turn): by then the predecessors have
// had a whole ticket's time to publish their own aggregates, so the
// look-back sums published entries instead of waiting for the slowest wave
// still coding.  The copy to the final place issues a batch of loads before
// any store.  (32 lanes: unit order 0 or single streams only, so both halves
// of a chain share its plane.)
template <int L, int PB, int S, int BPT, int NW, int HB>
__global__ __launch_bounds__(64 * NW, L == 32 ? 3 : 4) void encode_mc_kernel(EncodeArgs a) {
  constexpr int CH = S * BPT;                     // chains per lane
  constexpr int HALVES = L == 32 ? 2 : 1;
  constexpr int NS = CH * HALVES;                 // streams (units) per ticket
  constexpr int KB = 8 / BPT < 4 ? 4 : 8 / BPT;  // steps per prefetch batch (elements loaded ahead per block)
  constexpr uint32_t BLOCK = L == 32 ? 4096u : 8192u;
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t SSTAGE = CSTAGE / HALVES;    // a unit's words at most, rounded
  constexpr uint32_t H = 1u << HB;  // spill chunk words
  constexpr uint32_t H4 = H / 8;    // spill chunk uint4s
  constexpr uint32_t REG = H + 256u;  // LDS stage words per stream
  constexpr uint32_t SLOT = CH * CSTAGE;  // spill slot words
  constexpr int CB = 8;  // copy batch: uint4 loads per lane in flight
  constexpr int TP = L == 32 ? S : (S > 1 ? S : kMaxPlanes);  // tables held
  static_assert(BPT == 1 || (S < 4 && BLOCK / L % KB == 0), "blocks per ticket");
  static_assert(H >= 256 && H4 <= 64, "256 <= spill chunk <= 512 words");
  static_assert(CSTAGE % 16 == 0 && SSTAGE % 8 == 0 && REG % 8 == 0, "uint4-aligned areas");
  extern __shared__ uint4 smem[];
  __shared__ LanePlane s_pl[kMaxPlanes];
  __shared__ uint64_t s_ubase[kMaxPlanes + 1];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int P = a.d.num_planes;
  const uint32_t half = L == 32 ? (uint32_t)(lane >> 5) : 0u;
  const uint32_t lin = (uint32_t)lane & (L - 1);

  // [P][256] entries of enc64_entry, at a fixed LDS address so each chain's
  // base is an immediate offset of its reads
  __shared__ uint4 s_tab[TP * kNumSymbols];
  for (int i = t; i < P * (int)kNumSymbols; i += 64 * NW) {
    const uint4 e = a.table[i];
    s_tab[i] = enc64_entry(e.x, e.y, PB);
  }
  if (t < P) {
    s_pl[t].src = a.d.plane[t].src;
    s_pl[t].nsym = a.d.plane[t].nsym;
    s_pl[t].section = a.d.plane[t].section;
    s_pl[t].nb = a.d.plane[t].nb;
    s_ubase[t] = a.d.plane[t].unit_base;
  }
  if (t == 0) s_ubase[P] = a.d.total_units;
  __syncthreads();

  // staging: per wave, NS stages of REG words, then one dword-spaced dummy
  // slot per lane
  const uint32_t wave_u = __builtin_amdgcn_readfirstlane((uint32_t)wave);
  uint16_t* wstage = (uint16_t*)smem + wave_u * (NS * REG + 128);
  const uint32_t dummy_addr = lds_addr(wstage + NS * REG + 2u * (uint32_t)lane);
  // this wave's two spill slots of NS streams of SSTAGE words
  uint16_t* const spill = a.spill + ((uint64_t)blockIdx.x * NW + wave_u) * (2 * SLOT);
  uint32_t ptot[kMaxPlanes];  // words coded per plane by this wave
#pragma unroll
  for (int k = 0; k < kMaxPlanes; ++k) ptot[k] = 0;

  // the ticket coded last, whose offset is not resolved yet (wave-uniform)
  bool pend = false;
  uint32_t p_tk = 0, p_slot = 0;
  uint64_t p_agg = 0;
  uint32_t p_pad[NS];
  uint64_t p_bw[NS];  // byte offset of the unit's block-word pair, 0: inactive
#pragma unroll
  for (int c = 0; c < NS; ++c) p_pad[c] = 0, p_bw[c] = 0;

  // resolve the pending ticket's offset, write its units' word offsets and
  // copy its words from its slot to their final place
  auto finish_pending = [&]() {
    uint32_t qb[NS + 1];  // first uint4 of each stream in the ticket's output
    qb[0] = 0;
#pragma unroll
    for (int c = 0; c < NS; ++c) qb[c + 1] = qb[c] + p_pad[c] / 8;
    const uint32_t nq = qb[NS];
    // the slot was written by other lanes of this wave: their stores first
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    const u32x4* s4 = (const u32x4*)(spill + p_slot * SLOT);
    auto load = [&](u32x4 (&q)[CB], uint32_t j0) {
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        if (j0 + u * 64 < nq) {  // wave-uniform
          // lanes past the end load entry 0 (always in the slot)
          const uint32_t j = j0 + u * 64 + (uint32_t)lane;
          const uint32_t jj = j < nq ? j : 0u;
          // stream k's words sit at k * SSTAGE / 8 in the slot, at qb[k] here
          uint32_t d = 0;
#pragma unroll
          for (int k = 1; k < NS; ++k) d = jj >= qb[k] ? k * (SSTAGE / 8) - qb[k] : d;
          q[u] = s4[jj + d];
        }
      }
    };
    auto store = [&](const u32x4 (&q)[CB], u32x4* dst, uint32_t j0) {
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const uint32_t j = j0 + u * 64 + (uint32_t)lane;
        if (j < nq) dst[j] = q[u];
      }
    };
    // the first batch is in flight during the look-back
    u32x4 q[CB];
    load(q, 0);
    const uint64_t excl = resolve_prefix<4>(a.status, p_tk / NS, p_agg, lane);
    uint64_t y = excl;
#pragma unroll
    for (int c = 0; c < NS; ++c) {
      if (p_bw[c] && lane == 0) ((uint32_t*)(a.out + p_bw[c]))[1] = (uint32_t)y;
      y += p_pad[c];
    }
    u32x4* dst = (u32x4*)(a.out + a.d.data_off + 2ull * excl);
    store(q, dst, 0);
    for (uint32_t j0 = CB * 64; j0 < nq; j0 += CB * 64) {
      load(q, j0);
      store(q, dst, j0);
    }
    if ((uint64_t)p_tk + NS >= a.d.total_units && lane == 0) {
      const uint64_t total_words = excl + p_agg;
      const uint64_t size = a.d.data_off + 2ull * total_words;
      if (a.d_out_size32) *a.d_out_size32 = (uint32_t)size;
      if (a.d_out_size64) *a.d_out_size64 = size;
      if (a.frame) {
        ((uint64_t*)a.out)[6] = total_words;  // data_words at byte 48
        ((uint64_t*)a.out)[7] = size;         // total_size at byte 56
      }
    }
  };

  uint32_t slot = 0;
  uint32_t tk = 0;
  if (lane == 0) tk = atomicAdd(a.ticket, NS);
  tk = __builtin_amdgcn_readfirstlane(tk);
  while ((uint64_t)tk < a.d.total_units) {
    // per chain (this lane's half): unit, plane, block, symbols
    uint32_t x[CH], n[CH];
    const uint4* tab[CH];
    uint64_t blk[CH];
    int kp[CH];
    bool act[CH];
    const uint8_t* src[CH];
    // per stream: LDS byte address of the stage and of the next word, words
    // moved to the slot
    uint32_t cb[NS], sb[NS], fl[NS];
    uint16_t* const wslot = spill + slot * SLOT;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint64_t unit = (uint64_t)tk + half * CH + c;
      act[c] = unit < a.d.total_units;
      int k = 0;
      uint64_t b = 0;
      if (act[c]) {
        if (a.d.order == 0) {  // planes of one element array: P is 1, 2 or 4
          k = (int)((uint32_t)unit & (uint32_t)(P - 1));
          b = unit >> (P == 4 ? 2 : (P == 2 ? 1 : 0));
        } else {
          while (k + 1 < P && unit >= s_ubase[k + 1]) ++k;
          b = unit - s_ubase[k];
        }
      }
      kp[c] = k;
      blk[c] = b;
      const LanePlane pl = s_pl[k];
      const uint64_t start = b * BLOCK;
      n[c] = act[c] ? (uint32_t)((pl.nsym - start) < BLOCK ? (pl.nsym - start) : BLOCK) : 0u;
      if (L == 64) n[c] = __builtin_amdgcn_readfirstlane(n[c]);
      src[c] = pl.src;
      x[c] = kMinState;
      // S > 1 only on element arrays of S planes: chain c is plane c % S
      // (both halves: CH is a multiple of P)
      tab[c] = s_tab + (S > 1 ? c % S : (L == 32 ? 0 : k)) * kNumSymbols;
    }
#pragma unroll
    for (int c = 0; c < NS; ++c) {
      fl[c] = 0;
      cb[c] = lds_addr(wstage + c * REG);
      sb[c] = cb[c];
    }
    uint32_t n_min = n[0], n_max = n[0];
#pragma unroll
    for (int c = 1; c < CH; ++c) {
      n_min = n[c] < n_min ? n[c] : n_min;
      n_max = n[c] > n_max ? n[c] : n_max;
    }
    if (L == 32) {
      const uint32_t a0 = __builtin_amdgcn_readlane(n_min, 0), a1 = __builtin_amdgcn_readlane(n_min, 32);
      const uint32_t b0 = __builtin_amdgcn_readlane(n_max, 0), b1 = __builtin_amdgcn_readlane(n_max, 32);
      n_min = a0 < a1 ? a0 : a1;
      n_max = b0 > b1 ? b0 : b1;
    }
    const uint32_t all = (uint32_t)div_up(n_max, L);

    // stage upkeep: a stream with H or more unspilled words moves its first
    // H to the slot and the rest (fewer than 256: at most 64 words per step,
    // at most 4 steps apart) to the front
    auto spill_full = [&]() {
#pragma unroll
      for (int c = 0; c < NS; ++c) {
        const uint32_t un = (sb[c] - cb[c]) >> 1;  // unspilled words
        if (un >= H) {
          uint16_t* st = wstage + c * REG;
          if ((uint32_t)lane < H4) ((uint4*)(wslot + c * SSTAGE + fl[c]))[lane] = ((const uint4*)st)[lane];
          if (4u * (uint32_t)lane < un - H) ((uint2*)st)[lane] = ((const uint2*)(st + H))[lane];
          fl[c] += H;
          sb[c] -= 2u * H;
        }
      }
    };
    auto step_full = [&](int c, const uint4& e) {
      if constexpr (L == 64) enc_step64f(x[c], sb[c], dummy_addr, e);
      else enc_step32f(x[c], sb[c], sb[CH + c], half, e);
    };
    auto step_part = [&](int c, const uint4& e, bool v) {
      if constexpr (L == 64) enc_step64(x[c], sb[c], dummy_addr, e, v);
      else enc_step32(x[c], sb[c], sb[CH + c], half, dummy_addr, e, v);
    };

    {
      // Per step each lane takes one element of each of the ticket's BPT
      // blocks (of its half); byte b of block d's element feeds chain
      // d * S + b.  With S > 1 the first chain of a block is its plane 0,
      // whose pointer is the element array; with S == 1 each chain has its
      // own byte stream (single streams, concatenated planes).
      const uint8_t* eb[BPT];
#pragma unroll
      for (int d = 0; d < BPT; ++d) eb[d] = src[d * S] + blk[d * S] * BLOCK * S;
      uint32_t s = 0;
      // Full blocks: each lane loads its element of each block
      // (zero-extended: byte b of the register is byte b of the element).
      // Two register buffers: while one batch is coded the next is loading.
      using EL = typename ElemOf<S>::T;
      auto load = [&](uint32_t (&buf)[BPT][KB], uint32_t s0) {
#pragma unroll
        for (int d = 0; d < BPT; ++d) {
          const gptr<EL> bp = (gptr<EL>)eb[d] + lin + (uint64_t)s0 * L;
#pragma unroll
          for (int j = 0; j < KB; ++j) buf[d][j] = bp[j * L];
        }
      };
      // table entries of the chains for step j of a batch
      auto lookup = [&](uint4 (&e)[CH], const uint32_t (&buf)[BPT][KB], int j) {
#pragma unroll
        for (int d = 0; d < BPT; ++d) {
          const uint32_t w = buf[d][j];
          e[d * S] = *(const uint4*)((const char*)tab[d * S] + byte_x16<0>(w));
          if constexpr (S > 1) e[d * S + 1] = *(const uint4*)((const char*)tab[d * S + 1] + byte_x16<1>(w));
          if constexpr (S > 2) {
            e[d * S + 2] = *(const uint4*)((const char*)tab[d * S + 2] + byte_x16<2>(w));
            e[d * S + 3] = *(const uint4*)((const char*)tab[d * S + 3] + byte_x16<3>(w));
          }
        }
      };
      // The table entries of step j + 1 are read before step j is coded:
      // they depend on the symbols only, not on the states.  The stage is
      // checked every 4 steps: at most 4 * 64 words arrive in between.
      auto code = [&](const uint32_t (&buf)[BPT][KB]) {
        uint4 e[CH];
        lookup(e, buf, 0);
#pragma unroll
        for (int j = 0; j < KB; ++j) {
          uint4 en[CH];
          if (j + 1 < KB) lookup(en, buf, j + 1);
#pragma unroll
          for (int c = 0; c < CH; ++c) step_full(c, e[c]);
          if ((j & 3) == 3) spill_full();
          if (j + 1 < KB) {
#pragma unroll
            for (int c = 0; c < CH; ++c) e[c] = en[c];
          }
        }
      };
      static_assert(KB % 4 == 0, "stage checks at batch-relative steps 3, 7, ...");
      // The prefetch is unconditional (the last batch reloads itself) so the
      // wait before each batch only covers that batch's loads.  Partial blocks
      // (one per plane) take the byte path below, which never reads past the
      // end of the input.
      const uint32_t nbatch = n_min == BLOCK ? BLOCK / L / KB : 0u;
      if (nbatch > 0) {
        const uint32_t last = (nbatch - 1) * KB;
        uint32_t bufA[BPT][KB], bufB[BPT][KB];
        load(bufA, 0);
        for (;;) {
          load(bufB, s + KB <= last ? s + KB : last);
          code(bufA);
          s += KB;
          if (s > last) break;
          load(bufA, s + KB <= last ? s + KB : last);
          code(bufB);
          s += KB;
          if (s > last) break;
        }
      }
      for (; s < all; ++s) {
        const uint32_t i = s * L + lin;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          const bool v = i < n[c];
          const uint32_t sym = v ? (uint32_t)eb[c / S][(uint64_t)i * S + (c % S)] : 0u;
          step_part(c, tab[c][sym], v);
        }
        spill_full();
      }
    }

    // ---- per stream: pad, move the stage's remainder to the slot; per chain
    // and half: states and block words; ticket aggregate
    uint64_t agg = 0;
    uint32_t pad[NS];
    uint64_t bwo[NS];
#pragma unroll
    for (int c = 0; c < NS; ++c) {
      const uint32_t rem = (sb[c] - cb[c]) >> 1;  // words still in the stage
      const uint32_t cnt = fl[c] + rem;
      const uint32_t padded = (uint32_t)round_up(cnt, 8);
      if (cnt > SSTAGE && lane == 0) atomicOr(a.err, kErrStageOverflow);
      uint16_t* st = wstage + c * REG;
      if ((uint32_t)lane < padded - cnt) st[rem + lane] = 0;
      const uint32_t nr = (padded - fl[c]) / 8;
      for (uint32_t j = lane; j < nr; j += 64) ((uint4*)(wslot + c * SSTAGE + fl[c]))[j] = ((const uint4*)st)[j];
      pad[c] = padded;  // set to 0 below for inactive units
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const LanePlane& pl = s_pl[kp[c]];
      if (act[c]) ((uint32_t*)(a.out + pl.section + kSectionFixed))[blk[c] * L + lin] = x[c];
      // this lane's half: stream h * CH + c
      uint32_t cnt_h = 0;
#pragma unroll
      for (int h = 0; h < HALVES; ++h) {
        const int sidx = h * CH + c;
        const uint32_t cnt = fl[sidx] + ((sb[sidx] - cb[sidx]) >> 1);
        cnt_h = (uint32_t)h == half ? cnt : cnt_h;
      }
      const uint64_t bw = pl.section + kSectionFixed + 4ull * L * pl.nb + 8ull * blk[c];
      if (act[c] && lin == 0) ((uint32_t*)(a.out + bw))[0] = (n[c] << 16) | cnt_h;
#pragma unroll
      for (int h = 0; h < HALVES; ++h) {
        const int sidx = h * CH + c;
        const uint32_t src_lane = (uint32_t)h * 32u;
        const bool ah = L == 32 ? (bool)__builtin_amdgcn_readlane((uint32_t)act[c], src_lane) : act[c];
        const uint32_t kh = L == 32 ? __builtin_amdgcn_readlane((uint32_t)kp[c], src_lane) : (uint32_t)kp[c];
        const uint64_t bwh =
            L == 32 ? ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(bw >> 32), src_lane) << 32) |
                          __builtin_amdgcn_readlane((uint32_t)bw, src_lane)
                    : bw;
        pad[sidx] = ah ? pad[sidx] : 0u;
        bwo[sidx] = ah ? bwh : 0ull;
#pragma unroll
        for (int k = 0; k < kMaxPlanes; ++k) ptot[k] += kh == (uint32_t)k ? pad[sidx] : 0u;
      }
    }
#pragma unroll
    for (int c = 0; c < NS; ++c) agg += pad[c];
    publish_aggregate(a.status, tk / NS, agg, lane);
    if (pend) finish_pending();
    pend = true;
    p_tk = tk;
    p_slot = slot;
    p_agg = agg;
#pragma unroll
    for (int c = 0; c < NS; ++c) p_pad[c] = pad[c], p_bw[c] = bwo[c];
    slot ^= 1u;
    // the next ticket only once this one is published: tickets start in
    // ticket order, and every ticket a look-back waits on belongs to a wave
    // that is coding it
    __threadfence_block();
    if (lane == 0) tk = atomicAdd(a.ticket, NS);
    tk = __builtin_amdgcn_readfirstlane(tk);
  }
  if (pend) finish_pending();
  // total coded words of each plane (the section headers start at 0)
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kMaxPlanes; ++k)
      if (k < P && ptot[k]) atomicAdd((uint32_t*)(a.out + s_pl[k].section + 12), ptot[k]);
  }
}

// The whole-unit staged 32-lane kernel (encode_kernel): CH == S chains, no
// spill areas.  The multi-chain kernel (encode_mc_kernel): BPT blocks per
// half-wave ticket, spill areas.
template <int L, int PB, int S, int CH, int NW>
hipError_t launch_staged(const EncodeArgs& a, int cu, hipStream_t s, uint64_t* need_spill) {
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  const size_t lds = (size_t)a.d.num_planes * kNumSymbols * sizeof(uint4) + (size_t)NW * (CH * CSTAGE + 128) * 2u;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, encode_kernel<L, PB, S, CH, NW>, 64 * NW, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  const uint64_t waves = div_up(a.d.total_units, CH * 2u);
  uint64_t grid = div_up(waves, NW);
  const uint64_t cap = (uint64_t)cu * (uint64_t)per_cu;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  if (need_spill) {
    *need_spill = 0;
    return hipSuccess;
  }
  hipLaunchKernelGGL((encode_kernel<L, PB, S, CH, NW>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, a);
  return hipGetLastError();
}

template <int L, int PB, int S, int BPT, int NW, int HB>
hipError_t launch_mc(const EncodeArgs& a, int cu, hipStream_t s, uint64_t* need_spill) {
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr int CH = S * BPT, NS = CH * (L == 32 ? 2 : 1);
  constexpr uint32_t REG = (1u << HB) + 256u;
  // dynamic LDS: the staging areas (the tables are static)
  const size_t lds = (size_t)NW * (NS * REG + 128) * 2u;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, encode_mc_kernel<L, PB, S, BPT, NW, HB>,
                                                              64 * NW, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  const uint64_t waves = div_up(a.d.total_units, NS);
  uint64_t grid = div_up(waves, NW);
  const uint64_t cap = (uint64_t)cu * (uint64_t)per_cu;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  // two slots per wave of CH chains of 64 lanes
  const uint64_t spill_bytes = grid * NW * 2ull * CH * CSTAGE * 2ull;
  if (need_spill) {
    *need_spill = spill_bytes;
    return hipSuccess;
  }
  if (!a.spill || spill_bytes > a.spill_bytes) return hipErrorInvalidValue;
  hipLaunchKernelGGL((encode_mc_kernel<L, PB, S, BPT, NW, HB>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, a);
  return hipGetLastError();
}

// staged: the 32-lane whole-unit kernel (no spill areas: the allocation-free
// entry point, and ADM planes concatenated at 32 lanes).
template <int L, int PB>
hipError_t launch_s(const EncodeArgs& a, int cu, hipStream_t s, uint64_t* need_spill, bool staged) {
  const int P = a.d.num_planes;
  const uint32_t S = a.d.plane[0].stride;
  if constexpr (L == 64) {
    // four chains per lane, linear stages of 256 + 256 words: 4.3 KiB of
    // LDS per wave
    if (a.d.order == 0 && P == 2 && S == 2) return launch_mc<64, PB, 2, 2, 4, 8>(a, cu, s, need_spill);
    if (a.d.order == 0 && P == 4 && S == 4) return launch_mc<64, PB, 4, 1, 4, 8>(a, cu, s, need_spill);
    if (S == 1) return launch_mc<64, PB, 1, 4, 4, 8>(a, cu, s, need_spill);
  } else {
    if (!staged && a.d.order == 0) {
      // four chains per lane, two units per chain (one per half-wave)
      if (P == 2 && S == 2) return launch_mc<32, PB, 2, 2, 4, 8>(a, cu, s, need_spill);
      if (P == 4 && S == 4) return launch_mc<32, PB, 4, 1, 4, 8>(a, cu, s, need_spill);
      if (P == 1 && S == 1) return launch_mc<32, PB, 1, 4, 4, 8>(a, cu, s, need_spill);
    }
    if (a.d.order == 0 && P == 2 && S == 2) return launch_staged<L, PB, 2, 2, 2>(a, cu, s, need_spill);
    if (a.d.order == 0 && P == 4 && S == 4) return launch_staged<L, PB, 4, 4, 1>(a, cu, s, need_spill);
    if (S == 1) return launch_staged<L, PB, 1, 1, 4>(a, cu, s, need_spill);
  }
  return hipErrorInvalidValue;
}

template <int L>
hipError_t launch_p(const EncodeArgs& a, int cu, hipStream_t s, uint64_t* need_spill, bool staged) {
  switch (a.d.prob_bits) {
    case 9: return launch_s<L, 9>(a, cu, s, need_spill, staged);
    case 10: return launch_s<L, 10>(a, cu, s, need_spill, staged);
    case 11: return launch_s<L, 11>(a, cu, s, need_spill, staged);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t encode_spill_bytes(const EncodeArgs& a, int cu, uint64_t* bytes) {
  *bytes = 0;
  if (a.d.total_units == 0) return hipSuccess;
  return a.d.lanes == 64 ? launch_p<64>(a, cu, nullptr, bytes, false) : launch_p<32>(a, cu, nullptr, bytes, false);
}

// a.spill == nullptr selects the 32-lane whole-unit kernel (no spill areas).
hipError_t launch_encode(const EncodeArgs& a, int cu, hipStream_t s) {
  if (a.d.total_units == 0) return hipSuccess;
  if (a.d.lanes == 64) return launch_p<64>(a, cu, s, nullptr, false);
  return launch_p<32>(a, cu, s, nullptr, a.spill == nullptr);
}

}  // namespace mans
