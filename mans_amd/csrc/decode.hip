// decode.hip -- interleaved rANS decoder on gfx950 wavefronts.
//
// Reference: include/ans/GpuANSDecode.h:27-46 (packed LUT entry),
// :48-142 (decodeOneWarp / decodeOnePartialWarp), :144-302 (block loop,
// grid-stride over blocks), :304-377 (ansDecodeTable).  Differences in
// structure, not in results:
//  * every workgroup builds the 2^p-entry LUT of each plane in LDS straight
//    from the container's pdf table (no separate table launch);
//  * each lane runs CH chains, one per byte plane of the same elements (the
//    32-lane reference format puts two blocks on one wavefront, one per
//    half), so the serial per-chain dependency is interleaved CH ways and the
//    decoded bytes are assembled into whole elements: one coalesced store per
//    lane per step instead of one byte store per plane;
//  * the chain reads its words from LDS.  In the 64-lane format a chain keeps
//    only a ring of R words of its unit (plus a 64-word mirror of the ring's
//    head, so the words of one step never wrap): the next R/2 words are
//    loaded from HBM half a ring ahead of use and written to LDS just before
//    they are needed.  Small LDS per wave keeps many waves resident, which is
//    what hides the LDS and HBM latency of the serial chains.  The 32-lane
//    format stages each unit's words whole;
//  * every lane reads a word each step (lanes that take none discard it), so
//    there is no branch between the chains' steps;
//  * a block is valid only if it consumes exactly its words and every lane
//    ends in the start state 2^15; anything else raises an error flag.
// Per symbol: slot = x & (2^p - 1); e = LUT[slot] gives sym, pdf and
// slot - cdf; x = pdf * (x >> p) + slot - cdf; lanes with x < 2^15 read a word from the
// end of the block's remaining words, the highest such lane taking the last
// word; x = (x << 16) + word.
#include "adm_device.hpp"
#include "codec_device.hpp"

namespace mans {
namespace {

// LUT entry per format.  64 lanes: {pdf | sym << 24, slot - cdf} (8 bytes,
// the next state in one v_mad_u32_u24).  32 lanes: (slot - cdf) << 20 |
// pdf << 8 | sym (4 bytes; pdf <= 2^11 and slot - cdf < pdf fit 12 bits
// each): two more VALU operations per symbol, half the LDS bytes per read.
// Measured on 1 GiB: the 64-lane decode takes 0.46 ms with the 8-byte entry
// and 0.50 ms with the 4-byte one, the 32-lane decode 0.83 ms and 0.77 ms.
template <int L>
using LutEntryT = typename std::conditional<L == 32, uint32_t, uint2>::type;
__device__ __forceinline__ void lut_put(uint2& d, uint32_t sym, uint32_t pdf, uint32_t off) {
  d = make_uint2(pdf | (sym << 24), off);
}
__device__ __forceinline__ void lut_put(uint32_t& d, uint32_t sym, uint32_t pdf, uint32_t off) {
  d = (off << 20) | (pdf << 8) | sym;
}

struct DPlane {
  uint8_t* dst;
  uint64_t nsym;
  uint64_t section;
  uint64_t nb;
};

// L lanes per unit, PB probability bits, S element bytes, CH chains per lane
// (CH / S consecutive element blocks, 64-lane format only: chain d * S + b
// decodes byte b of block d's elements), NW wavefronts per group, RB: log2 of
// the ring in words (64-lane format), 0 = whole-unit staging.
// FUSE (64-lane ADM frames, one plane, one chain): each block's codes go to
// the wave's own BLOCK bytes of LDS instead of memory, and once the block is
// decoded the wave restores its elements (ADM lanes of 16 codes) from the
// codes, the already decoded bit plane, the offsets and the centers
// (adm_restore_lane), so the codes never leave the CU.
template <int L, int PB, int S, int CH, int NW, int RB, bool FUSE = false>
__global__ __launch_bounds__(64 * NW) void decode_kernel(DecodeArgs a) {
  using T = typename ElemOf<S>::T;
  constexpr int BPT = CH / S;  // element blocks per wave iteration
  constexpr uint32_t BLOCK = L == 32 ? 4096u : 8192u;
  constexpr uint32_t HALVES = L == 32 ? 2u : 1u;
  constexpr uint32_t UPW = CH * HALVES;  // units per wave iteration: one stream each
  constexpr int NSTR = CH * (int)HALVES;
  constexpr uint32_t W = 1u << PB;
  constexpr bool RING = RB > 0;
  constexpr uint32_t R = RING ? 1u << RB : 0u;  // ring words
  constexpr uint32_t H = RING ? R / 2 : 8u;     // refill chunk words
  constexpr uint32_t H4 = H / 8;                // refill chunk uint4s
  constexpr uint32_t MASK = RING ? R - 1u : 0xffffffffu;
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t HSTAGE = CSTAGE / HALVES;
  constexpr uint32_t REG = RING ? R + 64u : CSTAGE;  // LDS words per chain
  constexpr uint32_t RK = H >= 512 ? 4u : 2u;          // steps between ring checks
  static_assert(!RING || 2u * L * RK <= H, "a chunk's load starts a check before its commit");
  constexpr uint32_t UNIT_CAP = L * lane_max_words(BLOCK / L, PB);
  static_assert(CH % S == 0 && (L == 64 || RING || CH == S) && CH * 8 <= 32,
                "chain d * S + b decodes byte b of block d");
  static_assert(!RING || (H >= 128 && H4 <= 64), "ring: 128 <= R/2 <= 512 words");
  static_assert(!FUSE || (L == 64 && S == 1 && CH == 1 && RING), "fused ADM restore: 64-lane code plane");
  constexpr uint32_t WAVE_U16 = (RING ? NSTR : CH) * REG + (RING ? 0u : 64u);  // stage u16 per wave
  extern __shared__ uint4 smem[];
  __shared__ DPlane s_pl[kMaxPlanes];
  __shared__ uint64_t s_ubase[kMaxPlanes + 1];
  __shared__ uint32_t s_cdf[kNumSymbols + 1];
  __shared__ uint32_t s_tmp[4];
  __shared__ uint64_t s_total_units, s_data_off, s_data_words;
  __shared__ int s_bad;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int P = a.d.num_planes;
  using LutEntry = LutEntryT<L>;
  LutEntry* s_lut = (LutEntry*)smem;  // [P][W]

  if (t == 0) s_bad = 0;
  if (t < P) {
    s_pl[t].dst = a.d.plane[t].dst;
    s_pl[t].section = a.d.plane[t].section;
    s_pl[t].nsym = a.d.plane[t].nsym;
    s_pl[t].nb = a.d.plane[t].nb;
    s_ubase[t] = a.d.plane[t].unit_base;
  }
  __syncthreads();
  if (a.from_header && t == 0) {
    // single stream: block count and size come from the device header
    const uint32_t* h = (const uint32_t*)(a.in + s_pl[0].section);
    const uint32_t mv = h[0];
    const uint64_t nb = h[1], nsym = h[2];
    const Format f = format_for_lanes(L);
    if ((mv >> 16) != kAnsMagic || (mv & 0xffffu) != f.version) s_bad = kErrMagic;
    else if ((h[4] & 0xfu) != (uint32_t)PB) s_bad = kErrPrecision;
    else if (nb != div_up(nsym, BLOCK)) s_bad = kErrBlockWords;
    else if (a.out_capacity && nsym > a.out_capacity) s_bad = kErrCapacity;
    s_pl[0].nsym = nsym;
    s_pl[0].nb = nb;
    s_ubase[0] = 0;
    s_total_units = nb;
    s_data_off = s_pl[0].section + section_bytes(nb, f);
    s_data_words = h[3];
    if (a.in_bytes && s_data_off + 2ull * s_data_words > a.in_bytes) s_bad = kErrBlockWords;
  } else if (t == 0) {
    s_total_units = a.d.total_units;
    s_data_off = a.d.data_off;
    s_data_words = a.data_words;
  }
  if (t == 0) s_ubase[P] = s_total_units;
  __syncthreads();
  if (s_bad) {
    if (blockIdx.x == 0 && t == 0) atomicOr(a.err, (uint32_t)s_bad);
    return;
  }
  if (!a.from_header && t < P) {
    // frames: each plane section must carry the stream magic / version /
    // precision and the block count of the frame header
    const uint32_t* h = (const uint32_t*)(a.in + s_pl[t].section);
    const Format f = format_for_lanes(L);
    if ((h[0] >> 16) != kAnsMagic || (h[0] & 0xffffu) != f.version) atomicOr(&s_bad, (int)kErrMagic);
    else if ((h[4] & 0xfu) != (uint32_t)PB) atomicOr(&s_bad, (int)kErrPrecision);
    else if (h[1] != s_pl[t].nb || h[2] != s_pl[t].nsym) atomicOr(&s_bad, (int)kErrBlockWords);
  }
  __syncthreads();

  // ---- LUT per plane ------------------------------------------------------
  // exclusive scan of the 256 pdfs by the first 256 / PER threads (PER pdfs
  // each, at most 4 wavefronts), then one binary search per slot
  constexpr int PER = 64 * NW >= (int)kNumSymbols ? 1 : (int)kNumSymbols / (64 * NW);
  constexpr int SCAN_T = (int)kNumSymbols / PER;
  for (int k = 0; k < P; ++k) {
    const uint16_t* probs = (const uint16_t*)(a.in + s_pl[k].section + kAnsHeaderBytes);
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      v[j] = t < SCAN_T ? probs[t * PER + j] : 0u;
      sum += v[j];
    }
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63 && t < SCAN_T) s_tmp[wave] = inc;
    __syncthreads();
    if (t < SCAN_T) {
      uint32_t base = 0;
      for (int j = 0; j < wave; ++j) base += s_tmp[j];
      uint32_t run = base + inc - sum;
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        s_cdf[t * PER + j] = run;
        run += v[j];
      }
      if (t == SCAN_T - 1) s_cdf[kNumSymbols] = run;
    }
    __syncthreads();
    if (s_pl[k].nsym > 0 && s_cdf[kNumSymbols] != W) {
      if (t == 0) s_bad = kErrTable;
    } else {
      LutEntry* lut = s_lut + k * W;
      for (uint32_t slot = t; slot < W; slot += 64 * NW) {
        // largest symbol s with cdf[s] <= slot (that symbol has pdf > 0)
        uint32_t lo = 0, hi = kNumSymbols;
        while (hi - lo > 1) {
          uint32_t mid = (lo + hi) >> 1;
          if (s_cdf[mid] <= slot) lo = mid; else hi = mid;
        }
        const uint32_t c = s_cdf[lo], p = s_cdf[lo + 1] - c;
        lut_put(lut[slot], lo, p, slot - c);
      }
    }
    __syncthreads();
  }
  if (s_bad) {
    if (blockIdx.x == 0 && t == 0) atomicOr(a.err, (uint32_t)s_bad);
    return;
  }

  const uint32_t half = L == 32 ? (uint32_t)(lane >> 5) : 0u;
  const uint32_t lin = (uint32_t)lane & (L - 1);
  // per wave: CH regions of REG words (ring + mirror, or a whole unit), then
  // with whole-unit staging 64 words of slack (a lane that takes no word may
  // read one slot past its unit's words)
  const uint32_t wave_u = __builtin_amdgcn_readfirstlane((uint32_t)wave);
  uint16_t* const stage0 = (uint16_t*)(smem + div_up((uint64_t)P * W * sizeof(LutEntry), 16u));
  uint16_t* wstage = stage0 + wave_u * WAVE_U16;
  // FUSE: this wave's block of codes and 16 bytes of signal scratch per lane
  uint8_t* const wcodes = (uint8_t*)(stage0 + NW * WAVE_U16) + wave_u * (BLOCK + 64u * kAdmChunk);
  uint8_t* const wsig = wcodes + BLOCK + (uint32_t)lane * kAdmChunk;
  const uint64_t total_units = s_total_units;
  const uint64_t data_off = s_data_off, data_words = s_data_words;
  const uint64_t nwaves = (uint64_t)gridDim.x * NW;

  for (uint64_t wu = (uint64_t)blockIdx.x * NW + wave_u; wu * UPW < total_units; wu += nwaves) {
    uint32_t x[CH], ptr[CH], n[CH];
    uint32_t wds[CH], taken[CH];
    bool act[CH];
    uint64_t blk[CH];
    const LutEntry* lut[CH];
    const uint16_t* stage[CH];
    // ring state per stream (unit), wave-uniform: lowest word index held in
    // LDS, whether a refill is in flight, and its registers
    uint32_t lo[NSTR];
    bool pend[NSTR];
    uint4 rv[NSTR];
    const uint4* src4[CH];  // this lane's unit (staged units)
    const uint4* ssrc4[NSTR];  // each stream's words (rings)
    uint32_t swords[NSTR];
    uint8_t* dstp[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint64_t unit = wu * UPW + half * CH + c;
      act[c] = unit < total_units;
      int k = 0;
      uint64_t b = 0;
      if (act[c]) {
        if (a.d.order == 0) {
          k = (int)((uint32_t)unit & (uint32_t)(P - 1));
          b = unit >> (P == 4 ? 2 : (P == 2 ? 1 : 0));
        } else {
          while (k + 1 < P && unit >= s_ubase[k + 1]) ++k;
          b = unit - s_ubase[k];
        }
      }
      blk[c] = b;
      const DPlane pl = s_pl[k];
      dstp[c] = pl.dst;
      const uint8_t* sec = a.in + pl.section;
      const uint64_t start = b * BLOCK;
      uint32_t nn = 0, words = 0, y = 0;
      if (act[c]) {
        const uint2 bw = *(const uint2*)(sec + kSectionFixed + 4ull * L * pl.nb + 8ull * b);
        nn = bw.x >> 16;
        words = bw.x & 0xffffu;
        y = bw.y;
        const uint64_t expect = (pl.nsym - start) < BLOCK ? (pl.nsym - start) : BLOCK;
        if (nn != expect || words > UNIT_CAP || (!RING && words > HSTAGE) ||
            (uint64_t)y + round_up(words, 8) > data_words ||
            (a.in_bytes && data_off + 2ull * ((uint64_t)y + words) > a.in_bytes)) {
          if (lin == 0) atomicOr(a.err, kErrBlockWords);
          act[c] = false;
          nn = 0;
          words = 0;
        }
      }
      n[c] = nn;
      lut[c] = s_lut + k * W;
      // all lanes of a unit read the same block words: with L == 64 the
      // count is wave-uniform, and kept in a scalar register
      if (L == 64) {
        words = __builtin_amdgcn_readfirstlane(words);
        y = __builtin_amdgcn_readfirstlane(y);
      }
      src4[c] = (const uint4*)(a.in + data_off + 2ull * y);  // y is a multiple of 8 words
      // a ring per stream h * CH + c; whole units at half * HSTAGE of the chain's area
      uint16_t* reg = RING ? wstage + (half * CH + c) * REG : wstage + c * REG + half * HSTAGE;
      stage[c] = reg;
      if constexpr (RING) {
#pragma unroll
        for (int h = 0; h < (int)HALVES; ++h) {
          const int sidx = h * CH + c;
          swords[sidx] = L == 32 ? __builtin_amdgcn_readlane(words, (uint32_t)h * 32u) : words;
          const uint32_t yh = L == 32 ? __builtin_amdgcn_readlane(y, (uint32_t)h * 32u) : y;
          ssrc4[sidx] = (const uint4*)(a.in + data_off + 2ull * yh);
        }
      } else {
        // whole unit -> LDS (16-byte loads)
        uint4* d4 = (uint4*)reg;
        const uint32_t n4 = (uint32_t)div_up(words, 8);
        for (uint32_t j = lin; j < n4; j += L) d4[j] = src4[c][j];
      }
      x[c] = act[c] ? ((const uint32_t*)(sec + kSectionFixed))[b * L + lin] : kMinState;
      ptr[c] = words;
      wds[c] = words;
      taken[c] = 0;
    }
    if constexpr (RING) {
      // the top one or two chunks of each stream: chunk j holds words
      // [j H, (j + 1) H) at ring half j & 1; half 0's first 64 words are
      // mirrored after the ring (all 64 lanes load each stream's chunks)
#pragma unroll
      for (int sidx = 0; sidx < NSTR; ++sidx) {
        uint16_t* reg = wstage + sidx * REG;
        const uint32_t words = swords[sidx];
        lo[sidx] = 0;
        pend[sidx] = false;
        rv[sidx] = make_uint4(0, 0, 0, 0);
        if (words > 0) {
          const uint32_t jt = (words - 1) / H;
          const uint32_t j0 = jt > 0 ? jt - 1 : 0;
          lo[sidx] = j0 * H;
          for (uint32_t j = j0; j <= jt; ++j) {
            const uint32_t left = words - j * H;
            const uint32_t n4 = (uint32_t)div_up(left < H ? left : H, 8);
            if ((uint32_t)lane < n4) {
              const uint4 q = ssrc4[sidx][j * H4 + lane];
              ((uint4*)(reg + (j & 1) * H))[lane] = q;
              if ((j & 1) == 0 && lane < 8) ((uint4*)(reg + R))[lane] = q;
            }
          }
        }
      }
    }
    // wait for the state loads here, once: a first use inside the step loop
    // would put a wait for all outstanding loads (the ring refills) and
    // stores in every step
#pragma unroll
    for (int c = 0; c < CH; ++c) __asm__ volatile("" : "+v"(x[c]));
    __threadfence_block();

    // ring upkeep, at most RK steps apart: start loading the chunk below once
    // the one above the lowest is used up, and write it to LDS once fewer
    // than 64 RK words remain above lo (the next RK steps may take up to 64
    // each; the load was started at least one check earlier, since RK steps
    // take at most 64 RK <= H - 64 RK words).  (ptr is signed here: below 0
    // only on a corrupt stream.)
    auto refill = [&]() {
      if constexpr (RING) {
#pragma unroll
        for (int sidx = 0; sidx < NSTR; ++sidx) {
          const int c = sidx % CH;
          const int32_t p = (int32_t)(L == 32 ? __builtin_amdgcn_readlane(ptr[c], (uint32_t)(sidx / CH) * 32u)
                                              : ptr[c]);
          if (!pend[sidx] && lo[sidx] > 0 && p <= (int32_t)(lo[sidx] + H)) {
            const uint32_t j = lo[sidx] / H - 1;
            if ((uint32_t)lane < H4) rv[sidx] = ssrc4[sidx][j * H4 + lane];
            pend[sidx] = true;
          }
          if (pend[sidx] && p < (int32_t)(lo[sidx] + L * RK)) {
            const uint32_t j = lo[sidx] / H - 1;
            uint16_t* reg = wstage + sidx * REG;
            if ((uint32_t)lane < H4) ((uint4*)(reg + (j & 1) * H))[lane] = rv[sidx];
            if ((j & 1) == 0 && lane < 8) ((uint4*)(reg + R))[lane] = rv[sidx];
            lo[sidx] -= H;
            pend[sidx] = false;
          }
        }
      }
    };

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
    // the first chain of each block is its plane 0 whenever S > 1: its
    // destination is the element array
    __attribute__((address_space(1))) T* ep[BPT];
#pragma unroll
    for (int d = 0; d < BPT; ++d)
      ep[d] = (__attribute__((address_space(1))) T*)dstp[d * S] + blk[d * S] * BLOCK + lin;
    // block d's element of step s: bytes d * S .. d * S + S - 1 of el
    auto put = [&](uint32_t s, uint32_t el, const bool (&v)[CH]) {
      if constexpr (FUSE) {
        if (v[0]) wcodes[s * L + lin] = (uint8_t)el;
      } else {
#pragma unroll
        for (int d = 0; d < BPT; ++d)
          if (v[d * S]) ep[d][(uint64_t)s * L] = (T)(el >> (8 * S * d));
      }
    };
    const uint32_t full = n_min / L;
    const uint32_t all = (uint32_t)div_up(n_max, L);
    for (uint32_t s = all; s > full;) {
      --s;
      const uint32_t i = s * L + lin;
      bool v[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) v[c] = i < n[c];
      put(s, dec_steps<L, PB, CH, MASK, LutEntry, !RING>(x, ptr, taken, lut, stage, v, half), v);
      refill();
    }
    bool vall[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) vall[c] = true;
    // RK steps between ring checks (full is a multiple of RK for whole
    // blocks; a remainder takes single steps first)
    uint32_t s = full;
    while (s % RK) {
      --s;
      put(s, dec_steps<L, PB, CH, MASK, LutEntry, !RING>(x, ptr, taken, lut, stage, vall, half), vall);
      refill();
    }
    while (s > 0) {
      s -= RK;
      // one address per block for the RK steps (the steps' stores take
      // immediate offsets from it)
      __attribute__((address_space(1))) T* gp[BPT];
#pragma unroll
      for (int d = 0; d < BPT; ++d) gp[d] = ep[d] + (uint64_t)s * L;
#pragma unroll
      for (int k = (int)RK - 1; k >= 0; --k) {
        const uint32_t el = dec_steps<L, PB, CH, MASK, LutEntry, !RING>(x, ptr, taken, lut, stage, vall, half);
        if constexpr (FUSE) {
          wcodes[(s + (uint32_t)k) * L + lin] = (uint8_t)el;
        } else {
#pragma unroll
          for (int d = 0; d < BPT; ++d) gp[d][k * L] = (T)(el >> (8 * S * d));
        }
      }
      refill();
    }
    // a valid block uses every word and returns each lane to 2^15
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const bool bad = act[c] && ((RING ? ptr[c] != 0u : taken[c] != wds[c]) || x[c] != kMinState);
      if (__ballot(bad) && lin == 0 && act[c]) atomicOr(a.err, kErrFinalState);
    }
    if constexpr (FUSE) {
      // the block's elements: ADM lanes li0 .. li0 + BLOCK / 16 - 1 (this
      // lane: every 64th), codes from the wave's LDS (written by other lanes)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (act[0]) {
        const uint64_t li0 = blk[0] * (BLOCK / kAdmChunk);
        for (uint32_t k = lin; k < BLOCK / kAdmChunk; k += 64u) {
          const uint64_t base = (li0 + k) * kAdmChunk;
          if (base >= a.adm.n) break;
          const uint32_t cnt = (uint32_t)((a.adm.n - base) < kAdmChunk ? (a.adm.n - base) : kAdmChunk);
          const uint4 q = *(const uint4*)(wcodes + k * kAdmChunk);
          const uint32_t cw[4] = {q.x, q.y, q.z, q.w};
          adm_restore_lane(a.adm, li0 + k, cnt, cw, wsig);
        }
      }
      // (the next block's codes overwrite these)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    __threadfence_block();
  }
}

template <int L, int PB, int S, int CH, int NW, int RB, bool FUSE = false>
hipError_t launch_t(const DecodeArgs& a, int cu, hipStream_t s) {
  constexpr uint32_t CSTAGE = chain_stage_words<PB>();
  constexpr uint32_t REG = RB > 0 ? (1u << RB) + 64u : CSTAGE;
  constexpr uint32_t NSTR = CH * (L == 32 ? 2u : 1u);
  constexpr uint32_t BLOCK = L == 32 ? 4096u : 8192u;
  const size_t lds = (size_t)round_up((uint64_t)a.d.num_planes * (1u << PB) * sizeof(LutEntryT<L>), 16u) +
                     (size_t)NW * ((RB > 0 ? NSTR : CH) * REG + (RB > 0 ? 0u : 64u)) * 2u +
                     (FUSE ? (size_t)NW * (BLOCK + 64u * kAdmChunk) : 0u);
  int per_cu = 0;
  hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decode_kernel<L, PB, S, CH, NW, RB, FUSE>, 64 * NW, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  uint64_t grid = (uint64_t)cu * (uint64_t)per_cu;
  if (!a.from_header) {
    constexpr uint32_t UPW = CH * (L == 32 ? 2u : 1u);
    const uint64_t want = div_up(div_up(a.d.total_units, UPW), NW);
    if (want < grid) grid = want;
  }
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((decode_kernel<L, PB, S, CH, NW, RB, FUSE>), dim3((uint32_t)grid), dim3(64 * NW), lds, s, a);
  return hipGetLastError();
}

template <int L, int PB>
hipError_t launch_s(const DecodeArgs& a, int cu, hipStream_t s) {
  const int P = a.d.num_planes;
  const uint32_t S = a.d.plane[0].stride;
  if constexpr (L == 64) {
    // one chain per byte plane of one element block, rings of 1024 words
    // (512 with four chains): 2-9 KiB of LDS per wave.  (Two blocks per
    // wave, four chains, measured slower for uint16 and uint8 elements.)
    if (a.d.order == 0 && P == 2 && S == 2) return launch_t<L, PB, 2, 2, 8, 10>(a, cu, s);
    if (a.d.order == 0 && P == 4 && S == 4) return launch_t<L, PB, 4, 4, 8, 9>(a, cu, s);
    if (S == 1) return launch_t<L, PB, 1, 1, 8, 10>(a, cu, s);
  } else {
    // two units per chain (one per half-wave), a ring of 1024 words per unit
    if (a.d.order == 0 && P == 2 && S == 2) return launch_t<L, PB, 2, 2, 4, 10>(a, cu, s);
    if (a.d.order == 0 && P == 4 && S == 4) return launch_t<L, PB, 4, 4, 2, 9>(a, cu, s);
    if (a.d.order == 0 && P == 1 && S == 1) return launch_t<L, PB, 1, 2, 4, 10>(a, cu, s);
    // concatenated planes (ADM frames): whole-unit staging; wavefronts per
    // group chosen so LUTs + staging fit 64 KiB at p <= 10
    if (S == 1) return launch_t<L, PB, 1, 1, 4, 0>(a, cu, s);
  }
  return hipErrorInvalidValue;
}

template <int L>
hipError_t launch_p(const DecodeArgs& a, int cu, hipStream_t s) {
  switch (a.d.prob_bits) {
    case 9: return launch_s<L, 9>(a, cu, s);
    case 10: return launch_s<L, 10>(a, cu, s);
    case 11: return launch_s<L, 11>(a, cu, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_decode(const DecodeArgs& a, int cu, hipStream_t s) {
  if (!a.from_header && a.d.total_units == 0) return hipSuccess;
  return a.d.lanes == 64 ? launch_p<64>(a, cu, s) : launch_p<32>(a, cu, s);
}

hipError_t launch_decode_fused(const DecodeArgs& a, int cu, hipStream_t s) {
  if (a.d.total_units == 0) return hipSuccess;
  if (a.d.lanes != 64 || a.d.num_planes != 1 || a.d.order != 0 || a.d.plane[0].stride != 1 || a.from_header)
    return hipErrorInvalidValue;
  // four waves per workgroup, each holding 9 KiB of codes and signal
  // scratch and a 512-word ring (on 512 MiB of detector codes: 0.54 ms;
  // 1024-word rings 0.71 ms; two waves per workgroup 0.99 ms)
  switch (a.d.prob_bits) {
    case 9: return launch_t<64, 9, 1, 1, 4, 9, true>(a, cu, s);
    case 10: return launch_t<64, 10, 1, 1, 4, 9, true>(a, cu, s);
    case 11: return launch_t<64, 11, 1, 1, 4, 9, true>(a, cu, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mans
