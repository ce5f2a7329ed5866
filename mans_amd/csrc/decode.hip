// decode.hip -- interleaved rANS decoder, one gfx950 wavefront per 64 lanes.
//
// Reference: include/ans/GpuANSDecode.h:27-46 (packed LUT entry),
// :48-142 (decodeOneWarp / decodeOnePartialWarp), :144-302 (block loop,
// grid-stride over blocks), :304-377 (ansDecodeTable).  Differences in
// structure, not in results:
//  * every workgroup builds the 2^p-entry LUT of each plane in LDS straight
//    from the container's pdf table (no separate table launch, no global
//    table);
//  * a wavefront decodes two 32-lane reference blocks (one per half-wave) or
//    one 64-lane native block; the block's words are first copied to LDS with
//    16-byte loads, so the serial per-lane chain reads LDS, not HBM;
//  * a block is valid only if it consumes exactly its words and every lane
//    ends in the start state 2^15; anything else raises an error flag.
// Per symbol: slot = x & (2^p - 1); {sym, pdf, slot - cdf} = LUT[slot];
// x = pdf * (x >> p) + slot - cdf; lanes with x < 2^15 read a word from the
// end of the block's remaining words, the highest such lane taking the last
// word; x = (x << 16) + word.
#include "mans_internal.hpp"

namespace mans {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

struct DPlane {
  uint8_t* dst;
  uint64_t nsym;
  uint64_t section;
  uint64_t nb;
};

template <int L, int PB, int S>
__global__ __launch_bounds__(kThreads) void decode_kernel(DecodeArgs a) {
  constexpr uint32_t BLOCK = L == 32 ? 4096u : 8192u;
  constexpr uint32_t UPW = L == 32 ? 2u : 1u;
  constexpr uint32_t W = 1u << PB;
  constexpr uint32_t STAGE = (uint32_t)round_up(64u * lane_max_words(128u, PB), 16u);
  constexpr uint32_t STAGE_UNIT = STAGE / UPW;
  constexpr uint32_t UNIT_CAP = L * lane_max_words(BLOCK / L, PB);
  extern __shared__ uint4 smem[];
  __shared__ DPlane s_pl[kMaxPlanes];
  __shared__ uint64_t s_ubase[kMaxPlanes + 1];
  __shared__ uint32_t s_cdf[kNumSymbols + 1];
  __shared__ uint32_t s_tmp[kWaves];
  __shared__ uint64_t s_total_units, s_data_off, s_data_words;
  __shared__ int s_bad;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int P = a.d.num_planes;
  uint32_t* s_lut = (uint32_t*)smem;  // [P][W]

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

  // ---- LUT per plane: entry = (slot - cdf) << 20 | pdf << 8 | sym --------
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
  for (int k = 0; k < P; ++k) {
    const uint16_t* probs = (const uint16_t*)(a.in + s_pl[k].section + kAnsHeaderBytes);
    const uint32_t pdf = probs[t];
    uint32_t inc = pdf;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      uint32_t u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (int j = 0; j < wave; ++j) base += s_tmp[j];
    s_cdf[t] = base + inc - pdf;
    if (t == kThreads - 1) s_cdf[kNumSymbols] = base + inc;
    __syncthreads();
    if (s_pl[k].nsym > 0 && s_cdf[kNumSymbols] != W) {
      if (t == 0) s_bad = kErrTable;
    } else {
      uint32_t* lut = s_lut + k * W;
      for (uint32_t slot = t; slot < W; slot += kThreads) {
        // largest symbol s with cdf[s] <= slot (that symbol has pdf > 0)
        uint32_t lo = 0, hi = kNumSymbols;
        while (hi - lo > 1) {
          uint32_t mid = (lo + hi) >> 1;
          if (s_cdf[mid] <= slot) lo = mid; else hi = mid;
        }
        const uint32_t c = s_cdf[lo], p = s_cdf[lo + 1] - c;
        lut[slot] = ((slot - c) << 20) | (p << 8) | lo;
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
  uint16_t* stage = (uint16_t*)(smem + div_up((uint64_t)P * W * 4u, 16u)) + wave * STAGE + half * STAGE_UNIT;
  const uint64_t total_units = s_total_units;
  const uint64_t data_off = s_data_off, data_words = s_data_words;
  const uint64_t nwaves = (uint64_t)gridDim.x * kWaves;

  for (uint64_t wu = (uint64_t)blockIdx.x * kWaves + wave; wu * UPW < total_units; wu += nwaves) {
    const uint64_t unit = wu * UPW + half;
    bool active = unit < total_units;
    int k = 0;
    uint64_t b = 0;
    if (active) {
      if (a.d.order == 0) {
        k = (int)((uint32_t)unit & (uint32_t)(P - 1));
        b = unit >> (P == 4 ? 2 : (P == 2 ? 1 : 0));
      } else {
        while (k + 1 < P && unit >= s_ubase[k + 1]) ++k;
        b = unit - s_ubase[k];
      }
    }
    const DPlane pl = s_pl[k];
    const uint8_t* sec = a.in + pl.section;
    const uint64_t start = b * BLOCK;
    uint32_t n = 0, words = 0, y = 0;
    if (active) {
      const uint2 bw = *(const uint2*)(sec + kSectionFixed + 4ull * L * pl.nb + 8ull * b);
      n = bw.x >> 16;
      words = bw.x & 0xffffu;
      y = bw.y;
      const uint64_t expect = (pl.nsym - start) < BLOCK ? (pl.nsym - start) : BLOCK;
      if (n != expect || words > UNIT_CAP || words > STAGE_UNIT ||
          (uint64_t)y + round_up(words, 8) > data_words ||
          (a.in_bytes && data_off + 2ull * ((uint64_t)y + words) > a.in_bytes)) {
        if (lin == 0) atomicOr(a.err, kErrBlockWords);
        active = false;
        n = 0;
        words = 0;
      }
    }
    // words of the unit -> LDS (16-byte loads; y is a multiple of 8 words)
    {
      const uint4* src4 = (const uint4*)(a.in + data_off + 2ull * y);
      uint4* dst4 = (uint4*)stage;
      const uint32_t n4 = (uint32_t)div_up(words, 8);
      for (uint32_t j = lin; j < n4; j += L) dst4[j] = src4[j];
    }
    __threadfence_block();
    uint32_t x = active ? ((const uint32_t*)(sec + kSectionFixed))[b * L + lin] : kMinState;
    uint32_t ptr = words;
    uint32_t n_min, n_max;
    if (L == 32) {
      uint32_t n0 = __builtin_amdgcn_readlane(n, 0), n1 = __builtin_amdgcn_readlane(n, 32);
      n_min = n0 < n1 ? n0 : n1;
      n_max = n0 < n1 ? n1 : n0;
    } else {
      n_min = n_max = __builtin_amdgcn_readfirstlane(n);
    }
    const uint32_t* lut = s_lut + k * W;
    uint8_t* dp = pl.dst + (start + lin) * S;

    auto step = [&](uint32_t s, bool v) {
      const uint32_t e = lut[x & (W - 1u)];
      if (v) x = __umul24((e >> 8) & 0xfffu, x >> PB) + (e >> 20);
      const bool r = v && (x < kMinState);
      const uint64_t m = __ballot(r);
      const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)(m >> 32);
      uint32_t base, cnt;
      if (L == 64) {
        cnt = (uint32_t)__popcll(m);
        base = ptr - cnt;
      } else {
        cnt = half ? (uint32_t)__popc(mhi) : (uint32_t)__popc(mlo);
        base = half ? ptr - cnt - (uint32_t)__popc(mlo) : ptr - cnt;
      }
      const uint32_t idx = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, base));
      if (r) x = (x << 16) | (uint32_t)stage[idx];
      ptr -= cnt;
      if (v) dp[(uint64_t)s * L * S] = (uint8_t)e;
    };

    const uint32_t full = n_min / L;
    const uint32_t all = (uint32_t)div_up(n_max, L);
    for (uint32_t s = all; s > full;) {
      --s;
      step(s, s * L + lin < n);
    }
    for (uint32_t s = full; s > 0;) {
      --s;
      step(s, true);
    }
    // a valid block uses every word and returns each lane to 2^15
    const bool bad = active && (ptr != 0 || x != kMinState);
    if (__ballot(bad) && lin == 0 && active) atomicOr(a.err, kErrFinalState);
    __threadfence_block();
  }
}

template <int L, int PB, int S>
hipError_t launch_t(const DecodeArgs& a, int cu, hipStream_t s) {
  constexpr uint32_t STAGE = (uint32_t)round_up(64u * lane_max_words(128u, PB), 16u);
  const size_t lds = (size_t)round_up((uint64_t)a.d.num_planes * (1u << PB) * 4u, 16u) + (size_t)kWaves * STAGE * 2u;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, decode_kernel<L, PB, S>, kThreads, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  uint64_t grid = (uint64_t)cu * (uint64_t)per_cu;
  if (!a.from_header) {
    constexpr uint32_t UPW = L == 32 ? 2u : 1u;
    const uint64_t want = div_up(div_up(a.d.total_units, UPW), kWaves);
    if (want < grid) grid = want;
  }
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((decode_kernel<L, PB, S>), dim3((uint32_t)grid), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

template <int L, int PB>
hipError_t launch_s(const DecodeArgs& a, int cu, hipStream_t s) {
  switch (a.d.plane[0].stride) {
    case 1: return launch_t<L, PB, 1>(a, cu, s);
    case 2: return launch_t<L, PB, 2>(a, cu, s);
    case 4: return launch_t<L, PB, 4>(a, cu, s);
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

}  // namespace mans
