// encode.hip -- interleaved rANS encoder, one gfx950 wavefront per 64 lanes.
//
// Reference: include/ans/GpuANSEncode.h:24-164 (encodeOne / ansEncodeBatch),
// :166-178 + BatchPrefixSum.h:29-187 (prefix of padded word counts) and
// :180-284 (ansEncodeCoalesceBatch).  The reference encodes each 4 KiB block
// into an uncoalesced staging buffer in HBM, scans the word counts with up to
// three launches (limit 512^2 blocks) and copies every block again into the
// final container.  Here one persistent launch does all of it:
//  * a wavefront takes a ticket (atomic counter) for its units: two 32-lane
//    reference blocks (one per half-wave) or one 64-lane native block;
//  * emitted 16-bit words go to an LDS staging area (bounded by
//    lane_max_words), in lane order within each step (64-bit ballot + mbcnt);
//  * the wave publishes its padded word count and resolves its output offset
//    by decoupled look-back over earlier tickets (tickets are taken by running
//    waves, so every wave waited on is resident and finishes);
//  * it then writes states, block words and the LDS words straight to their
//    final place, 16 bytes per lane per store.
// The arithmetic per symbol is the reference's: emit x & 0xffff when
// x >= pdf * 2^(31-p); q = (umulhi(x, magic) + x) >> shift = x / pdf;
// x = q * 2^p + (x - q * pdf) + cdf, computed as x + q * (2^p - pdf) + cdf.
#include "mans_internal.hpp"

namespace mans {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagPrefix = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;

struct LanePlane {   // per-plane fields copied to LDS (dynamic index by lane)
  const uint8_t* src;
  uint64_t nsym;
  uint64_t section;
  uint64_t nb;
};

template <int L, int PB, int S>
__global__ __launch_bounds__(kThreads) void encode_kernel(EncodeArgs a) {
  extern __shared__ uint4 smem[];
  __shared__ LanePlane s_pl[kMaxPlanes];
  __shared__ uint64_t s_ubase[kMaxPlanes + 1];
  constexpr uint32_t BLOCK = L == 32 ? 4096u : 8192u;
  constexpr uint32_t UPW = L == 32 ? 2u : 1u;
  constexpr uint32_t STAGE = (uint32_t)round_up(64u * lane_max_words(128u, PB), 16u);
  constexpr uint32_t STAGE_UNIT = STAGE / UPW;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int P = a.d.num_planes;

  uint4* s_tab = smem;  // [P][256] {thresh, magic, shift | cdf << 8, 2^PB - pdf}
  for (int i = t; i < P * (int)kNumSymbols; i += kThreads) {
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
  uint16_t* stage = (uint16_t*)(smem + P * kNumSymbols) + wave * STAGE + half * STAGE_UNIT;

  for (;;) {
    uint32_t tk = 0;
    if (lane == 0) tk = atomicAdd(a.ticket, UPW);
    tk = __builtin_amdgcn_readfirstlane(tk);
    if ((uint64_t)tk >= a.d.total_units) break;
    const uint64_t unit = (uint64_t)tk + half;
    const bool active = unit < a.d.total_units;
    int k = 0;
    uint64_t b = 0;
    if (active) {
      if (a.d.order == 0) {  // planes of one element array: P is 1, 2 or 4
        k = (int)((uint32_t)unit & (uint32_t)(P - 1));
        b = unit >> (P == 4 ? 2 : (P == 2 ? 1 : 0));
      } else {
        while (k + 1 < P && unit >= s_ubase[k + 1]) ++k;
        b = unit - s_ubase[k];
      }
    }
    const LanePlane pl = s_pl[k];
    const uint64_t start = b * BLOCK;
    const uint32_t n = active ? (uint32_t)((pl.nsym - start) < BLOCK ? (pl.nsym - start) : BLOCK) : 0u;
    uint32_t n_min, n_max;
    if (L == 32) {
      uint32_t n0 = __builtin_amdgcn_readlane(n, 0), n1 = __builtin_amdgcn_readlane(n, 32);
      n_min = n0 < n1 ? n0 : n1;
      n_max = n0 < n1 ? n1 : n0;
    } else {
      n_min = n_max = __builtin_amdgcn_readfirstlane(n);
    }
    const uint8_t* sp = pl.src + (start + lin) * S;
    const uint4* tab = s_tab + k * kNumSymbols;
    uint32_t x = kMinState;
    uint32_t cnt = 0;  // words emitted by this lane's unit (uniform per half)

    auto step = [&](uint32_t sym, bool v) {
      const uint4 e = tab[sym];
      const bool w = v && (x >= e.x);
      const uint64_t m = __ballot(w);
      const uint32_t mlo = (uint32_t)m, mhi = (uint32_t)(m >> 32);
      uint32_t base, add;
      if (L == 64) {
        base = cnt;
        add = (uint32_t)__popcll(m);
      } else {
        base = half ? cnt - (uint32_t)__popc(mlo) : cnt;
        add = half ? (uint32_t)__popc(mhi) : (uint32_t)__popc(mlo);
      }
      const uint32_t pos = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, base));
      if (w) {
        stage[pos] = (uint16_t)x;
        x >>= 16;
      }
      cnt += add;
      if (v) {
        const uint32_t q = (__umulhi(x, e.y) + x) >> (e.z & 31u);
        x = __umul24(q, e.w) + x + (e.z >> 8);
      }
    };

    const uint32_t full = n_min / L;
    const uint32_t all = (uint32_t)div_up(n_max, L);
    uint32_t s = 0;
    for (; s + 8 <= full; s += 8) {
      uint32_t sy[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) sy[j] = sp[(uint64_t)(s + j) * L * S];
#pragma unroll
      for (int j = 0; j < 8; ++j) step(sy[j], true);
    }
    for (; s < full; ++s) step(sp[(uint64_t)s * L * S], true);
    for (; s < all; ++s) {
      const bool v = s * L + lin < n;
      const uint32_t sym = v ? (uint32_t)sp[(uint64_t)s * L * S] : 0u;
      step(sym, v);
    }

    const uint32_t words = cnt;
    const uint32_t padded = (uint32_t)round_up(words, 8);
    if (words > STAGE_UNIT && lin == 0) atomicOr(a.err, kErrStageOverflow);
    if (lin < padded - words) stage[words + lin] = 0;
    uint8_t* sec = a.out + pl.section;
    if (active) ((uint32_t*)(sec + kSectionFixed))[b * L + lin] = x;

    // ---- decoupled look-back over tickets --------------------------------
    // Each status word carries its flag and its value together, and no other
    // memory is handed over through it, so relaxed device-scope atomics are
    // enough.  (Release / acquire at agent scope would write back / invalidate
    // the whole L2 of the XCD on every publish and poll.)
    const uint32_t a_lo = __builtin_amdgcn_readlane(padded, 0);
    const uint32_t a_hi = L == 32 ? __builtin_amdgcn_readlane(padded, 32) : 0u;
    const uint64_t agg = (uint64_t)a_lo + a_hi;
    const uint64_t chain = tk / UPW;
    uint64_t excl = 0;
    if (lane == 0) {
      uint64_t* st = a.status + chain;
      if (chain == 0) {
        __hip_atomic_store(st, kFlagPrefix | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(st, kFlagAgg | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t j = chain - 1;
        for (;;) {
          const uint64_t v = __hip_atomic_load(a.status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t f = v >> 62;
          if (f == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          excl += v & kValueMask;
          if (f == 2) break;
          --j;
        }
        __hip_atomic_store(st, kFlagPrefix | (excl + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    excl = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(excl >> 32)) << 32) |
           (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)excl);
    const uint64_t y = excl + (half ? a_lo : 0u);

    if (active && lin == 0) {
      uint32_t* bw = (uint32_t*)(sec + kSectionFixed + 4ull * L * pl.nb + 8ull * b);
      bw[0] = (n << 16) | words;
      bw[1] = (uint32_t)y;
      atomicAdd((uint32_t*)(sec + 12), padded);
    }
    __threadfence_block();
    uint4* dst = (uint4*)(a.out + a.d.data_off + 2ull * y);
    const uint4* src4 = (const uint4*)stage;
    for (uint32_t j = lin; j < padded / 8; j += L) dst[j] = src4[j];

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
  }
}

template <int L, int PB, int S>
hipError_t launch_t(const EncodeArgs& a, int cu, hipStream_t s) {
  constexpr uint32_t STAGE = (uint32_t)round_up(64u * lane_max_words(128u, PB), 16u);
  const size_t lds = (size_t)a.d.num_planes * kNumSymbols * sizeof(uint4) + (size_t)kWaves * STAGE * 2u;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, encode_kernel<L, PB, S>, kThreads, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  constexpr uint32_t UPW = L == 32 ? 2u : 1u;
  const uint64_t waves = div_up(a.d.total_units, UPW);
  uint64_t grid = div_up(waves, kWaves);
  const uint64_t cap = (uint64_t)cu * (uint64_t)per_cu;
  if (grid > cap) grid = cap;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((encode_kernel<L, PB, S>), dim3((uint32_t)grid), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

template <int L, int PB>
hipError_t launch_s(const EncodeArgs& a, int cu, hipStream_t s) {
  switch (a.d.plane[0].stride) {
    case 1: return launch_t<L, PB, 1>(a, cu, s);
    case 2: return launch_t<L, PB, 2>(a, cu, s);
    case 4: return launch_t<L, PB, 4>(a, cu, s);
  }
  return hipErrorInvalidValue;
}

template <int L>
hipError_t launch_p(const EncodeArgs& a, int cu, hipStream_t s) {
  switch (a.d.prob_bits) {
    case 9: return launch_s<L, 9>(a, cu, s);
    case 10: return launch_s<L, 10>(a, cu, s);
    case 11: return launch_s<L, 11>(a, cu, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_encode(const EncodeArgs& a, int cu, hipStream_t s) {
  if (a.d.total_units == 0) return hipSuccess;
  return a.d.lanes == 64 ? launch_p<64>(a, cu, s) : launch_p<32>(a, cu, s);
}

}  // namespace mans
