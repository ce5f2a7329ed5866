// adm.hip -- ADM byte mapping of uint16 data (code plane + bit-signal plane).
//
// Reference: mapping/cpu/mappingcpuV1.cpp:33-276 (semantics) and
// mapping/nv/mappingV15_fused_uint16.cu:24-121, 124-398, 400-730 (GPU
// kernels).  Group g = elements [512g, 512g + 512) = 32 lanes x 16 elements.
// center = floor(sum / count); per element d = |x - c|, len = 1 if x == c else
// ceil(d / 126), code = 1 if x == c else 2(d - 126(len - 1)) (+1 if x < c);
// each lane sets bit `bo` (MSB first) then advances bo by len; the group's
// lanes are padded to M_g = ceil(max L / 8) bytes (the byte holding bit L gets
// 0xFF >> (L % 8)), lanes without elements stay zero (mappingcpuV1.cpp:146);
// lane l's bytes go to 32 O_g + l M_g with O = exclusive prefix of M.
//
// Structure: a wavefront takes 32 groups per ticket, two at a time (one
// per half-wave).  Pass A writes codes and centers, finds each group's byte
// length and publishes the ticket's sum; after pass A of the wave's next
// ticket, the look-back gives the offset and pass B rebuilds the bits (lane
// bit buffers in LDS, ds_or per element) and writes them at their offsets.
// O_g is resolved by decoupled look-back over tickets (any number of groups;
// the reference's fused path is limited to 1024 groups,
// mappingV15_fused_uint16.cu:851), the whole wave reading 512 predecessors
// per round trip.  Unmap parses each lane's bytes with clz over 32-bit words.
#include "adm_device.hpp"
#include "codec_device.hpp"

namespace mans {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ uint32_t half_sum(uint32_t v) {
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint32_t half_max(uint32_t v) {
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) {
    uint32_t u = __shfl_xor(v, o, 64);
    v = v > u ? v : u;
  }
  return v;
}

// Group pairs per ticket.  A ticket of 32 groups (16384 elements) keeps the
// look-back chain short and the ticket counter's traffic low: 2-group
// tickets (256K per 512 MiB) made the scan, not the data, the bound, and on
// 512 MiB 16-group tickets took 0.55 ms, 32 0.50 ms, 64 0.53 ms.
constexpr uint32_t kPairs = kAdmTicketGroups / 2;
static_assert(kPairs % 2 == 0, "pairs are mapped two at a time");
// Lane bit buffers are padded to 9 words so the 64 lanes' words of one
// index fall in distinct LDS banks.
constexpr uint32_t kLaneWords = kAdmLaneBytes / 4 + 1;

// The 16 elements of one lane (half-wave = one group): raw_lane loads them
// (two 16-byte loads for a whole lane, element loads for the last, partial
// lane of the input); unpack_center spreads them and returns the group's
// center (all lanes of the half-wave).  Split so a pair's loads can be in
// flight while the previous pair is mapped.
struct LaneRaw {
  uint4 q0, q1;
  uint32_t cnt;
  uint64_t base;
};
// FULL: every lane of the ticket is whole (tickets before the input's end):
// no bounds, no branch around the loads.
template <bool FULL>
__device__ __forceinline__ LaneRaw raw_lane(const AdmMapArgs& a, uint64_t g, uint32_t lin) {
  LaneRaw r;
  r.base = g * kAdmGroup + (uint64_t)lin * kAdmChunk;
  if constexpr (FULL) {
    r.cnt = kAdmChunk;
    const uint4* p = (const uint4*)(a.in + r.base);
    r.q0 = p[0];
    r.q1 = p[1];
    return r;
  }
  const bool gactive = g < a.gsize;
  r.cnt = gactive && r.base < a.n ? (uint32_t)((a.n - r.base) < kAdmChunk ? (a.n - r.base) : kAdmChunk) : 0u;
  // (a partial lane, only at the end of the input, is loaded element by
  // element in unpack_center: no element loads here, so nothing makes the
  // wave wait for these loads before their pair is mapped)
  if (r.cnt == kAdmChunk) {
    const uint4* p = (const uint4*)(a.in + r.base);
    r.q0 = p[0];
    r.q1 = p[1];
  } else {
    r.q0 = make_uint4(0, 0, 0, 0);
    r.q1 = make_uint4(0, 0, 0, 0);
  }
  return r;
}
template <bool FULL>
__device__ __forceinline__ uint32_t unpack_center(const AdmMapArgs& a, const LaneRaw& r, uint32_t (&v)[kAdmChunk]) {
  uint32_t w[8] = {r.q0.x, r.q0.y, r.q0.z, r.q0.w, r.q1.x, r.q1.y, r.q1.z, r.q1.w};
  if (!FULL && r.cnt > 0 && r.cnt < kAdmChunk) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t lo = (uint32_t)(2 * i) < r.cnt ? (uint32_t)a.in[r.base + 2 * i] : 0u;
      const uint32_t hi = (uint32_t)(2 * i + 1) < r.cnt ? (uint32_t)a.in[r.base + 2 * i + 1] : 0u;
      w[i] = lo | (hi << 16);
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    v[2 * i] = w[i] & 0xffffu;
    v[2 * i + 1] = w[i] >> 16;
  }
  uint32_t lsum = 0;
#pragma unroll
  for (int i = 0; i < (int)kAdmChunk; ++i) lsum += (uint32_t)i < r.cnt ? v[i] : 0u;
  const uint32_t gsum = half_sum(lsum), gcnt = half_sum(r.cnt);
  return gcnt ? gsum / gcnt : 0u;  // mappingcpuV1.cpp:60-66
}

// Per element (mappingcpuV1.cpp:68-146): d = |x - c|, len = 1 if x == c else
// ceil(d / 126), code = 1 if x == c else 2 (d - 126 (len - 1)) (+1 if x < c);
// a one at bit `bo` (MSB first), then bo += len.  With `bits`, sets those
// bits in the lane's LDS buffer; returns bo (the lane's bit length).
template <bool PACK>
__device__ __forceinline__ uint32_t map_lane(const uint32_t (&v)[kAdmChunk], uint32_t cnt, uint32_t c,
                                             uint32_t (&codes)[4], uint32_t* bits) {
  uint32_t bo = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) codes[i] = 0;
  // (branch-free over the elements: an element past cnt adds no code byte
  // and no length)
#pragma unroll
  for (int i = 0; i < (int)kAdmChunk; ++i) {
    const bool ok = (uint32_t)i < cnt;
    const uint32_t x = v[i];
    const uint32_t d = x > c ? x - c : c - x;
    const uint32_t len = x == c ? 1u : (d + kAdmMod - 1) / kAdmMod;
    const uint32_t code = x == c ? 1u : ((d - kAdmMod * (len - 1)) * 2u + (x > c ? 0u : 1u));
    codes[i >> 2] |= (ok ? code & 0xffu : 0u) << (8 * (i & 3));
    if (PACK && ok && bo < 8 * kAdmLaneBytes)
      atomicOr(bits + (bo >> 5), 1u << (((bo >> 3) & 3u) * 8u + 7u - (bo & 7u)));
    bo += ok ? len : 0u;
  }
  return bo;
}

// A wavefront resolves a ticket's offset only after pass A of its next
// ticket (the byte lengths of two tickets in LDS, used in turn): by then the
// predecessors have had a whole pass A to publish their lengths, so the
// look-back sums published entries instead of waiting for the slowest wave
// still in pass A.  A wave waits only on lower tickets, each of which
// publishes its aggregate right after its pass A, so the launch drains.
__global__ __launch_bounds__(kThreads) void adm_map_kernel(AdmMapArgs a) {
  __shared__ uint32_t s_bits[kWaves][64][kLaneWords];
  __shared__ uint32_t s_m[kWaves][2][2 * kPairs];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t half = (uint32_t)lane >> 5, lin = (uint32_t)lane & 31u;
  uint32_t* mybits = s_bits[wave][lane];
  const uint8_t* wavebytes = (const uint8_t*)s_bits[wave];
  constexpr uint32_t kTicketGroups = 2 * kPairs;

  // the ticket whose pass A is done and whose offset is not yet resolved
  // (wave-uniform), its byte lengths in s_m[wave][slot ^ 1]; lane g < 16:
  // group g's exclusive byte offset within it
  bool pend = false;
  uint32_t p_tk = 0, p_exlo = 0, slot = 0;
  uint64_t p_agg = 0;

  // offsets and pass B of the pending ticket: each pair's bit plane, lane
  // l's m bytes at 32 og + l m
  auto finish_pending = [&]() {
    const uint32_t* wm = s_m[wave][slot ^ 1u];
    const uint64_t chain = p_tk / kTicketGroups;
    uint64_t excl = resolve_prefix(a.status, chain, p_agg, lane);
    excl = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(excl >> 32)) << 32) |
           (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)excl);
    {
      const uint64_t g = (uint64_t)p_tk + lane;
      if ((uint32_t)lane < kTicketGroups && g < a.gsize) a.offsets[g] = (int32_t)(excl + p_exlo);
    }
    if ((uint64_t)p_tk + kTicketGroups >= a.gsize && lane == 0) a.offsets[a.gsize] = (int32_t)(excl + p_agg);
    auto pair_b = [&](uint32_t pr, const LaneRaw& cur, auto full) {
      constexpr bool FULL = decltype(full)::value;
      const uint64_t g = (uint64_t)p_tk + 2 * pr + half;
      uint32_t v[kAdmChunk], codes[4];
      const uint32_t cnt = cur.cnt;
      const uint32_t c = unpack_center<FULL>(a, cur, v);
#pragma unroll
      for (int i = 0; i < (int)kLaneWords; ++i) mybits[i] = 0;
      __threadfence_block();
      const uint32_t bo = map_lane<true>(v, cnt, c, codes, mybits);
      const uint32_t m = wm[2 * pr + half];
      if (cnt && bo < 8u * m) {
        const uint32_t pad = (bo & 7u) == 0 ? 0xffu : (0xffu >> (bo & 7u));
        atomicOr(mybits + (bo >> 5), pad << (((bo >> 3) & 3u) * 8u));
      }
      __threadfence_block();
      const uint32_t og_lo = __shfl(p_exlo, (int)(2 * pr), 64), og_hi = __shfl(p_exlo, (int)(2 * pr + 1), 64);
      const uint64_t og = excl + (half ? og_hi : og_lo);
      if (g < a.gsize && m > 0) {
        const uint64_t dst0 = 32ull * og;
        const uint32_t total = 32u * m;
        const uint32_t inv = (65536u + m - 1u) / m;
        const uint64_t valid_lanes = a.n > g * kAdmGroup ? div_up(a.n - g * kAdmGroup, kAdmChunk) : 0;
        if (dst0 + total <= a.bits_capacity) {
          for (uint32_t q = lin; q < total; q += 32u) {
            const uint32_t l = (q * inv) >> 16, j = q - l * m;
            const uint8_t byte = l < valid_lanes ? wavebytes[(half * 32u + l) * (4u * kLaneWords) + j] : (uint8_t)0;
            a.bits[dst0 + q] = byte;
          }
        } else if (lin == 0) {
          atomicOr(a.err, kErrCapacity);
        }
      }
      __threadfence_block();
    };
    // two buffers used in turn (no register copy between them, so the next
    // pair's loads stay in flight while a pair is mapped)
    auto pass_b = [&](auto full) {
      constexpr bool FULL = decltype(full)::value;
      LaneRaw ra = raw_lane<FULL>(a, (uint64_t)p_tk + half, lin), rb;
      for (uint32_t pr = 0; pr < kPairs; pr += 2) {
        rb = raw_lane<FULL>(a, (uint64_t)p_tk + 2 * (pr + 1) + half, lin);
        pair_b(pr, ra, full);
        if (pr + 2 < kPairs) ra = raw_lane<FULL>(a, (uint64_t)p_tk + 2 * (pr + 2) + half, lin);
        pair_b(pr + 1, rb, full);
      }
    };
    if ((uint64_t)(p_tk + kTicketGroups) * kAdmGroup <= a.n) pass_b(std::true_type{});
    else pass_b(std::false_type{});
  };

  for (;;) {
    uint32_t tk = 0;
    if (lane == 0) tk = atomicAdd(a.ticket, kTicketGroups);
    tk = __builtin_amdgcn_readfirstlane(tk);
    const bool more = (uint64_t)tk < a.gsize;
    uint32_t exlo = 0;
    uint64_t agg = 0;
    if (more) {
      // ---- pass A: codes, centers and each group's byte length M_g -------
      uint32_t* wm = s_m[wave][slot];
      bool over = false;
      auto pair_a = [&](uint32_t pr, const LaneRaw& cur, auto full) {
        constexpr bool FULL = decltype(full)::value;
        const uint64_t g = (uint64_t)tk + 2 * pr + half;
        uint32_t v[kAdmChunk], codes[4];
        const uint32_t cnt = cur.cnt;
        const uint64_t base = cur.base;
        const uint32_t c = unpack_center<FULL>(a, cur, v);
        const uint32_t bo = map_lane<false>(v, cnt, c, codes, nullptr);
        if (bo > 8 * kAdmLaneBytes) over = true;
        uint32_t m = half_max(cnt ? (bo + 7u) / 8u : 0u);
        if (m > kAdmLaneBytes) m = kAdmLaneBytes;
        if (cnt == kAdmChunk) {
          *(uint4*)(a.codes + base) = make_uint4(codes[0], codes[1], codes[2], codes[3]);
        } else {
          for (uint32_t i = 0; i < cnt; ++i) a.codes[base + i] = (uint8_t)(codes[i >> 2] >> (8 * (i & 3)));
        }
        if (g < a.gsize && lin == 0) a.centers[g] = (uint16_t)c;
        if (lin == 0) wm[2 * pr + half] = m;
      };
      auto pass_a = [&](auto full) {
        constexpr bool FULL = decltype(full)::value;
        LaneRaw ra = raw_lane<FULL>(a, (uint64_t)tk + half, lin), rb;
        for (uint32_t pr = 0; pr < kPairs; pr += 2) {
          rb = raw_lane<FULL>(a, (uint64_t)tk + 2 * (pr + 1) + half, lin);
          pair_a(pr, ra, full);
          if (pr + 2 < kPairs) ra = raw_lane<FULL>(a, (uint64_t)tk + 2 * (pr + 2) + half, lin);
          pair_a(pr + 1, rb, full);
        }
      };
      // (tickets before the input's end: whole lanes only)
      if ((uint64_t)(tk + kTicketGroups) * kAdmGroup <= a.n) pass_a(std::true_type{});
      else pass_a(std::false_type{});
      if (__ballot(over) && lane == 0) atomicOr(a.err, kErrAdmDomain);
      __threadfence_block();
      // the ticket's groups: exclusive byte offsets within it, and their sum
      const uint32_t mine = (uint32_t)lane < kTicketGroups ? wm[lane] : 0u;
      uint32_t incl = mine;
#pragma unroll
      for (int o = 1; o < (int)kTicketGroups; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
      }
      agg = __builtin_amdgcn_readlane(incl, kTicketGroups - 1);
      exlo = incl - mine;
      publish_aggregate(a.status, tk / kTicketGroups, agg, lane);
    }
    if (pend) finish_pending();
    if (!more) break;
    pend = true;
    p_tk = tk;
    p_exlo = exlo;
    p_agg = agg;
    slot ^= 1u;
  }
}

__global__ __launch_bounds__(kThreads) void adm_unmap_kernel(AdmUnmapArgs a) {
  __shared__ uint8_t s_sig[kThreads][kAdmChunk];
  const int t = threadIdx.x;
  const uint64_t nlanes = (uint64_t)gridDim.x * kThreads;
  for (uint64_t li = (uint64_t)blockIdx.x * kThreads + t; li < a.gsize * kAdmLanes; li += nlanes) {
    const uint64_t base = li * kAdmChunk;
    if (base >= a.n) continue;
    const uint32_t cnt = (uint32_t)((a.n - base) < kAdmChunk ? (a.n - base) : kAdmChunk);
    const uint8_t* cp = a.codes + base;
    uint32_t cw[4];
    if (cnt == kAdmChunk) {
      const uint4 q = *(const uint4*)cp;
      cw[0] = q.x; cw[1] = q.y; cw[2] = q.z; cw[3] = q.w;
    } else {
      for (int i = 0; i < 4; ++i) cw[i] = 0;
      for (uint32_t i = 0; i < cnt; ++i) cw[i >> 2] |= (uint32_t)cp[i] << (8 * (i & 3));
    }
    adm_restore_lane(a, li, cnt, cw, s_sig[t]);
  }
}

}  // namespace

hipError_t launch_adm_map(const AdmMapArgs& a, int cu, hipStream_t s) {
  if (a.gsize == 0) return hipSuccess;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adm_map_kernel, kThreads, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  uint64_t grid = div_up(div_up(a.gsize, kAdmTicketGroups), kWaves);
  const uint64_t cap = (uint64_t)cu * per_cu;
  if (grid > cap) grid = cap;
  hipLaunchKernelGGL(adm_map_kernel, dim3((uint32_t)grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adm_unmap(const AdmUnmapArgs& a, hipStream_t s) {
  if (a.gsize == 0) return hipSuccess;
  uint64_t grid = div_up(a.gsize * kAdmLanes, kThreads);
  if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(adm_unmap_kernel, dim3((uint32_t)grid), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace mans
