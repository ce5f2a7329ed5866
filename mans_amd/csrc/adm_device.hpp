// adm_device.hpp -- the ADM restore of one 16-element lane (the unmap
// kernel's per-thread work), split so that the bit bytes, center and codes
// may come from global memory or LDS.  mappingcpuV1.cpp:195-276;
// mappingV15_fused_uint16.cu:24-121.
#pragma once
#include "mans_internal.hpp"

namespace mans {

// Restores the cnt elements at out from the lane's m bytes of the bit plane
// at src (global memory or LDS): parses them into per-element signals and
// sets x = c + d or c - d with d = code / 2 + 126 signal (code & 1: below
// the center).  cw holds the 16 code bytes, little endian; sig is 16 bytes
// of this lane's own LDS scratch.
__device__ inline void adm_restore_core(uint32_t m, const uint8_t* src, uint32_t c, uint32_t cnt,
                                        const uint32_t (&cw)[4], uint8_t* sig, uint16_t* out) {
  uint32_t w[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint32_t i = 4u * r + b;
      v = (v << 8) | (i < m ? (uint32_t)src[i] : 0u);
    }
    w[r] = v;  // bit position 32r + j is bit (31 - j) of w[r]
  }
#pragma unroll
  for (int i = 0; i < (int)kAdmChunk; ++i) sig[i] = 0;
  // mappingcpuV1.cpp:233-243: a one advances the element index (stop at
  // 16), zeros after the i-th one count into signal[i] (uint8)
  int idx = -1;
  uint32_t prev = 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    uint32_t v = w[r];
    while (v != 0 && idx < (int)kAdmChunk) {
      const uint32_t z = (uint32_t)__clz((int)v);
      const uint32_t pos = 32u * r + z;
      if (idx >= 0) sig[idx] = (uint8_t)(pos - prev - 1u);
      ++idx;
      prev = pos;
      v &= ~(0x80000000u >> z);
    }
  }
  if (idx >= 0 && idx < (int)kAdmChunk) sig[idx] = (uint8_t)(8u * m - prev - 1u);
  uint32_t o[8];
#pragma unroll
  for (int i = 0; i < (int)kAdmChunk; ++i) {
    const uint32_t code = (cw[i >> 2] >> (8 * (i & 3))) & 0xffu;
    const uint32_t d = (code >> 1) + kAdmMod * (uint32_t)sig[i];
    const uint32_t x = (code & 1u) ? (c - d) & 0xffffu : (c + d) & 0xffffu;
    if (i & 1) o[i >> 1] |= x << 16; else o[i >> 1] = x;
  }
  if (cnt == kAdmChunk) {
    uint4* op = (uint4*)out;
    op[0] = make_uint4(o[0], o[1], o[2], o[3]);
    op[1] = make_uint4(o[4], o[5], o[6], o[7]);
  } else {
    for (uint32_t i = 0; i < cnt; ++i) out[i] = (uint16_t)(o[i >> 1] >> (16 * (i & 1)));
  }
}

// Offsets o0 = O_g, o1 = O_{g+1} of a group are usable when the lane bytes
// they give fit the lane buffer and the bit plane.
__device__ __forceinline__ bool adm_offsets_ok(int32_t o0, int32_t o1, uint64_t bits_len) {
  return o0 >= 0 && o1 >= o0 && o1 - o0 <= (int32_t)kAdmLaneBytes && 32ull * (uint64_t)o1 <= bits_len;
}

// Lane li of the frame (elements [16 li, 16 li + cnt), group li / 32), all
// inputs from global memory.  Inconsistent offsets raise kErrBlockWords and
// store nothing.
__device__ inline void adm_restore_lane(const AdmUnmapArgs& a, uint64_t li, uint32_t cnt, const uint32_t (&cw)[4],
                                        uint8_t* sig) {
  const uint64_t g = li / kAdmLanes;
  const uint32_t lin = (uint32_t)(li % kAdmLanes);
  const int32_t o0 = a.offsets[g], o1 = a.offsets[g + 1];
  if (!adm_offsets_ok(o0, o1, a.bits_len)) {
    atomicOr(a.err, kErrBlockWords);
    return;
  }
  const uint32_t m = (uint32_t)(o1 - o0);
  adm_restore_core(m, a.bits + 32ull * (uint64_t)o0 + (uint64_t)lin * m, a.centers[g], cnt, cw, sig,
                   a.out + li * kAdmChunk);
}

}  // namespace mans
