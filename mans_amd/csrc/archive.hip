// archive.hip -- the multi-buffer container: the frames of a batch packed
// back to back behind one header and an index, and the header gather used
// when a batch decode takes element counts and widths from the frames.
//
// The reference has no container above one stream (GpuANSCodec.h:68-209 is
// single-buffer, and its ADM output is not self-describing,
// mapping/nv/mappingV15_fused_uint16.cu:933-951).  Layout (little-endian):
//   0   u32 magic "MANA", u8 version 1, u8[3] 0
//   8   u32 count, u32 0
//   16  u64 total size, u64 index offset (64), u64 data offset, u64[3] 0
//   64  count x {u64 offset, u64 frame bytes, u64 elements, u32 element
//       bytes, u32 0}
//   frame i at its offset (16-byte aligned), zero padding between frames.
#include "mans_internal.hpp"

namespace mans {
namespace {

constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kCopyChunk = 1u << 18;  // bytes of one frame one workgroup copies

// One workgroup: exclusive scan of the frames' 16-byte-rounded sizes in
// chunks of 1024 frames, writing the index entries and the header.
__global__ __launch_bounds__(kScanThreads) void archive_index_kernel(ArchiveArgs a) {
  __shared__ uint64_t s_part[kScanThreads / 64];
  __shared__ uint64_t s_base;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  if (t == 0) s_base = a.data_off;
  __syncthreads();
  for (uint32_t i0 = 0; i0 < a.count; i0 += kScanThreads) {
    const uint32_t i = i0 + t;
    const uint64_t sz = i < a.count ? a.sizes[i] : 0;
    const uint64_t r = round_up(sz, 16);
    uint64_t inc = r;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t u = __shfl_up(inc, o, 64);
      if (lane >= (uint32_t)o) inc += u;
    }
    if (lane == 63) s_part[wave] = inc;
    __syncthreads();
    uint64_t before = s_base;
    for (uint32_t w = 0; w < wave; ++w) before += s_part[w];
    const uint64_t off = before + inc - r;
    if (i < a.count) {
      uint64_t* e = (uint64_t*)(a.out + 64 + 32ull * i);
      e[0] = off;
      e[1] = sz;
      e[2] = a.bufs[i].n;
      e[3] = a.bufs[i].eb;
      a.offsets[i] = off;
    }
    __syncthreads();
    if (t == kScanThreads - 1) s_base = off + r;
    __syncthreads();
  }
  if (t < 8) {
    uint64_t v = 0;
    if (t == 0) v = (uint64_t)kArchiveMagic | (1ull << 32);
    if (t == 1) v = a.count;
    if (t == 2) v = s_base;
    if (t == 3) v = 64;
    if (t == 4) v = a.data_off;
    ((uint64_t*)a.out)[t] = v;
  }
  if (t == 0 && a.d_size) *a.d_size = s_base;
}

// Workgroup (x, y) copies chunk x of frame y to its offset, and zeroes the
// padding up to the next 16-byte boundary.
__global__ __launch_bounds__(256) void archive_copy_kernel(ArchiveArgs a) {
  const uint32_t f = blockIdx.y;
  const uint64_t sz = a.sizes[f];
  const uint64_t c0 = (uint64_t)blockIdx.x * kCopyChunk;
  const uint64_t r = round_up(sz, 16);
  if (c0 >= r) return;
  const uint64_t c1 = c0 + kCopyChunk < r ? c0 + kCopyChunk : r;
  const uint8_t* src = a.bufs[f].out;
  uint8_t* dst = a.out + a.offsets[f];
  // whole uint4s of the frame (frames start 16-byte aligned)
  const uint64_t full = sz / 16;
  for (uint64_t q = c0 / 16 + threadIdx.x; q < c1 / 16; q += 256) {
    uint4 v;
    if (q < full) {
      v = ((const uint4*)src)[q];
    } else {
      // the last, partial uint4: frame bytes then zeros
      uint8_t b[16];
      for (uint32_t k = 0; k < 16; ++k) b[k] = 16 * q + k < sz ? src[16 * q + k] : 0;
      v = *(const uint4*)b;
    }
    ((uint4*)dst)[q] = v;
  }
}

// Thread i copies the 128-byte header of frame i.
__global__ void gather_headers_kernel(const uint8_t* const* src, uint32_t count, uint4* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint4* h = (const uint4*)src[i];
#pragma unroll
  for (int k = 0; k < 8; ++k) out[8ull * i + k] = h[k];
}

}  // namespace

hipError_t launch_archive_pack(const ArchiveArgs& a, uint64_t max_frame, hipStream_t s) {
  hipLaunchKernelGGL(archive_index_kernel, dim3(1), dim3(kScanThreads), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || a.count == 0) return e;
  const uint64_t gx = div_up(round_up(max_frame, 16), kCopyChunk);
  hipLaunchKernelGGL(archive_copy_kernel, dim3((uint32_t)(gx < 1 ? 1 : gx), a.count), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_gather_headers(const uint8_t* const* src, uint32_t count, uint4* out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_headers_kernel, dim3((uint32_t)div_up(count, 256)), dim3(256), 0, s, src, count, out);
  return hipGetLastError();
}

}  // namespace mans
