// hist.hip -- per-byte-plane histograms with the probability normalisation
// fused into the last workgroup to finish.
//
// Reference: include/ans/GpuANSStatistics.h:23-151 (histogram),
// :186-406 (normalisation), :409-468 (launches).  The reference runs
// histogram -> quantizeWeights as two launches on one byte stream; here one
// launch counts every plane of the input (byte k of each element -> plane k),
// and the last workgroup (device-scope done counter) normalises each plane,
// writes the tables and the container's plane headers, and resets the
// accumulators so the next call needs no memset.
#include "mans_internal.hpp"

namespace mans {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over the 256 threads of the block (all threads get the result).
__device__ uint32_t block_sum(uint32_t v, uint32_t* s_tmp) {
  const int t = threadIdx.x;
  v = wave_sum(v);
  __syncthreads();
  if ((t & 63) == 0) s_tmp[t >> 6] = v;
  __syncthreads();
  uint32_t r = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) r += s_tmp[w];
  return r;
}

// Exclusive prefix sum over the 256 threads.
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  __syncthreads();
  if (lane == 63) s_tmp[w] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (int j = 0; j < w; ++j) base += s_tmp[j];
  return base + inc - v;
}

// Normalise one plane; thread t owns symbol t.  Restates
// GpuANSStatistics.h:186-385 exactly:
//  * q = (u32)(float(2^p) * (float(count) / float(total))), IEEE round to
//    nearest on both operations, truncating conversion; present symbols get
//    at least 1;
//  * rank = position of (q << 16 | s) in descending order;
//  * sum below 2^p: q[s] += diff / 256 + (s < diff % 256) (the reference's
//    repeated "+1 to symbol ids < min(diff, 256)" loop);
//  * sum above 2^p: repeatedly take G = #{q > 1}, it = min(diff, G) and
//    subtract one at the ranks [G - it, G) of the original sorted order;
//  * cdf = exclusive scan in symbol order; shift = 32 - clz(pdf - 1);
//    magic = 2^32 (2^shift - pdf) / pdf + 1 (0 for pdf == 0).
__device__ uint4 normalize_symbol(uint32_t count, uint32_t total, int p, uint32_t* s_keys,
                                  uint32_t* s_tmp, bool& bad) {
  const int t = threadIdx.x;
  if (total == 0) return make_uint4(0, 0, 0, 0);  // :202-204 (left unwritten there)
  const uint32_t W = 1u << p;
  float ratio = __fdiv_rn((float)count, (float)total);
  float prod = __fmul_rn((float)W, ratio);
  uint32_t q = (uint32_t)prod;
  if (count > 0 && q == 0) q = 1;
  const uint32_t qsum = block_sum(q, s_tmp);
  const uint32_t key = (q << 16) | (uint32_t)t;
  s_keys[t] = key;
  __syncthreads();
  int rank = 0;
  for (int j = 0; j < (int)kNumSymbols; ++j) rank += s_keys[j] > key;
  int diff = (int)W - (int)qsum;
  if (diff > 0) {
    q += (uint32_t)(diff >> 8) + ((t < (diff & 255)) ? 1u : 0u);
  } else if (diff < 0) {
    int d = -diff;
    while (d > 0) {
      int g = (int)block_sum((uint32_t)__popcll(__ballot(q > 1)) * ((t & 63) == 0 ? 1u : 0u), s_tmp);
      int it = d < g ? d : g;
      if (it <= 0) { bad = true; break; }
      if (rank >= g - it && rank < g) q -= 1;
      d -= it;
    }
  }
  const uint32_t pdf = q;
  const uint32_t cdf = block_exclusive_scan(pdf, s_tmp);
  uint32_t shift = 0, magic = 0;
  if (pdf > 0) {
    shift = 32u - (uint32_t)__clz((int)(pdf - 1u));
    uint64_t m = ((1ull << 32) * ((1ull << shift) - pdf)) / pdf + 1ull;
    magic = (uint32_t)m;
  }
  return make_uint4(pdf, cdf, magic, shift);
}

// Counting.  The workgroup shares one set of counters, [P][256 bins][R]
// 32-bit words with R = 32 replicas for 1 or 2 planes and 16 for 3 or 4
// (64 KiB), thread t adding into replica t % R: with R = 32 the 32 lanes of
// an LDS pass hit 32 different banks whatever the bytes, so a skewed plane
// (one or two byte values) never serialises on a word.  Every thread loads
// bytes no other thread loads and counts all of their planes (byte b of an
// element-aligned uint4 is plane plane0 + b % S).  (Measured on 1 GiB of
// uint16 codes: 0.30 ms; a wave per plane with 16 private replicas each,
// both waves loading the same bytes, 0.33 ms; 32 KiB of 16 / 8 replicas at
// five workgroups per CU, 0.43 ms.)
constexpr int kCountWords = 16384;
constexpr int kLoads = 8;                                  // 16-byte loads in flight per thread per buffer
constexpr uint32_t kRowU4 = (uint32_t)kThreads * kLoads;  // uint4s per row (32 KiB)
static_assert(4 * 256 * 16 <= kCountWords && 2 * 256 * 32 <= kCountWords, "counters fit");

// c: this thread's replica of plane plane0 (plane plane0 + j at c + 256 R j)
template <int S, int R>
__device__ __forceinline__ void count_u4(uint32_t* c, const uint4& v) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int m = 0; m < 4; ++m)
      atomicAdd(c + (uint32_t)((4 * k + m) % S) * 256u * R + __builtin_amdgcn_ubfe(w[k], 8u * (uint32_t)m, 8u) * R, 1u);
}

// Whole rows [r0, r1) of a source, counted by the workgroup: row r is the
// kRowU4 uint4s from r * kRowU4, thread t loading uint4s t, t + kThreads, ...
// of it.  Two register buffers, so one row loads while the other is
// counted; row indices are uniform (scalar row base, one thread offset for
// every load).
template <int S, int R>
__device__ __forceinline__ void count_rows(const uint4* body, uint32_t r0, uint32_t r1, uint32_t* c) {
  const uint32_t t = threadIdx.x;
  uint4 bufA[kLoads], bufB[kLoads];
  // (32-bit row indices: their compares stay scalar)
  auto load = [&](uint4 (&buf)[kLoads], uint32_t r) {
    const uint4* rp = body + (uint64_t)r * kRowU4;
#pragma unroll
    for (int k = 0; k < kLoads; ++k) buf[k] = rp[t + (uint32_t)kThreads * (uint32_t)k];
    // issue the row's loads here, ahead of the counting that follows (the
    // scheduler would otherwise move counting of the row in flight above them)
    __builtin_amdgcn_sched_barrier(0);
  };
  auto count = [&](const uint4 (&buf)[kLoads]) {
#pragma unroll
    for (int k = 0; k < kLoads; ++k) count_u4<S, R>(c, buf[k]);
  };
  // (the prefetch past the last row re-reads that row: straight-line code,
  // so the compiler keeps it in flight while the current row is counted)
  uint32_t r = r0;
  load(bufA, r);
  for (;;) {
    load(bufB, r + 1 < r1 ? r + 1 : r);
    count(bufA);
    if (++r >= r1) break;
    load(bufA, r + 1 < r1 ? r + 1 : r);
    count(bufB);
    if (++r >= r1) break;
  }
}

// Rows per grab.  Workgroups take their rows in chunks from a device-wide
// counter instead of a fixed share: workgroups on different CUs stream at
// different rates, so a static split leaves the launch waiting on its
// slowest workgroup.
constexpr uint32_t kChunkRows = 4;

// One source: the workgroup grabs a chunk, counts it, and takes the next
// grab while the current chunk is counted (s_grab double-buffered, one
// barrier per chunk).  The edges (the uint4s after the whole rows, the
// unaligned head and tail bytes) are counted by workgroup 0.  c: this
// thread's replica of plane plane0.
template <int S, int R>
__device__ __forceinline__ void count_source(const HistSource& src, uint32_t* c, bool edges, uint32_t* grab,
                             uint32_t* s_grab) {
  const uint8_t* p = src.ptr;
  const uint64_t n = src.nbytes;
  const uint32_t mis = (uint32_t)((uintptr_t)p & 15u);
  const uint64_t head = mis ? ((16u - mis) < n ? (16u - mis) : n) : 0;
  const uint64_t n4 = (n - head) / 16u;
  const uint64_t tail0 = head + 16u * n4;
  const uint4* body = (const uint4*)(p + head);
  const uint32_t t = threadIdx.x;
  const uint64_t nrows = n4 / kRowU4;  // whole rows; the rest with the edges
  const uint64_t nchunks = div_up(nrows, (uint64_t)kChunkRows);
  if (t == 0) s_grab[0] = atomicAdd(grab, 1u);
  __syncthreads();
  for (uint32_t k = 0;; ++k) {
    const uint32_t cb = s_grab[k & 1u];
    if (cb >= nchunks) break;
    if (t == 0) s_grab[(k + 1) & 1u] = atomicAdd(grab, 1u);
    const uint64_t r1 = ((uint64_t)cb + 1) * kChunkRows;
    count_rows<S, R>(body, cb * kChunkRows, (uint32_t)(r1 < nrows ? r1 : nrows), c);
    __syncthreads();
  }
  // (every thread has read the last grab before s_grab is reused)
  __syncthreads();
  if (edges) {
    for (uint64_t i = nrows * kRowU4 + t; i < n4; i += kThreads) count_u4<S, R>(c, body[i]);
    // head and tail bytes: plane plane0 + (offset from the element-aligned
    // source start) % S
    if (t < head) atomicAdd(c + (t % S) * 256u * R + p[t] * R, 1u);
    const uint64_t j = tail0 + t;
    if (t < 16 && j < n) atomicAdd(c + (uint32_t)(j % S) * 256u * R + p[j] * R, 1u);
  }
}

template <int R>
__device__ __forceinline__ void count_all(const HistArgs& a, uint32_t* cnt, uint32_t* s_grab) {
  const uint32_t t = threadIdx.x;
  for (int s = 0; s < a.nsrc; ++s) {
    const HistSource& src = a.src[s];
    const bool edges = blockIdx.x == 0;
    uint32_t* const c = cnt + (uint32_t)src.plane0 * 256u * R + t % R;
    if (src.stride == 1) count_source<1, R>(src, c, edges, a.grab + s, s_grab);
    else if (src.stride == 2) count_source<2, R>(src, c, edges, a.grab + s, s_grab);
    else count_source<4, R>(src, c, edges, a.grab + s, s_grab);
  }
  __syncthreads();
  // merge: thread t sums bin t (t + kThreads, ...) of the P x 256, its
  // replicas read from a rotated start so a pass spreads over the banks
  for (uint32_t i = t; i < (uint32_t)a.num_planes * kNumSymbols; i += kThreads) {
    const uint32_t* row = cnt + i * R;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)R; ++q) v += row[(q + i) % R];
    if (v) atomicAdd(a.hist + i, v);
  }
}

__global__ __launch_bounds__(kThreads) void hist_kernel(HistArgs a) {
  // 64 KiB of counters; the finisher reuses the front
  extern __shared__ uint32_t smem[];
  uint32_t* s_keys = smem;                 // [256]  (finisher)
  uint32_t* s_tmp = smem + kNumSymbols;    // [kWaves]
  __shared__ int s_last_flag;
  int* s_last = &s_last_flag;
  __shared__ uint32_t s_grab[2];
  const int t = threadIdx.x;
  const int P = a.num_planes;

  // Clear the encode launch's look-back words (the previous user of this
  // workspace has finished: kernels on one stream run in order).
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + t; i < a.status_count;
       i += (uint64_t)gridDim.x * kThreads)
    a.status[i] = 0;

  for (int i = t; i < kCountWords / 4; i += kThreads) ((uint4*)smem)[i] = make_uint4(0, 0, 0, 0);
  __syncthreads();
  if (P <= 2) count_all<32>(a, smem, s_grab);
  else count_all<16>(a, smem, s_grab);
  __threadfence();
  __syncthreads();
  if (t == 0) *s_last = atomicAdd(a.done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!*s_last) return;
  __threadfence();

  // ---- finisher: normalise each plane, write tables and headers ----------
  bool bad = false;
  for (int k = 0; k < P; ++k) {
    uint32_t c = __hip_atomic_load(a.hist + k * kNumSymbols + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a.hist[k * kNumSymbols + t] = 0;
    if (a.hist_out) a.hist_out[k * kNumSymbols + t] = c;
    uint4 e = normalize_symbol(c, (uint32_t)a.plane_nsym[k], a.prob_bits, s_keys, s_tmp, bad);
    a.table[k * kNumSymbols + t] = e;
    if (a.write_headers) {
      uint8_t* h = a.out + a.plane_section[k];
      ((uint16_t*)(h + kAnsHeaderBytes))[t] = (uint16_t)e.x;
      if (t < 8) {
        uint32_t f = 0;
        if (t == 0) f = (kAnsMagic << 16) | a.version;
        if (t == 1) f = (uint32_t)a.plane_nb[k];
        if (t == 2) f = (uint32_t)a.plane_nsym[k];
        if (t == 4) f = (uint32_t)a.prob_bits;
        ((uint32_t*)h)[t] = f;  // totalCompressedWords (t == 3) accumulates in encode
      }
      const uint64_t nb = a.plane_nb[k];
      if ((nb & 1) && t < 2) {
        uint32_t lanes = a.version == 2 ? 64u : 32u;
        ((uint32_t*)(h + kSectionFixed + 4ull * lanes * nb + 8ull * nb))[t] = 0;
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    *a.done = 0;
    a.grab[0] = 0;
    a.grab[1] = 0;
    if (a.ticket) *a.ticket = 0;
  }
  if (a.write_headers && a.has_frame && t < 32) ((uint32_t*)a.out)[t] = a.frame_hdr[t];
  if (a.total_units == 0 && t == 0) {
    if (a.d_out_size32) *a.d_out_size32 = (uint32_t)a.empty_size;
    if (a.d_out_size64) *a.d_out_size64 = a.empty_size;
  }
  (void)bad;
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

hipError_t launch_hist(const HistArgs& a, int cu, hipStream_t s) {
  uint64_t n4 = 0;
  for (int i = 0; i < a.nsrc; ++i) n4 = n4 > a.src[i].nbytes / 16u ? n4 : a.src[i].nbytes / 16u;
  const size_t lds = (size_t)kCountWords * sizeof(uint32_t);
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, hist_kernel, kThreads, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  const uint64_t want = div_up(n4, (uint64_t)kRowU4 * kChunkRows);
  const uint64_t cap = (uint64_t)cu * (uint64_t)per_cu;
  const uint32_t grid = (uint32_t)(want < 1 ? 1 : (want > cap ? cap : want));
  hipLaunchKernelGGL(hist_kernel, dim3(grid), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t grid = (uint32_t)(div_up(n, 256) < 1024 ? div_up(n, 256) : 1024);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t grid = (uint32_t)(div_up(n, 256) < 1024 ? div_up(n, 256) : 1024);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace mans
