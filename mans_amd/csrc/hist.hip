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
      int g = __syncthreads_count(q > 1);
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

// Per-thread run-length merging: a thread keeps (current byte, run length)
// per plane and adds the run to LDS only when the byte changes.  Planes that
// are nearly one value (high bytes of quantisation codes, ADM bit planes)
// would otherwise send all 64 lanes of a wavefront to one LDS bin per step;
// uniform planes pay two extra VALU ops per byte.
template <int S>
struct Runs {
  uint32_t cur[S];
  uint32_t cnt[S];
};

template <int S>
__device__ __forceinline__ void run_add(uint32_t* h, Runs<S>& r, int p, uint32_t b) {
  const bool flush = b != r.cur[p];
  if (flush) atomicAdd(h + (p << 8) + r.cur[p], r.cnt[p]);  // adds 0 for the initial empty run
  r.cnt[p] = flush ? 1u : r.cnt[p] + 1u;
  r.cur[p] = b;
}

template <int S>
__device__ __forceinline__ void count_word(uint32_t* h, Runs<S>& r, uint32_t v) {
  // byte b of a 4-byte word belongs to plane b % S (elements are S-aligned)
  run_add<S>(h, r, 0 % S, v & 0xffu);
  run_add<S>(h, r, 1 % S, (v >> 8) & 0xffu);
  run_add<S>(h, r, 2 % S, (v >> 16) & 0xffu);
  run_add<S>(h, r, 3 % S, v >> 24);
}

template <int S>
__device__ void count_source(const HistSource& src, uint32_t* h /* wave's [P][256] + plane0 */) {
  Runs<S> r;
#pragma unroll
  for (int p = 0; p < S; ++p) {
    r.cur[p] = 0;
    r.cnt[p] = 0;
  }
  const uint8_t* p = src.ptr;
  const uint64_t n = src.nbytes;
  const uint32_t mis = (uint32_t)((uintptr_t)p & 15u);
  const uint64_t head = mis ? ((16u - mis) < n ? (16u - mis) : n) : 0;
  const uint64_t n4 = (n - head) / 16u;
  const uint64_t tail0 = head + 16u * n4;
  const uint4* body = (const uint4*)(p + head);
  const uint64_t tid = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint64_t nthreads = (uint64_t)gridDim.x * kThreads;
  uint64_t i = tid;
  for (; i + nthreads < n4; i += 2 * nthreads) {
    uint4 a = body[i];
    uint4 b = body[i + nthreads];
    count_word<S>(h, r, a.x); count_word<S>(h, r, a.y); count_word<S>(h, r, a.z); count_word<S>(h, r, a.w);
    count_word<S>(h, r, b.x); count_word<S>(h, r, b.y); count_word<S>(h, r, b.z); count_word<S>(h, r, b.w);
  }
  if (i < n4) {
    uint4 a = body[i];
    count_word<S>(h, r, a.x); count_word<S>(h, r, a.y); count_word<S>(h, r, a.z); count_word<S>(h, r, a.w);
  }
#pragma unroll
  for (int p = 0; p < S; ++p)
    if (r.cnt[p]) atomicAdd(h + (p << 8) + r.cur[p], r.cnt[p]);
  if (blockIdx.x == 0) {
    // unaligned head (< 16 bytes) and tail (< 16 bytes); offsets from the
    // element-aligned source start give the plane
    if (threadIdx.x < head) atomicAdd(h + ((threadIdx.x % S) << 8) + p[threadIdx.x], 1u);
    uint64_t j = tail0 + threadIdx.x;
    if (threadIdx.x < 16 && j < n) atomicAdd(h + ((j % S) << 8) + p[j], 1u);
  }
}

__global__ __launch_bounds__(kThreads) void hist_kernel(HistArgs a) {
  __shared__ uint32_t s_hist[kWaves][kMaxPlanes][kNumSymbols];
  __shared__ uint32_t s_keys[kNumSymbols];
  __shared__ uint32_t s_tmp[kWaves];
  __shared__ int s_last;
  const int t = threadIdx.x;
  const int P = a.num_planes;

  // Clear the encode launch's look-back words (the previous user of this
  // workspace has finished: kernels on one stream run in order).
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + t; i < a.status_count;
       i += (uint64_t)gridDim.x * kThreads)
    a.status[i] = 0;

  for (int w = 0; w < kWaves; ++w)
    for (int k = 0; k < P; ++k) s_hist[w][k][t] = 0;
  __syncthreads();

  uint32_t* wh = &s_hist[t >> 6][0][0];
  for (int s = 0; s < a.nsrc; ++s) {
    const HistSource& src = a.src[s];
    uint32_t* h = wh + src.plane0 * kNumSymbols;
    if (src.stride == 1) count_source<1>(src, h);
    else if (src.stride == 2) count_source<2>(src, h);
    else count_source<4>(src, h);
  }
  __syncthreads();
  for (int k = 0; k < P; ++k) {
    uint32_t v = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) v += s_hist[w][k][t];
    if (v) atomicAdd(a.hist + k * kNumSymbols + t, v);
  }
  __threadfence();
  __syncthreads();
  if (t == 0) s_last = atomicAdd(a.done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
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
  for (int i = 0; i < a.nsrc; ++i) n4 += a.src[i].nbytes / 16u;
  uint64_t want = div_up(n4, 2ull * kThreads);
  uint64_t cap = (uint64_t)cu * 8u;
  uint32_t grid = (uint32_t)(want < 1 ? 1 : (want > cap ? cap : want));
  hipLaunchKernelGGL(hist_kernel, dim3(grid), dim3(kThreads), 0, s, a);
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
