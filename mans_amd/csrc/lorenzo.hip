// lorenzo.hip -- dual-quantisation 3-D Lorenzo front end: the producer of
// uint16 quantisation codes for the codec, and its inverse.
//
// The reference has no such kernel; its mapping programs take the codes as
// given (mapping/cpu/mappingcpuV1.cpp reads a uint16 file).  This is the
// error-bounded step that makes them from a float32 field
// f[nz][ny][nx] with absolute error bound eb:
//   q = rint(f * (1 / (2 eb)))                    (prequantisation, float32)
//   L2(z, y, x) = q(z,y,x) - q(z,y-1,x) - q(z,y,x-1) + q(z,y-1,x-1)
//   r(z, y, x)  = L2(z, y, x) - L2(z-1, y, x)     (3-D Lorenzo residual; q = 0
//                                                 outside the field)
//   code = 32768 + r if |r| < radius, else 0 and (index, r) is an outlier.
// The inverse rebuilds q as the 3-D prefix sum of r and f' = q * (2 eb), so
// |f - f'| <= eb up to float rounding.
//
// Forward: a workgroup of 256 threads takes a tile of 256 x-columns x 16 rows
// x 16 planes (tiles taken by an atomic ticket, in ticket order).  A thread
// keeps its column's L2 of the previous plane for its 16 rows in registers,
// so each element's f is read once (plus one halo row and the plane below
// the tile), a plane's 17 rows loaded at once; q(z, y, x - 1) comes from the
// neighbouring lane.  Outliers are
// compacted in tile order (tile ticket, thread, plane, row) with a decoupled
// look-back over tiles; a thread holds up to kHold of them in registers, and
// a tile with more goes on a list, with each thread's offset inside the
// tile and its counts before planes 4, 8 and 12, for the fix-up launch to
// write (four plane groups of the tile at once).
#include "codec_device.hpp"

namespace mans {
namespace {

constexpr uint32_t kTX = 256, kTY = 16, kTZ = 16;
constexpr int kHold = kLorenzoHold;
constexpr int32_t kQLimit = 1 << 28;  // |q| bound: residuals stay inside int32

__device__ __forceinline__ int32_t quant(float f, float inv, bool& bad) {
  const float v = rintf(f * inv);
  if (!(fabsf(v) < (float)kQLimit)) {  // also NaN / inf
    bad = true;
    return 0;
  }
  return (int32_t)v;
}

struct TileGeom {
  uint32_t nxb, nyb, nzb;
};

__device__ __forceinline__ void tile_of(const LorenzoArgs& a, uint32_t t, uint32_t& x0, uint32_t& y0, uint32_t& z0) {
  const uint32_t nxb = (uint32_t)div_up(a.nx, kTX), nyb = (uint32_t)div_up(a.ny, kTY);
  x0 = (t % nxb) * kTX;
  y0 = ((t / nxb) % nyb) * kTY;
  z0 = (t / (nxb * nyb)) * kTZ;
}

// Codes of planes [dz0, dz1) of one tile (a thread per x-column: column
// threadIdx.x % 256); outliers go to `hold` (first kHold) and are counted in
// `cnt`, and `grp` gets the count before planes 4, 8 and 12 (a byte each;
// at most 192).  With WRITE_ALL, every outlier is written at a.out_idx /
// a.out_val from position `pos` on instead (the fix-up pass).  FULL: nx is a
// multiple of 256 and ny of 16, so every column and row of every tile is
// inside the field (only the row above the first row of tiles is not).
template <bool WRITE_ALL, bool FULL>
__device__ void code_tile(const LorenzoArgs& a, uint32_t x0, uint32_t y0, uint32_t z0, uint32_t dz0, uint32_t dz1,
                          uint32_t& cnt, uint32_t& grp, uint32_t (&hidx)[kHold], int32_t (&hval)[kHold], uint64_t pos,
                          bool& bad) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t x = x0 + threadIdx.x % kTX;
  const bool xin = FULL || x < a.nx;
  const uint64_t plane = (uint64_t)a.ny * a.nx;
  const float inv = a.inv2eb;
  const uint32_t ny1 = FULL ? kTY : (a.ny - y0 < kTY ? a.ny - y0 : kTY);
  const uint32_t nz0 = a.nz - z0 < kTZ ? a.nz - z0 : kTZ;
  const uint32_t nz1 = dz1 < nz0 ? dz1 : nz0;
  // q of plane z at rows y0 - 1 .. y0 + kTY - 1 (row j - 1 of the tile in
  // slot j; rows outside the field or the tile are 0) for this column and,
  // on lane 0, the column to its left (the other lanes take it from their
  // neighbour).  All of a plane's loads are issued before any is used.
  auto plane_q = [&](uint32_t z, int32_t (&qc)[kTY + 1], int32_t (&ql)[kTY + 1]) {
    float fc[kTY + 1], fl[kTY + 1];
#pragma unroll
    for (uint32_t j = 0; j <= kTY; ++j) {
      const int64_t y = (int64_t)y0 - 1 + (int64_t)j;
      const bool row = FULL ? (j > 0 || y0 > 0) : (y >= 0 && (j == 0 || j - 1 < ny1));
      const uint64_t r = row ? z * plane + (uint64_t)y * a.nx : 0u;
      fc[j] = row && xin ? a.field[r + x] : 0.0f;
      fl[j] = 0.0f;
      if (lane == 0 && row && x > 0 && (FULL || x - 1 < a.nx)) fl[j] = a.field[r + x - 1];
    }
#pragma unroll
    for (uint32_t j = 0; j <= kTY; ++j) {
      const int64_t y = (int64_t)y0 - 1 + (int64_t)j;
      const bool row = FULL ? (j > 0 || y0 > 0) : (y >= 0 && (j == 0 || j - 1 < ny1));
      qc[j] = row && xin ? quant(fc[j], inv, bad) : 0;
      int32_t nb = __shfl_up(qc[j], 1, 64);
      if (lane == 0) nb = row && x > 0 && (FULL || x - 1 < a.nx) ? quant(fl[j], inv, bad) : 0;
      ql[j] = nb;
    }
  };
  int32_t prev[kTY];
#pragma unroll
  for (uint32_t j = 0; j < kTY; ++j) prev[j] = 0;
  // the plane below the first one: its L2 per row
  if (z0 + dz0 > 0 && dz0 < nz1) {
    int32_t qc[kTY + 1], ql[kTY + 1];
    plane_q(z0 + dz0 - 1, qc, ql);
#pragma unroll
    for (uint32_t j = 0; j < kTY; ++j)
      if (j < ny1) prev[j] = qc[j + 1] - qc[j] - ql[j + 1] + ql[j];
  }
  for (uint32_t dz = dz0; dz < nz1; ++dz) {
    const uint32_t z = z0 + dz;
    if (!WRITE_ALL && dz > 0 && (dz & 3u) == 0) grp |= cnt << (8u * (dz / 4u - 1u));
    int32_t qc[kTY + 1], ql[kTY + 1];
    plane_q(z, qc, ql);
#pragma unroll
    for (uint32_t j = 0; j < kTY; ++j) {
      if (j < ny1) {
        const uint32_t y = y0 + j;
        const int32_t l2 = qc[j + 1] - qc[j] - ql[j + 1] + ql[j];
        const int32_t r = l2 - prev[j];
        prev[j] = l2;
        if (xin) {
          const uint64_t i = z * plane + (uint64_t)y * a.nx + x;
          const bool in = r > -(int32_t)a.radius && r < (int32_t)a.radius;
          if (!WRITE_ALL) a.codes[i] = in ? (uint16_t)(32768 + r) : (uint16_t)0;
          if (!in) {
            if (WRITE_ALL) {
              if (pos + cnt < a.out_cap) {
                a.out_idx[pos + cnt] = (uint32_t)i;
                a.out_val[pos + cnt] = r;
              }
            } else {
#pragma unroll
              for (int k = 0; k < kHold; ++k)
                if ((uint32_t)k == cnt) {
                  hidx[k] = (uint32_t)i;
                  hval[k] = r;
                }
            }
            ++cnt;
          }
        }
      }
    }
  }
}

// Exclusive scan of v over the workgroup's 256 threads; *total gets the sum.
__device__ __forceinline__ uint32_t block_scan256(uint32_t v, uint32_t* s_w, uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(inc, o, 64);
    if (lane >= (uint32_t)o) inc += u;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
  for (uint32_t k = 0; k < kTX / 64; ++k) {
    if (k < w) before += s_w[k];
    total += s_w[k];
  }
  __syncthreads();
  return before + inc - v;
}

template <bool FULL>
__global__ __launch_bounds__(kTX) void lorenzo_kernel(LorenzoArgs a) {
  __shared__ uint32_t s_t, s_k, s_w[kTX / 64];
  __shared__ uint64_t s_off;
  __shared__ int s_bad;
  if (threadIdx.x == 0) {
    s_t = atomicAdd(a.ticket, 1u);
    s_bad = 0;
  }
  __syncthreads();
  const uint32_t t = s_t;
  uint32_t x0, y0, z0;
  tile_of(a, t, x0, y0, z0);
  uint32_t cnt = 0, grp = 0;
  uint32_t hidx[kHold];
  int32_t hval[kHold];
  bool bad = false;
  code_tile<false, FULL>(a, x0, y0, z0, 0, kTZ, cnt, grp, hidx, hval, 0, bad);
  if (bad) s_bad = 1;
  uint32_t total = 0;
  const uint32_t mine = block_scan256(cnt, s_w, total);
  // wave 0: the tile's offset among the outliers of earlier tiles
  if (threadIdx.x < 64) {
    const uint64_t excl = look_back(a.status, t, total, (int)threadIdx.x);
    if (threadIdx.x == 0) {
      s_off = excl;
      a.tile_off[t] = excl;
      if (t + 1 == a.ntiles) *a.count = excl + total;
    }
  }
  __syncthreads();
  const uint64_t pos = s_off + mine;
#pragma unroll
  for (int k = 0; k < kHold; ++k)
    if ((uint32_t)k < cnt && pos + k < a.out_cap) {
      a.out_idx[pos + k] = hidx[k];
      a.out_val[pos + k] = hval[k];
    }
  // a thread with more outliers than it holds: the tile goes on the fix-up
  // list with every thread's offset inside the tile, and the fix-up launch
  // writes the tile's outliers; outliers past the capacity: flagged
  const uint32_t over = __syncthreads_or(cnt > (uint32_t)kHold);
  if (over) {  // uniform over the workgroup
    if (threadIdx.x == 0) s_k = atomicAdd(a.fix_list, 1u);
    __syncthreads();
    const uint32_t k = s_k;
    if (k < a.fix_cap) {
      if (threadIdx.x == 0) a.fix_list[1 + k] = t;
      a.fix_pos[(uint64_t)k * kTX + threadIdx.x] = mine;
      a.fix_grp[(uint64_t)k * kTX + threadIdx.x] = grp;
    } else if (threadIdx.x == 0) {
      // more than fix_cap tiles over kHold: more than out_cap outliers
      atomicOr(a.err, kErrCapacity);
    }
  }
  if (threadIdx.x == 0) {
    if (s_off + total > a.out_cap) atomicOr(a.err, kErrCapacity);
    if (s_bad) atomicOr(a.err, kErrAdmDomain);
  }
}

// Writes every outlier of the tiles on the fix-up list.  A workgroup of
// kFixGroups x 256 threads takes a listed tile, thread (g, column) codes
// planes [4g, 4g + 4) of its column from the offsets the main launch left:
// the column's first outlier in the tile plus its count before plane 4g.
constexpr uint32_t kFixGroups = kTZ / 4;
__global__ __launch_bounds__(kTX * kFixGroups) void lorenzo_fixup_kernel(LorenzoArgs a) {
  const uint32_t nfix = a.fix_list[0] < a.fix_cap ? a.fix_list[0] : a.fix_cap;
  const uint32_t col = threadIdx.x % kTX, g = threadIdx.x / kTX;
  for (uint32_t k = blockIdx.x; k < nfix; k += gridDim.x) {
    const uint32_t t = a.fix_list[1 + k];
    uint32_t x0, y0, z0;
    tile_of(a, t, x0, y0, z0);
    uint32_t cnt = 0, grp = 0, dummy_i[kHold];
    int32_t dummy_v[kHold];
    bool bad = false;
    const uint64_t e = (uint64_t)k * kTX + col;
    const uint32_t before = g ? (a.fix_grp[e] >> (8u * (g - 1u))) & 0xffu : 0u;
    const uint64_t pos = a.tile_off[t] + a.fix_pos[e] + before;
    code_tile<true, false>(a, x0, y0, z0, 4u * g, 4u * g + 4u, cnt, grp, dummy_i, dummy_v, pos, bad);
  }
}

// ---- inverse ---------------------------------------------------------------
// q of outliers first (scattered), then per plane: x-prefix of each row plus
// the y-prefix (running per column), then per column: the z-prefix and
// f' = q * 2 eb.
__global__ void lorenzo_scatter_kernel(LorenzoArgs a) {
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n_out; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = a.in_idx[k];
    if ((uint64_t)i < (uint64_t)a.nx * a.ny * a.nz) a.q[i] = a.in_val[k];
    else atomicOr(a.err, kErrBlockWords);
  }
}

constexpr uint32_t kPlaneChunk = 1024;  // elements of a row a workgroup scans at a time
constexpr uint32_t kRowGroup = 8;       // rows of a plane loaded, scanned and synchronised together

// One workgroup per plane: per row, the x-prefix of r (wave scans, then the
// waves' totals through LDS), summed down each column (the y-prefix).  A
// thread takes E adjacent elements of a row (E = 2: one 4-byte load of two
// codes, one 8-byte store of two sums; needs an even row length and 4-byte
// aligned codes).  Rows go kRowGroup at a time: their loads are all issued
// before any is scanned, and one barrier serves the group (the wave totals
// alternate between two LDS buffers, so a buffer is rewritten only after the
// barrier that follows its last read).  Rows longer than kPlaneChunk go
// chunk by chunk with the row's running total carried between chunks.
template <uint32_t E>
__global__ __launch_bounds__(kPlaneChunk / E) __attribute__((amdgpu_waves_per_eu(8)))
void lorenzo_plane_kernel(LorenzoArgs a) {
  constexpr uint32_t kT = kPlaneChunk / E, kW = kT / 64;  // threads, waves
  constexpr uint32_t kMaxChunks = 8;                     // nx <= 8192
  static_assert(E == 1 || E == 2, "one or two elements a thread");
  __shared__ int32_t s_w[2][kRowGroup][kW];
  __shared__ int32_t s_col[kMaxChunks][kPlaneChunk];  // column sums of each chunk's elements
  const uint32_t z = blockIdx.x;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t wu = __builtin_amdgcn_readfirstlane(w);
  const uint32_t lw = lane & (kW - 1);
  const uint64_t plane = (uint64_t)a.ny * a.nx;
  const uint32_t nchunks = (uint32_t)div_up(a.nx, kPlaneChunk);
  for (uint32_t c = 0; c < nchunks; ++c)
#pragma unroll
    for (uint32_t e = 0; e < E; ++e) s_col[c][threadIdx.x * E + e] = 0;
  uint32_t it = 0;
  for (uint32_t y0 = 0; y0 < a.ny; y0 += kRowGroup) {
    const uint32_t nrow = a.ny - y0 < kRowGroup ? a.ny - y0 : kRowGroup;
    int32_t carry[kRowGroup];
#pragma unroll
    for (uint32_t g = 0; g < kRowGroup; ++g) carry[g] = 0;
#pragma unroll 1
    for (uint32_t c = 0; c < nchunks; ++c) {
      // this thread's first element; with E = 2 the row length is even, so
      // x < nx means x + 1 < nx
      const uint32_t x = c * kPlaneChunk + threadIdx.x * E;
      const bool xin = x < a.nx;
      // row g of the group starts at element row0 + g * nx (uniform; the
      // lane adds x)
      const uint64_t row0 = z * plane + (uint64_t)y0 * a.nx;
      // cells outside the field or the group read as code 32768 (r = 0)
      uint32_t code[kRowGroup];
#pragma unroll
      for (uint32_t g = 0; g < kRowGroup; ++g) {
        const uint16_t* p = a.codes + row0 + (uint64_t)g * a.nx + x;
        if (E == 1) code[g] = xin && g < nrow ? *p : 32768u;
        else code[g] = xin && g < nrow ? *(const uint32_t*)p : 0x80008000u;
      }
      int32_t v[kRowGroup][E], inc[kRowGroup];
#pragma unroll
      for (uint32_t g = 0; g < kRowGroup; ++g) {
        inc[g] = 0;
#pragma unroll
        for (uint32_t e = 0; e < E; ++e) {
          const uint32_t ce = (code[g] >> (16 * e)) & 0xFFFFu;
          v[g][e] = ce ? (int32_t)ce - 32768 : (a.q + row0 + (uint64_t)g * a.nx)[x + e];
          inc[g] += v[g][e];
        }
      }
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (uint32_t g = 0; g < kRowGroup; ++g) {
          const int32_t u = __shfl_up(inc[g], o, 64);
          if (lane >= (uint32_t)o) inc[g] += u;
        }
      }
      const uint32_t buf = it & 1u;
      ++it;
      if (lane == 63) {
#pragma unroll
        for (uint32_t g = 0; g < kRowGroup; ++g) s_w[buf][g][w] = inc[g];
      }
      __syncthreads();
      int32_t colsum[E];
#pragma unroll
      for (uint32_t e = 0; e < E; ++e) colsum[e] = s_col[c][threadIdx.x * E + e];
#pragma unroll
      for (uint32_t g = 0; g < kRowGroup; ++g) {
        // scan of the kW wave totals over each kW-lane group of the wave
        int32_t t = s_w[buf][g][lw];
#pragma unroll
        for (uint32_t o = 1; o < kW; o <<= 1) {
          const int32_t u = __shfl_up(t, o, kW);
          if (lw >= o) t += u;
        }
        // all waves' total and the total of the waves before this one
        // (uniform: read from one lane)
        const int32_t all = __builtin_amdgcn_readlane(t, kW - 1);
        const int32_t prior = wu ? __builtin_amdgcn_readlane(t, wu - 1) : 0;
        int32_t run = carry[g] + prior + inc[g];  // x-prefix through this thread's last element
#pragma unroll
        for (uint32_t e = 0; e < E; ++e) run -= v[g][e];
#pragma unroll
        for (uint32_t e = 0; e < E; ++e) {
          run += v[g][e];
          colsum[e] += run;  // x-prefix of this row, summed over rows
        }
        carry[g] += all;
        if (xin && g < nrow) {
          int32_t* q = a.q + row0 + (uint64_t)g * a.nx + x;
          if (E == 1) *q = colsum[0];
          else *(int2*)q = make_int2(colsum[0], colsum[E - 1]);
        }
      }
#pragma unroll
      for (uint32_t e = 0; e < E; ++e) s_col[c][threadIdx.x * E + e] = colsum[e];
    }
  }
}

constexpr uint32_t kColumnPlanes = 16;  // planes of a column loaded before any is summed

// Per (y, x) column: the z-prefix of the plane sums and f' = q * 2 eb.
__global__ __launch_bounds__(256) void lorenzo_column_kernel(LorenzoArgs a) {
  const uint64_t plane = (uint64_t)a.ny * a.nx;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < plane; j += (uint64_t)gridDim.x * blockDim.x) {
    int32_t s = 0;
    for (uint32_t z0 = 0; z0 < a.nz; z0 += kColumnPlanes) {
      int32_t v[kColumnPlanes];
#pragma unroll
      for (uint32_t k = 0; k < kColumnPlanes; ++k)
        v[k] = z0 + k < a.nz ? a.q[(uint64_t)(z0 + k) * plane + j] : 0;
#pragma unroll
      for (uint32_t k = 0; k < kColumnPlanes; ++k) {
        s += v[k];
        if (z0 + k < a.nz) a.recon[(uint64_t)(z0 + k) * plane + j] = (float)s * a.twoeb;
      }
    }
  }
}

}  // namespace

hipError_t launch_lorenzo(const LorenzoArgs& a, hipStream_t s) {
  if (a.ntiles == 0) return hipSuccess;
  if (a.nx % kTX == 0 && a.ny % kTY == 0)
    hipLaunchKernelGGL(lorenzo_kernel<true>, dim3(a.ntiles), dim3(kTX), 0, s, a);
  else
    hipLaunchKernelGGL(lorenzo_kernel<false>, dim3(a.ntiles), dim3(kTX), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the fix-up list is short (a tile is on it only with more than kHold
  // outliers in one thread): a small grid walks it
  const uint32_t g = a.fix_cap < 1024u ? a.fix_cap : 1024u;
  hipLaunchKernelGGL(lorenzo_fixup_kernel, dim3(g), dim3(kTX * kFixGroups), 0, s, a);
  return hipGetLastError();
}

uint32_t lorenzo_tiles(uint32_t nz, uint32_t ny, uint32_t nx) {
  return (uint32_t)(div_up(nx, kTX) * div_up(ny, kTY) * div_up(nz, kTZ));
}

hipError_t launch_lorenzo_inverse(const LorenzoArgs& a, int cu, hipStream_t s) {
  const uint64_t n = (uint64_t)a.nx * a.ny * a.nz;
  if (n == 0) return hipSuccess;
  if (a.n_out) {
    const uint32_t g = (uint32_t)std::min<uint64_t>(div_up(a.n_out, 256), 4096);
    hipLaunchKernelGGL(lorenzo_scatter_kernel, dim3(g), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // two elements a thread where rows are even and the codes 4-byte aligned
  // (the q scratch is a whole allocation)
  if (a.nx % 2 == 0 && ((uintptr_t)a.codes & 3u) == 0)
    hipLaunchKernelGGL(lorenzo_plane_kernel<2>, dim3(a.nz), dim3(kPlaneChunk / 2), 0, s, a);
  else
    hipLaunchKernelGGL(lorenzo_plane_kernel<1>, dim3(a.nz), dim3(kPlaneChunk), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t plane = (uint64_t)a.ny * a.nx;
  const uint32_t g = (uint32_t)std::min<uint64_t>(div_up(plane, 256), (uint64_t)cu * 16);
  hipLaunchKernelGGL(lorenzo_column_kernel, dim3(g), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace mans
