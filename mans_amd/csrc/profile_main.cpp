// profile_main.cpp -- `mans_profile`: a host program with no Python or torch
// in it, for hardware-counter runs (rocprofv3 --pmc) of the frame codec.
//
//   mans_profile <codes.u2> [--steps K] [--lanes 32|64] [--prob-bits P]
//                [--elem-bytes 1|2|4] [--mode planes|adm] [--no-copy]
//
// Reads the element file into device memory, runs K frame encodes and K frame
// decodes on one stream (checking the round trip once), then one plain device
// copy of the input (a kernel of known traffic: it reads and writes exactly
// the input size), used to scale the FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/mans.h"

#define CHECK_HIP(x)                                                        \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "HIP error %s at %s\n", hipGetErrorString(e_), #x);   \
      return 2;                                                             \
    }                                                                       \
  } while (0)

__global__ void copy_u4(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr,
            "usage: %s <codes> [--steps K] [--lanes L] [--prob-bits P] [--elem-bytes B] [--mode planes|adm] "
            "[--no-copy]\n",
            argv[0]);
    return 1;
  }
  int steps = 3, lanes = 64, pb = 10, eb = 2;
  int32_t mode = MANS_MODE_PLANES;
  bool copy = true;
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--steps" && i + 1 < argc) steps = atoi(argv[++i]);
    else if (a == "--lanes" && i + 1 < argc) lanes = atoi(argv[++i]);
    else if (a == "--prob-bits" && i + 1 < argc) pb = atoi(argv[++i]);
    else if (a == "--elem-bytes" && i + 1 < argc) eb = atoi(argv[++i]);
    else if (a == "--mode" && i + 1 < argc) {
      const std::string m = argv[++i];
      if (m == "adm") mode = MANS_MODE_ADM;
      else if (m != "planes") {
        fprintf(stderr, "unknown mode %s\n", m.c_str());
        return 1;
      }
    } else if (a == "--no-copy") copy = false;
  }
  std::ifstream f(argv[1], std::ios::binary | std::ios::ate);
  if (!f) {
    fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  const size_t bytes = (size_t)f.tellg();
  f.seekg(0);
  std::vector<uint8_t> host(bytes);
  f.read((char*)host.data(), (std::streamsize)bytes);
  const uint64_t n = bytes / (size_t)eb;
  mans_params p = {pb, lanes, eb, mode};
  mans_workspace* ws = nullptr;
  if (mans_workspace_create(&ws, 0) != MANS_OK) return 2;
  mans_workspace_reserve(ws, bytes, &p);
  const uint64_t cap = mans_frame_max_size(n, &p);
  uint8_t *d_in = nullptr, *d_enc = nullptr, *d_dec = nullptr;
  uint64_t* d_size = nullptr;
  CHECK_HIP(hipMalloc(&d_in, bytes));
  CHECK_HIP(hipMalloc(&d_dec, bytes));
  CHECK_HIP(hipMalloc(&d_enc, cap));
  CHECK_HIP(hipMalloc(&d_size, 8));
  CHECK_HIP(hipMemcpy(d_in, host.data(), bytes, hipMemcpyHostToDevice));
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));
  uint64_t csize = 0;
  mans_status st = mans_frame_encode(&p, d_in, n, d_enc, cap, d_size, &csize, ws, s);
  if (st != MANS_OK) {
    fprintf(stderr, "encode: %s\n", mans_status_string(st));
    return 2;
  }
  std::vector<uint8_t> header(128);
  CHECK_HIP(hipMemcpy(header.data(), d_enc, 128, hipMemcpyDeviceToHost));
  // one untimed decode, then K encodes and K decodes with per-launch events
  mans_frame_decode(d_enc, csize, header.data(), d_dec, bytes, nullptr, ws, s);
  CHECK_HIP(hipStreamSynchronize(s));
  mans_set_profiling(ws, 1);
  mans_last_kernel_times(ws, nullptr, nullptr, 0);
  for (int k = 0; k < steps; ++k) mans_frame_encode(&p, d_in, n, d_enc, cap, d_size, nullptr, ws, s);
  for (int k = 0; k < steps; ++k) mans_frame_decode(d_enc, csize, header.data(), d_dec, bytes, nullptr, ws, s);
  CHECK_HIP(hipStreamSynchronize(s));
  {
    std::vector<double> ms(4096);
    std::vector<const char*> names(4096);
    const int got = mans_last_kernel_times(ws, ms.data(), names.data(), (int)ms.size());
    mans_set_profiling(ws, 0);
    std::vector<std::string> order;
    std::vector<double> sum;
    std::vector<int> cnt;
    for (int i = 0; i < got; ++i) {
      size_t j = 0;
      while (j < order.size() && order[j] != names[i]) ++j;
      if (j == order.size()) {
        order.push_back(names[i]);
        sum.push_back(0.0);
        cnt.push_back(0);
      }
      sum[j] += ms[i];
      cnt[j] += 1;
    }
    for (size_t j = 0; j < order.size(); ++j)
      printf("kernel %-16s launches %3d avg_ms %.4f\n", order[j].c_str(), cnt[j], sum[j] / cnt[j]);
  }
  if (mans_workspace_check(ws, s, nullptr) != MANS_OK) {
    fprintf(stderr, "decode flagged a corrupt stream\n");
    return 3;
  }
  std::vector<uint8_t> back(bytes);
  CHECK_HIP(hipMemcpy(back.data(), d_dec, bytes, hipMemcpyDeviceToHost));
  const bool ok = memcmp(back.data(), host.data(), bytes) == 0;
  if (copy) {
    hipLaunchKernelGGL(copy_u4, dim3(4096), dim3(256), 0, s, (const uint4*)d_in, (uint4*)d_dec, bytes / 16);
    CHECK_HIP(hipStreamSynchronize(s));
  }
  printf("elements %llu bytes %zu compressed %llu ratio %.4f steps %d round_trip %s copy_bytes %zu\n",
         (unsigned long long)n, bytes, (unsigned long long)csize, (double)bytes / (double)csize, steps,
         ok ? "bit-exact" : "MISMATCH", copy ? bytes / 16 * 16 : (size_t)0);
  mans_workspace_destroy(ws);
  return ok ? 0 : 4;
}
