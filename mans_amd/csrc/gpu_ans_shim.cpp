// gpu_ans_shim.cpp -- the reference's multibyte_ans C++ API over the C ABI.
// Reference: src/GpuANS.cu:16-44, include/ans/GpuANSCodec.h:237-246.
#include "../../include/gpu_ans.h"

#include <iostream>
#include <map>
#include <mutex>
#include <utility>

#include "../../include/mans.h"

namespace multibyte_ans {
namespace {

std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, mans_workspace*> g_ws;

// One workspace per (device, stream): calls on one stream are ordered, so
// they may share scratch; different streams never do.
mans_workspace* workspace_for(hipStream_t stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(dev, stream);
  auto it = g_ws.find(key);
  if (it != g_ws.end()) return it->second;
  mans_workspace* ws = nullptr;
  if (mans_workspace_create(&ws, dev) != MANS_OK) return nullptr;
  g_ws[key] = ws;
  return ws;
}

bool precision_ok(int precision) {
  if (precision >= 9 && precision <= 11) return true;
  std::cout << "unhandled pdf precision " << precision << std::endl;
  return false;
}

}  // namespace

uint32_t getMaxCompressedSize(uint32_t uncompressedBytes) {
  // Reference formula (GpuANSCodec.h:237-246): overhead(4096 blocks) +
  // roundUp(4096 * 1.25, 16) per block, rounded to 16 bytes.
  const uint64_t blocks = (uncompressedBytes + 4095ull) / 4096ull;
  uint64_t ref = 32ull + 512ull + 128ull * 4096ull + 8ull * 4096ull;
  ref += 5120ull * blocks;
  ref = (ref + 15ull) / 16ull * 16ull;
  mans_params p = {11, 32, 1, MANS_MODE_PLANES};
  const uint64_t exact = mans_max_compressed_size(uncompressedBytes, &p);
  const uint64_t m = ref > exact ? ref : exact;
  return m > 0xffffffffull ? 0xffffffffu : (uint32_t)m;
}

void ansEncodeBatch(int precision, uint8_t* in, uint32_t inSize, uint8_t* out, uint32_t* outSize,
                    hipStream_t stream) {
  if (!precision_ok(precision)) return;
  mans_workspace* ws = workspace_for(stream);
  if (!ws) {
    std::cerr << "ansEncodeBatch: no workspace" << std::endl;
    return;
  }
  mans_params p = {precision, 32, 1, MANS_MODE_PLANES};
  const uint64_t cap = mans_max_compressed_size(inSize, &p);
  mans_status st = mans_encode(&p, in, inSize, out, cap, outSize, nullptr, ws, stream);
  if (st != MANS_OK) std::cerr << "ansEncodeBatch: " << mans_status_string(st) << std::endl;
}

void ansDecodeBatch(int precision, uint8_t* in, uint8_t* out, hipStream_t stream) {
  if (!precision_ok(precision)) return;
  mans_workspace* ws = workspace_for(stream);
  if (!ws) {
    std::cerr << "ansDecodeBatch: no workspace" << std::endl;
    return;
  }
  mans_params p = {precision, 32, 1, MANS_MODE_PLANES};
  mans_status st = mans_decode(&p, in, 0, out, 0, ws, stream);
  if (st != MANS_OK) std::cerr << "ansDecodeBatch: " << mans_status_string(st) << std::endl;
}

}  // namespace multibyte_ans
