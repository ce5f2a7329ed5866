/*
 * gpu_ans.h -- reference-compatible C++ API of libmans.so.
 *
 * Same names, argument order and behaviour as the reference's library ABI
 * (src/GpuANS.cu:16-44, include/ans/GpuANSCodec.h:237-246), with
 * hipStream_t in place of cudaStream_t.  Implemented over the C ABI in
 * mans.h with one cached workspace per (device, stream):
 *   - all data pointers are device pointers; *outSize is written by the
 *     device (uint32), as in the reference;
 *   - an unsupported precision prints "unhandled pdf precision N" and runs
 *     nothing (GpuANSEncode.h:358-359, GpuANSDecode.h:443-444);
 *   - the output is the reference container (32 lanes, 4 KiB blocks), with
 *     the bytes the reference leaves uninitialised written as zero.
 */
#ifndef MANS_GPU_ANS_H
#define MANS_GPU_ANS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace multibyte_ans {

void ansEncodeBatch(int precision, uint8_t* in, uint32_t inSize, uint8_t* out, uint32_t* outSize,
                    hipStream_t stream);

void ansDecodeBatch(int precision, uint8_t* in, uint8_t* out, hipStream_t stream);

/* At least the reference's value (GpuANSCodec.h:237-246) and never below the
 * true worst case of the format. */
uint32_t getMaxCompressedSize(uint32_t uncompressedBytes);

}  // namespace multibyte_ans

#endif
