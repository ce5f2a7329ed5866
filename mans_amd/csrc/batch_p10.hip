// batch_p10.hip -- the batched kernels at probability precision 10, every
// launch shape.
#include "batch_kernels.hpp"

namespace mans {
hipError_t launch_batch_encode_p10(const BatchEncodeArgs& a, BatchShape sh, int cu, hipStream_t s,
                                   uint64_t* need_spill) {
  return launch_enc<10>(a, sh, cu, s, need_spill);
}
hipError_t launch_batch_decode_p10(const BatchDecodeArgs& a, BatchShape sh, int cu, hipStream_t s) {
  return launch_dec<10>(a, sh, cu, s);
}
}  // namespace mans
