// batch_p11.hip -- the batched kernels at probability precision 11, every
// launch shape.
#include "batch_kernels.hpp"

namespace mans {
hipError_t launch_batch_encode_p11(const BatchEncodeArgs& a, BatchShape sh, int cu, hipStream_t s,
                                   uint64_t* need_spill) {
  return launch_enc<11>(a, sh, cu, s, need_spill);
}
hipError_t launch_batch_decode_p11(const BatchDecodeArgs& a, BatchShape sh, int cu, hipStream_t s) {
  return launch_dec<11>(a, sh, cu, s);
}
}  // namespace mans
