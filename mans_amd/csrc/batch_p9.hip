// batch_p9.hip -- the batched kernels at probability precision 9, every
// launch shape.
#include "batch_kernels.hpp"

namespace mans {
hipError_t launch_batch_encode_p9(const BatchEncodeArgs& a, BatchShape sh, int cu, hipStream_t s,
                                   uint64_t* need_spill) {
  return launch_enc<9>(a, sh, cu, s, need_spill);
}
hipError_t launch_batch_decode_p9(const BatchDecodeArgs& a, BatchShape sh, int cu, hipStream_t s) {
  return launch_dec<9>(a, sh, cu, s);
}
}  // namespace mans
