// batch.hip -- precision dispatch of the batched kernels (batch_kernels.hpp,
// instantiated per precision in batch_p9.hip, batch_p10.hip, batch_p11.hip).
#include "mans_internal.hpp"

namespace mans {

bool batch_shape_ok(BatchShape sh, int pb, bool decode) {
  if (sh.nw != 4 && sh.nw != 8 && sh.nw != 16) return false;
  if (sh.kb != 4 && sh.kb != 8 && sh.kb != 16 && !(decode && sh.kb == 1)) return false;
  if (pb < 9 || pb > 11) return false;
  return (decode ? batch_decode_lds(sh, pb) : batch_encode_lds(sh)) <= kLdsBytes;
}

hipError_t launch_batch_encode(const BatchEncodeArgs& a, int pb, BatchShape sh, int cu, hipStream_t s,
                               uint64_t* need_spill) {
  if (!batch_shape_ok(sh, pb, false)) return hipErrorInvalidConfiguration;
  switch (pb) {
    case 9: return launch_batch_encode_p9(a, sh, cu, s, need_spill);
    case 10: return launch_batch_encode_p10(a, sh, cu, s, need_spill);
    case 11: return launch_batch_encode_p11(a, sh, cu, s, need_spill);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_decode(const BatchDecodeArgs& a, int pb, BatchShape sh, int cu, hipStream_t s) {
  if (!batch_shape_ok(sh, pb, true)) return hipErrorInvalidConfiguration;
  switch (pb) {
    case 9: return launch_batch_decode_p9(a, sh, cu, s);
    case 10: return launch_batch_decode_p10(a, sh, cu, s);
    case 11: return launch_batch_decode_p11(a, sh, cu, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace mans
