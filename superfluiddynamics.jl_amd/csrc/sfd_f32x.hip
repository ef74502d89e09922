// sfd_f32x.hip -- the precision-dependent routines of libsfdhip.so for
// complex FP32 fields with FP64 arithmetic (SFD_F32X).
#include "sfd_impl.hpp"

namespace sfd {
const PrecOps kOpsF32X = make_ops<PF32X>();
}  // namespace sfd
