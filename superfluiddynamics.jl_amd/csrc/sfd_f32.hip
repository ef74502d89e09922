// sfd_f32.hip -- the precision-dependent routines of libsfdhip.so for
// complex FP32 fields and FP32 arithmetic (SFD_F32).
#include "sfd_impl.hpp"

namespace sfd {
const PrecOps kOpsF32 = make_ops<PF32>();
}  // namespace sfd
