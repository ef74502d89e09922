// sfd_f64.hip -- the precision-dependent routines of libsfdhip.so for
// complex FP64 fields and arithmetic (SFD_F64).
#include "sfd_impl.hpp"

namespace sfd {
const PrecOps kOpsF64 = make_ops<PF64>();
}  // namespace sfd
