// step_fused4 instantiations (set B) for FP64 storage.
#include "lib_fused.h"

#pragma GCC visibility push(hidden)
namespace wff {
WF_FUSED4_F64_B(WF_FUSED4_INSTANTIATE)
}  // namespace wff
#pragma GCC visibility pop
