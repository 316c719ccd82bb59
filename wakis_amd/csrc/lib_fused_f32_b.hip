// step_fused4 instantiations (set B) for FP32 storage.
#include "lib_fused.h"

#pragma GCC visibility push(hidden)
namespace wff {
WF_FUSED4_F32_B(WF_FUSED4_INSTANTIATE)
}  // namespace wff
#pragma GCC visibility pop
