// Non-temporal variants of the batch-contiguous register-resident kernels
// that Measure planning picks for the axis-0 pass (the last pass of an N-D
// forward, whose data is not read again soon): 1024^3 and 512^3, FP32 and
// FP64.  Table only (no grouped or segmented form); lowest priority, so only
// the planner's timing selects them.
#include "pfx_internal.hpp"

namespace pfx {

#define PFX_NT_ENTRY(T, TPT, C, ...)                                                        \
  {prec_of<T>(), KIND_C2C, prod<__VA_ARGS__>(), 1, TPT, C,                                   \
   (int)(C * prod<__VA_ARGS__>() * sizeof(T)), 0, "wide-nt:" #__VA_ARGS__,                    \
   &launch_c2c_wide_nt<T, TPT, C, false, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr, nullptr},

static const KernelEntry kTable[] = {
    PFX_NT_ENTRY(float, 16, 32, 8, 8, 8)
    PFX_NT_ENTRY(double, 64, 16, 8, 8, 8)
    PFX_NT_ENTRY(double, 64, 16, 16, 16, 4)
};

const KernelEntry *table_nt(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
