// Table entries of the streaming (non-temporal) batch-contiguous kernels of
// pfx_nt_list.inc.  Their HBM loads and stores do not keep the data in L2;
// they have the lowest priority, so only the planner's timing selects them.
#include "pfx_internal.hpp"
#include "pfx_nt_list.inc"

namespace pfx {

#define PFX_NT_ENTRY(T, TPT, C, ...)                                                        \
  {prec_of<T>(), KIND_C2C, prod<__VA_ARGS__>(), 1, TPT, C,                                   \
   (int)(C * prod<__VA_ARGS__>() * sizeof(T)), 0, "wide-nt:" #__VA_ARGS__,                    \
   &launch_c2c_wide_nt<T, TPT, C, false, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr, nullptr},

static const KernelEntry kTable[] = {PFX_NT_LIST(PFX_NT_ENTRY)};

const KernelEntry *table_nt(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
