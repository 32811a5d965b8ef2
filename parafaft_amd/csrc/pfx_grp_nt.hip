// Grouped launchers of the streaming batch-contiguous kernels, in the order
// of pfx_nt_list.inc (a unit of their own, like every grouped list).
#include "pfx_launch_extra.hpp"
#include "pfx_nt_list.inc"

namespace pfx {

#define PFX_NT_GROUP(T, TPT, C, ...) &group_c2c_wide_nt<T, TPT, C, false, __VA_ARGS__>,

static const GroupLaunchFn kTable[] = {PFX_NT_LIST(PFX_NT_GROUP)};

const GroupLaunchFn *group_nt(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
