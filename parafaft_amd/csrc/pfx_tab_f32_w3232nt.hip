// The non-temporal form of the 1024-thread radix-32 column kernel, in a unit
// of its own for the same register-allocation reason as pfx_tab_f32_w3232.hip.
#include "pfx_internal.hpp"

namespace pfx {

static const KernelEntry kTable[] = {
    {prec_of<float>(), KIND_C2C, 1024, 1, 32, 32, (int)(32 * 1024 * sizeof(float)), 0,
     "wide-nt:32, 32", &launch_c2c_wide_nt<float, 32, 32, false, 32, 32>, radices<32, 32>(),
     nullptr, nullptr},
};

const KernelEntry *table_f32_w3232nt(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
