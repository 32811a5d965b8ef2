// Split-precision MFMA kernels (complex<float> data): radix-16 passes on
// v_mfma_f32_16x16x16 with FP16 hi/lo or BF16 hi/mid/lo operand splits.
#include "pfx_internal.hpp"
#include "pfx_split.hpp"
#include "pfx_tiles.hpp"

namespace pfx {

template <int MODE, int TPT, int C, bool COL, int... Rs>
void launch_split(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2c_split_kernel<MODE, TPT, C, COL, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}

#define PFX_SPLIT_ENTRY(MODE, COL, TILE, TPT, ...)                                         \
  {MODE, KIND_C2C, prod<__VA_ARGS__>(), COL, TPT, TILE,                                     \
   (int)(TILE * prod<__VA_ARGS__>() * 8 + prod<__VA_ARGS__>() * 8), 1, "split:" #__VA_ARGS__, \
   &launch_split<MODE, TPT, TILE, COL, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr},
#define PFX_SPLIT_PLAN(MODE, TPT, ...)                                                     \
  PFX_SPLIT_ENTRY(MODE, 0, (row_tile<prod<__VA_ARGS__>(), TPT, 8>()), TPT, __VA_ARGS__)     \
  PFX_SPLIT_ENTRY(MODE, 1, (col_tile<prod<__VA_ARGS__>(), TPT, 8, kLdsBudget>()), TPT, __VA_ARGS__)
#define PFX_SPLIT_MODE(MODE)                 \
  PFX_SPLIT_PLAN(MODE, 16, 16, 16)           \
  PFX_SPLIT_PLAN(MODE, 64, 16, 16, 4)        \
  PFX_SPLIT_PLAN(MODE, 128, 16, 16, 8)       \
  PFX_SPLIT_PLAN(MODE, 256, 16, 16, 16)

static const KernelEntry kTable[] = {
    PFX_SPLIT_MODE(SPLIT_F16)
    PFX_SPLIT_MODE(SPLIT_BF16)
};

const KernelEntry *table_split(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
