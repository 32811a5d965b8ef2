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

template <int MODE, int R3, int WPB>
void launch_split_row(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + WPB - 1) / WPB;
  c2c_split_row_kernel<MODE, R3, WPB><<<dim3((unsigned)nb), dim3(64 * WPB), 0, s>>>(a);
}
template <int MODE, int R3, int W, int CG, bool NT = false>
void launch_split_col(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + 16 * CG - 1) / (16 * CG);
  c2c_split_col_kernel<MODE, R3, W, CG, NT><<<dim3((unsigned)nb), dim3(64 * W * CG), 0, s>>>(a);
}
template <int MODE, int R3, int W, int CG>
void launch_split_colr(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + 16 * CG - 1) / (16 * CG);
  c2c_split_colr_kernel<MODE, R3, W, CG><<<dim3((unsigned)nb), dim3(64 * W * CG), 0, s>>>(a);
}
template <int R3> constexpr Radices split_radices() {
  return R3 == 1 ? radices<16, 16>() : radices<16, 16, R3>();
}

// Register-resident kernels (passes 16, 16, R3): one wave per row transform
// (WPB per workgroup), or 16 CG batch-contiguous transforms over W waves.
// Batch-contiguous variants holding 32 values per lane at 1024 threads
// (R3 = 4 with W = 8, CG = 2) exceed 128 registers and spill, so they are
// not listed; N = 2048 uses the round-exchange kernel.
#define PFX_SPLIT_ROW(MODE, R3, WPB)                                                        \
  {MODE, KIND_C2C, 256 * R3, 0, 64, WPB, WPB * 256 * R3 * 8, 2, "split-reg:16, 16, " #R3,  \
   &launch_split_row<MODE, R3, WPB>, split_radices<R3>(), nullptr},
#define PFX_SPLIT_COL(MODE, R3, W, CG)                                                      \
  {MODE, KIND_C2C, 256 * R3, 1, 4 * W, 16 * CG, 256 * R3 * 16 * CG * 8, 2,                  \
   "split-reg:16, 16, " #R3 "/w" #W, &launch_split_col<MODE, R3, W, CG>, split_radices<R3>(), \
   nullptr},
// ... with streaming (non-temporal) HBM accesses, for axis-0 passes
#define PFX_SPLIT_COLNT(MODE, R3, W, CG)                                                    \
  {MODE, KIND_C2C, 256 * R3, 1, 4 * W, 16 * CG, 256 * R3 * 16 * CG * 8, 0,                  \
   "split-reg-nt:16, 16, " #R3 "/w" #W, &launch_split_col<MODE, R3, W, CG, true>,           \
   split_radices<R3>(), nullptr},
// ... exchanging in R3 rounds through two 256 x 16 CG buffers (N = 2048)
#define PFX_SPLIT_COLR(MODE, R3, W, CG)                                                     \
  {MODE, KIND_C2C, 256 * R3, 1, 4 * W, 16 * CG, (R3 > 1 ? 2 : 1) * 256 * 16 * CG * 8, 2,     \
   "split-reg:16, 16, " #R3 "/w" #W "/rounds", &launch_split_colr<MODE, R3, W, CG>,          \
   split_radices<R3>(), nullptr},
#define PFX_SPLIT_REG(MODE)                                                                 \
  PFX_SPLIT_ROW(MODE, 1, 4) PFX_SPLIT_ROW(MODE, 2, 4) PFX_SPLIT_ROW(MODE, 4, 4)             \
  PFX_SPLIT_ROW(MODE, 4, 2) PFX_SPLIT_ROW(MODE, 8, 4)                                       \
  PFX_SPLIT_COL(MODE, 1, 4, 2) PFX_SPLIT_COL(MODE, 2, 8, 2) PFX_SPLIT_COL(MODE, 4, 16, 1)   \
  PFX_SPLIT_COL(MODE, 4, 8, 1) PFX_SPLIT_COLR(MODE, 8, 16, 1)                               \
  PFX_SPLIT_COLNT(MODE, 4, 16, 1) PFX_SPLIT_COLNT(MODE, 4, 8, 1)

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
    PFX_SPLIT_REG(SPLIT_F16)
    PFX_SPLIT_REG(SPLIT_BF16)
    PFX_SPLIT_MODE(SPLIT_F16)
    PFX_SPLIT_MODE(SPLIT_BF16)
};

const KernelEntry *table_split(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
