// Radix-32 split-precision kernels (complex<float> data, N = 1024): two
// passes on v_mfma_f32_32x32x16 over tiles of 32 batch-contiguous transforms.
#include "pfx_internal.hpp"
#include "pfx_split32.hpp"

namespace pfx {

template <int MODE, int W, int PD, bool NT>
void launch_split32_col(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + 31) / 32;
  c2c_split32_col_kernel<MODE, W, PD, NT><<<dim3((unsigned)nb), dim3(64 * W), 0, s>>>(a);
}

#define PFX_SPLIT32_COL(MODE, W, PD, NT, PREF, TAG)                                        \
  {MODE, KIND_C2C, 1024, 1, 2 * W, 32, Split32Lds<MODE>::bytes, PREF,                       \
   "split32" TAG ":32, 32/w" #W "/p" #PD, &launch_split32_col<MODE, W, PD, NT>,             \
   radices<32, 32>(), nullptr},
#define PFX_SPLIT32(MODE, PD1, PD2)                                                        \
  PFX_SPLIT32_COL(MODE, 8, PD1, false, 1, "") PFX_SPLIT32_COL(MODE, 8, PD2, false, 1, "")  \
  PFX_SPLIT32_COL(MODE, 8, PD1, true, 0, "-nt") PFX_SPLIT32_COL(MODE, 8, PD2, true, 0, "-nt")

// FP16 only: the BF16 form's three-term operand splits spill 65 or more
// VGPRs at 512 threads (122 or more at 256), so BF16 keeps the radix-16
// kernels.  Rows keep the radix-16 kernels too: a radix-32 row kernel (one
// wave per row, 8 KB exchange per wave) was measured slower (0.100 against
// 0.093 ms for 32768 rows of 1024), since rows already run in 256-byte runs.
static const KernelEntry kTable[] = {
    PFX_SPLIT32(SPLIT_F16, 4, 2)
};

const KernelEntry *table_split32(int *count) {
  *count = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}

}  // namespace pfx
