#define PFX_T float
#define PFX_T_IS_DOUBLE 0
#define PFX_MODE 2
#define PFX_TABLE_FN seg_f32_w3232
#define PFX_ONE PFX_WIDE_ENTRY(1, 32, 32, 32, 32)
#include "pfx_tab_one.inc"
