#define PFX_T double
#define PFX_T_IS_DOUBLE 1
#define PFX_MODE 2
#define PFX_TABLE_FN seg_f64_col
#include "pfx_tab_col.inc"
