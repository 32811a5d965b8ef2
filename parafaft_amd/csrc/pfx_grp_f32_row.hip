#define PFX_T float
#define PFX_T_IS_DOUBLE 0
#define PFX_MODE 1
#define PFX_TABLE_FN group_f32_row
#include "pfx_tab_row.inc"
