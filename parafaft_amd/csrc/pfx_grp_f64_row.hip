#define PFX_T double
#define PFX_T_IS_DOUBLE 1
#define PFX_MODE 1
#define PFX_TABLE_FN group_f64_row
#include "pfx_tab_row.inc"
