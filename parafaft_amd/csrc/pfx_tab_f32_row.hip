#define PFX_T float
#define PFX_T_IS_DOUBLE 0
#define PFX_MODE 0
#define PFX_TABLE_FN table_f32_row
#include "pfx_tab_row.inc"
