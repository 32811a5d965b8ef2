#define PFX_T float
#define PFX_TABLE_FN table_f32_col
#include "pfx_tab_col.inc"
