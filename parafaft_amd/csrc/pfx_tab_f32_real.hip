#define PFX_T float
#define PFX_TABLE_FN table_f32_real
#include "pfx_tab_real.inc"
