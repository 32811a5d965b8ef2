#define PFX_T double
#define PFX_TABLE_FN table_f64_col
#include "pfx_tab_col.inc"
