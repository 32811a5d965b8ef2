#define PFX_T double
#define PFX_TABLE_FN table_f64_row
#include "pfx_tab_row.inc"
