#define PFX_T double
#define PFX_T_IS_DOUBLE 1
#define PFX_MODE 0
#define PFX_TABLE_FN table_f64_row
#include "pfx_tab_row.inc"
