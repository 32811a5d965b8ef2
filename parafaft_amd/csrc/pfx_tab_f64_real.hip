#define PFX_T double
#define PFX_TABLE_FN table_f64_real
#include "pfx_tab_real.inc"
