#define PFX_T double
#define PFX_T_IS_DOUBLE 1
#define PFX_CHIRP_FN chirp_fused_f64
#include "pfx_chirp.inc"
