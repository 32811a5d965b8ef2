#define PFX_T float
#define PFX_T_IS_DOUBLE 0
#define PFX_CHIRP_FN chirp_fused_f32
#include "pfx_chirp.inc"
