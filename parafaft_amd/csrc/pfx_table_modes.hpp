// pfx_table_modes.hpp -- one entry list, three compilations.  Each kernel
// list (pfx_tab_row.inc, pfx_tab_col.inc, pfx_tab_real.inc) is compiled with
//   PFX_MODE 0: the KernelEntry table with the plain launchers;
//   PFX_MODE 1: the grouped launchers of the same entries, same order;
//   PFX_MODE 2: the segmented launchers (fused exchange), same order;
// and the registry merges the three by index.
//
// Entry macros: PFX_E_C2C (tiled row / batch-contiguous kernel), PFX_E_WIDE
// (register-resident kernel), PFX_E_R2C, PFX_E_C2R.  C and LDS are
// parenthesised constant expressions.
#pragma once
#include "pfx_internal.hpp"
#include "pfx_tiles.hpp"
#if PFX_MODE != 0
#include "pfx_launch_extra.hpp"
#endif

#if PFX_MODE == 0
#define PFX_E_C2C(PREF, NAME, TPT, C, COL, LDS, ...)                                     \
  {prec_of<T>(), KIND_C2C, prod<__VA_ARGS__>(), COL ? 1 : 0, TPT, C, (int)LDS, PREF, NAME, \
   &launch_c2c<T, TPT, C, COL, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr, nullptr},
#define PFX_E_WIDE(PREF, NAME, TPT, C, ROWL, LDS, ...)                                      \
  {prec_of<T>(), KIND_C2C, prod<__VA_ARGS__>(), ROWL ? 0 : 1, TPT, C, (int)LDS, PREF, NAME,  \
   &launch_c2c_wide<T, TPT, C, ROWL, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr, nullptr},
#define PFX_E_R2C(NAME, TPT, C, LDS, ...)                                                  \
  {prec_of<T>(), KIND_R2C, prod<__VA_ARGS__>(), 0, TPT, C, (int)LDS, 1, NAME,             \
   &launch_r2c<T, TPT, C, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr, nullptr},
#define PFX_E_C2R(NAME, TPT, C, LDS, ...)                                                  \
  {prec_of<T>(), KIND_C2R, prod<__VA_ARGS__>(), 0, TPT, C, (int)LDS, 1, NAME,             \
   &launch_c2r<T, TPT, C, __VA_ARGS__>, radices<__VA_ARGS__>(), nullptr, nullptr},
#define PFX_TABLE_TYPE KernelEntry
#elif PFX_MODE == 1
#define PFX_E_C2C(PREF, NAME, TPT, C, COL, LDS, ...) &group_c2c<T, TPT, C, COL, __VA_ARGS__>,
#define PFX_E_WIDE(PREF, NAME, TPT, C, ROWL, LDS, ...) \
  &group_c2c_wide<T, TPT, C, ROWL, __VA_ARGS__>,
#define PFX_E_R2C(NAME, TPT, C, LDS, ...) &group_r2c<T, TPT, C, __VA_ARGS__>,
#define PFX_E_C2R(NAME, TPT, C, LDS, ...) &group_c2r<T, TPT, C, __VA_ARGS__>,
#define PFX_TABLE_TYPE GroupLaunchFn
#else
#define PFX_E_C2C(PREF, NAME, TPT, C, COL, LDS, ...) &seg_c2c<T, TPT, C, COL, __VA_ARGS__>,
#define PFX_E_WIDE(PREF, NAME, TPT, C, ROWL, LDS, ...) \
  &seg_c2c_wide<T, TPT, C, ROWL, __VA_ARGS__>,
#define PFX_E_R2C(NAME, TPT, C, LDS, ...) &seg_r2c<T, TPT, C, __VA_ARGS__>,
#define PFX_E_C2R(NAME, TPT, C, LDS, ...) nullptr,
#define PFX_TABLE_TYPE LaunchFn
#endif

#define PFX_DEFINE_TABLE_FN                                  \
  const PFX_TABLE_TYPE *PFX_TABLE_FN(int *count) {          \
    *count = (int)(sizeof(kTable) / sizeof(kTable[0]));      \
    return kTable;                                           \
  }
