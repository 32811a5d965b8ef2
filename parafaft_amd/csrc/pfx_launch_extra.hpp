// pfx_launch_extra.hpp -- grouped and segmented launchers.  Included only by
// the pfx_grp_* and pfx_seg_* units, so their kernels never share a unit
// with the plain kernels of the same body.
#pragma once
#include "pfx_internal.hpp"

namespace pfx {

template <typename T, int TPT, int C, bool COL, int... Rs>
void group_c2c(const GroupArgs &g, int64_t blocks, hipStream_t s) {
  c2c_group_kernel<T, TPT, C, COL, Rs...><<<dim3((unsigned)blocks), dim3(TPT * C), 0, s>>>(g);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
void group_c2c_wide(const GroupArgs &g, int64_t blocks, hipStream_t s) {
  c2c_wide_group_kernel<T, TPT, C, ROWL, Rs...><<<dim3((unsigned)blocks), dim3(TPT * C), 0, s>>>(g);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
void group_c2c_wide_nt(const GroupArgs &g, int64_t blocks, hipStream_t s) {
  c2c_wide_nt_group_kernel<T, TPT, C, ROWL, Rs...><<<dim3((unsigned)blocks), dim3(TPT * C), 0, s>>>(g);
}
template <typename T, int TPT, int C, int... Rs>
void group_r2c(const GroupArgs &g, int64_t blocks, hipStream_t s) {
  r2c_group_kernel<T, TPT, C, Rs...><<<dim3((unsigned)blocks), dim3(TPT * C), 0, s>>>(g);
}
template <typename T, int TPT, int C, int... Rs>
void group_c2r(const GroupArgs &g, int64_t blocks, hipStream_t s) {
  c2r_group_kernel<T, TPT, C, Rs...><<<dim3((unsigned)blocks), dim3(TPT * C), 0, s>>>(g);
}

template <typename T, int TPT, int C, bool COL, int... Rs>
void seg_c2c(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2c_kernel<T, TPT, C, COL, true, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
void seg_c2c_wide(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2c_wide_kernel<T, TPT, C, ROWL, true, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}
template <typename T, int TPT, int C, int... Rs>
void seg_r2c(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  r2c_kernel<T, TPT, C, true, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}

}  // namespace pfx
