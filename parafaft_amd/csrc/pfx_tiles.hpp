// pfx_tiles.hpp -- workgroup tile shapes (transforms per workgroup) for the
// tiled kernels, chosen at compile time from the FFT length, the threads per
// transform and the element size.
#pragma once
#include <stdint.h>

namespace pfx {

constexpr int kLdsBudget = 81920;   // two workgroups per CU (160 KiB of LDS each CU)
constexpr int kLdsBig = 163840;     // one large-tile workgroup per CU

constexpr int imin(int a, int b) { return a < b ? a : b; }
constexpr int imax(int a, int b) { return a > b ? a : b; }

// Every tiled kernel also stages its per-pass twiddle table (N entries) in LDS.
// Row layout: ~256 threads per workgroup, padded rows must fit the budget.
template <int N, int TPT, int ESZ> constexpr int row_tile() {
  int c = imax(1, 256 / TPT);
  while (c > 1 && c * (N + N / 16 + 1) * ESZ + N * ESZ > kLdsBudget) c /= 2;
  return c;
}

// Batch-contiguous layout: 16 transforms (a 128-byte run of complex<float>)
// per element index; at least 128 threads; bounded by LDS and 1024 threads.
template <int N, int TPT, int ESZ, int BUDGET> constexpr int col_tile() {
  int c = 16;
  while (TPT * c < 128 && c < 64) c *= 2;
  while (c > 1 && (c * N * ESZ + N * ESZ > BUDGET || TPT * c > 1024)) c /= 2;
  return c;
}

}  // namespace pfx
