// pfx_planerow.hip -- the first half of a split axis-0 transform fused into
// the pass over the last axis, for the one-GPU N-D plan.
//
// Axis 0 of length n0 = q G splits as x0 = x1 + q x2 (x1 < q, x2 < G) and
// k0 = k2 + G k1 (k2 < G, k1 < q):
//   X[k2 + G k1] = sum_x1 W_q^(x1 k1) [W_n0^(x1 k2) sum_x2 x[x1 + q x2] W_G^(x2 k2)]
// The bracket is this kernel's part.  For one x1 and one row y (an index of
// the axes between axis 0 and the last), a tile loads row y of the G planes
// x1 + q x2, takes the G-point DFT along x2 of every column, multiplies
// output k2 by W_n0^(x1 k2), and transforms each of the G rows along the
// last axis (Stockham, LDS-resident).  The three steps act on different
// axes and the twiddle is constant along a row, so their order is free: the
// DFT along x2 comes first because its loads are G full rows, one coalesced
// run each.  Output k2 goes to plane G x1 + k2 of the destination, so the q
// planes a chunk of consecutive x1 writes are consecutive, and the q-point
// DFTs along x1 that finish axis 0 read and write planes k2 + G m (m < q) at
// one stride: an ordinary batch-contiguous plan from the scratch into the
// caller's array.
//
// Backward: the same kernel with the re/im swap on its loads and stores, which
// turns each of its steps (and the whole split schedule, run in the same
// order) into its conjugate, the inverse DFT.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pfx_internal.hpp"
#include "pfx_fsrun.hpp"

namespace pfx {

namespace {

template <typename T, int G, int TPT, int... Rs>
__global__ void __launch_bounds__(TPT * G) plane_row_kernel(PlaneRowArgs a) {
  constexpr int N = (Rs * ... * 1);  // length of the last axis
  constexpr int NT = TPT * G;
  static_assert(N % NT == 0, "whole columns per thread");
  constexpr int E = N / NT;           // columns per thread in the plane step
  using L = LdsLayout<T, N, G, false>;
  using B = LdsBlock<T, V2<T>, N, L::size>;
  static_assert(B::bytes <= kLdsMax, "tile beyond the LDS of a CU");
  __shared__ __align__(16) unsigned char smem[B::bytes];
  V2<T>* lds = B::data(smem);
  V2<T>* stw = B::tw(smem);
  TwiddleStage<T, N, sizeof...(Rs), NT> ts;
  ts.fetch(a.ptw);
  const int tid = lane_tid();
  const int64_t xr = blockIdx.x / a.rows;
  const int64_t y = blockIdx.x - xr * a.rows;
  const int64_t x1 = a.x1_begin + xr;
  const bool bw = a.backward != 0;
  const V2<T>* src = static_cast<const V2<T>*>(a.src) + (x1 * a.in_a * a.plane + y * N);
  const int64_t ls = a.in_b * a.plane;  // between the tile's G rows
  const V2<T>* w = static_cast<const V2<T>*>(a.tw);
  // Streaming loads: the source is not read again before the last pass
  // rewrites it, so its lines need not take Infinity Cache space from the
  // scratch chunk this kernel writes and the middle axes read next.
  typedef T E2 __attribute__((ext_vector_type(2)));
  using GE2 = __attribute__((address_space(1))) E2;
  V2<T> v[E][G];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int z = tid + e * NT;
#pragma unroll
    for (int m = 0; m < G; ++m) {
      const E2 x2 = __builtin_nontemporal_load((const GE2*)(src + (m * ls + z)));
      const V2<T> x = mk<T>(x2.x, x2.y);
      v[e][m] = bw ? swp<T>(x) : x;
    }
  }
  // W_n0^(x1 m), m < G: x1 m <= (q - 1)(G - 1) < n0
  V2<T> t[G];
#pragma unroll
  for (int m = 1; m < G; ++m) t[m] = w[x1 * m];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int z = tid + e * NT;
    bfly<T, G>(v[e]);
#pragma unroll
    for (int m = 1; m < G; ++m) v[e][m] = cmul<T>(v[e][m], t[m]);
#pragma unroll
    for (int m = 0; m < G; ++m) lds[L::at(z, m)] = v[e][m];
  }
  ts.put(stw);
  __syncthreads();
  int j, c;
  thread_coords<TPT, G, false>(j, c);
  auto ld = [&](int e) -> V2<T> { return lds[L::at(e, c)]; };
  struct Store {
    V2<T>* dst;
    bool bw;
    __device__ void before() const {}
    __device__ void after() const {}
    __device__ void operator()(int e, V2<T> v) const { dst[e] = bw ? swp<T>(v) : v; }
  };
  V2<T>* dst = static_cast<V2<T>*>(a.dst) + ((x1 * a.out_a + c * a.out_b) * a.plane + y * N);
  FsRun<T, N, TPT, G, false, 1, Rs...>::run(j, c, lds, stw, ld, Store{dst, bw});
}

template <typename T, int G, int TPT, int... Rs>
void launch_one(const PlaneRowArgs& a, hipStream_t s) {
  const int64_t blocks = a.nx1 * a.rows;
  plane_row_kernel<T, G, TPT, Rs...><<<dim3((unsigned)blocks), dim3(TPT * G), 0, s>>>(a);
}

struct PlaneRowEntry {
  int prec;  // 0: complex<float>, 1: complex<double>
  int n;
  int g;
  Radices rad;
  void (*launch)(const PlaneRowArgs&, hipStream_t);
};

// The last-axis radices are those of the tiled row kernels of the same
// length (threads per row, largest radix first).
#define PFX_PR(PREC, T, G, TPT, ...) \
  {PREC, prod<__VA_ARGS__>(), G, radices<__VA_ARGS__>(), &launch_one<T, G, TPT, __VA_ARGS__>}

const PlaneRowEntry kPlaneRow[] = {
    PFX_PR(0, float, 4, 64, 8, 8, 8),
    PFX_PR(0, float, 8, 64, 8, 8, 8),
    PFX_PR(0, float, 4, 64, 16, 16, 4),
    PFX_PR(0, float, 8, 64, 16, 16, 4),
    PFX_PR(0, float, 16, 64, 16, 16, 4),
    PFX_PR(0, float, 4, 128, 16, 16, 8),
    PFX_PR(0, float, 8, 128, 16, 16, 8),
    PFX_PR(1, double, 4, 64, 8, 8, 8),
    PFX_PR(1, double, 8, 64, 8, 8, 8),
    PFX_PR(1, double, 4, 64, 16, 16, 4),
    PFX_PR(1, double, 8, 64, 16, 16, 4),
};

const PlaneRowEntry* find_entry(int prec, int64_t n, int g) {
  for (const PlaneRowEntry& e : kPlaneRow)
    if (e.prec == prec && e.n == n && e.g == g) return &e;
  return nullptr;
}

}  // namespace

bool plane_row_radices(int prec, int64_t n, int g, Radices* rad) {
  const PlaneRowEntry* e = find_entry(prec, n, g);
  if (e && rad) *rad = e->rad;
  return e != nullptr;
}

hipError_t launch_plane_row(int prec, int64_t n, int g, const PlaneRowArgs& a, hipStream_t s) {
  const PlaneRowEntry* e = find_entry(prec, n, g);
  if (!e) return hipErrorInvalidValue;
  if (a.nx1 <= 0 || a.rows <= 0) return hipSuccess;
  e->launch(a, s);
  return hipGetLastError();
}

}  // namespace pfx
