// pfx_split.hpp -- split-precision MFMA form of the batched C2C transform.
//
// Radix-16 passes run as dense complex DFT products on the matrix cores:
// Y = F16 X for groups of 16 butterflies, F16[k][m] = exp(-2 pi i k m / 16),
// split into real products Yr = Fr Xr + (-Fi) Xi, Yi = Fi Xr + Fr Xi on
// v_mfma_f32_16x16x16 (FP32 accumulation).  FP32 accuracy is won back by
// splitting both operands:
//   FP16: x = hi + lo (hi = fp16(x), lo = fp16(x - hi)), keeping the three
//         products hi*hi + hi*lo + lo*hi; before the split the 16 inputs
//         of every butterfly are scaled by a power of two so their largest
//         magnitude sits just under 2^15 (exact, undone on the FP32
//         accumulators).
//   BF16: x = hi + mid + lo (8-bit mantissas each), keeping the six products
//         of total order <= 2; BF16 has FP32's range, so no scaling.
// Twiddles between passes and passes of other radices stay in FP32.
//
// The matrix cores' FP32 accumulation rounds with a bias toward -infinity
// (measured: the mean error of every output over a large batch is negative,
// at tens of standard errors, where plain FP32 arithmetic shows none).  A
// bias that is the same for every transform adds up coherently in a later
// axis' sums: in a 1024^3 transform the low-frequency bins of the axis-0/1
// DC line came out ~1e-4 off.  Each butterfly therefore runs on negated
// inputs and negates its outputs when the parity of the set bits of its
// transform and butterfly indices is odd (a Thue-Morse sign): the bias flips
// sign with it, and cancels over any aligned power-of-two range of either
// index.  For FP16 the sign rides on the power-of-two scale factors.
//
// Out of place (the tiled passes of a long plan), every split kernel loads
// from StockhamArgs::src at the offsets of `data` (src_dist and
// src_dist_outer equal to dist and dist_outer, which is how those passes are
// launched) and stores into `data`; StockhamArgs::dst is not used.
//
// Operand layout of v_mfma_f32_16x16x16 (A: M x K, B: K x N, D: M x N, lane l):
//   A[i][k]: i = l % 16, k = 4 (l / 16) + r   (r = 0..3, one 4-vector per lane)
//   B[k][n]: n = l % 16, k = 4 (l / 16) + r
//   D[i][n]: n = l % 16, i = 4 (l / 16) + r
// so lane l feeds inputs m = 4 (l/16) + r of butterfly l % 16 and receives
// outputs k = 4 (l/16) + r of the same butterfly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pfx_stockham.hpp"

namespace pfx {

enum SplitMode { SPLIT_F16 = 2, SPLIT_BF16 = 3 };

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(short h) {
  return __uint_as_float(((uint32_t)(uint16_t)h) << 16);
}

// max(v of lane l, v of lane l ^ 32) in every lane, by v_permlane32_swap
// (lanes 32-63 of one copy trade places with lanes 0-31 of the other, so the
// two results hold the pair in some order): a cross-lane move in the VALU,
// no LDS round trip.
__device__ __forceinline__ float max_lane_xor32(float v) {
  const auto r =
      __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
}
// the same for lane l ^ 16 (v_permlane16_swap: odd 16-lane rows of one copy
// trade places with even rows of the other)
__device__ __forceinline__ float max_lane_xor16(float v) {
  const auto r =
      __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  return fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
}

// Largest value over the 16 inputs of the butterfly in lane column
// u = lane % 16 (lanes u, u + 16, u + 32, u + 48 hold them).
__device__ __forceinline__ float butterfly_max(float v) {
  return max_lane_xor32(max_lane_xor16(v));
}

// cos(2 pi t / 16), t = 0..15, by selects (no indexed table, which would
// live in scratch memory)
__device__ __forceinline__ float cos16(int t) {
  int a = t & 15;
  if (a > 8) a = 16 - a;  // cos(2 pi (16 - a) / 16) = cos(2 pi a / 16)
  const bool neg = a > 4;
  if (neg) a = 8 - a;     // cos(pi - x) = -cos x
  const float c = a == 0 ? 1.f : a == 1 ? float(k::COS_PI8) : a == 2 ? float(k::SQRT1_2)
                : a == 3 ? float(k::SIN_PI8) : 0.f;
  return neg ? -c : c;
}
// cos / sin of 2 pi t / 16
__device__ __forceinline__ void root16(int t, float& c, float& s) {
  c = cos16(t);
  s = cos16(t + 12);  // sin x = cos(x - pi/2)
}

// Thue-Morse sign of a butterfly's pass through the matrix cores.
__device__ __forceinline__ bool split_sign(int64_t b, int jj) {
  return ((__popcll((unsigned long long)b) + __popc((unsigned)jj)) & 1) != 0;
}

template <int MODE> struct MfmaDft16;

template <> struct MfmaDft16<SPLIT_F16> {
  half4_t frh, frl, fih, fil, nfih, nfil;
  __device__ MfmaDft16() {
    const int lane = threadIdx.x & 63;
    const int kout = lane & 15, m0 = 4 * (lane >> 4);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float c, s;
      root16(kout * (m0 + r), c, s);
      const float fr = c, fi = -s;
      const _Float16 frh_ = (_Float16)fr, fih_ = (_Float16)fi;
      frh[r] = frh_;
      frl[r] = (_Float16)(fr - (float)frh_);
      fih[r] = fih_;
      fil[r] = (_Float16)(fi - (float)fih_);
      nfih[r] = -fih[r];
      nfil[r] = -fil[r];
    }
  }
  // x: this lane's 4 inputs (m = 4 (lane/16) + r); y: its 4 outputs.
  // Every lane of the wave must call this together.  Each butterfly gets its
  // own power-of-two scale, so precision is relative to the butterfly's
  // largest input.
  __device__ __forceinline__ void apply(const float2* x, float2* y, bool neg) const {
    float amax = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) amax = fmaxf(amax, fmaxf(fabsf(x[r].x), fabsf(x[r].y)));
    amax = butterfly_max(amax);
    // amax < 2^ex, from the exponent field (ex = -126 for zero and
    // subnormals); the shift keeps the scaled maximum below 2^15 and both
    // factors normal
    const int ex = (int)(__float_as_uint(amax) >> 23) - 126;
    const int sh = ex - 15 < -100 ? -100 : ex - 15;
    const uint32_t sbit = neg ? 0x80000000u : 0u;
    const float sin_ = __uint_as_float(((uint32_t)(127 - sh) << 23) | sbit);
    const float sout = __uint_as_float(((uint32_t)(127 + sh) << 23) | sbit);
    // hi = fp16(x s); lo = fp16(x s - hi), with x s - hi exact (s is a power
    // of two) and formed by one mixed-precision fma (v_fma_mixlo_f16)
    half4_t xrh, xrl, xih, xil;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const _Float16 hr = (_Float16)(x[r].x * sin_), hi = (_Float16)(x[r].y * sin_);
      xrh[r] = hr;
      xrl[r] = (_Float16)__builtin_fmaf(x[r].x, sin_, -(float)hr);
      xih[r] = hi;
      xil[r] = (_Float16)__builtin_fmaf(x[r].y, sin_, -(float)hi);
    }
    float4_t zr = {0.f, 0.f, 0.f, 0.f}, zi = zr;
    // small products first, then the leading ones
    zr = __builtin_amdgcn_mfma_f32_16x16x16f16(frh, xrl, zr, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x16f16(frl, xrh, zr, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x16f16(nfih, xil, zr, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x16f16(nfil, xih, zr, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x16f16(frh, xrh, zr, 0, 0, 0);
    zr = __builtin_amdgcn_mfma_f32_16x16x16f16(nfih, xih, zr, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x16f16(fih, xrl, zi, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x16f16(fil, xrh, zi, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x16f16(frh, xil, zi, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x16f16(frl, xih, zi, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x16f16(fih, xrh, zi, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_16x16x16f16(frh, xih, zi, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = mk<float>(zr[r] * sout, zi[r] * sout);
  }
};

template <> struct MfmaDft16<SPLIT_BF16> {
  bf16x4_t f[3][3];  // [matrix: Fr, Fi, -Fi][term: hi, mid, lo]
  __device__ static __forceinline__ void split3(float v, short& h, short& m, short& l) {
    h = bf16_rne(v);
    const float r1 = v - bf16_to_f(h);
    m = bf16_rne(r1);
    l = bf16_rne(r1 - bf16_to_f(m));
  }
  __device__ MfmaDft16() {
    const int lane = threadIdx.x & 63;
    const int kout = lane & 15, m0 = 4 * (lane >> 4);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float c, s;
      root16(kout * (m0 + r), c, s);
      const float vals[3] = {c, -s, s};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        short h, m, l;
        split3(vals[q], h, m, l);
        f[q][0][r] = h;
        f[q][1][r] = m;
        f[q][2][r] = l;
      }
    }
  }
  __device__ static __forceinline__ float4_t prod(const bf16x4_t* a, const bf16x4_t* b, float4_t z) {
    // terms of order <= 2 in the split: lo*hi, mid*mid, hi*lo, then mid*hi,
    // hi*mid, then hi*hi
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[2], b[0], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b[1], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[2], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b[0], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[1], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[0], z, 0, 0, 0);
    return z;
  }
  __device__ __forceinline__ void apply(const float2* x, float2* y, bool neg) const {
    const uint32_t sbit = neg ? 0x80000000u : 0u;
    bf16x4_t xr[3], xi[3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      short h, m, l;
      split3(__uint_as_float(__float_as_uint(x[r].x) ^ sbit), h, m, l);
      xr[0][r] = h;
      xr[1][r] = m;
      xr[2][r] = l;
      split3(__uint_as_float(__float_as_uint(x[r].y) ^ sbit), h, m, l);
      xi[0][r] = h;
      xi[1][r] = m;
      xi[2][r] = l;
    }
    float4_t zr = {0.f, 0.f, 0.f, 0.f}, zi = zr;
    zr = prod(f[2], xi, zr);  // (-Fi) Xi
    zr = prod(f[0], xr, zr);  // Fr Xr
    zi = prod(f[0], xi, zi);  // Fr Xi
    zi = prod(f[1], xr, zi);  // Fi Xr
#pragma unroll
    for (int r = 0; r < 4; ++r)
      y[r] = mk<float>(__uint_as_float(__float_as_uint(zr[r]) ^ sbit),
                       __uint_as_float(__float_as_uint(zi[r]) ^ sbit));
  }
};

// One radix-16 Stockham pass over the workgroup's LDS tile on the matrix
// cores: groups of 16 butterflies of one transform per MFMA product.
template <int MODE, int N, int TPT, int C, bool COL, int Ns>
__device__ __forceinline__ void mfma_pass16(V2<float>* lds, const V2<float>* stw, const MfmaDft16<MODE>& dft,
                                            int64_t b0) {
  using L = LdsLayout<float, N, C, COL>;
  constexpr int NB = N / 16;     // butterflies per transform
  constexpr int GPT = NB / 16;   // groups per transform
  constexpr int G = C * GPT;     // groups per workgroup
  constexpr int NW = TPT * C / 64;
  static_assert(NB % 16 == 0 && G % NW == 0, "tile does not split into MFMA groups");
  constexpr int GW = G / NW;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m0 = 4 * (lane >> 4);
  float2 xin[GW][4];
#pragma unroll
  for (int gi = 0; gi < GW; ++gi) {
    const int g = wave + gi * NW;
    const int c = g / GPT, jj = (g % GPT) * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + r;
      float2 v = lds[L::at(jj + m * NB, c)];
      if constexpr (Ns > 1) {
        if (m > 0) v = cmul<float>(v, stw[m * Ns + jj % Ns - 1]);
      }
      xin[gi][r] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int gi = 0; gi < GW; ++gi) {
    const int g = wave + gi * NW;
    const int c = g / GPT, jj = (g % GPT) * 16 + (lane & 15);
    float2 y[4];
    dft.apply(xin[gi], y, split_sign(b0 + c, jj));
    const int ob = (jj / Ns) * Ns * 16 + jj % Ns;
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[L::at(ob + (m0 + r) * Ns, c)] = y[r];
  }
  __syncthreads();
}

template <int MODE, int N, int TPT, int C, bool COL, int Ns, int... Rs>
struct SplitPasses;

template <int MODE, int N, int TPT, int C, bool COL, int Ns>
struct SplitPasses<MODE, N, TPT, C, COL, Ns> {
  __device__ static __forceinline__ void run(int, int, V2<float>*, const V2<float>*,
                                             const MfmaDft16<MODE>&, int64_t) {}
};

template <int MODE, int N, int TPT, int C, bool COL, int Ns, int R, int... Rest>
struct SplitPasses<MODE, N, TPT, C, COL, Ns, R, Rest...> {
  using L = LdsLayout<float, N, C, COL>;
  // b0: index of the workgroup's first transform
  __device__ static __forceinline__ void run(int j, int c, V2<float>* lds, const V2<float>* stw,
                                    const MfmaDft16<MODE>& dft, int64_t b0) {
    if constexpr (R == 16) {
      mfma_pass16<MODE, N, TPT, C, COL, Ns>(lds, stw, dft, b0);
    } else {
      auto load = [&](int e) -> V2<float> { return lds[L::at(e, c)]; };
      struct StoreLds {
        V2<float>* lds;
        int c;
        __device__ void before() const { __syncthreads(); }
        __device__ void after() const { __syncthreads(); }
        __device__ void operator()(int e, V2<float> v) const { lds[L::at(e, c)] = v; }
      };
      stockham_pass<float, N, R, Ns, TPT>(j, stw, load, StoreLds{lds, c});
    }
    SplitPasses<MODE, N, TPT, C, COL, Ns * R, Rest...>::run(j, c, lds, stw, dft, b0);
  }
};

template <int MODE, int TPT, int C, bool COL, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_split_kernel(StockhamArgs a) {
  constexpr int N = (Rs * ... * 1);
  constexpr int E = N / TPT;
  using L = LdsLayout<float, N, C, COL>;
  using B = LdsBlock<float, V2<float>, N, L::size>;
  __shared__ __align__(16) unsigned char smem[B::bytes];
  V2<float>* lds = B::data(smem);
  V2<float>* stw = B::tw(smem);
  const V2<float>* tw = stage_twiddles<float, N, sizeof...(Rs), TPT * C>(stw, a.ptw);
  const MfmaDft16<MODE> dft;
  int j, c;
  thread_coords<TPT, C, COL>(j, c);
  int64_t tile = blockIdx.x;
  if (COL) tile = tile_order(tile, gridDim.x, a.xcd_remap);
  const int64_t b = tile * C + c;
  const bool valid = b < a.total;
  const int64_t off = valid ? xform_offset(b, a) : 0;
  V2<float>* p = static_cast<V2<float>*>(a.data) + off;
  // out of place (a long plan's tiled pass): loads from `src` at the same
  // offsets, stores into `data`
  const V2<float>* q = static_cast<const V2<float>*>(a.src ? a.src : a.data) + off;
  const bool sw = a.backward != 0;
  // HBM -> LDS
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int e = j + m * TPT;
    V2<float> x = mk<float>(0.f, 0.f);
    if (valid) x = q[(int64_t)e * a.stride];
    lds[L::at(e, c)] = sw ? swp<float>(x) : x;
  }
  __syncthreads();
  SplitPasses<MODE, N, TPT, C, COL, 1, Rs...>::run(j, c, lds, tw, dft, tile * C);
  // LDS -> HBM
  const float sc = (float)a.scale;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int e = j + m * TPT;
    V2<float> x = lds[L::at(e, c)];
    if (a.use_scale) x = scl<float>(x, sc);
    if (sw) x = swp<float>(x);
    if (valid) p[(int64_t)e * a.stride] = x;
  }
}

// ---------------------------------------------------------------------------
// Register-resident split kernels for N = 256 R3 (passes 16, 16, R3 with
// R3 = 1, 2, 4 or 8).  Both radix-16 passes run on the matrix cores with the
// butterfly's 16 inputs in the B operand (lane l holds inputs 4 (l/16) + r,
// r < 4, of the butterfly in lane column u = l % 16, and receives the same
// outputs):
//  - pass 1 loads straight from HBM into that layout;
//  - one exchange through LDS regroups its outputs for pass 2;
//  - a lane holds the pass-2 butterflies b + 16 g, g < R3, whose outputs m
//    are exactly the R3 inputs of pass-3 butterfly b + 16 m, so pass 3 runs
//    in registers and stores straight to HBM.
// Pass 2 (span 16) reads twiddle W_256^{b m}, pass 3 (span 256) W_N^{jj g},
// both from the per-pass table in global memory (entry m Ns + kk - 1).
// The direction is a template parameter of the body (one uniform branch per
// kernel instead of a select per element), and every unrolled load and store
// offset is a fixed combination of a few precomputed multiples of the stride.
// ---------------------------------------------------------------------------
template <int R3> constexpr int log2_of() { return R3 == 1 ? 0 : R3 == 2 ? 1 : R3 == 4 ? 2 : 3; }

// Pass 3 (radix R3, span 256) on values a lane holds: v[g] is input g of
// butterfly jj; twiddle W_N^{jj g}, then the butterfly; output g lands at
// jj + 256 g.
template <int R3>
__device__ __forceinline__ void split_last_pass(V2<float>* v, const V2<float>* tw, int jj) {
  if constexpr (R3 > 1) {
#pragma unroll
    for (int g = 1; g < R3; ++g) v[g] = cmul<float>(v[g], tw[g * 256 + jj - 1]);
    bfly<float, R3>(v);
  }
}

// The same last pass with its twiddles already in registers
// (t[g - 1] = W^{g jj}, loaded earlier).
template <int R3>
__device__ __forceinline__ void split_last_pass_tw(V2<float>* v, const V2<float>* t) {
  if constexpr (R3 > 1) {
#pragma unroll
    for (int g = 1; g < R3; ++g) v[g] = cmul<float>(v[g], t[g - 1]);
    bfly<float, R3>(v);
  }
}

// Pass-2 twiddle W_256^{b m} for input m (1 for m = 0; the table index is
// kept in range for every lane).
__device__ __forceinline__ V2<float> split_pass2_twiddle(const V2<float>* tw, int m, int b) {
  const V2<float> t = tw[m > 0 ? m * 16 + b - 1 : 0];
  return m > 0 ? t : mk<float>(1.f, 0.f);
}

// One complex<float> from / to HBM, streaming (non-temporal) when NT.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
template <bool NT> __device__ __forceinline__ V2<float> split_ld(const V2<float>* p) {
  if constexpr (NT) {
    const f32x2_t x = __builtin_nontemporal_load(reinterpret_cast<const f32x2_t*>(p));
    return mk<float>(x.x, x.y);
  } else {
    return *p;
  }
}
template <bool NT> __device__ __forceinline__ void split_st(V2<float>* p, V2<float> v) {
  if constexpr (NT) {
    f32x2_t x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, reinterpret_cast<f32x2_t*>(p));
  } else {
    *p = v;
  }
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row layout: one wave per transform, lane column u indexes 16 adjacent
// butterflies, and the exchange stays inside the wave (no workgroup barrier).
// LDS slot of element p: p with bits 0-1 and 4 XOR-ed by bits 5-6 and L of
// p, so both the pass-1 stores (p = 256 g + 16 u + 4 q + r) and the pass-2
// loads (p = 16 g + u + (N/16)(4 q + r)) spread a wave's 64 lanes over all
// 32 complex slots of a bank row, two lanes each.
template <int R3> __device__ __forceinline__ int split_row_slot(int p) {
  constexpr int L = 6 + log2_of<R3>();
  return p ^ ((p >> 5) & 3) ^ (((p >> L) & 1) << 4);
}

template <int MODE, int R3, int WPB, bool SW>
__device__ __forceinline__ void split_row_body(const StockhamArgs& a, V2<float>* lds, int64_t b) {
  constexpr int N = 256 * R3, NB = N / 16;
  const int lane = threadIdx.x & 63, u = lane & 15, q = lane >> 4;
  const int64_t off = xform_offset(b, a), st = a.stride;
  const V2<float>* src = static_cast<const V2<float>*>(a.src ? a.src : a.data) + off;
  V2<float>* dst = static_cast<V2<float>*>(a.data) + off;
  const V2<float>* tw = static_cast<const V2<float>*>(a.ptw);
  const MfmaDft16<MODE> dft;
  V2<float> v[R3][4];
  // pass 1: butterfly 16 i + u, input m = 4 q + r at element 16 i + u + NB m
  {
    const V2<float>* ld = src + (int64_t)(u + 4 * NB * q) * st;
    const int64_t s16 = 16 * st, sNB = NB * st;
#pragma unroll
    for (int i = 0; i < R3; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const V2<float> x = ld[i * s16 + r * sNB];
        v[i][r] = SW ? swp<float>(x) : x;
      }
  }
  // pass-2 and pass-3 twiddles, loaded while the data loads are in flight
  constexpr int G3 = R3 > 1 ? R3 - 1 : 1;
  V2<float> t2[4], t3[4][G3];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    t2[r] = split_pass2_twiddle(tw, 4 * q + r, u);
    const int jj = u + 16 * (4 * q + r);
#pragma unroll
    for (int g = 1; g < R3; ++g) t3[r][g - 1] = tw[g * 256 + jj - 1];
  }
#pragma unroll
  for (int i = 0; i < R3; ++i) {
    V2<float> y[4];
    dft.apply(v[i], y, split_sign(b, 16 * i + u));
    // output m of butterfly jj lands at 16 jj + m
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[split_row_slot<R3>(256 * i + 16 * u + 4 * q + r)] = y[r];
  }
  wave_lds_sync();
  // pass 2: butterfly 16 i + u (span 16), input m = 4 q + r at 16 i + u + NB m
#pragma unroll
  for (int i = 0; i < R3; ++i) {
    V2<float> x[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      x[r] = cmul<float>(lds[split_row_slot<R3>(16 * i + u + NB * (4 * q + r))], t2[r]);
    dft.apply(x, v[i], split_sign(b, 16 * i + u));
  }
  // pass 3: butterfly jj = u + 16 m holds output m of pass-2 butterflies
  // 16 i + u, i < R3, which land at 256 i + u + 16 m; outputs to HBM (a
  // scale of 1 multiplies exactly)
  const float sc = (float)a.scale;
  V2<float>* sp = dst + (int64_t)(u + 64 * q) * st;
  const int64_t s16 = 16 * st, s256 = 256 * st;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    V2<float> t[R3];
#pragma unroll
    for (int i = 0; i < R3; ++i) t[i] = v[i][r];
    split_last_pass_tw<R3>(t, t3[r]);
#pragma unroll
    for (int i = 0; i < R3; ++i) {
      const V2<float> x = scl<float>(t[i], sc);
      sp[r * s16 + i * s256] = SW ? swp<float>(x) : x;
    }
  }
}

template <int MODE, int R3, int WPB>
__global__ void __launch_bounds__(64 * WPB) c2c_split_row_kernel(StockhamArgs a) {
  __shared__ __align__(16) V2<float> lds_all[WPB * 256 * R3];
  const int wave = threadIdx.x >> 6;
  const int64_t b = (int64_t)blockIdx.x * WPB + wave;
  if (b >= a.total) return;  // whole waves only; no workgroup barrier follows
  V2<float>* lds = lds_all + wave * 256 * R3;
  if (a.backward) split_row_body<MODE, R3, WPB, true>(a, lds, b);
  else split_row_body<MODE, R3, WPB, false>(a, lds, b);
}

// Batch-contiguous layout: lane column u indexes 16 adjacent transforms (a
// 128-byte run per element), W waves share a group of 16 transforms and CG
// groups make a workgroup.  The exchange goes through an LDS tile of
// N x 16 CG complex values: element e of transform c sits at e * C + c with
// bit 4 of that index (the half of a 64-bank row) XOR-ed by bit 2 and bit L
// of e, which splits the lanes of both the pass-1 stores (e = 16 jj + 4 q +
// r) and the pass-2 loads (e = jj + (N/16)(4 q + r)) evenly over the two
// halves.
template <int N, int C> struct SplitColLds {
  static constexpr int L = (N == 256 ? 6 : N == 512 ? 7 : N == 1024 ? 8 : 9);  // log2(N / 4)
  static constexpr int size = N * C;
  __device__ static __forceinline__ int at(int e, int c) {
    const int key = ((e >> 2) ^ (e >> L)) & 1;
    return (e * C + c) ^ (key << 4);
  }
};

template <int MODE, int R3, int W, int CG, bool SW, bool NT>
__device__ __forceinline__ void split_col_body(const StockhamArgs& a, V2<float>* lds) {
  constexpr int N = 256 * R3, NB = N / 16, C = 16 * CG;
  constexpr int P1 = NB / W;  // pass-1 butterflies per lane
  constexpr int P2 = 16 / W;  // pass-2 butterfly columns per lane
  static_assert(16 % W == 0 && P1 == P2 * R3, "waves must split the butterflies evenly");
  using L = SplitColLds<N, C>;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int u = lane & 15, q = lane >> 4;
  const int cg = wave % CG, w = wave / CG;
  const int c = 16 * cg + u;
  const int64_t tile = tile_order(blockIdx.x, gridDim.x, a.xcd_remap);
  const int64_t b = tile * C + c;
  const bool valid = b < a.total;
  // lanes past the end of the batch read the last transform and store nothing
  const int64_t off = xform_offset(valid ? b : a.total - 1, a), st = a.stride;
  const V2<float>* src = static_cast<const V2<float>*>(a.src ? a.src : a.data) + off;
  V2<float>* dst = static_cast<V2<float>*>(a.data) + off;
  const V2<float>* tw = static_cast<const V2<float>*>(a.ptw);
  const MfmaDft16<MODE> dft;
  V2<float> v[P1][4];
  // pass 1: butterfly w + W i, input m = 4 q + r at element jj + NB m
  {
    const V2<float>* ld = src + (int64_t)(w + 4 * NB * q) * st;
    const int64_t sW = W * st, sNB = NB * st;
#pragma unroll
    for (int i = 0; i < P1; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const V2<float> x = split_ld<NT>(ld + (i * sW + r * sNB));
        v[i][r] = SW ? swp<float>(x) : x;
      }
  }
#pragma unroll
  for (int i = 0; i < P1; ++i) {
    V2<float> y[4];
    dft.apply(v[i], y, split_sign(b, w + W * i));
    // output m of butterfly jj lands at 16 jj + m
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[L::at(16 * (w + W * i) + 4 * q + r, c)] = y[r];
  }
  // The pass-2 and pass-3 twiddles do not depend on the exchange: their
  // loads are issued before the barrier, so they are in flight during it
  // (the pass-1 values are dead by now, so they cost no extra registers).
  constexpr int G3 = R3 > 1 ? R3 - 1 : 1;
  V2<float> t2[P2][4], t3[P2][4][G3];
#pragma unroll
  for (int k = 0; k < P2; ++k) {
    const int bb = w + W * k;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      t2[k][r] = split_pass2_twiddle(tw, 4 * q + r, bb);
      const int jj = bb + 16 * (4 * q + r);
#pragma unroll
      for (int g = 1; g < R3; ++g) t3[k][r][g - 1] = tw[g * 256 + jj - 1];
    }
  }
  __syncthreads();
  // pass 2 takes butterflies b + 16 k2, b = w + W k, k2 < R3, at slot k R3 + k2
#pragma unroll
  for (int k = 0; k < P2; ++k)
#pragma unroll
    for (int k2 = 0; k2 < R3; ++k2)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        v[k * R3 + k2][r] = lds[L::at(w + W * k + 16 * k2 + NB * (4 * q + r), c)];
  const float sc = (float)a.scale;
  V2<float>* sp = dst + (int64_t)(w + 64 * q) * st;
  const int64_t sW = W * st, s16 = 16 * st, s256 = 256 * st;
#pragma unroll
  for (int k = 0; k < P2; ++k) {
    const int bb = w + W * k;
    // pass 2 (span 16): twiddle W_256^{bb m}
#pragma unroll
    for (int k2 = 0; k2 < R3; ++k2) {
      V2<float> x[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) x[r] = cmul<float>(v[k * R3 + k2][r], t2[k][r]);
      dft.apply(x, v[k * R3 + k2], split_sign(b, bb + 16 * k2));
    }
    // pass 3 in registers, then HBM at element jj + 256 k2
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      V2<float> t[R3];
#pragma unroll
      for (int k2 = 0; k2 < R3; ++k2) t[k2] = v[k * R3 + k2][r];
      split_last_pass_tw<R3>(t, t3[k][r]);
      if (valid) {
#pragma unroll
        for (int k2 = 0; k2 < R3; ++k2) {
          const V2<float> x = scl<float>(t[k2], sc);
          split_st<NT>(sp + (k * sW + r * s16 + k2 * s256), SW ? swp<float>(x) : x);
        }
      }
    }
  }
}

// NT: streaming HBM loads and stores, for an axis-0 pass (data not read
// again soon); a planner candidate beside the ordinary form.
template <int MODE, int R3, int W, int CG, bool NT>
__global__ void __launch_bounds__(64 * W * CG) c2c_split_col_kernel(StockhamArgs a) {
  __shared__ __align__(16) V2<float> lds[SplitColLds<256 * R3, 16 * CG>::size];
  if (a.backward) split_col_body<MODE, R3, W, CG, true, NT>(a, lds);
  else split_col_body<MODE, R3, W, CG, false, NT>(a, lds);
}

// Batch-contiguous layout with the exchange in R3 rounds.  Input m of
// pass-2 butterfly b + 16 k2 is output b of pass-1 butterfly R3 m + k2, so
// round k2 runs the sixteen pass-1 butterflies R3 t + k2, exchanges their
// outputs as a 16 x 16 transpose per transform (slot (t, m') written, slot
// (m, b) read), and runs the sixteen pass-2 butterflies b + 16 k2.  Two
// alternating buffers of 256 x 16 CG complex values stand in for the whole
// N x 16 CG tile (64 KB instead of 256 KB at N = 2048, 16 transforms), which
// gives N = 2048 a register-resident batch-contiguous kernel; one barrier
// per round (a buffer is rewritten two rounds later, after every lane has
// passed the barrier between).  At N = 1024 the smaller LDS would let two
// 512-thread workgroups share a CU, but such a kernel needs 150 VGPRs and
// spills when held to 128, and was measured no faster.  Slot (t, m)
// has bit 4 of its index (the half of a 64-bank row) XOR-ed by bit 2 of t
// and of m, which spreads the four lane groups of both the stores (m' =
// 4 q + r) and the loads (t = 4 q + r) over both halves.
template <int C> __device__ __forceinline__ int split_round_slot(int t, int m, int c) {
  const int key = ((m >> 2) ^ (t >> 2)) & 1;
  return ((t * 16 + m) * C + c) ^ (key << 4);
}

template <int MODE, int R3, int W, int CG, bool SW>
__device__ __forceinline__ void split_colr_body(const StockhamArgs& a, V2<float>* lds) {
  constexpr int N = 256 * R3, NB = N / 16, C = 16 * CG;
  constexpr int G = 16 / W;  // butterflies per lane and round, in each pass
  static_assert(16 % W == 0, "waves must split the sixteen butterflies of a round");
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int u = lane & 15, q = lane >> 4;
  const int cg = wave % CG, w = wave / CG;
  const int c = 16 * cg + u;
  const int64_t tile = tile_order(blockIdx.x, gridDim.x, a.xcd_remap);
  const int64_t b = tile * C + c;
  const bool valid = b < a.total;
  // lanes past the end of the batch read the last transform and store nothing
  const int64_t off = xform_offset(valid ? b : a.total - 1, a), st = a.stride;
  const V2<float>* src = static_cast<const V2<float>*>(a.src ? a.src : a.data) + off;
  V2<float>* dst = static_cast<V2<float>*>(a.data) + off;
  const V2<float>* tw = static_cast<const V2<float>*>(a.ptw);
  const MfmaDft16<MODE> dft;
  V2<float> v[R3 * G][4];
  // every pass-1 input up front: slot k2 G + i holds butterfly R3 t + k2,
  // t = w + W i, inputs m = 4 q + r at element R3 t + k2 + NB m
  {
    const V2<float>* ld = src + (int64_t)(R3 * w + 4 * NB * q) * st;
    const int64_t sT = R3 * W * st, sNB = NB * st;
#pragma unroll
    for (int k2 = 0; k2 < R3; ++k2)
#pragma unroll
      for (int i = 0; i < G; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const V2<float> x = ld[k2 * st + i * sT + r * sNB];
          v[k2 * G + i][r] = SW ? swp<float>(x) : x;
        }
  }
#pragma unroll
  for (int k2 = 0; k2 < R3; ++k2) {
    V2<float>* buf = lds + (k2 & 1) * 256 * C;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int t = w + W * i;
      V2<float> y[4];
      dft.apply(v[k2 * G + i], y, split_sign(b, R3 * t + k2));
#pragma unroll
      for (int r = 0; r < 4; ++r) buf[split_round_slot<C>(t, 4 * q + r, c)] = y[r];
    }
    __syncthreads();
    // pass 2 (span 16): butterfly bb + 16 k2, input m from slot (m, bb),
    // twiddle W_256^{bb m}
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int bb = w + W * i;
      V2<float> x[4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
        x[r] = cmul<float>(buf[split_round_slot<C>(4 * q + r, bb, c)],
                           split_pass2_twiddle(tw, 4 * q + r, bb));
      dft.apply(x, v[k2 * G + i], split_sign(b, bb + 16 * k2));
    }
  }
  // pass 3 in registers: butterfly jj = bb + 16 m takes output m of pass-2
  // butterflies bb + 16 k2 (slots k2 G + i); output k2 lands at jj + 256 k2
  const float sc = (float)a.scale;
  V2<float>* sp = dst + (int64_t)(w + 64 * q) * st;
  const int64_t sW = W * st, s16 = 16 * st, s256 = 256 * st;
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const int bb = w + W * i;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int jj = bb + 16 * (4 * q + r);
      V2<float> t[R3];
#pragma unroll
      for (int k2 = 0; k2 < R3; ++k2) t[k2] = v[k2 * G + i][r];
      split_last_pass<R3>(t, tw, jj);
      if (valid) {
#pragma unroll
        for (int k2 = 0; k2 < R3; ++k2) {
          const V2<float> x = scl<float>(t[k2], sc);
          sp[i * sW + r * s16 + k2 * s256] = SW ? swp<float>(x) : x;
        }
      }
    }
  }
}

template <int MODE, int R3, int W, int CG>
__global__ void __launch_bounds__(64 * W * CG) c2c_split_colr_kernel(StockhamArgs a) {
  __shared__ __align__(16) V2<float> lds[(R3 > 1 ? 2 : 1) * 256 * 16 * CG];
  if (a.backward) split_colr_body<MODE, R3, W, CG, true>(a, lds);
  else split_colr_body<MODE, R3, W, CG, false>(a, lds);
}

}  // namespace pfx
