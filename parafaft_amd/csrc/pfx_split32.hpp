// pfx_split32.hpp -- radix-32 sub-DFTs of the split-precision form on the
// 32 x 32 matrix cores: N = 1024 = 32 x 32 in two matrix-core passes over a
// tile of 32 batch-contiguous transforms.
//
// Element n1 + 32 n2 of a transform (n1, n2 < 32):
//   pass 1: Y[n1][k2] = sum_n2 W_32^{n2 k2} x[n1 + 32 n2], then * W_1024^{n1 k2}
//   pass 2: X[k2 + 32 k1] = sum_n1 W_32^{n1 k1} Y[n1][k2]
// (the Stockham passes (Ns, R) = (1, 32), (32, 32), whose per-pass twiddle
// table holds W_1024^{k2 n1} at entry 32 n1 + k2 - 1).
//
// Each pass is a product D = F B on v_mfma_f32_32x32x16 (two K = 16 halves)
// with F[k][n] = W_32^{kn} as the A operand and the 32 transforms of the tile
// as the 32 columns of B, so a lane's loads and stores run along the
// contiguous batch axis (256-byte runs).  Operand layout (lane l, register r):
//   A[i][k], B[k][j]: i or j = l % 32, k = 8 (l / 32) + r           (r < 8)
//   D[i][j]:          j = l % 32, i = 8 (r / 4) + 4 (l / 32) + r % 4 (r < 16)
// so lane l holds inputs n = 16 h + 8 (l / 32) + r (K half h) and receives
// outputs k = 8 (r / 4) + 4 (l / 32) + r % 4 of the butterfly in column l % 32.
//
// The operand splits, per-butterfly power-of-two scaling (FP16) and
// Thue-Morse output signs are those of the radix-16 form (pfx_split.hpp).
// Pass 2 needs every pass-1 output of its k2, so the pass-1 outputs go through
// an LDS buffer in two rounds of 16 k2 values (128 KB); the F operand splits
// and the pass-2 twiddles are staged in LDS ahead of it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pfx_split.hpp"

namespace pfx {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef short short8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

namespace k32 {
constexpr double C1 = 0.98078528040323044912618223613423903;  // cos(pi / 16)
constexpr double C3 = 0.83146961230254523707878837761790576;  // cos(3 pi / 16)
constexpr double C5 = 0.55557023301960222474283081394853288;  // cos(5 pi / 16)
constexpr double C7 = 0.19509032201612826784828486847702224;  // cos(7 pi / 16)
}  // namespace k32

// cos(2 pi t / 32), t < 32, as one indexed load per entry from a constant
// table (a select chain compiles to a branching search here)
static __constant__ float kCos32[32] = {
    1.f, float(k32::C1), float(k::COS_PI8), float(k32::C3), float(k::SQRT1_2), float(k32::C5),
    float(k::SIN_PI8), float(k32::C7), 0.f, -float(k32::C7), -float(k::SIN_PI8), -float(k32::C5),
    -float(k::SQRT1_2), -float(k32::C3), -float(k::COS_PI8), -float(k32::C1), -1.f,
    -float(k32::C1), -float(k::COS_PI8), -float(k32::C3), -float(k::SQRT1_2), -float(k32::C5),
    -float(k::SIN_PI8), -float(k32::C7), 0.f, float(k32::C7), float(k::SIN_PI8), float(k32::C5),
    float(k::SQRT1_2), float(k32::C3), float(k::COS_PI8), float(k32::C1)};
__device__ __forceinline__ float cos32(int t) { return kCos32[t & 31]; }

// LDS block of the kernel: F operand splits, pass-2 twiddles, exchange buffer
// (small tables first, so their addresses fit ds_read's 16-bit offset).
template <int MODE> struct Split32Lds {
  // operands per K half: FP16 {Fr, Fi, -Fi} x {hi, lo}; BF16 x {hi, mid, lo}
  static constexpr int terms = MODE == SPLIT_F16 ? 2 : 3;
  static constexpr int ops = 2 * 3 * terms;
  static constexpr int f_bytes = ops * 64 * 16;
  static constexpr int ntw = 31 * 32;  // W_1024^{k2 n1}, n1 = 1..31
  static constexpr int tw_bytes = ntw * 8;
  static constexpr int buf_elems = 16 * 32 * 32;  // one round: 16 k2 x 32 n1 x 32 columns
  static constexpr int bytes = f_bytes + tw_bytes + buf_elems * 8;
  static_assert(bytes <= kLdsMax, "radix-32 split tile exceeds LDS");
  // operand index: K half h, matrix m (0 Fr, 1 Fi, 2 -Fi), split term t
  __device__ static __forceinline__ int op(int h, int m, int t) { return (h * 3 + m) * terms + t; }
};

// Fill the F operand table: one (K half, matrix, lane) vector of 8 entries
// per task, 256 tasks over the workgroup's NT threads.
template <int MODE>
__device__ __forceinline__ void split32_fill_task(u32x4_t* F, int task);
template <int MODE, int NT>
__device__ __forceinline__ void split32_fill_f(u32x4_t* F) {
#pragma unroll
  for (int task = threadIdx.x; task < 256; task += NT) split32_fill_task<MODE>(F, task);
}
template <int MODE>
__device__ __forceinline__ void split32_fill_task(u32x4_t* F, int task) {
  using S = Split32Lds<MODE>;
  const int h = task >> 7, m = (task >> 6) & 1, L = task & 63;
  const int i = L & 31, k0 = 16 * h + 8 * (L >> 5);
  if constexpr (MODE == SPLIT_F16) {
    half8_t hi, lo;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int t = i * (k0 + r);
      const float v = m == 0 ? cos32(t) : -cos32(t + 24);  // Fi = -sin(2 pi t / 32)
      const _Float16 vh = (_Float16)v;
      hi[r] = vh;
      lo[r] = (_Float16)(v - (float)vh);
    }
    F[S::op(h, m, 0) * 64 + L] = __builtin_bit_cast(u32x4_t, hi);
    F[S::op(h, m, 1) * 64 + L] = __builtin_bit_cast(u32x4_t, lo);
    if (m == 1) {
      F[S::op(h, 2, 0) * 64 + L] = __builtin_bit_cast(u32x4_t, -hi);
      F[S::op(h, 2, 1) * 64 + L] = __builtin_bit_cast(u32x4_t, -lo);
    }
  } else {
    short8_t s[3];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int t = i * (k0 + r);
      const float v = m == 0 ? cos32(t) : -cos32(t + 24);
      short a, b, c;
      MfmaDft16<SPLIT_BF16>::split3(v, a, b, c);
      s[0][r] = a;
      s[1][r] = b;
      s[2][r] = c;
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      F[S::op(h, m, t) * 64 + L] = __builtin_bit_cast(u32x4_t, s[t]);
      // the split of -v is the negated split of v (round to nearest even is
      // symmetric)
      if (m == 1) F[S::op(h, 2, t) * 64 + L] = __builtin_bit_cast(u32x4_t, s[t] ^ (short)0x8000);
    }
  }
}

template <int MODE> struct MfmaDft32;

template <> struct MfmaDft32<SPLIT_F16> {
  using S = Split32Lds<SPLIT_F16>;
  __device__ static __forceinline__ half8_t f(const u32x4_t* F, int h, int m, int t, int lane) {
    return __builtin_bit_cast(half8_t, F[S::op(h, m, t) * 64 + lane]);
  }
  // x[8 h + r]: input n = 16 h + 8 (lane / 32) + r; y[r]: output
  // k = 8 (r / 4) + 4 (lane / 32) + r % 4, times `post`.  Every lane of the
  // wave calls this together.  The butterfly's 32 inputs sit in lanes l and
  // l ^ 32.
  __device__ static __forceinline__ void apply(const V2<float>* x, V2<float>* y, bool neg,
                                               const u32x4_t* F, float post = 1.f) {
    const int lane = threadIdx.x & 63;
    float amax = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) amax = fmaxf(amax, fmaxf(fabsf(x[r].x), fabsf(x[r].y)));
    // (ds_bpermute: v_permlane32_swap needs a second register here, where
    // the kernel sits at its 256-VGPR limit, and spilled one)
    amax = fmaxf(amax, __shfl_xor(amax, 32, 64));
    // amax < 2^ex; the scaled maximum stays below 2^15 (radix-16 form)
    const int ex = (int)(__float_as_uint(amax) >> 23) - 126;
    const int sh = ex - 15 < -100 ? -100 : ex - 15;
    const uint32_t sbit = neg ? 0x80000000u : 0u;
    const float sin_ = __uint_as_float(((uint32_t)(127 - sh) << 23) | sbit);
    // a power of two times `post`: the same rounding as scaling by the power
    // of two and then by `post`
    const float sout = __uint_as_float(((uint32_t)(127 + sh) << 23) | sbit) * post;
    float16_t zr = {}, zi = {};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      half8_t xrh, xrl, xih, xil;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const V2<float> v = x[8 * h + r];
        const _Float16 hr = (_Float16)(v.x * sin_), hi = (_Float16)(v.y * sin_);
        xrh[r] = hr;
        xrl[r] = (_Float16)__builtin_fmaf(v.x, sin_, -(float)hr);
        xih[r] = hi;
        xil[r] = (_Float16)__builtin_fmaf(v.y, sin_, -(float)hi);
      }
      // Yr = Fr Xr + (-Fi) Xi, Yi = Fi Xr + Fr Xi, each as hi*lo + lo*hi + hi*hi
      const half8_t frh = f(F, h, 0, 0, lane), frl = f(F, h, 0, 1, lane);
      zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(frh, xrl, zr, 0, 0, 0);
      zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(frh, xil, zi, 0, 0, 0);
      zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(frl, xrh, zr, 0, 0, 0);
      zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(frl, xih, zi, 0, 0, 0);
      zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(frh, xrh, zr, 0, 0, 0);
      zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(frh, xih, zi, 0, 0, 0);
      const half8_t fih = f(F, h, 1, 0, lane), fil = f(F, h, 1, 1, lane);
      zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fih, xrl, zi, 0, 0, 0);
      zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fil, xrh, zi, 0, 0, 0);
      zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(fih, xrh, zi, 0, 0, 0);
      const half8_t nfih = f(F, h, 2, 0, lane), nfil = f(F, h, 2, 1, lane);
      zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(nfih, xil, zr, 0, 0, 0);
      zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(nfil, xih, zr, 0, 0, 0);
      zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(nfih, xih, zr, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) y[r] = mk<float>(zr[r] * sout, zi[r] * sout);
  }
};

template <> struct MfmaDft32<SPLIT_BF16> {
  using S = Split32Lds<SPLIT_BF16>;
  __device__ static __forceinline__ bf16x8_t f(const u32x4_t* F, int h, int m, int t, int lane) {
    return __builtin_bit_cast(bf16x8_t, F[S::op(h, m, t) * 64 + lane]);
  }
  // terms of order <= 2 in the split, smallest first
  __device__ static __forceinline__ float16_t prod(const u32x4_t* F, int h, int m, int lane,
                                                   const bf16x8_t* b, float16_t z) {
    const bf16x8_t a0 = f(F, h, m, 0, lane), a1 = f(F, h, m, 1, lane), a2 = f(F, h, m, 2, lane);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b[0], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b[1], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b[2], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b[0], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b[1], z, 0, 0, 0);
    z = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b[0], z, 0, 0, 0);
    return z;
  }
  __device__ static __forceinline__ void apply(const V2<float>* x, V2<float>* y, bool neg,
                                               const u32x4_t* F, float post = 1.f) {
    const int lane = threadIdx.x & 63;
    const uint32_t sbit = neg ? 0x80000000u : 0u;
    float16_t zr = {}, zi = {};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      short8_t sr[3], si[3];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const V2<float> v = x[8 * h + r];
        short a, b, c;
        MfmaDft16<SPLIT_BF16>::split3(__uint_as_float(__float_as_uint(v.x) ^ sbit), a, b, c);
        sr[0][r] = a;
        sr[1][r] = b;
        sr[2][r] = c;
        MfmaDft16<SPLIT_BF16>::split3(__uint_as_float(__float_as_uint(v.y) ^ sbit), a, b, c);
        si[0][r] = a;
        si[1][r] = b;
        si[2][r] = c;
      }
      bf16x8_t xr[3], xi[3];
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        xr[t] = __builtin_bit_cast(bf16x8_t, sr[t]);
        xi[t] = __builtin_bit_cast(bf16x8_t, si[t]);
      }
      zr = prod(F, h, 2, lane, xi, zr);  // (-Fi) Xi
      zr = prod(F, h, 0, lane, xr, zr);  // Fr Xr
      zi = prod(F, h, 0, lane, xi, zi);  // Fr Xi
      zi = prod(F, h, 1, lane, xr, zi);  // Fi Xr
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      y[r] = mk<float>(__uint_as_float(__float_as_uint(zr[r]) ^ sbit) * post,
                       __uint_as_float(__float_as_uint(zi[r]) ^ sbit) * post);
  }
};

// The kernel body.  W waves (8 or 16); wave w runs the pass-1 butterflies
// n1 = w + W g (g < 32 / W) and, in round rho, the pass-2 butterflies
// k2 = 16 rho + w + W j (j < 16 / W).  Exchange buffer slot of (k2, n1,
// column c) in round rho: ((k2 - 16 rho) 32 + n1) 32 + c, so the 32 lanes of
// a half wave touch 256 contiguous bytes in both directions.
template <int MODE, int W, int PD, bool SW, bool NT>
__device__ __forceinline__ void split32_col_body(const StockhamArgs& a, unsigned char* smem) {
  using S = Split32Lds<MODE>;
  constexpr int G = 32 / W, J = 16 / W;
  static_assert(PD >= 1 && PD <= G, "prefetch depth out of range");
  u32x4_t* F = reinterpret_cast<u32x4_t*>(smem);
  V2<float>* stw = reinterpret_cast<V2<float>*>(smem + S::f_bytes);
  V2<float>* buf = reinterpret_cast<V2<float>*>(smem + S::f_bytes + S::tw_bytes);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c = lane & 31, hh = lane >> 5;
  const int64_t tile = tile_order(blockIdx.x, gridDim.x, a.xcd_remap);
  const int64_t b = tile * 32 + c;
  const bool valid = b < a.total;
  // lanes past the end of the batch read the last transform and store nothing
  const int64_t off = xform_offset(valid ? b : a.total - 1, a), st = a.stride;
  const V2<float>* src = static_cast<const V2<float>*>(a.src ? a.src : a.data) + off;
  V2<float>* dst = static_cast<V2<float>*>(a.data) + off;
  // pass-1 inputs: butterfly n1 = wave + W g, input n = 16 h + 8 hh + r at
  // element n1 + 32 n.  The loads of groups g < PD are issued up front, those
  // of group g + PD just before group g is transformed (PD = G: all at once).
  V2<float> v[G][16];
  const V2<float>* ld = src + (int64_t)(wave + 256 * hh) * st;
  auto load_group = [&](int g) {
    const int64_t s32 = 32 * st, s512 = 512 * st, og = (int64_t)g * W * st;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const V2<float> x = split_ld<NT>(ld + (og + h * s512 + r * s32));
        v[g][8 * h + r] = SW ? swp<float>(x) : x;
      }
  };
#pragma unroll
  for (int g = 0; g < PD; ++g) load_group(g);
  // tables, while the loads are in flight
  split32_fill_f<MODE, 64 * W>(F);
  {
    const V2<float>* tw = static_cast<const V2<float>*>(a.ptw) + 31;
    for (int e = threadIdx.x; e < S::ntw; e += 64 * W) stw[e] = tw[e];
  }
  __syncthreads();
  // pass 1; outputs k2 < 16 (registers r < 8) go to the buffer now, the rest
  // after round 0
  V2<float> hold[G][8];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (g + PD < G) load_group(g + PD);
    const int n1 = wave + W * g;
    V2<float> y[16];
    MfmaDft32<MODE>::apply(v[g], y, split_sign(b, n1), F);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int k2 = 8 * (r >> 2) + 4 * hh + (r & 3);
      if (n1 > 0) y[r] = cmul<float>(y[r], stw[(n1 - 1) * 32 + k2]);
      if (r < 8) buf[((k2 * 32) + n1) * 32 + c] = y[r];
      else hold[g][r - 8] = y[r];
    }
  }
  const float sc = (float)a.scale;
#pragma unroll
  for (int rho = 0; rho < 2; ++rho) {
    if (rho == 1) {
      __syncthreads();  // every round-0 read is done
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int n1 = wave + W * g;
#pragma unroll
        for (int r = 8; r < 16; ++r) {
          const int k2 = 8 * (r >> 2) + 4 * hh + (r & 3) - 16;
          buf[((k2 * 32) + n1) * 32 + c] = hold[g][r - 8];
        }
      }
    }
    __syncthreads();
    // pass 2: butterfly k2, input n1 = 16 h + 8 hh + r; output k1 =
    // 8 (r / 4) + 4 hh + r % 4 at element k2 + 32 k1
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int k2l = wave + W * j, k2 = 16 * rho + k2l;
      V2<float> x[16], y[16];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 8; ++r) x[8 * h + r] = buf[((k2l * 32) + 16 * h + 8 * hh + r) * 32 + c];
      MfmaDft32<MODE>::apply(x, y, split_sign(b, k2), F, sc);
      if (valid) {
        V2<float>* sp = dst + (int64_t)(k2 + 128 * hh) * st;
        const int64_t s32 = 32 * st, s256 = 256 * st;
#pragma unroll
        for (int r = 0; r < 16; ++r)
          split_st<NT>(sp + ((r >> 2) * s256 + (r & 3) * s32), SW ? swp<float>(y[r]) : y[r]);
      }
    }
  }
}

template <int MODE, int W, int PD, bool NT>
__global__ void __launch_bounds__(64 * W) c2c_split32_col_kernel(StockhamArgs a) {
  __shared__ __align__(16) unsigned char smem[Split32Lds<MODE>::bytes];
  if (a.backward) split32_col_body<MODE, W, PD, true, NT>(a, smem);
  else split32_col_body<MODE, W, PD, false, NT>(a, smem);
}

}  // namespace pfx
