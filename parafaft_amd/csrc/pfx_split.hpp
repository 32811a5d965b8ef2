// pfx_split.hpp -- split-precision MFMA form of the batched C2C transform.
//
// Radix-16 passes run as dense complex DFT products on the matrix cores:
// Y = F16 X for groups of 16 butterflies, F16[k][m] = exp(-2 pi i k m / 16),
// split into real products Yr = Fr Xr + (-Fi) Xi, Yi = Fi Xr + Fr Xi on
// v_mfma_f32_16x16x16 (FP32 accumulation).  FP32 accuracy is won back by
// splitting both operands:
//   FP16: x = hi + lo (hi = fp16(x), lo = fp16(x - hi)), keeping the three
//         products hi*hi + hi*lo + lo*hi; before the split every group of 16
//         butterflies is scaled by a power of two so its largest magnitude
//         sits just under 2^15 (exact, undone on the FP32 accumulators).
//   BF16: x = hi + mid + lo (8-bit mantissas each), keeping the six products
//         of total order <= 2; BF16 has FP32's range, so no scaling.
// Twiddles between passes and passes of other radices stay in FP32.
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

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// cos / sin of 2 pi t / 16, t = 0..15
__device__ __forceinline__ void root16(int t, float& c, float& s) {
  const float a = float(k::COS_PI8), b = float(k::SIN_PI8), r = float(k::SQRT1_2);
  const float cs[16] = {1.f, a, r, b, 0.f, -b, -r, -a, -1.f, -a, -r, -b, 0.f, b, r, a};
  c = cs[t & 15];
  s = cs[(t + 12) & 15];  // sin x = cos(x - pi/2)
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
  // Every lane of the wave must call this together.
  __device__ void apply(const float2* x, float2* y) const {
    float amax = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) amax = fmaxf(amax, fmaxf(fabsf(x[r].x), fabsf(x[r].y)));
    amax = wave_max(amax);
    int e = 0;
    if (amax > 0.f) frexpf(amax, &e);  // amax < 2^e
    const int sh = e - 15;             // scaled values below 2^15
    const float sin_ = ldexpf(1.f, -sh), sout = ldexpf(1.f, sh);
    half4_t xrh, xrl, xih, xil;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float xr = x[r].x * sin_, xi = x[r].y * sin_;
      xrh[r] = (_Float16)xr;
      xrl[r] = (_Float16)(xr - (float)xrh[r]);
      xih[r] = (_Float16)xi;
      xil[r] = (_Float16)(xi - (float)xih[r]);
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
  __device__ void apply(const float2* x, float2* y) const {
    bf16x4_t xr[3], xi[3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      short h, m, l;
      split3(x[r].x, h, m, l);
      xr[0][r] = h;
      xr[1][r] = m;
      xr[2][r] = l;
      split3(x[r].y, h, m, l);
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
    for (int r = 0; r < 4; ++r) y[r] = mk<float>(zr[r], zi[r]);
  }
};

// One radix-16 Stockham pass over the workgroup's LDS tile on the matrix
// cores: groups of 16 butterflies of one transform per MFMA product.
template <int MODE, int N, int TPT, int C, bool COL, int Ns>
__device__ __forceinline__ void mfma_pass16(V2<float>* lds, const V2<float>* stw, const MfmaDft16<MODE>& dft) {
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
    dft.apply(xin[gi], y);
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
  __device__ static __forceinline__ void run(int, int, V2<float>*, const V2<float>*, const MfmaDft16<MODE>&) {}
};

template <int MODE, int N, int TPT, int C, bool COL, int Ns, int R, int... Rest>
struct SplitPasses<MODE, N, TPT, C, COL, Ns, R, Rest...> {
  using L = LdsLayout<float, N, C, COL>;
  __device__ static __forceinline__ void run(int j, int c, V2<float>* lds, const V2<float>* stw,
                                    const MfmaDft16<MODE>& dft) {
    if constexpr (R == 16) {
      mfma_pass16<MODE, N, TPT, C, COL, Ns>(lds, stw, dft);
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
    SplitPasses<MODE, N, TPT, C, COL, Ns * R, Rest...>::run(j, c, lds, stw, dft);
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
  const V2<float>* tw = stage_twiddles<float, N, sizeof...(Rs)>(stw, a.ptw);
  const MfmaDft16<MODE> dft;
  int j, c;
  thread_coords<TPT, C, COL>(j, c);
  int64_t tile = blockIdx.x;
  if (COL) tile = tile_order(tile, gridDim.x, a.xcd_remap);
  const int64_t b = tile * C + c;
  const bool valid = b < a.total;
  V2<float>* p = static_cast<V2<float>*>(a.data) + (valid ? xform_offset(b, a) : 0);
  const bool sw = a.backward != 0;
  // HBM -> LDS
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int e = j + m * TPT;
    V2<float> x = mk<float>(0.f, 0.f);
    if (valid) x = p[(int64_t)e * a.stride];
    lds[L::at(e, c)] = sw ? swp<float>(x) : x;
  }
  __syncthreads();
  SplitPasses<MODE, N, TPT, C, COL, 1, Rs...>::run(j, c, lds, tw, dft);
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

}  // namespace pfx
