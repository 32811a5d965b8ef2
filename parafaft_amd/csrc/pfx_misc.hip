// pfx_misc.hip -- twiddle generation, the O(n^2) fallback DFT, pack/unpack,
// 2D copy and scale kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pfx_internal.hpp"

namespace pfx {

// ---------------------------------------------------------------------------
// Twiddles: W_L^i = exp(-2 pi i i / L) = (cospi(2i/L), -sinpi(2i/L)).
// The angle is reduced to the first quadrant exactly in integer arithmetic
// (i -> i mod L, then by quarter turns when 4 | L) so sincospi sees small,
// exactly representable arguments wherever possible.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void twiddle_kernel(V2<T>* out, int64_t L) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  double s, c;
  if ((L & 3) == 0) {
    const int64_t q = L / 4;
    const int64_t quad = i / q, r = i - quad * q;  // angle = quad*pi/2 + 2 pi r / L
    double sr, cr;
    sincospi(2.0 * (double)r / (double)L, &sr, &cr);
    switch (quad & 3) {
      case 0: s = sr; c = cr; break;
      case 1: s = cr; c = -sr; break;
      case 2: s = -sr; c = -cr; break;
      default: s = -cr; c = sr; break;
    }
  } else {
    sincospi(2.0 * (double)i / (double)L, &s, &c);
  }
  out[i] = mk<T>(T(c), T(-s));
}

// exp(-2 pi i e / L) for 0 <= e < L with quarter-turn reduction.
__device__ inline void root_of_unity(int64_t e, int64_t L, double& c, double& s) {
  if ((L & 3) == 0) {
    const int64_t q = L / 4;
    const int64_t quad = e / q, r = e - quad * q;
    double sr, cr;
    sincospi(2.0 * (double)r / (double)L, &sr, &cr);
    switch (quad & 3) {
      case 0: s = sr; c = cr; break;
      case 1: s = cr; c = -sr; break;
      case 2: s = -sr; c = -cr; break;
      default: s = -cr; c = sr; break;
    }
  } else {
    sincospi(2.0 * (double)e / (double)L, &s, &c);
  }
}

struct PassTwArgs {
  int n;        // number of passes
  int ns[8];    // span before pass p
  int r[8];     // radix of pass p
  int N;
};

template <typename T>
__global__ void pass_twiddle_kernel(V2<T>* out, PassTwArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N - 1) return;
  const int t = i + 1;  // = m * Ns + kk for the last pass with Ns <= t
  int p = 0;
  for (int q = 0; q < a.n; ++q)
    if (a.ns[q] <= t) p = q;
  const int Ns = a.ns[p], L = Ns * a.r[p];
  const int m = t / Ns, kk = t - m * Ns;
  double c, s;
  root_of_unity((int64_t)kk * m % L, L, c, s);
  out[i] = mk<T>(T(c), T(-s));
}

hipError_t make_pass_twiddles(int prec, const Radices& rad, void* dst, hipStream_t s) {
  PassTwArgs a;
  a.n = rad.n;
  int ns = 1;
  for (int p = 0; p < rad.n; ++p) {
    a.ns[p] = ns;
    a.r[p] = rad.r[p];
    ns *= rad.r[p];
  }
  a.N = ns;
  if (a.N < 2) return hipSuccess;
  const unsigned nb = (unsigned)((a.N - 1 + 255) / 256);
  if (prec == 1) pass_twiddle_kernel<double><<<nb, 256, 0, s>>>(static_cast<double2*>(dst), a);
  else pass_twiddle_kernel<float><<<nb, 256, 0, s>>>(static_cast<float2*>(dst), a);
  return hipGetLastError();
}

hipError_t make_twiddles(int prec, int64_t L, void* dst, hipStream_t s) {
  const int bs = 256;
  const unsigned nb = (unsigned)((L + bs - 1) / bs);
  if (prec == 1) twiddle_kernel<double><<<nb, bs, 0, s>>>(static_cast<double2*>(dst), L);
  else twiddle_kernel<float><<<nb, bs, 0, s>>>(static_cast<float2*>(dst), L);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fallback: direct DFT of one transform per workgroup, input staged in LDS.
// C2C: X[k] = sum_j x[j] W^{jk};  R2C: real row -> k = 0..n/2;
// C2R: x[j] = Re(sum over the Hermitian extension).  Accumulates in T.
// ---------------------------------------------------------------------------
constexpr int kNaiveLdsBytes = 65536;

template <typename T>
__global__ void __launch_bounds__(256) naive_kernel(NaiveArgs a) {
  extern __shared__ unsigned char smem[];
  V2<T>* xs = reinterpret_cast<V2<T>*>(smem);
  T* rs = reinterpret_cast<T*>(smem);
  const int n = a.n;
  const int64_t b = blockIdx.x;
  const int64_t o = b / a.inner, in = b - o * a.inner;
  const int64_t off = o * a.dist_outer + in * a.dist;  // complex units (C2C) or scalars (real)
  const V2<T>* tw = static_cast<const V2<T>*>(a.tw);
  const T sc = a.use_scale ? T(a.scale) : T(1);
  if (a.kind == KIND_C2C) {
    V2<T>* p = static_cast<V2<T>*>(a.data) + off;
    for (int j = threadIdx.x; j < n; j += blockDim.x) xs[j] = p[(int64_t)j * a.stride];
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      T sr = 0, si = 0;
      int64_t e = 0;
      for (int j = 0; j < n; ++j) {
        V2<T> w = tw[e];
        if (a.backward) w.y = -w.y;
        sr += xs[j].x * w.x - xs[j].y * w.y;
        si += xs[j].x * w.y + xs[j].y * w.x;
        e += k;
        if (e >= n) e -= n;
      }
      xs[n + k] = mk<T>(sr * sc, si * sc);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) p[(int64_t)k * a.stride] = xs[n + k];
  } else if (a.kind == KIND_R2C) {
    T* row = static_cast<T*>(a.data) + off;
    V2<T>* crow = reinterpret_cast<V2<T>*>(row);
    for (int j = threadIdx.x; j < n; j += blockDim.x) rs[j] = row[(int64_t)j * a.stride];
    __syncthreads();
    const int h = n / 2 + 1;
    V2<T>* outs = reinterpret_cast<V2<T>*>(smem + ((n * sizeof(T) + 15) / 16) * 16);
    for (int k = threadIdx.x; k < h; k += blockDim.x) {
      T sr = 0, si = 0;
      int64_t e = 0;
      for (int j = 0; j < n; ++j) {
        V2<T> w = tw[e];
        sr += rs[j] * w.x;
        si += rs[j] * w.y;
        e += k;
        if (e >= n) e -= n;
      }
      outs[k] = mk<T>(sr * sc, si * sc);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < h; k += blockDim.x) crow[(int64_t)k * a.stride] = outs[k];
  } else {  // C2R
    T* row = static_cast<T*>(a.data) + off;
    V2<T>* crow = reinterpret_cast<V2<T>*>(row);
    const int h = n / 2 + 1;
    for (int k = threadIdx.x; k < h; k += blockDim.x) xs[k] = crow[(int64_t)k * a.stride];
    __syncthreads();
    T* outs = reinterpret_cast<T*>(smem + (size_t)h * sizeof(V2<T>));
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      T acc = xs[0].x;
      if ((n & 1) == 0) acc += ((j & 1) ? -xs[n / 2].x : xs[n / 2].x);
      const int kmax = (n - 1) / 2;
      int64_t e = j % n;
      for (int k = 1; k <= kmax; ++k) {
        V2<T> w = tw[e];  // W^{jk}; need Re(X[k] conj(W^{jk}))
        acc += T(2) * (xs[k].x * w.x + xs[k].y * w.y);
        e += j;
        if (e >= n) e -= n;
      }
      outs[j] = acc * sc;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) row[(int64_t)j * a.stride] = outs[j];
  }
}

int naive_max_length(int prec) {
  // C2C needs 2n complex values in LDS.
  return kNaiveLdsBytes / (2 * (prec == 1 ? 16 : 8));
}

void launch_naive(int prec, const NaiveArgs& a, hipStream_t s) {
  const size_t esz = prec == 1 ? 16 : 8;
  const size_t lds = 2 * (size_t)(a.n + 2) * esz + 32;
  if (prec == 1)
    naive_kernel<double><<<dim3((unsigned)a.total), dim3(256), lds, s>>>(a);
  else
    naive_kernel<float><<<dim3((unsigned)a.total), dim3(256), lds, s>>>(a);
}

// ---------------------------------------------------------------------------
// Pack / unpack: segment s copies `rows` rows of row_bytes bytes.  Pack:
// source rows at src_off + r*src_pitch, destination contiguous at
// dst_off + r*row_bytes.  Unpack: the mirror.  Elements are moved in units
// of U bytes (the widest power of two dividing every offset, pitch and row
// length), so the copy is bit-exact.  blockIdx.y = segment.
// ---------------------------------------------------------------------------
constexpr int kMaxSeg = 64;
struct PackArgs {
  char* dst;
  const char* src;
  uint64_t rows;
  uint64_t dst_pitch;  // 0: rows are contiguous (pitch = row_bytes)
  uint64_t src_pitch;  // 0: rows are contiguous
  PackSeg seg[kMaxSeg];
};

template <typename U>
__global__ void __launch_bounds__(256) pack_kernel(PackArgs a) {
  const PackSeg sg = a.seg[blockIdx.y];
  const uint64_t upr = sg.row_bytes / sizeof(U);      // units per row
  const uint64_t total = upr * a.rows;
  const uint64_t dp = a.dst_pitch ? a.dst_pitch : sg.row_bytes;
  const uint64_t sp = a.src_pitch ? a.src_pitch : sg.row_bytes;
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < total;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = u / upr, c = u - r * upr;
    const U v = *reinterpret_cast<const U*>(a.src + sg.src_off + r * sp + c * sizeof(U));
    *reinterpret_cast<U*>(a.dst + sg.dst_off + r * dp + c * sizeof(U)) = v;
  }
}

static uint64_t unit_for(uint64_t x) {
  if (x % 16 == 0) return 16;
  if (x % 8 == 0) return 8;
  if (x % 4 == 0) return 4;
  if (x % 2 == 0) return 2;
  return 1;
}

hipError_t launch_pack(void* dst, const void* src, const PackSeg* segs, int nseg, uint64_t rows,
                       uint64_t dst_pitch_mode, uint64_t src_pitch_mode, uint64_t dst_pitch,
                       uint64_t src_pitch, hipStream_t s) {
  (void)dst_pitch_mode;
  (void)src_pitch_mode;
  for (int base = 0; base < nseg; base += kMaxSeg) {
    const int cnt = (nseg - base) < kMaxSeg ? (nseg - base) : kMaxSeg;
    PackArgs a;
    a.dst = static_cast<char*>(dst);
    a.src = static_cast<const char*>(src);
    a.rows = rows;
    a.dst_pitch = dst_pitch;
    a.src_pitch = src_pitch;
    uint64_t g = unit_for((uint64_t)(uintptr_t)dst) < unit_for((uint64_t)(uintptr_t)src)
                     ? unit_for((uint64_t)(uintptr_t)dst)
                     : unit_for((uint64_t)(uintptr_t)src);
    uint64_t maxunits = 0;
    for (int i = 0; i < cnt; ++i) {
      a.seg[i] = segs[base + i];
      uint64_t m = unit_for(a.seg[i].dst_off);
      if (unit_for(a.seg[i].src_off) < m) m = unit_for(a.seg[i].src_off);
      if (unit_for(a.seg[i].row_bytes) < m) m = unit_for(a.seg[i].row_bytes);
      if (g > m) g = m;
    }
    if (dst_pitch && unit_for(dst_pitch) < g) g = unit_for(dst_pitch);
    if (src_pitch && unit_for(src_pitch) < g) g = unit_for(src_pitch);
    for (int i = 0; i < cnt; ++i) {
      uint64_t u = a.seg[i].row_bytes / g * rows;
      if (u > maxunits) maxunits = u;
    }
    if (maxunits == 0) continue;
    uint64_t blocks = (maxunits + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    dim3 grid((unsigned)blocks, (unsigned)cnt);
    switch (g) {
      case 16: pack_kernel<uint4><<<grid, 256, 0, s>>>(a); break;
      case 8: pack_kernel<uint2><<<grid, 256, 0, s>>>(a); break;
      case 4: pack_kernel<uint32_t><<<grid, 256, 0, s>>>(a); break;
      case 2: pack_kernel<uint16_t><<<grid, 256, 0, s>>>(a); break;
      default: pack_kernel<uint8_t><<<grid, 256, 0, s>>>(a); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_copy2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                         size_t height, hipStream_t s) {
  PackSeg seg{0, 0, (uint64_t)width};
  return launch_pack(dst, src, &seg, 1, height, 0, 0, dpitch, spitch, s);
}

// ---------------------------------------------------------------------------
// Scale
// ---------------------------------------------------------------------------
template <typename T>
__global__ void scale_kernel(T* p, int64_t n, T s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] *= s;
}

hipError_t launch_scale(void* data, int64_t count, int prec, double s, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 262144) blocks = 262144;
  if (prec == 1)
    scale_kernel<double><<<(unsigned)blocks, 256, 0, st>>>(static_cast<double*>(data), count, s);
  else
    scale_kernel<float><<<(unsigned)blocks, 256, 0, st>>>(static_cast<float*>(data), count, (float)s);
  return hipGetLastError();
}

}  // namespace pfx
