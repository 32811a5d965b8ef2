// pfx_fourstep.hip -- the element-wise passes of the plans for lengths
// beyond one kernel: the four-step plan (n = n1 * n2, both factors with
// tiled kernels), Bluestein's chirp convolution for the lengths that do not
// split, and the long R2C / C2R plans built on C2C.
//
// Four-step, for element j = j1 + n1 j2 of a transform (j1 < n1, j2 < n2)
// and output k = n2 k1 + k2:
//   X[n2 k1 + k2] = sum_j1 W_n1^(j1 k1) W_n^(j1 k2) sum_j2 x[j1 + n1 j2] W_n2^(j2 k2)
// 1. length-n2 DFTs along j2 (element stride n1 s), out of place into the
//    plan's scratch at the same offsets: Y[j1][k2] at j1 + n1 k2;
// 2. the twiddle Y[j1][k2] *= W_n^(j1 k2) (this file);
// 3. length-n1 DFTs along j1 (element stride s), in place on the scratch:
//    Z[k1][k2] at k1 + n1 k2;
// 4. the transpose into the caller's array: X at n2 k1 + k2 (this file),
//    with the caller's scale.
// Positions are in elements of the transform, times the stride s, plus b
// dist for transform b of the batch.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pfx_internal.hpp"
#include "pfx_tiles.hpp"
#include "pfx_fsrun.hpp"

namespace pfx {

namespace {

// Step 2.  s == 1: j1 fastest across threads (contiguous); else the batch
// index fastest (contiguous when dist == 1, the batch-contiguous layout).
template <typename T>
__global__ void __launch_bounds__(256) fourstep_twiddle_kernel(V2<T>* buf, const V2<T>* tw,
                                                               FourStepArgs a) {
  const int64_t n = a.n1 * a.n2, total = n * a.batch;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t b, j1, k2;
    if (a.stride == 1) {
      j1 = t % a.n1;
      const int64_t r = t / a.n1;
      k2 = r % a.n2;
      b = r / a.n2;
    } else {
      b = t % a.batch;
      const int64_t r = t / a.batch;
      j1 = r % a.n1;
      k2 = r / a.n1;
    }
    V2<T>* p = buf + b * a.dist + (j1 + a.n1 * k2) * a.stride;
    const V2<T> w = tw[j1 * k2];  // j1 k2 <= (n1 - 1)(n2 - 1) < n
    *p = a.backward ? cmulc<T>(*p, w) : cmul<T>(*p, w);
  }
}

// Step 4, rows (s == 1): per transform an n2 x n1 matrix (row k2, column k1)
// to n1 x n2, through 32 x 32 LDS tiles so both sides are coalesced.
template <typename T>
__global__ void __launch_bounds__(256) fourstep_transpose_rows(const V2<T>* src, V2<T>* dst,
                                                               FourStepArgs a) {
  __shared__ V2<T> tile[32][33];
  const int64_t t1 = (a.n1 + 31) / 32, t2 = (a.n2 + 31) / 32;
  const int64_t nt = t1 * t2 * a.batch;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const T sc = T(a.scale);
  for (int64_t blk = blockIdx.x; blk < nt; blk += gridDim.x) {
    const int64_t b = blk / (t1 * t2), r = blk - b * (t1 * t2);
    const int64_t k1_0 = (r % t1) * 32, k2_0 = (r / t1) * 32;
    const V2<T>* s = src + b * a.dist;
    V2<T>* d = dst + b * a.dist;
    for (int i = ty; i < 32; i += 8) {
      const int64_t k2 = k2_0 + i, k1 = k1_0 + tx;
      if (k1 < a.n1 && k2 < a.n2) tile[i][tx] = s[k1 + a.n1 * k2];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
      const int64_t k1 = k1_0 + i, k2 = k2_0 + tx;
      if (k1 < a.n1 && k2 < a.n2) {
        V2<T> v = tile[tx][i];
        if (a.use_scale) v = scl<T>(v, sc);
        d[a.n2 * k1 + k2] = v;
      }
    }
    __syncthreads();
  }
}

// Step 4, strided transforms and short rows: one element per thread, the
// batch index fastest when dist == 1 (contiguous), else the output element
// fastest (contiguous for rows; a short row's gather stays inside it).
template <typename T>
__global__ void __launch_bounds__(256) fourstep_transpose_strided(const V2<T>* src, V2<T>* dst,
                                                                  FourStepArgs a) {
  const int64_t n = a.n1 * a.n2, total = n * a.batch;
  const T sc = T(a.scale);
  const bool bfast = a.dist == 1;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = bfast ? t % a.batch : t / n;
    const int64_t o = bfast ? t / a.batch : t - b * n;  // output element o = n2 k1 + k2
    const int64_t k1 = o / a.n2, k2 = o - k1 * a.n2;
    V2<T> v = src[b * a.dist + (k1 + a.n1 * k2) * a.stride];
    if (a.use_scale) v = scl<T>(v, sc);
    dst[b * a.dist + o * a.stride] = v;
  }
}

// Long R2C, after the length-h complex FFT of each padded row read as h
// complex pairs z[m] = x[2m] + i x[2m+1]: pairs (k, h - k), k <= h / 2,
//   X[k] = (s - i W_2h^k d) / 2,  s = Z[k] + conj Z[h-k],  d = Z[k] - conj Z[h-k]
// (Z[h] = Z[0]); X[h] goes to the row's spare slot h.  tw: W_2h^k, k <= h.
template <typename T>
__global__ void __launch_bounds__(256) real_post_kernel(V2<T>* data, const V2<T>* tw,
                                                        RealPassArgs a) {
  const int64_t np = a.h / 2 + 1, total = np * a.batch;
  const T half = T(0.5), sc = T(a.scale);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / np, k = t - b * np;
    V2<T>* row = data + b * a.dist;
    const int64_t km = (a.h - k) % a.h;
    const V2<T> A = row[k], Bm = row[km];
    {
      const V2<T> B = conj<T>(Bm);
      const V2<T> s = add<T>(A, B), d = sub<T>(A, B);
      const V2<T> wd = cmul<T>(tw[k], d);
      V2<T> x = mk<T>(half * (s.x + wd.y), half * (s.y - wd.x));
      if (a.use_scale) x = scl<T>(x, sc);
      row[k] = x;
    }
    if (2 * k != a.h) {  // X[h - k]; k = 0 gives X[h]
      const int64_t kk = a.h - k;
      const V2<T> B2 = conj<T>(A);
      const V2<T> s = add<T>(Bm, B2), d = sub<T>(Bm, B2);
      const V2<T> wd = cmul<T>(tw[kk], d);
      V2<T> x = mk<T>(half * (s.x + wd.y), half * (s.y - wd.x));
      if (a.use_scale) x = scl<T>(x, sc);
      row[kk] = x;
    }
  }
}

// Long C2R, before the inverse length-h complex FFT: from the half spectrum
// X[0..h] (Im X[0] and Im X[h] ignored, a Hermitian input),
//   Z[k] = s + i W_2h^-k d,  s = X[k] + conj X[h-k],  d = X[k] - conj X[h-k],
// for k < h; the unnormalised inverse then gives z[m] = x[2m] + i x[2m+1].
template <typename T>
__global__ void __launch_bounds__(256) real_pre_kernel(V2<T>* data, const V2<T>* tw,
                                                       RealPassArgs a) {
  const int64_t np = a.h / 2 + 1, total = np * a.batch;
  const T sc = T(a.scale);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / np, k = t - b * np;
    V2<T>* row = data + b * a.dist;
    const int64_t km = a.h - k;
    V2<T> Xk = row[k], Xm = row[km];
    if (k == 0) {
      Xk.y = T(0);
      Xm.y = T(0);
    }
    {
      const V2<T> B = conj<T>(Xm);
      const V2<T> s = add<T>(Xk, B), d = sub<T>(Xk, B);
      const V2<T> wd = cmulc<T>(d, tw[k]);  // W^-k d
      V2<T> z = mk<T>(s.x - wd.y, s.y + wd.x);
      if (a.use_scale) z = scl<T>(z, sc);
      row[k] = z;
    }
    if (k != 0 && 2 * k != a.h) {
      const V2<T> B = conj<T>(Xk);
      const V2<T> s = add<T>(Xm, B), d = sub<T>(Xm, B);
      const V2<T> wd = cmulc<T>(d, tw[km]);
      V2<T> z = mk<T>(s.x - wd.y, s.y + wd.x);
      if (a.use_scale) z = scl<T>(z, sc);
      row[km] = z;
    }
  }
}

// ---------------------------------------------------------------- Bluestein
// X[k] = w[k] sum_j (x[j] w[j]) conj(w[k - j]),  w[m] = exp(-pi i m^2 / n)
// (backward: w -> conj w), the sum a circular convolution of length
// M >= 2n - 1 done with M-point FFTs.  m^2 is reduced mod 2n in integers,
// so sincospi sees an argument in [0, 2).
template <typename T>
__global__ void __launch_bounds__(256) chirp_kernel(V2<T>* w, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = (int64_t)(((uint64_t)j * (uint64_t)j) % (uint64_t)(2 * n));  // n < 2^31
    double sn, cs;
    sincospi((double)r / (double)n, &sn, &cs);
    w[j] = mk<T>(T(cs), T(-sn));
  }
}

// The convolution filter of one direction, before its FFT: b[m] and b[M - m]
// = conj w[m] (forward) or w[m] (backward) for m < n, zero between.
template <typename T>
__global__ void __launch_bounds__(256) chirp_filter_kernel(V2<T>* b, const V2<T>* w, int64_t n,
                                                           int64_t M, int backward) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = i < n ? i : (i > M - n ? M - i : -1);
    V2<T> v = mk<T>(T(0), T(0));
    if (m >= 0) v = backward ? w[m] : conj<T>(w[m]);
    b[i] = v;
  }
}

__device__ __forceinline__ int64_t xform_base(int64_t t, const ChirpArgs& a) {
  const int64_t o = t / a.inner, i = t - o * a.inner;
  return o * a.dist_outer + i * a.dist;
}

// scratch row t = (x_(t0+t)[j] w'[j] for j < n, then zeros), M per row
template <typename T>
__global__ void __launch_bounds__(256) chirp_pre_kernel(const V2<T>* x, V2<T>* scr,
                                                        const V2<T>* w, ChirpArgs a) {
  const int64_t total = a.total * a.M;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / a.M, j = e - t * a.M;
    V2<T> v = mk<T>(T(0), T(0));
    if (j < a.n) {
      const V2<T> xv = x[xform_base(a.t0 + t, a) + j * a.stride];
      v = a.backward ? cmulc<T>(xv, w[j]) : cmul<T>(xv, w[j]);
    }
    scr[e] = v;
  }
}

// scratch row t *= B (the filter's spectrum)
template <typename T>
__global__ void __launch_bounds__(256) chirp_mul_kernel(V2<T>* scr, const V2<T>* B, ChirpArgs a) {
  const int64_t total = a.total * a.M;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    scr[e] = cmul<T>(scr[e], B[e % a.M]);
}

// x_(t0+t)[k] = scale / M * w'[k] * scratch row t [k], k < n
template <typename T>
__global__ void __launch_bounds__(256) chirp_post_kernel(const V2<T>* scr, V2<T>* x,
                                                         const V2<T>* w, ChirpArgs a) {
  const int64_t total = a.total * a.n;
  const T sc = T(a.scale / (double)a.M);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / a.n, k = e - t * a.n;
    const V2<T> s = scr[t * a.M + k];
    const V2<T> v = a.backward ? cmulc<T>(s, w[k]) : cmul<T>(s, w[k]);
    x[xform_base(a.t0 + t, a) + k * a.stride] = scl<T>(v, sc);
  }
}

// ------------------------------------------- real transforms through C2C
// Odd real lengths beyond the direct-DFT limit: the whole row as a length-n
// complex transform on a scratch of n complex values per row.
// R2C: scratch row = (x[j], 0); after the C2C, the row takes X[k], k <= n/2.
template <typename T>
__global__ void __launch_bounds__(256) real_full_load(const T* rows, V2<T>* scr, RealFullArgs a) {
  const int64_t total = a.batch * a.n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / a.n, j = e - b * a.n;
    scr[e] = mk<T>(rows[b * a.dist + j], T(0));
  }
}
template <typename T>
__global__ void __launch_bounds__(256) real_full_store(const V2<T>* scr, T* rows, RealFullArgs a) {
  const int64_t h = a.n / 2 + 1, total = a.batch * h;
  const T sc = T(a.scale);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / h, k = e - b * h;
    V2<T> v = scr[b * a.n + k];
    if (a.use_scale) v = scl<T>(v, sc);
    rows[b * a.dist + 2 * k] = v.x;
    rows[b * a.dist + 2 * k + 1] = v.y;
  }
}
// C2R: scratch row = the Hermitian extension of X[0..n/2] (Im X[0], and
// Im X[n/2] for even n, ignored); after the inverse C2C, the row takes the
// real parts.
template <typename T>
__global__ void __launch_bounds__(256) herm_full_load(const T* rows, V2<T>* scr, RealFullArgs a) {
  const int64_t total = a.batch * a.n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / a.n, k = e - b * a.n;
    const bool mirror = k > a.n / 2;
    const int64_t kk = mirror ? a.n - k : k;
    V2<T> v = mk<T>(rows[b * a.dist + 2 * kk], rows[b * a.dist + 2 * kk + 1]);
    if (kk == 0 || 2 * kk == a.n) v.y = T(0);
    scr[e] = mirror ? conj<T>(v) : v;
  }
}
template <typename T>
__global__ void __launch_bounds__(256) herm_full_store(const V2<T>* scr, T* rows, RealFullArgs a) {
  const int64_t total = a.batch * a.n;
  const T sc = T(a.scale);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / a.n, j = e - b * a.n;
    T v = scr[e].x;
    if (a.use_scale) v *= sc;
    rows[b * a.dist + j] = v;
  }
}

// Fused first pass of the four-step plan for batch-contiguous callers
// (element k of transform b at k st + b): the length-n2 DFTs along j2 of
// x[j1 + n1 j2], times the twiddle W_n^(+-j1 k2), stored at (n2 j1 + k2) st
// + b of the scratch.  The second pass (the length-n1 DFTs along j1, stride
// n2 st: an ordinary out-of-place batch-contiguous plan) reads them there
// and writes X[n2 k1 + k2] into the caller's array at the same offsets, so
// the plan makes two passes instead of four.  A tile is C consecutive
// transforms b of one j1: its loads and stores run along the contiguous
// batch axis.  The passes run in LDS (Stockham, natural order in and out).
template <typename T, int TPT, int C, int... Rs>
__global__ void __launch_bounds__(TPT * C) fourstep_cols_pass1(FsColArgs a) {
  constexpr int N = (Rs * ... * 1);  // n2
  constexpr int E = N / TPT;
  static_assert(N % TPT == 0, "whole elements per thread");
  using L = LdsLayout<T, N, C, true>;
  using B = LdsBlock<T, V2<T>, N, L::size>;
  __shared__ __align__(16) unsigned char smem[B::bytes];
  V2<T>* lds = B::data(smem);
  const V2<T>* stw = stage_twiddles<T, N, sizeof...(Rs), TPT * C>(B::tw(smem), a.ptw);
  int j, c;
  thread_coords<TPT, C, true>(j, c);
  const int64_t j1 = blockIdx.x / a.tiles_per_j1;
  const int64_t b = (blockIdx.x - j1 * a.tiles_per_j1) * C + c;
  const bool valid = b < a.batch;
  const bool bw = a.backward != 0;
  const V2<T>* src = static_cast<const V2<T>*>(a.src) + (j1 * a.st + (valid ? b : 0));
  const int64_t ls = a.n1 * a.st;
  auto ld = [&](int e) -> V2<T> {
    V2<T> x = mk<T>(T(0), T(0));
    if (valid) x = src[e * ls];
    return bw ? swp<T>(x) : x;
  };
  struct Store {
    V2<T>* dst;
    const V2<T>* w;
    int64_t st, j1;
    bool bw, valid;
    __device__ void before() const {}
    __device__ void after() const {}
    __device__ void operator()(int k2, V2<T> y) const {
      if (bw) y = swp<T>(y);
      const V2<T> t = w[j1 * k2];  // j1 k2 <= (n1 - 1)(n2 - 1) < n
      y = bw ? cmulc<T>(y, t) : cmul<T>(y, t);
      if (valid) dst[k2 * st] = y;
    }
  };
  V2<T>* dst = static_cast<V2<T>*>(a.dst) + ((int64_t)N * j1 * a.st + (valid ? b : 0));
  FsRun<T, N, TPT, C, true, 1, Rs...>::run(
      j, c, lds, stw, ld, Store{dst, static_cast<const V2<T>*>(a.tw), a.st, j1, bw, valid});
}

template <typename T, int TPT, int... Rs>
void launch_cols_pass1(const FsColArgs& a, hipStream_t s) {
  constexpr int N = (Rs * ... * 1);
  // two workgroups per CU up to 512 points (16 columns, 128-byte runs);
  // beyond, one with the CU's whole LDS keeps 8-16 columns (32-column tiles
  // of the short lengths measured no faster)
  constexpr int C = col_tile<N, TPT, (int)sizeof(V2<T>), (N >= 1024 ? kLdsBig : kLdsBudget)>();
  FsColArgs b = a;
  b.tiles_per_j1 = (a.batch + C - 1) / C;
  const int64_t blocks = a.n1 * b.tiles_per_j1;
  fourstep_cols_pass1<T, TPT, C, Rs...><<<dim3((unsigned)blocks), dim3(TPT * C), 0, s>>>(b);
}

// Columns per entry of the first row pass's table W_n^(c k2) (tiles hold at
// most 16 adjacent j1).
constexpr int kRowsTabC = 16;

// Two-pass four-step for rows (element j1 + n1 j2 of row r at r dist +
// (j1 + n1 j2) st).  Pass 1: the length-n2 DFTs of C adjacent j1 of one row
// in LDS (lanes across j1, so a load of one j2 is C consecutive elements),
// each output times W_n^(j1 k2), stored to the scratch at the same offsets
// (j1 + n1 k2).  Pass 2: the length-n1 DFTs of C adjacent k2 from the
// scratch (one contiguous run of C n1 elements per tile, loaded linearly),
// stored transposed to n2 k1 + k2 of the caller's array (lanes across k2:
// runs of C elements) with the caller's scale.
template <typename T, int TPT, int C, int... Rs>
__global__ void __launch_bounds__(TPT * C) fourstep_rows_pass1(FsRowArgs a) {
  constexpr int N = (Rs * ... * 1);  // n2
  constexpr int E = N / TPT;
  static_assert(N % TPT == 0, "whole elements per thread");
  static_assert(C <= kRowsTabC && kRowsTabC % C == 0, "the twiddle table's columns");
  using L = LdsLayout<T, N, C, true>;
  using B = LdsBlock<T, V2<T>, N, L::size>;
  __shared__ __align__(16) unsigned char smem[B::bytes];
  V2<T>* lds = B::data(smem);
  const V2<T>* stw = stage_twiddles<T, N, sizeof...(Rs), TPT * C>(B::tw(smem), a.ptw);
  int j, c;
  thread_coords<TPT, C, true>(j, c);
  // workgroup order as dispatched (adjacent tiles of a row on different
  // XCDs at the same time): the XCD-contiguous order measured 4% faster at
  // 2^15 points and 15% slower at 2^20
  const int64_t r = blockIdx.x / a.tiles;
  const int64_t j1 = (blockIdx.x - r * a.tiles) * C + c;
  const bool valid = j1 < a.n1;
  const bool bw = a.backward != 0;
  const int64_t off = r * a.dist + (valid ? j1 : 0) * a.st;
  const V2<T>* src = static_cast<const V2<T>*>(a.src) + off;
  V2<T>* dst = static_cast<V2<T>*>(a.dst) + off;
  const int64_t ls = a.n1 * a.st;
  // W_n^(j1 k2) = W_n^(c k2) W_n^(j10 k2), j1 = j10 + c: the first from a
  // table laid out [k2][16] (the lanes' c are consecutive, so one load is
  // one run), the second the same for the 16 lanes of one k2 (the lanes' j1
  // differ, so W_n^(j1 k2) from the table of n would scatter their loads)
  const V2<T>* wn = static_cast<const V2<T>*>(a.tw);
  const V2<T>* wc = static_cast<const V2<T>*>(a.twc);
  const int64_t j10 = j1 - c;
  auto ld = [&](int e) -> V2<T> {
    V2<T> x = mk<T>(T(0), T(0));
    if (valid) x = src[e * ls];
    return bw ? swp<T>(x) : x;
  };
  struct Store {
    V2<T>* dst;
    const V2<T>* wn;
    const V2<T>* wc;
    int64_t ls, j10;
    int c;
    bool bw, valid;
    __device__ void before() const {}
    __device__ void after() const {}
    __device__ void operator()(int k2, V2<T> y) const {
      if (bw) y = swp<T>(y);
      // j10 k2 <= (n1 - C)(n2 - 1) < n, c k2 < 16 n2 <= n
      const V2<T> t = cmul<T>(wc[k2 * kRowsTabC + c], wn[j10 * k2]);
      y = bw ? cmulc<T>(y, t) : cmul<T>(y, t);
      if (valid) dst[k2 * ls] = y;
    }
  };
  FsRun<T, N, TPT, C, true, 1, Rs...>::run(j, c, lds, stw, ld,
                                           Store{dst, wn, wc, ls, j10, c, bw, valid});
}

template <typename T, int TPT, int C, int... Rs>
__global__ void __launch_bounds__(TPT * C) fourstep_rows_pass2(FsRowArgs a) {
  constexpr int N = (Rs * ... * 1);  // n1
  constexpr int E = N / TPT;
  static_assert(N % TPT == 0, "whole elements per thread");
  using L = LdsLayout<T, N, C, false>;
  using B = LdsBlock<T, V2<T>, N, L::size>;
  __shared__ __align__(16) unsigned char smem[B::bytes];
  V2<T>* lds = B::data(smem);
  const V2<T>* stw = stage_twiddles<T, N, sizeof...(Rs), TPT * C>(B::tw(smem), a.ptw);
  const int tid = lane_tid();
  const int64_t r = blockIdx.x / a.tiles;
  const int64_t k20 = (blockIdx.x - r * a.tiles) * C;
  const bool bw = a.backward != 0;
  int j, c;
  thread_coords<TPT, C, false>(j, c);
  // element e of the tile's transform c sits at (n1 (k20 + c) + e) st
  const bool valid = k20 + c < a.n2;
  const V2<T>* src = static_cast<const V2<T>*>(a.src) + r * a.dist +
                     a.n1 * (valid ? k20 + c : 0) * a.st;
  auto ld = [&](int e) -> V2<T> {
    V2<T> x = mk<T>(T(0), T(0));
    if (valid) x = src[e * a.st];
    return bw ? swp<T>(x) : x;
  };
  // the outputs stay in LDS: the stores below run across the tile's
  // transforms (adjacent k2) for one k1
  typename FsPasses<T, N, TPT, C, false, 1, Rs...>::StoreLds to_lds{lds, c};
  FsRun<T, N, TPT, C, false, 1, Rs...>::run(j, c, lds, stw, ld, to_lds);
  V2<T>* dst = static_cast<V2<T>*>(a.dst) + r * a.dist;
  const bool use_scale = a.scale != 1.0;
  const T sc = T(a.scale);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int i = tid + m * TPT * C;
    const int t = i % C, k1 = i / C;
    V2<T> y = lds[L::at(k1, t)];
    if (bw) y = swp<T>(y);
    if (use_scale) y = scl<T>(y, sc);
    if (k20 + t < a.n2) dst[(a.n2 * k1 + k20 + t) * a.st] = y;
  }
}

// Tiles: pass 1 as the batch-contiguous first pass (16 columns up to 512
// points, 8-16 beyond with the CU's whole LDS); pass 2 16 transforms per
// tile (128-byte runs of the transposed stores), halved while the padded
// rows exceed the LDS or the tile 1024 threads.
template <int N, int TPT, int ESZ> constexpr int rows_pass2_tile() {
  int c = 16;
  while (c > 1 && (c * (N + N / 16 + 1) * ESZ + N * ESZ > kLdsBig || TPT * c > 1024)) c /= 2;
  return c;
}

template <typename T, int TPT, int... Rs>
void launch_rows_pass1(const FsRowArgs& a, hipStream_t s) {
  constexpr int N = (Rs * ... * 1);
  constexpr int C0 = col_tile<N, TPT, (int)sizeof(V2<T>), (N >= 1024 ? kLdsBig : kLdsBudget)>();
  constexpr int C = C0 < kRowsTabC ? C0 : kRowsTabC;
  FsRowArgs b = a;
  b.tiles = (a.n1 + C - 1) / C;
  fourstep_rows_pass1<T, TPT, C, Rs...><<<dim3((unsigned)(a.rows * b.tiles)), dim3(TPT * C), 0, s>>>(b);
}

template <typename T, int TPT, int... Rs>
void launch_rows_pass2(const FsRowArgs& a, hipStream_t s) {
  constexpr int N = (Rs * ... * 1);
  constexpr int C = rows_pass2_tile<N, TPT, (int)sizeof(V2<T>)>();
  FsRowArgs b = a;
  b.tiles = (a.n2 + C - 1) / C;
  fourstep_rows_pass2<T, TPT, C, Rs...><<<dim3((unsigned)(a.rows * b.tiles)), dim3(TPT * C), 0, s>>>(b);
}

int grid_for(int64_t work) {
  const int64_t blocks = (work + 255) / 256;
  return (int)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
}

}  // namespace

template <typename T>
__global__ void __launch_bounds__(256) rows_pass1_table_kernel(V2<T>* out, const V2<T>* tw,
                                                               int64_t n2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2 * kRowsTabC) out[i] = tw[(i % kRowsTabC) * (i / kRowsTabC)];
}

hipError_t make_fourstep_rows_table(int prec, void* out, const void* tw, int64_t n2,
                                    hipStream_t s) {
  const int g = (int)((n2 * kRowsTabC + 255) / 256);
  if (prec)
    rows_pass1_table_kernel<double><<<g, 256, 0, s>>>(static_cast<V2<double>*>(out),
                                                      static_cast<const V2<double>*>(tw), n2);
  else
    rows_pass1_table_kernel<float><<<g, 256, 0, s>>>(static_cast<V2<float>*>(out),
                                                     static_cast<const V2<float>*>(tw), n2);
  return hipGetLastError();
}

hipError_t launch_fourstep_twiddle(int prec, void* buf, const void* tw, const FourStepArgs& a,
                                   hipStream_t s) {
  const int g = grid_for(a.n1 * a.n2 * a.batch);
  if (prec)
    fourstep_twiddle_kernel<double><<<g, 256, 0, s>>>(static_cast<V2<double>*>(buf),
                                                      static_cast<const V2<double>*>(tw), a);
  else
    fourstep_twiddle_kernel<float><<<g, 256, 0, s>>>(static_cast<V2<float>*>(buf),
                                                     static_cast<const V2<float>*>(tw), a);
  return hipGetLastError();
}

hipError_t launch_fourstep_transpose(int prec, const void* src, void* dst, const FourStepArgs& a,
                                     hipStream_t s) {
  if (a.stride == 1 && a.n1 >= 32 && a.n2 >= 32) {
    const int64_t tiles = ((a.n1 + 31) / 32) * ((a.n2 + 31) / 32) * a.batch;
    const int g = (int)(tiles < 65536 ? tiles : 65536);
    if (prec)
      fourstep_transpose_rows<double><<<g, 256, 0, s>>>(static_cast<const V2<double>*>(src),
                                                        static_cast<V2<double>*>(dst), a);
    else
      fourstep_transpose_rows<float><<<g, 256, 0, s>>>(static_cast<const V2<float>*>(src),
                                                       static_cast<V2<float>*>(dst), a);
  } else {
    const int g = grid_for(a.n1 * a.n2 * a.batch);
    if (prec)
      fourstep_transpose_strided<double><<<g, 256, 0, s>>>(static_cast<const V2<double>*>(src),
                                                           static_cast<V2<double>*>(dst), a);
    else
      fourstep_transpose_strided<float><<<g, 256, 0, s>>>(static_cast<const V2<float>*>(src),
                                                          static_cast<V2<float>*>(dst), a);
  }
  return hipGetLastError();
}

hipError_t launch_real_post(int prec, void* data, const void* tw, const RealPassArgs& a,
                            hipStream_t s) {
  const int g = grid_for((a.h / 2 + 1) * a.batch);
  if (prec)
    real_post_kernel<double><<<g, 256, 0, s>>>(static_cast<V2<double>*>(data),
                                               static_cast<const V2<double>*>(tw), a);
  else
    real_post_kernel<float><<<g, 256, 0, s>>>(static_cast<V2<float>*>(data),
                                              static_cast<const V2<float>*>(tw), a);
  return hipGetLastError();
}

hipError_t launch_real_pre(int prec, void* data, const void* tw, const RealPassArgs& a,
                           hipStream_t s) {
  const int g = grid_for((a.h / 2 + 1) * a.batch);
  if (prec)
    real_pre_kernel<double><<<g, 256, 0, s>>>(static_cast<V2<double>*>(data),
                                              static_cast<const V2<double>*>(tw), a);
  else
    real_pre_kernel<float><<<g, 256, 0, s>>>(static_cast<V2<float>*>(data),
                                             static_cast<const V2<float>*>(tw), a);
  return hipGetLastError();
}

hipError_t make_chirp(int prec, void* w, int64_t n, hipStream_t s) {
  if (prec) chirp_kernel<double><<<grid_for(n), 256, 0, s>>>(static_cast<V2<double>*>(w), n);
  else chirp_kernel<float><<<grid_for(n), 256, 0, s>>>(static_cast<V2<float>*>(w), n);
  return hipGetLastError();
}

hipError_t make_chirp_filter(int prec, void* b, const void* w, int64_t n, int64_t M, int backward,
                             hipStream_t s) {
  if (prec)
    chirp_filter_kernel<double><<<grid_for(M), 256, 0, s>>>(
        static_cast<V2<double>*>(b), static_cast<const V2<double>*>(w), n, M, backward);
  else
    chirp_filter_kernel<float><<<grid_for(M), 256, 0, s>>>(
        static_cast<V2<float>*>(b), static_cast<const V2<float>*>(w), n, M, backward);
  return hipGetLastError();
}

hipError_t launch_chirp_pre(int prec, const void* x, void* scr, const void* w, const ChirpArgs& a,
                            hipStream_t s) {
  const int g = grid_for(a.total * a.M);
  if (prec)
    chirp_pre_kernel<double><<<g, 256, 0, s>>>(static_cast<const V2<double>*>(x),
                                               static_cast<V2<double>*>(scr),
                                               static_cast<const V2<double>*>(w), a);
  else
    chirp_pre_kernel<float><<<g, 256, 0, s>>>(static_cast<const V2<float>*>(x),
                                              static_cast<V2<float>*>(scr),
                                              static_cast<const V2<float>*>(w), a);
  return hipGetLastError();
}

hipError_t launch_chirp_mul(int prec, void* scr, const void* B, const ChirpArgs& a, hipStream_t s) {
  const int g = grid_for(a.total * a.M);
  if (prec)
    chirp_mul_kernel<double><<<g, 256, 0, s>>>(static_cast<V2<double>*>(scr),
                                               static_cast<const V2<double>*>(B), a);
  else
    chirp_mul_kernel<float><<<g, 256, 0, s>>>(static_cast<V2<float>*>(scr),
                                              static_cast<const V2<float>*>(B), a);
  return hipGetLastError();
}

hipError_t launch_chirp_post(int prec, const void* scr, void* x, const void* w, const ChirpArgs& a,
                             hipStream_t s) {
  const int g = grid_for(a.total * a.n);
  if (prec)
    chirp_post_kernel<double><<<g, 256, 0, s>>>(static_cast<const V2<double>*>(scr),
                                                static_cast<V2<double>*>(x),
                                                static_cast<const V2<double>*>(w), a);
  else
    chirp_post_kernel<float><<<g, 256, 0, s>>>(static_cast<const V2<float>*>(scr),
                                               static_cast<V2<float>*>(x),
                                               static_cast<const V2<float>*>(w), a);
  return hipGetLastError();
}

hipError_t launch_real_full(int prec, int which, void* rows, void* scr, const RealFullArgs& a,
                            hipStream_t s) {
  const int64_t work = a.batch * a.n;
  const int g = grid_for(work);
  if (prec) {
    double* r = static_cast<double*>(rows);
    V2<double>* c = static_cast<V2<double>*>(scr);
    switch (which) {
      case 0: real_full_load<double><<<g, 256, 0, s>>>(r, c, a); break;
      case 1: real_full_store<double><<<g, 256, 0, s>>>(c, r, a); break;
      case 2: herm_full_load<double><<<g, 256, 0, s>>>(r, c, a); break;
      default: herm_full_store<double><<<g, 256, 0, s>>>(c, r, a); break;
    }
  } else {
    float* r = static_cast<float*>(rows);
    V2<float>* c = static_cast<V2<float>*>(scr);
    switch (which) {
      case 0: real_full_load<float><<<g, 256, 0, s>>>(r, c, a); break;
      case 1: real_full_store<float><<<g, 256, 0, s>>>(c, r, a); break;
      case 2: herm_full_load<float><<<g, 256, 0, s>>>(r, c, a); break;
      default: herm_full_store<float><<<g, 256, 0, s>>>(c, r, a); break;
    }
  }
  return hipGetLastError();
}

// The fused first pass for n2 (radices of the length's default plan), both
// precisions; false when n2 has none (the plan then makes four passes).
namespace {
struct ColsPass1 {
  int n2;
  Radices rad;
  void (*f32)(const FsColArgs&, hipStream_t);
  void (*f64)(const FsColArgs&, hipStream_t);
};
#define PFX_FS_COLS(TPT, ...)                                                               \
  {prod<__VA_ARGS__>(), radices<__VA_ARGS__>(), &launch_cols_pass1<float, TPT, __VA_ARGS__>, \
   &launch_cols_pass1<double, TPT, __VA_ARGS__>},
const ColsPass1 kColsPass1[] = {
    PFX_FS_COLS(8, 8, 8) PFX_FS_COLS(8, 16, 8) PFX_FS_COLS(16, 16, 16)
    PFX_FS_COLS(64, 8, 8, 8) PFX_FS_COLS(64, 16, 16, 4) PFX_FS_COLS(128, 16, 16, 8)
    // smooth factors (the tiled plans' radices)
    PFX_FS_COLS(5, 16, 5) PFX_FS_COLS(12, 8, 4, 3) PFX_FS_COLS(15, 8, 3, 5)
    PFX_FS_COLS(20, 8, 4, 5) PFX_FS_COLS(12, 16, 4, 3) PFX_FS_COLS(15, 16, 3, 5)
    PFX_FS_COLS(20, 16, 4, 5) PFX_FS_COLS(24, 16, 8, 3) PFX_FS_COLS(40, 16, 8, 5)
    PFX_FS_COLS(48, 16, 16, 3) PFX_FS_COLS(80, 16, 16, 5) PFX_FS_COLS(48, 32, 16, 3)
    PFX_FS_COLS(128, 32, 16, 5) PFX_FS_COLS(192, 16, 16, 4, 3) PFX_FS_COLS(20, 5, 5, 4)
    PFX_FS_COLS(125, 8, 5, 5, 5)};
#undef PFX_FS_COLS
const ColsPass1* cols_pass1(int64_t n2) {
  for (const ColsPass1& e : kColsPass1)
    if (e.n2 == n2) return &e;
  return nullptr;
}
}  // namespace

namespace {
struct RowsPass {
  int n;
  Radices rad;
  void (*p1f32)(const FsRowArgs&, hipStream_t);
  void (*p1f64)(const FsRowArgs&, hipStream_t);
  void (*p2f32)(const FsRowArgs&, hipStream_t);
  void (*p2f64)(const FsRowArgs&, hipStream_t);
};
#define PFX_FS_ROWS(TPT, ...)                                                               \
  {prod<__VA_ARGS__>(), radices<__VA_ARGS__>(), &launch_rows_pass1<float, TPT, __VA_ARGS__>, \
   &launch_rows_pass1<double, TPT, __VA_ARGS__>, &launch_rows_pass2<float, TPT, __VA_ARGS__>, \
   &launch_rows_pass2<double, TPT, __VA_ARGS__>},
const RowsPass kRowsPass[] = {
    PFX_FS_ROWS(8, 8, 8) PFX_FS_ROWS(8, 16, 8) PFX_FS_ROWS(16, 16, 16)
    PFX_FS_ROWS(64, 8, 8, 8) PFX_FS_ROWS(64, 16, 16, 4) PFX_FS_ROWS(128, 16, 16, 8)
    // smooth factors (the tiled plans' radices)
    PFX_FS_ROWS(5, 16, 5) PFX_FS_ROWS(12, 8, 4, 3) PFX_FS_ROWS(15, 8, 3, 5)
    PFX_FS_ROWS(20, 8, 4, 5) PFX_FS_ROWS(12, 16, 4, 3) PFX_FS_ROWS(15, 16, 3, 5)
    PFX_FS_ROWS(20, 16, 4, 5) PFX_FS_ROWS(24, 16, 8, 3) PFX_FS_ROWS(40, 16, 8, 5)
    PFX_FS_ROWS(48, 16, 16, 3) PFX_FS_ROWS(80, 16, 16, 5) PFX_FS_ROWS(48, 32, 16, 3)
    PFX_FS_ROWS(128, 32, 16, 5) PFX_FS_ROWS(192, 16, 16, 4, 3) PFX_FS_ROWS(20, 5, 5, 4)
    PFX_FS_ROWS(125, 8, 5, 5, 5)};
#undef PFX_FS_ROWS
const RowsPass* rows_pass(int64_t n) {
  for (const RowsPass& e : kRowsPass)
    if (e.n == n) return &e;
  return nullptr;
}
}  // namespace

bool fourstep_rows_radices(int64_t n, Radices* rad) {
  const RowsPass* e = rows_pass(n);
  if (e && rad) *rad = e->rad;
  return e != nullptr;
}

hipError_t launch_fourstep_rows_pass(int prec, int pass, const FsRowArgs& a, hipStream_t s) {
  const RowsPass* e = rows_pass(pass == 1 ? a.n2 : a.n1);
  if (!e || (pass != 1 && pass != 2)) return hipErrorInvalidValue;
  if (a.rows == 0) return hipSuccess;
  if (pass == 1) (prec ? e->p1f64 : e->p1f32)(a, s);
  else (prec ? e->p2f64 : e->p2f32)(a, s);
  return hipGetLastError();
}

bool fourstep_cols_pass1_radices(int64_t n2, Radices* rad) {
  const ColsPass1* e = cols_pass1(n2);
  if (e && rad) *rad = e->rad;
  return e != nullptr;
}

hipError_t launch_fourstep_cols_pass1(int prec, const FsColArgs& a, hipStream_t s) {
  const ColsPass1* e = cols_pass1(a.n2);
  if (!e) return hipErrorInvalidValue;
  if (a.n1 * a.batch == 0) return hipSuccess;
  (prec ? e->f64 : e->f32)(a, s);
  return hipGetLastError();
}

}  // namespace pfx
