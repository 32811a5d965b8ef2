/*
 * pfx_oracle.c -- CPU oracle for the MI355X FFT backend.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only the test-suite, the smoke check and the CPU-baseline
 * leg of the benchmark use it, and only as the checker.
 *
 * It restates, in plain FP64 C, what the reference computes on the hot path:
 *
 *   - the batched strided 1D DFT that FFTWBackend<double>::execute_stage runs
 *     (backend/fftw3/fft_backend_fftw.hpp:387-424 geometry, :419-422 signs,
 *     :562-572 execution): forward exp(-2 pi i jk/n), backward exp(+...),
 *     both unnormalised;
 *   - the in-place padded R2C / C2R (fft_backend_fftw.hpp:439-549): rows of
 *     2(n/2+1) scalars, input dist counted in scalars, output dist/2 complex;
 *   - the plan geometry and stage loop of ParaFaFT_C2C / ParaFaFT_R2C on one
 *     rank (parafaft_c2c.hpp:249-287, 539-694, 781-843, 866-897;
 *     parafaft_r2c.hpp:360-542, 1069-1245);
 *   - decompose / init_exchange_geometry / the pack and unpack loops
 *     (parafaft_common.hpp:73-80, 345-393, 576-590, 654-668) and, as an
 *     independent check of those loops, a per-element enumerator of the
 *     C-order subarray that MPI_Type_create_subarray describes (:108-160);
 *   - chunk_range (backend/parallel_for.hpp:53-62);
 *   - the deterministic test generators of test/test_helpers.hpp:106-270.
 *
 * The FFT is an independent recursive mixed-radix decimation-in-time written
 * for this oracle; it shares no code or index scheme with the GPU kernels.
 * Twiddles come from long-double cosl/sinl, rounded once to double.
 *
 * How each function is pinned (golden vectors from numpy's pocketfft,
 * brute-force long-double DFTs, closed forms, the reference's own unit-test
 * known answers) is listed in the test-suite that exercises it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The scalar of the FFT core: double for the oracle; the CPU baseline of the
 * benchmark also builds the core alone with ORC_REAL=float (ORC_FFT_ONLY),
 * the FFTWBackend<float> counterpart. */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef struct { ORC_REAL re, im; } cplx;

/* ------------------------------------------------------------------------ */
/* Twiddle table: w[k] = exp(sign * 2 pi i k / n), long double then rounded. */
/* ------------------------------------------------------------------------ */
static cplx *make_twiddles(int64_t n, int sign) {
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int64_t k = 0; k < n; ++k) {
    /* reduce k/n to the first octant-free form: angle = 2 pi k / n */
    long double a = 2.0L * pi * (long double)k / (long double)n;
    w[k].re = (ORC_REAL)cosl(a);
    w[k].im = (ORC_REAL)(sign * sinl(a));
  }
  return w;
}

static int64_t smallest_factor(int64_t n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int64_t p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/*
 * Recursive DIT: out[0..n) = DFT_n(in[0], in[is], in[2 is], ...).
 * n0 / tw describe the root table of the top-level length; the twiddle
 * exp(sign 2 pi i r k / n) is tw[r k (n0/n) mod n0].
 * scratch must hold at least n0 entries of workspace per recursion level is
 * not needed: the combine step works in a small local buffer.
 */
static void fft_rec(const cplx *in, int64_t is, cplx *out, int64_t n,
                    const cplx *tw, int64_t n0, cplx *tmp) {
  if (n == 1) { out[0] = in[0]; return; }
  int64_t p = smallest_factor(n);
  int64_t m = n / p;
  /* sub-transforms: Y_r = DFT_m(x[r], x[r+p], ...) stored at out[r*m ..] */
  for (int64_t r = 0; r < p; ++r)
    fft_rec(in + r * is, is * p, out + r * m, m, tw, n0, tmp);
  /* combine: X[k + q m] = sum_r W_n^{r k} Y_r[k] W_p^{r q} */
  const int64_t step = n0 / n;      /* W_n^x = tw[x * step] */
  const int64_t pstep = n0 / p;     /* W_p^x = tw[x * pstep] */
  for (int64_t k = 0; k < m; ++k) {
    for (int64_t r = 0; r < p; ++r) {
      cplx y = out[r * m + k];
      int64_t e = (r * k) % n;
      cplx w = tw[e * step];
      tmp[r].re = y.re * w.re - y.im * w.im;
      tmp[r].im = y.re * w.im + y.im * w.re;
    }
    for (int64_t q = 0; q < p; ++q) {
      ORC_REAL sr = 0, si = 0;
      for (int64_t r = 0; r < p; ++r) {
        cplx w = tw[((r * q) % p) * pstep];
        sr += tmp[r].re * w.re - tmp[r].im * w.im;
        si += tmp[r].re * w.im + tmp[r].im * w.re;
      }
      out[q * m + k].re = sr;
      out[q * m + k].im = si;
    }
  }
}

static int64_t largest_prime_factor(int64_t n) {
  int64_t best = 1;
  for (int64_t p = 2; p * p <= n; ++p)
    while (n % p == 0) { best = p; n /= p; }
  if (n > 1 && n > best) best = n;
  return best;
}

/*
 * orc_c2c_batched: in-place batched strided complex DFT on interleaved
 * double data.  Element j of transform b is data[b*dist + j*stride].
 * sign = -1 forward, +1 backward (unnormalised).  Mirrors the semantics of
 * FFTWBackend<double>::create_stage_plan + execute_stage
 * (fft_backend_fftw.hpp:387-424, 562-572).  nthreads <= 0: all threads.
 */
int orc_c2c_batched(ORC_REAL *data, int64_t length, int64_t batch, int64_t stride,
                    int64_t dist, int sign, int nthreads) {
  if (length <= 0 || batch < 0) return -1;
  if (batch == 0) return 0;
  cplx *tw = make_twiddles(length, sign < 0 ? -1 : 1);
  cplx *x = (cplx *)data;
  int64_t lpf = largest_prime_factor(length);
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    cplx *in = (cplx *)malloc(sizeof(cplx) * (size_t)length);
    cplx *out = (cplx *)malloc(sizeof(cplx) * (size_t)length);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(lpf + 4));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t b = 0; b < batch; ++b) {
      cplx *base = x + b * dist;
      for (int64_t j = 0; j < length; ++j) in[j] = base[j * stride];
      fft_rec(in, 1, out, length, tw, length, tmp);
      for (int64_t j = 0; j < length; ++j) base[j * stride] = out[j];
    }
    free(in); free(out); free(tmp);
  }
  free(tw);
  return 0;
}

#ifndef ORC_FFT_ONLY
/* Brute-force O(n^2) DFT in long double; out-of-place, contiguous. */
int orc_dft_brute(const double *in, double *out, int64_t n, int sign) {
  const long double pi = 3.141592653589793238462643383279502884L;
  const cplx *x = (const cplx *)in;
  cplx *y = (cplx *)out;
  for (int64_t k = 0; k < n; ++k) {
    long double sr = 0.0L, si = 0.0L;
    for (int64_t j = 0; j < n; ++j) {
      int64_t e = (j * k) % n;
      long double a = 2.0L * pi * (long double)e / (long double)n;
      long double c = cosl(a), s = (sign < 0 ? -1.0L : 1.0L) * sinl(a);
      sr += (long double)x[j].re * c - (long double)x[j].im * s;
      si += (long double)x[j].re * s + (long double)x[j].im * c;
    }
    y[k].re = (double)sr;
    y[k].im = (double)si;
  }
  return 0;
}

/*
 * In-place padded R2C (fft_backend_fftw.hpp:439-488): row b starts at scalar
 * b*dist; its n reals are at stride `stride` scalars; the n/2+1 complex
 * outputs land at complex index b*(dist/2) + k*stride.  Computed as a full
 * complex DFT of the real row (no half-length trick), then truncated.
 */
int orc_r2c_inplace(double *padded, int64_t n, int64_t batch, int64_t stride,
                    int64_t dist, int nthreads) {
  if (n <= 0 || batch < 0) return -1;
  cplx *tw = make_twiddles(n, -1);
  int64_t lpf = largest_prime_factor(n);
  int64_t h = n / 2 + 1;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    cplx *in = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *out = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(lpf + 4));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t b = 0; b < batch; ++b) {
      double *row = padded + b * dist;
      for (int64_t j = 0; j < n; ++j) { in[j].re = row[j * stride]; in[j].im = 0.0; }
      fft_rec(in, 1, out, n, tw, n, tmp);
      cplx *crow = (cplx *)(padded) + b * (dist / 2);
      for (int64_t k = 0; k < h; ++k) crow[k * stride] = out[k];
    }
    free(in); free(out); free(tmp);
  }
  free(tw);
  return 0;
}

/*
 * In-place padded C2R (fft_backend_fftw.hpp:503-549): n/2+1 complex inputs at
 * complex index b*(dist/2) + k*stride; n real outputs at scalar b*dist +
 * j*stride.  Output = Re(unnormalised inverse DFT of the Hermitian
 * extension), which ignores Im of the DC and (even n) Nyquist inputs.
 */
int orc_c2r_inplace(double *padded, int64_t n, int64_t batch, int64_t stride,
                    int64_t dist, int nthreads) {
  if (n <= 0 || batch < 0) return -1;
  cplx *tw = make_twiddles(n, +1);
  int64_t lpf = largest_prime_factor(n);
  int64_t h = n / 2 + 1;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
  {
    cplx *in = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *out = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(lpf + 4));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t b = 0; b < batch; ++b) {
      cplx *crow = (cplx *)(padded) + b * (dist / 2);
      for (int64_t k = 0; k < h; ++k) in[k] = crow[k * stride];
      for (int64_t k = h; k < n; ++k) { in[k].re = in[n - k].re; in[k].im = -in[n - k].im; }
      fft_rec(in, 1, out, n, tw, n, tmp);
      double *row = padded + b * dist;
      for (int64_t j = 0; j < n; ++j) row[j * stride] = out[j].re;
    }
    free(in); free(out); free(tmp);
  }
  free(tw);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Geometry restatements                                                     */
/* ------------------------------------------------------------------------ */

/* parafaft_common.hpp:73-80 */
void orc_decompose(int N, int M, int p, int *n, int *s) {
  int q = N / M, r = N % M;
  *n = (r > p) ? (q + 1) : q;
  *s = (r > p) ? (*n * p) : (q * p + r);
}

/* backend/parallel_for.hpp:53-62 */
void orc_chunk_range(uint64_t total, int tid, int nthreads, uint64_t *begin,
                     uint64_t *end) {
  uint64_t t = (uint64_t)tid, n = (uint64_t)nthreads;
  uint64_t quot = total / n, rem = total % n;
  *begin = t * quot + (t < rem ? t : rem);
  *end = *begin + quot + (t < rem ? 1 : 0);
}

/*
 * ParaFaFT_C2C::setup_stage_shapes (parafaft_c2c.hpp:781-843) for a rank at
 * grid coordinates `coords` of a (D-1)-dim grid `dims`.
 * shapes[stage*D + axis]; global_start[D].
 */
void orc_stage_shapes_c2c(int D, const int *global_shape, const int *dims,
                          const int *coords, int *shapes, int *global_start) {
  int local_n[16], local_s[16];
  for (int i = 0; i < D - 1; ++i)
    orc_decompose(global_shape[i], dims[i], coords[i], &local_n[i], &local_s[i]);
  for (int i = 0; i < D - 1; ++i) global_start[i] = local_s[i];
  global_start[D - 1] = 0;
  for (int stage = 0; stage < D; ++stage) {
    for (int axis = 0; axis < D; ++axis) {
      int v;
      if (axis < D - 1 - stage) {
        v = local_n[axis];
      } else if (axis == D - 1 - stage) {
        v = global_shape[axis];
      } else {
        int grid_dim = (D - 1 - stage) + (axis - (D - stage));
        v = (grid_dim >= 0 && grid_dim < D - 1) ? local_n[grid_dim] : global_shape[axis];
      }
      shapes[stage * D + axis] = v;
    }
  }
}

/*
 * ParaFaFT_R2C::setup_stage_shapes (parafaft_r2c.hpp:1069-1151).
 * in_shapes / out_shapes [stage*D + axis]; starts: real[D], complex[D].
 */
void orc_stage_shapes_r2c(int D, const int *global_real, const int *dims,
                          const int *coords, int *in_shapes, int *out_shapes,
                          int *real_start, int *complex_start) {
  int gc[16], local_n[16], local_s[16];
  for (int i = 0; i < D; ++i) gc[i] = global_real[i];
  gc[D - 1] = global_real[D - 1] / 2 + 1;
  for (int i = 0; i < D - 1; ++i)
    orc_decompose(global_real[i], dims[i], coords[i], &local_n[i], &local_s[i]);
  for (int i = 0; i < D - 1; ++i) real_start[i] = local_s[i];
  real_start[D - 1] = 0;
  for (int a = 0; a < D - 1; ++a) {
    in_shapes[a] = local_n[a];
    out_shapes[a] = local_n[a];
  }
  in_shapes[D - 1] = global_real[D - 1];
  out_shapes[D - 1] = gc[D - 1];
  for (int stage = 1; stage < D; ++stage) {
    for (int axis = 0; axis < D; ++axis) {
      int v = 0;
      if (axis < D - 1 - stage) {
        v = local_n[axis];
      } else if (axis == D - 1 - stage) {
        v = gc[axis];
      } else if (axis > D - 1 - stage) {
        int grid_dim = (D - 1 - stage) + (axis - (D - stage));
        if (grid_dim >= 0 && grid_dim < D - 1) {
          int n, s;
          orc_decompose(gc[axis], dims[grid_dim], coords[grid_dim], &n, &s);
          v = n;
        } else {
          v = gc[axis];
        }
      }
      in_shapes[stage * D + axis] = v;
      out_shapes[stage * D + axis] = v;
    }
  }
  complex_start[0] = 0;
  for (int i = 1; i < D; ++i) {
    int n, s;
    orc_decompose(gc[i], dims[i - 1], coords[i - 1], &n, &s);
    complex_start[i] = s;
  }
}

/*
 * init_exchange_geometry (parafaft_common.hpp:345-393).  Outputs per peer:
 * src_n, src_s, dst_n, dst_s, send_counts, send_displs, recv_counts,
 * recv_displs; scalars in lt[6] = {src_leading, src_trailing, dst_leading,
 * dst_trailing, src_axis_extent, dst_axis_extent}.
 */
void orc_exchange_geometry(int nparts, int ndims, const int *src_sizes,
                           int src_axis, const int *dst_sizes, int dst_axis,
                           int *src_n, int *src_s, int *dst_n, int *dst_s,
                           uint64_t *send_counts, uint64_t *send_displs,
                           uint64_t *recv_counts, uint64_t *recv_displs,
                           uint64_t *lt) {
  uint64_t sl = 1, st = 1, dl = 1, dt = 1;
  for (int i = 0; i < src_axis; ++i) sl *= (uint64_t)src_sizes[i];
  for (int i = src_axis + 1; i < ndims; ++i) st *= (uint64_t)src_sizes[i];
  for (int i = 0; i < dst_axis; ++i) dl *= (uint64_t)dst_sizes[i];
  for (int i = dst_axis + 1; i < ndims; ++i) dt *= (uint64_t)dst_sizes[i];
  uint64_t off = 0;
  for (int p = 0; p < nparts; ++p) {
    orc_decompose(src_sizes[src_axis], nparts, p, &src_n[p], &src_s[p]);
    uint64_t c = sl * (uint64_t)src_n[p] * st;
    send_counts[p] = c; send_displs[p] = off; off += c;
  }
  off = 0;
  for (int p = 0; p < nparts; ++p) {
    orc_decompose(dst_sizes[dst_axis], nparts, p, &dst_n[p], &dst_s[p]);
    uint64_t c = dl * (uint64_t)dst_n[p] * dt;
    recv_counts[p] = c; recv_displs[p] = off; off += c;
  }
  lt[0] = sl; lt[1] = st; lt[2] = dl; lt[3] = dt;
  lt[4] = (uint64_t)src_sizes[src_axis];
  lt[5] = (uint64_t)dst_sizes[dst_axis];
}

/*
 * Pack loop of exchange_packed (parafaft_common.hpp:576-590): for each peer
 * p, src_leading rows of src_n[p]*src_trailing elements, read from
 * src + src_s[p]*src_trailing with pitch src_extent*src_trailing, written
 * contiguously at pack + send_displs[p].  esz = element size in bytes.
 */
void orc_pack(void *pack, const void *src, int nparts, const int *src_n,
              const int *src_s, const uint64_t *send_displs, uint64_t src_leading,
              uint64_t src_trailing, uint64_t src_extent, uint64_t esz) {
  char *d = (char *)pack;
  const char *s = (const char *)src;
  uint64_t pitch = src_extent * src_trailing * esz;
  for (int p = 0; p < nparts; ++p) {
    uint64_t row = (uint64_t)src_n[p] * src_trailing * esz;
    char *dp = d + send_displs[p] * esz;
    const char *sp = s + (uint64_t)src_s[p] * src_trailing * esz;
    for (uint64_t r = 0; r < src_leading; ++r)
      memcpy(dp + r * row, sp + r * pitch, row);
  }
}

/* Unpack loop (parafaft_common.hpp:654-668), the mirror of orc_pack. */
void orc_unpack(void *dst, const void *recv, int nparts, const int *dst_n,
                const int *dst_s, const uint64_t *recv_displs,
                uint64_t dst_leading, uint64_t dst_trailing, uint64_t dst_extent,
                uint64_t esz) {
  char *d = (char *)dst;
  const char *s = (const char *)recv;
  uint64_t pitch = dst_extent * dst_trailing * esz;
  for (int p = 0; p < nparts; ++p) {
    uint64_t row = (uint64_t)dst_n[p] * dst_trailing * esz;
    char *dp = d + (uint64_t)dst_s[p] * dst_trailing * esz;
    const char *sp = s + recv_displs[p] * esz;
    for (uint64_t r = 0; r < dst_leading; ++r)
      memcpy(dp + r * pitch, sp + r * row, row);
  }
}

/*
 * Independent check of the pack layout: serialise the C-order subarray of
 * partition p (MPI_Type_create_subarray with MPI_ORDER_C, as built by
 * subarray() at parafaft_common.hpp:108-160) element by element, walking the
 * full D-dim multi-index.  Writes the elements in MPI's wire order to out.
 */
uint64_t orc_subarray_serialize(void *out, const void *src, int ndims,
                                const int *sizes, int axis, int nparts, int p,
                                uint64_t esz) {
  int sub[16], start[16], idx[16];
  for (int i = 0; i < ndims; ++i) { sub[i] = sizes[i]; start[i] = 0; idx[i] = 0; }
  int n, s;
  orc_decompose(sizes[axis], nparts, p, &n, &s);
  sub[axis] = n; start[axis] = s;
  uint64_t total = 1;
  for (int i = 0; i < ndims; ++i) total *= (uint64_t)sub[i];
  char *o = (char *)out;
  const char *x = (const char *)src;
  for (uint64_t e = 0; e < total; ++e) {
    uint64_t flat = 0;
    for (int i = 0; i < ndims; ++i) flat = flat * (uint64_t)sizes[i] + (uint64_t)(start[i] + idx[i]);
    memcpy(o + e * esz, x + flat * esz, esz);
    for (int i = ndims - 1; i >= 0; --i) {
      if (++idx[i] < sub[i]) break;
      idx[i] = 0;
    }
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* One-rank stage loop (ParaFaFT_C2C::forward/backward and ParaFaFT_R2C).    */
/* ------------------------------------------------------------------------ */

/*
 * ParaFaFT_C2C<D>::forward / backward on one rank, complex<double>.
 * Plan geometry per stage from create_backend_plans (parafaft_c2c.hpp:249-287),
 * execution per perform_fft_on_buffer (:866-897).  On one rank every
 * exchange_local is an identity copy (parafaft_common.hpp:528-535), so the
 * data stays in place.  dir: -1 forward, +1 backward.
 */
int orc_parafaft_c2c(int D, const int *shape, double *data, int dir, int nthreads) {
  if (D < 1 || D > 16) return -1;
  int stages[16];
  if (dir < 0) for (int s = 0; s < D; ++s) stages[s] = s;
  else for (int s = 0; s < D; ++s) stages[s] = D - 1 - s;
  cplx *x = (cplx *)data;
  for (int si = 0; si < D; ++si) {
    int stage = stages[si];
    int axis = D - 1 - stage;
    int64_t length = shape[axis];
    int64_t lead = 1, trail = 1;
    for (int i = 0; i < axis; ++i) lead *= shape[i];
    for (int i = axis + 1; i < D; ++i) trail *= shape[i];
    if (axis == D - 1) {
      orc_c2c_batched(data, length, lead, 1, length, dir, nthreads);
    } else if (axis == 0) {
      orc_c2c_batched(data, length, trail, trail, 1, dir, nthreads);
    } else {
      for (int64_t i = 0; i < lead; ++i)
        orc_c2c_batched((double *)(x + i * length * trail), length, trail, trail, 1,
                        dir, nthreads);
    }
  }
  return 0;
}

/*
 * ParaFaFT_R2C<D>::forward_in_place / backward_in_place on one rank
 * (parafaft_r2c.hpp:360-430, 473-542): stage 0 is the padded in-place R2C
 * along the last axis (dist = 2(n/2+1) scalars, :1175-1178); stages 1..D-1 are
 * C2C along axes D-2..0 of the complex shape [..., n/2+1].
 * dir: -1 forward (R2C first), +1 backward (C2R last).
 */
int orc_parafaft_r2c(int D, const int *real_shape, double *padded, int dir,
                     int nthreads) {
  if (D < 1 || D > 16) return -1;
  int cs[16];
  for (int i = 0; i < D; ++i) cs[i] = real_shape[i];
  cs[D - 1] = real_shape[D - 1] / 2 + 1;
  int64_t batch0 = 1;
  for (int i = 0; i < D - 1; ++i) batch0 *= real_shape[i];
  int64_t pd = 2 * (int64_t)cs[D - 1];
  if (dir < 0) orc_r2c_inplace(padded, real_shape[D - 1], batch0, 1, pd, nthreads);
  cplx *x = (cplx *)padded;
  for (int k = 1; k < D; ++k) {
    int stage = (dir < 0) ? k : D - k;
    int axis = D - 1 - stage;
    int64_t length = cs[axis];
    int64_t lead = 1, trail = 1;
    for (int i = 0; i < axis; ++i) lead *= cs[i];
    for (int i = axis + 1; i < D; ++i) trail *= cs[i];
    if (axis == 0) {
      orc_c2c_batched(padded, length, trail, trail, 1, dir, nthreads);
    } else {
      for (int64_t i = 0; i < lead; ++i)
        orc_c2c_batched((double *)(x + i * length * trail), length, trail, trail, 1,
                        dir, nthreads);
    }
  }
  if (dir > 0) orc_c2r_inplace(padded, real_shape[D - 1], batch0, 1, pd, nthreads);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Deterministic generators                                                  */
/* ------------------------------------------------------------------------ */

/* SplitMix64 -> u in [0,1) with 53 bits -> 2u-1 in [-1,1). */
void orc_uniform(double *out, uint64_t n, uint64_t seed) {
  uint64_t s = seed;
  for (uint64_t i = 0; i < n; ++i) {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    out[i] = 2.0 * u - 1.0;
  }
}

/*
 * test_helpers.hpp:106-221 generators on a D-dim hypercube of side N, real
 * values (shape_id 0 Gaussian sigma=4, 1 step sphere r=N/4, 2 LCG polynomial
 * with seed 42+7D+13N).  out has N^D doubles (row-major).
 */
int orc_generate_shape(double *out, int D, int N, int shape_id) {
  uint64_t total = 1;
  for (int d = 0; d < D; ++d) total *= (uint64_t)N;
  double coeffs[1 + 3 * 16];
  int nc = 1 + 3 * D;
  unsigned int seed = 42u + (unsigned)D * 7u + (unsigned)N * 13u;
  for (int i = 0; i < nc; ++i) {
    seed = seed * 1103515245u + 12345u;
    coeffs[i] = ((seed >> 16) & 0x7FFF) / 32768.0 - 0.5;
  }
  const double center = N / 2.0, sigma = 4.0, radius = N / 4.0;
  int idx[16] = {0};
  for (uint64_t e = 0; e < total; ++e) {
    double v;
    if (shape_id == 0 || shape_id == 1) {
      double r2 = 0.0;
      for (int d = 0; d < D; ++d) { double x = idx[d] - center; r2 += x * x; }
      v = (shape_id == 0) ? exp(-r2 / (2.0 * sigma * sigma))
                          : (r2 <= radius * radius ? 1.0 : 0.0);
    } else {
      double inv = 1.0 / N;
      v = coeffs[0];
      for (int d = 0; d < D; ++d) {
        double x = idx[d] * inv;
        v += coeffs[1 + d] * x;
        v += coeffs[1 + D + d] * x * x;
        v += coeffs[1 + 2 * D + d] * x * x * x;
      }
    }
    out[e] = v;
    for (int d = D - 1; d >= 0; --d) {
      if (++idx[d] < N) break;
      idx[d] = 0;
    }
  }
  return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
#endif /* ORC_FFT_ONLY */
