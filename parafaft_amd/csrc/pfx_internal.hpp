// pfx_internal.hpp -- kernel registry shared by the instantiation units and
// the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "pfx_stockham.hpp"

namespace pfx {

using LaunchFn = void (*)(const StockhamArgs &, hipStream_t);
using GroupLaunchFn = void (*)(const GroupArgs &, int64_t blocks, hipStream_t);

struct Radices {
  int n;
  int r[8];
};
template <int... Rs> constexpr Radices radices() { return Radices{(int)sizeof...(Rs), {Rs...}}; }

struct KernelEntry {
  int prec;       // PFX_F32 / PFX_F64 / PFX_F32_SPLIT_F16 / PFX_F32_SPLIT_BF16
  int kind;       // KIND_C2C / KIND_R2C / KIND_C2R
  int n;          // complex FFT length held in LDS (R2C/C2R: half the real length)
  int col;        // 1: batch-contiguous layout, 0: row layout
  int tpt;        // threads per transform
  int per_block;  // transforms per workgroup
  int lds_bytes;
  int preferred;  // priority among entries with the same key (highest is the default)
  const char *radices;
  LaunchFn launch;
  Radices rad;    // pass radices, first pass first
  GroupLaunchFn launch_group;  // null: no grouped form
  LaunchFn launch_seg;         // segmented form (fused exchange); null: none
};

template <int... Rs> constexpr int prod() { return (Rs * ... * 1); }

// Plain launchers (in place, or out of place through StockhamArgs::src).
// The grouped and segmented launchers live in pfx_launch_extra.hpp and are
// compiled in units of their own: a unit that holds two kernels built from
// the same body (plain and grouped, or plain and segmented) was measured to
// allocate registers differently and spill in the plain kernel.
template <typename T, int TPT, int C, bool COL, int... Rs>
void launch_c2c(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2c_kernel<T, TPT, C, COL, false, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
void launch_c2c_wide(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2c_wide_kernel<T, TPT, C, ROWL, false, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
void launch_c2c_wide_nt(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2c_wide_nt_kernel<T, TPT, C, ROWL, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}
template <typename T, int TPT, int C, int... Rs>
void launch_r2c(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  r2c_kernel<T, TPT, C, false, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}
template <typename T, int TPT, int C, int... Rs>
void launch_c2r(const StockhamArgs &a, hipStream_t s) {
  const int64_t nb = (a.total + C - 1) / C;
  c2r_kernel<T, TPT, C, Rs...><<<dim3((unsigned)nb), dim3(TPT * C), 0, s>>>(a);
}

template <typename T> constexpr int prec_of();
template <> constexpr int prec_of<float>() { return 0; }
template <> constexpr int prec_of<double>() { return 1; }

// Registry tables, one per instantiation unit.  The grouped and segmented
// launcher lists run parallel to the entry list of the same name (same
// order and count); the registry merges them into the entries.
const KernelEntry *table_f32_row(int *count);
const KernelEntry *table_f32_col(int *count);
const KernelEntry *table_f32_real(int *count);
const KernelEntry *table_f64_row(int *count);
const KernelEntry *table_f64_col(int *count);
const KernelEntry *table_f64_real(int *count);
const KernelEntry *table_split(int *count);
const KernelEntry *table_split32(int *count);
const KernelEntry *table_f32_w3232(int *count);
const KernelEntry *table_nt(int *count);
const GroupLaunchFn *group_nt(int *count);
const KernelEntry *table_f32_w3232nt(int *count);
const GroupLaunchFn *group_f32_w3232(int *count);
const LaunchFn *seg_f32_w3232(int *count);
const GroupLaunchFn *group_f32_row(int *count);
const GroupLaunchFn *group_f32_col(int *count);
const GroupLaunchFn *group_f32_real(int *count);
const GroupLaunchFn *group_f64_row(int *count);
const GroupLaunchFn *group_f64_col(int *count);
const GroupLaunchFn *group_f64_real(int *count);
const LaunchFn *seg_f32_row(int *count);
const LaunchFn *seg_f32_col(int *count);
const LaunchFn *seg_f32_real(int *count);
const LaunchFn *seg_f64_row(int *count);
const LaunchFn *seg_f64_col(int *count);
const LaunchFn *seg_f64_real(int *count);

// Fallback kernels (any length that fits LDS): direct O(n^2) DFT.
struct NaiveArgs {
  void *data;
  const void *tw;   // W_n^i, i < n (forward sign)
  int64_t total, inner, stride, dist, dist_outer;
  int n;
  int kind;         // KIND_*
  int backward;
  int use_scale;
  double scale;
};
void launch_naive(int prec, const NaiveArgs &a, hipStream_t s);
int naive_max_length(int prec);

// Twiddle table generation on the device: W_L^i = exp(-2 pi i i / L), i < L,
// computed with sincospi in double and rounded to the table precision.
hipError_t make_twiddles(int prec, int64_t L, void *dst, hipStream_t s);
// Per-pass Stockham table for an FFT of length N = prod(rad): entry
// m * Ns + kk - 1 = W_{Ns R}^{kk m} for the pass (Ns, R), N - 1 entries.
hipError_t make_pass_twiddles(int prec, const Radices &rad, void *dst, hipStream_t s);

// Copy kernels.
struct PackSeg {
  uint64_t dst_off, src_off;  // bytes
  uint64_t row_bytes;
};
hipError_t launch_pack(void *dst, const void *src, const PackSeg *segs, int nseg,
                       uint64_t rows, uint64_t dst_pitch_mode, uint64_t src_pitch_mode,
                       uint64_t dst_pitch, uint64_t src_pitch, hipStream_t s);
hipError_t launch_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch,
                         size_t width, size_t height, hipStream_t s);
hipError_t launch_scale(void *data, int64_t count, int prec, double s, hipStream_t st);

// Element-wise passes of the four-step plan for long lengths (pfx_fourstep.hip):
// transform b of `batch` at b * dist, element j at j * stride, n = n1 * n2.
struct FourStepArgs {
  int64_t n1, n2, stride, batch, dist;
  int backward;
  int use_scale;
  double scale;
};
// buf[j1 + n1 k2] *= W_n^(+-j1 k2), tw: W_n^i, i < n
hipError_t launch_fourstep_twiddle(int prec, void *buf, const void *tw, const FourStepArgs &a,
                                   hipStream_t s);
// dst[n2 k1 + k2] = scale * src[k1 + n1 k2]
hipError_t launch_fourstep_transpose(int prec, const void *src, void *dst,
                                     const FourStepArgs &a, hipStream_t s);
// Fused first pass of the four-step plan for batch-contiguous callers
// (element k of transform b at k st + b): dst[(n2 j1 + k2) st + b] =
// W_n^(+-j1 k2) * DFT_n2 over j2 of src[(j1 + n1 j2) st + b]; tw: W_n^i,
// i < n; ptw: the per-pass table of the n2-point radix plan.
struct FsColArgs {
  const void *src;
  void *dst;
  const void *ptw;
  const void *tw;
  int64_t n1, n2, st, batch;
  int64_t tiles_per_j1;  // set by the launcher
  int backward;
};
// The radices of the fused first pass for n2; false: none for this n2.
bool fourstep_cols_pass1_radices(int64_t n2, Radices *rad);
hipError_t launch_fourstep_cols_pass1(int prec, const FsColArgs &a, hipStream_t s);
// Two-pass four-step of rows (row r at r dist, element stride st): pass 1
// the length-n2 DFTs times W_n^(j1 k2) from the caller's array into the
// scratch at the same offsets, pass 2 the length-n1 DFTs from the scratch,
// stored transposed into the caller's array with the scale.
struct FsRowArgs {
  const void *src;  // pass 1: the caller's array; pass 2: the scratch
  void *dst;        // pass 1: the scratch; pass 2: the caller's array
  const void *ptw;  // the pass's per-pass twiddle table (n2, resp. n1)
  const void *tw;   // pass 1: W_n^m, m < n
  const void *twc;  // pass 1: W_n^(c k2) at 16 k2 + c, c < 16, k2 < n2
  int64_t n1, n2, st, dist, rows;
  int64_t tiles;    // tiles per row, set by the launcher
  int backward;
  double scale;     // pass 2
};
// The radices of either pass for length n; false: none (four passes).
bool fourstep_rows_radices(int64_t n, Radices *rad);
// The first row pass's table W_n^(c k2) (16 n2 entries) from the table of n.
hipError_t make_fourstep_rows_table(int prec, void *out, const void *tw, int64_t n2, hipStream_t s);
hipError_t launch_fourstep_rows_pass(int prec, int pass, const FsRowArgs &a, hipStream_t s);
// Long R2C post-pass / C2R pre-pass on padded rows (row b at b * dist complex
// slots, h + 1 slots each); tw: W_2h^k, k <= h.
struct RealPassArgs {
  int64_t h, batch, dist;
  int use_scale;
  double scale;
};
hipError_t launch_real_post(int prec, void *data, const void *tw, const RealPassArgs &a,
                            hipStream_t s);
hipError_t launch_real_pre(int prec, void *data, const void *tw, const RealPassArgs &a,
                           hipStream_t s);
// Bluestein (pfx_fourstep.hip): `total` transforms from t0 on, transform t
// at (t / inner) dist_outer + (t % inner) dist, element j at j * stride;
// scratch rows of M >= 2n - 1 (row 0 for transform t0).
struct ChirpArgs {
  int64_t n, M, t0, total, inner, dist, dist_outer, stride;
  int backward;
  double scale;
};
// w[j] = exp(-pi i j^2 / n), j < n
hipError_t make_chirp(int prec, void *w, int64_t n, hipStream_t s);
// b[m] = b[M - m] = conj w[m] (backward: w[m]), m < n; zero between
hipError_t make_chirp_filter(int prec, void *b, const void *w, int64_t n, int64_t M,
                             int backward, hipStream_t s);
hipError_t launch_chirp_pre(int prec, const void *x, void *scr, const void *w,
                            const ChirpArgs &a, hipStream_t s);
hipError_t launch_chirp_mul(int prec, void *scr, const void *B, const ChirpArgs &a,
                            hipStream_t s);
hipError_t launch_chirp_post(int prec, const void *scr, void *x, const void *w,
                             const ChirpArgs &a, hipStream_t s);
// The chirp convolution in one kernel for a convolution length M with a
// tiled radix plan (pfx_chirp.inc): x in place, w the chirp (n), B the
// forward filter's spectrum (M), ptw the per-pass table of `rad`; t0 = 0,
// total = every transform of the plan; row: transforms along rows
// (row-layout tiles), col: batch-contiguous tiles.
using ChirpFusedFn = void (*)(void *x, const void *w, const void *B, const void *ptw,
                              const ChirpArgs &a, hipStream_t s);
struct ChirpFused {
  int M;
  Radices rad;
  ChirpFusedFn row, col;
};
const ChirpFused *chirp_fused_f32(int64_t M);
const ChirpFused *chirp_fused_f64(int64_t M);
// Odd real lengths through a length-n C2C on scratch rows of n: which = 0
// R2C load (reals -> complex), 1 R2C store (X[0..n/2] -> padded row), 2 C2R
// load (Hermitian extension), 3 C2R store (real parts -> row).  Row b at
// scalar b * dist.
struct RealFullArgs {
  int64_t n, batch, dist;
  int use_scale;
  double scale;
};
hipError_t launch_real_full(int prec, int which, void *rows, void *scr, const RealFullArgs &a,
                            hipStream_t s);

// Split axis 0 of the one-GPU N-D plan (pfx_planerow.hip), n0 = q G: for
// x1 in [x1_begin, x1_begin + nx1) and every row y of the planes, the G-point
// DFTs along the planes x1 in_a + m in_b (m < G) of row y, the twiddle
// W_n0^(x1 m) on output m, and the FFT of each row along the last axis,
// stored to planes x1 out_a + m out_b (backward: the same, conjugated through
// the re/im swap).  A plane holds `plane` elements, `rows` rows of the last
// axis.
struct PlaneRowArgs {
  const void *src;
  void *dst;
  const void *tw;   // W_n0^i, i < n0
  const void *ptw;  // per-pass table of the last axis' radices (plane_row_radices)
  int64_t plane, rows;
  int64_t x1_begin, nx1;
  int64_t in_a, in_b, out_a, out_b;
  int backward;
};
// The last-axis radices of the fused kernel for length n and G planes
// (prec: 0 FP32, 1 FP64); false: none built.
bool plane_row_radices(int prec, int64_t n, int g, Radices *rad);
hipError_t launch_plane_row(int prec, int64_t n, int g, const PlaneRowArgs &a, hipStream_t s);

}  // namespace pfx
