// ParaFaFT_C2C<D, MI355Backend<T>> and ParaFaFT_R2C<D, MI355Backend<T>> on one
// rank over the reference's integration matrix (test/CMakeLists.txt:215-271):
// shape in {Gaussian, step, LCG polynomial} x {D2 N32, D3 N32, D4 N16, D5 N8},
// forward compared with the oracle at the reference tolerances (max-abs <
// 1e-10 for double, test_fftw_c2c_nd.cpp:119; < 1e-4 * peak for float,
// test_fftw_c2c_nd_float.cpp:101-107), then backward with 1/N^D scaling
// against the input (test_fftw_c2c_roundtrip_nd.cpp:73-93).  Both the
// faithful driver and the one-GPU fast path run, and the split-precision
// MFMA backends (MI355Backend<float, PFX_F32_SPLIT_F16 / _BF16>) on lengths
// that have radix-16 passes.  Returns the failure count.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "orc.h"
#include "parafaft_mi355/parafaft_one_rank.hpp"

using parafaft_mi355::Mode;

static int failures = 0;
static const char *kShape[] = {"gaussian", "step", "poly"};

static void report(bool ok, const char *kind, const char *prec, int D, int N, int s, Mode m,
                   double e1, double e2) {
  std::printf("  %s %s D%d N%-3d %-8s %-8s fwd=%.2e roundtrip=%.2e %s\n", kind, prec, D, N,
              kShape[s], m == Mode::Fast ? "fast" : "faithful", e1, e2, ok ? "[PASS]" : "[FAIL]");
  if (!ok) ++failures;
}

static void sync_copy(void *dst, const void *src, size_t bytes) {
  parafaft::check_pfx(pfx_memcpy_async(dst, src, bytes, nullptr), "copy");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
}

template <typename T> constexpr int plain_mode() { return sizeof(T) == 8 ? PFX_F64 : PFX_F32; }
static const char *prec_name(int sm) {
  return sm == PFX_F64 ? "double" : sm == PFX_F32 ? "float" : sm == PFX_F32_SPLIT_F16 ? "spl-f16"
                                                                                     : "spl-bf16";
}

template <typename T, int D, int SM = plain_mode<T>()>
static void c2c_case(int N, int shape_id, Mode mode,
                     parafaft::FFTPlanFlag flag = parafaft::FFTPlanFlag::Estimate) {
  int shape[D];
  size_t tot = 1;
  for (int i = 0; i < D; ++i) {
    shape[i] = N;
    tot *= N;
  }
  std::vector<double> x(tot);
  orc_generate_shape(x.data(), D, N, shape_id);
  std::vector<std::complex<double>> ref(tot);
  for (size_t i = 0; i < tot; ++i) ref[i] = x[i];
  orc_parafaft_c2c(D, shape, reinterpret_cast<double *>(ref.data()), -1, 0);
  double peak = 0;
  for (auto &v : ref) peak = std::max(peak, std::abs(v));

  using B = parafaft::MI355Backend<T, SM>;
  parafaft_mi355::ParaFaFT_C2C<D, B> fft(shape, flag, mode);
  typename B::ComplexBuffer d(fft.get_required_output_size());
  std::vector<std::complex<T>> h(tot);
  for (size_t i = 0; i < tot; ++i) h[i] = std::complex<T>(T(x[i]), 0);
  sync_copy(d.data(), h.data(), tot * sizeof(h[0]));
  fft.forward(d.data());
  sync_copy(h.data(), d.data(), tot * sizeof(h[0]));
  double e1 = 0;
  for (size_t i = 0; i < tot; ++i) e1 = std::max(e1, std::abs(std::complex<double>(h[i]) - ref[i]));
  fft.backward(d.data(), 1.0 / double(tot));
  sync_copy(h.data(), d.data(), tot * sizeof(h[0]));
  double e2 = 0;
  for (size_t i = 0; i < tot; ++i) e2 = std::max(e2, std::abs(std::complex<double>(h[i]) - x[i]));
  const bool dbl = sizeof(T) == 8;
  const bool ok = dbl ? (e1 < 1e-10 && e2 < 1e-10) : (e1 < 1e-4 * peak && e2 < 1e-4);
  report(ok, "C2C", prec_name(SM), D, N, shape_id, mode, e1, e2);
}

template <typename T, int D, int SM = plain_mode<T>()>
static void r2c_case(int N, int shape_id, Mode mode,
                     parafaft::FFTPlanFlag flag = parafaft::FFTPlanFlag::Estimate) {
  int shape[D];
  size_t rows = 1;
  for (int i = 0; i < D; ++i) shape[i] = N;
  for (int i = 0; i < D - 1; ++i) rows *= N;
  const int h = N / 2 + 1, pd = 2 * h;
  std::vector<double> x(rows * N);
  orc_generate_shape(x.data(), D, N, shape_id);
  std::vector<double> ref(rows * pd, 0.0);
  std::vector<T> pad(rows * pd, T(0));
  for (size_t r = 0; r < rows; ++r)
    for (int i = 0; i < N; ++i) {
      pad[r * pd + i] = T(x[r * N + i]);
      ref[r * pd + i] = x[r * N + i];
    }
  orc_parafaft_r2c(D, shape, ref.data(), -1, 0);
  double peak = 0;
  for (double v : ref) peak = std::max(peak, std::abs(v));

  using B = parafaft::MI355Backend<T, SM>;
  parafaft_mi355::ParaFaFT_R2C<D, B> fft(shape, flag, mode);
  {  // one rank of a (D-1)-dimensional grid (parafaft_r2c.hpp:712-790)
    int dec[D], coords[D > 1 ? D - 1 : 1];
    fft.get_domain_decomposition(dec);
    fft.get_grid_coords(coords);
    bool acc_ok = decltype(fft)::get_grid_ndims() == D - 1 && fft.get_buffer() == nullptr &&
                  fft.get_real_buffer() == nullptr;
    for (int i = 0; i < D; ++i) acc_ok = acc_ok && dec[i] == 1;
    for (int i = 0; i < D - 1; ++i) acc_ok = acc_ok && coords[i] == 0;
    if (!acc_ok) {
      std::printf("    R2C accessors wrong\n");
      ++failures;
    }
  }
  typename B::Buffer d(fft.get_required_output_size());
  sync_copy(d.data(), pad.data(), pad.size() * sizeof(T));
  fft.forward_in_place(d.data());
  std::vector<T> out(pad.size());
  sync_copy(out.data(), d.data(), out.size() * sizeof(T));
  double e1 = 0;
  for (size_t i = 0; i < out.size(); ++i) e1 = std::max(e1, std::abs(double(out[i]) - ref[i]));
  fft.backward_in_place(d.data(), 1.0 / double(rows * N));
  sync_copy(out.data(), d.data(), out.size() * sizeof(T));
  double e2 = 0;
  for (size_t r = 0; r < rows; ++r)
    for (int i = 0; i < N; ++i) e2 = std::max(e2, std::abs(double(out[r * pd + i]) - x[r * N + i]));
  // out-of-place forms (parafaft_r2c.hpp:333-339, 446-453) give the same
  // bits as the padding copy + in-place transform above
  bool oop_ok = true;
  {
    std::vector<T> real(rows * N);
    for (size_t i = 0; i < real.size(); ++i) real[i] = T(x[i]);
    typename B::Buffer dr(real.size());
    typename B::ComplexBuffer dc(fft.get_required_output_size() / 2);
    sync_copy(dr.data(), real.data(), real.size() * sizeof(T));
    fft.forward(dr.data(), dc.data());
    std::vector<T> spec(pad.size()), spec_ip(pad.size());
    sync_copy(spec.data(), dc.data(), spec.size() * sizeof(T));
    sync_copy(d.data(), pad.data(), pad.size() * sizeof(T));
    fft.forward_in_place(d.data());
    sync_copy(spec_ip.data(), d.data(), spec_ip.size() * sizeof(T));
    oop_ok = spec == spec_ip;
    fft.backward(dc.data(), dr.data(), 1.0 / double(rows * N));
    std::vector<T> back(real.size());
    sync_copy(back.data(), dr.data(), back.size() * sizeof(T));
    fft.backward_in_place(d.data(), 1.0 / double(rows * N));
    sync_copy(out.data(), d.data(), out.size() * sizeof(T));
    for (size_t r = 0; r < rows && oop_ok; ++r)
      for (int i = 0; i < N; ++i) oop_ok = oop_ok && back[r * N + i] == out[r * pd + i];
    if (!oop_ok) std::printf("    out-of-place forward/backward differ from copy + in place\n");
  }
  const bool dbl = sizeof(T) == 8;
  const bool ok = oop_ok && (dbl ? (e1 < 1e-10 && e2 < 1e-10) : (e1 < 1e-4 * peak && e2 < 1e-4));
  report(ok, "R2C", prec_name(SM), D, N, shape_id, mode, e1, e2);
}

// Fast mode with Measure planning (temporary planning buffer, kernels and
// schedule timed); a 1 MB chunk budget makes 64^3 complex<double> and the
// 4D case take the chunked schedule into the comparison.
template <typename T> static void measured() {
  setenv("PFX_ND_CHUNK_MB", "1", 1);
  c2c_case<T, 3>(64, 1, Mode::Fast, parafaft::FFTPlanFlag::Measure);
  c2c_case<T, 4>(16, 2, Mode::Fast, parafaft::FFTPlanFlag::Patient);
  r2c_case<T, 3>(64, 0, Mode::Fast, parafaft::FFTPlanFlag::Measure);
  unsetenv("PFX_ND_CHUNK_MB");
}

template <typename T> static void matrix() {
  measured<T>();
  for (Mode m : {Mode::Faithful, Mode::Fast})
    for (int s = 0; s < 3; ++s) {
      c2c_case<T, 2>(32, s, m);
      c2c_case<T, 3>(32, s, m);
      c2c_case<T, 4>(16, s, m);
      c2c_case<T, 5>(8, s, m);
      r2c_case<T, 2>(32, s, m);
      r2c_case<T, 3>(32, s, m);
      r2c_case<T, 4>(16, s, m);
      r2c_case<T, 5>(8, s, m);
    }
}

template <class B> struct Tag { using type = B; };

// The split backend really runs its own kernels on both driver paths: its
// forward differs bit-wise from the plain float backend's on the same input.
template <int SM, Mode M> static void split_is_not_plain() {
  constexpr int D = 3, N = 256;
  int shape[D] = {N, N, N};
  const size_t tot = size_t(N) * N * N;
  std::vector<double> x(tot);
  orc_generate_shape(x.data(), D, N, 2);
  std::vector<std::complex<float>> h(tot), r[2];
  for (size_t i = 0; i < tot; ++i) h[i] = std::complex<float>(float(x[i]), 0.f);
  auto run = [&](auto backend_tag, std::vector<std::complex<float>> &out) {
    using B = typename decltype(backend_tag)::type;
    parafaft_mi355::ParaFaFT_C2C<D, B> fft(shape, parafaft::FFTPlanFlag::Estimate, M);
    typename B::ComplexBuffer d(fft.get_required_output_size());
    sync_copy(d.data(), h.data(), tot * sizeof(h[0]));
    fft.forward(d.data());
    out.resize(tot);
    sync_copy(out.data(), d.data(), tot * sizeof(h[0]));
  };
  run(Tag<parafaft::MI355Backend<float, SM>>{}, r[0]);
  run(Tag<parafaft::MI355Backend<float>>{}, r[1]);
  size_t diff = 0;
  for (size_t i = 0; i < tot; ++i) diff += r[0][i] != r[1][i];
  const bool ok = diff > 0;
  std::printf("  C2C %s D3 N256 %s differs from plain float in %zu of %zu elements %s\n",
              prec_name(SM), M == Mode::Fast ? "fast" : "faithful", diff, tot,
              ok ? "[PASS]" : "[FAIL]");
  if (!ok) ++failures;
}

// Split-precision backends: 256-point axes run their radix-16 passes on the
// matrix cores (the R2C stage itself stays plain FP32); same tolerances as
// float.
template <int SM> static void split() {
  split_is_not_plain<SM, Mode::Faithful>();
  split_is_not_plain<SM, Mode::Fast>();
  for (Mode m : {Mode::Faithful, Mode::Fast}) {
    c2c_case<float, 2, SM>(256, 0, m);
    c2c_case<float, 3, SM>(256, 2, m);
    r2c_case<float, 3, SM>(256, 1, m);
  }
  c2c_case<float, 3, SM>(256, 1, Mode::Fast, parafaft::FFTPlanFlag::Measure);
}

int main() {
  try {
    matrix<double>();
    matrix<float>();
    split<PFX_F32_SPLIT_F16>();
    split<PFX_F32_SPLIT_BF16>();
  } catch (const std::exception &e) {
    std::printf("exception: %s\n", e.what());
    return 100;
  }
  std::printf("%d failure(s)\n", failures);
  return failures;
}
