// Unit tests of MI355Backend, mirroring test/unit/test_hipfft_backend.hip:
// contiguous layout, strided layout, multiple stages, one plan on several
// pointers, R2C/C2R in place -- roundtrips at 1e-10 (double) as there, plus a
// direct comparison with the oracle.  Unlike the reference's unit tests this
// executable returns the number of failures.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "orc.h"
#include "parafaft_mi355/fft_backend_mi355.hpp"

using parafaft::FFTDirection;

static int failures = 0;

static void check(bool ok, const char *what, double err) {
  std::printf("  %-44s err=%.3e %s\n", what, err, ok ? "[PASS]" : "[FAIL]");
  if (!ok) ++failures;
}

template <typename T> using cvec = std::vector<std::complex<T>>;

template <typename T> static void h2d(void *d, const cvec<T> &h) {
  parafaft::check_pfx(pfx_memcpy_async(d, h.data(), h.size() * sizeof(h[0]), nullptr), "h2d");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
}
template <typename T> static void d2h(cvec<T> &h, const void *d) {
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  parafaft::check_pfx(pfx_memcpy_async(h.data(), d, h.size() * sizeof(h[0]), nullptr), "d2h");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
}

template <typename T> static double max_err(const cvec<T> &a, const std::vector<std::complex<double>> &b,
                                            double scale = 1.0) {
  double e = 0;
  for (size_t i = 0; i < a.size(); ++i)
    e = std::max(e, std::abs(std::complex<double>(a[i]) * scale - b[i]));
  return e;
}

template <typename T> static double tol() { return sizeof(T) == 8 ? 1e-10 : 2e-5; }

// Test 1: contiguous layout (stride=1, dist=N), Gaussian rows, roundtrip
template <typename T> static void test_contiguous() {
  const int N = 256, batch = 3;
  cvec<T> h(N * batch);
  for (int b = 0; b < batch; ++b)
    for (int i = 0; i < N; ++i) {
      double x = (i - N / 2) / 16.0;
      h[b * N + i] = std::complex<T>(std::exp(-x * x / 2.0), 0);
    }
  std::vector<std::complex<double>> orig(h.begin(), h.end());
  parafaft::MI355Backend<T> be(1);
  typename parafaft::MI355Backend<T>::ComplexBuffer d(N * batch);
  h2d(d.data(), h);
  be.create_stage_plan(0, N, batch, d.data(), 1, N);
  be.execute_stage(0, FFTDirection::Forward, d.data());
  be.sync();
  cvec<T> f(N * batch);
  d2h(f, d.data());
  std::vector<std::complex<double>> ref(orig);
  orc_c2c_batched(reinterpret_cast<double *>(ref.data()), N, batch, 1, N, -1, 1);
  check(max_err(f, ref) < tol<T>() * N, "contiguous forward vs oracle", max_err(f, ref));
  be.execute_stage(0, FFTDirection::Backward, d.data());
  be.sync();
  d2h(h, d.data());
  check(max_err(h, orig, 1.0 / N) < tol<T>(), "contiguous roundtrip", max_err(h, orig, 1.0 / N));
}

// Test 2: strided layout (stride=batch, dist=1)
template <typename T> static void test_strided() {
  const int N = 128, batch = 4;
  cvec<T> h(N * batch);
  for (int i = 0; i < N; ++i)
    for (int b = 0; b < batch; ++b) {
      double ph = 2.0 * M_PI * b / batch + 0.01 * i;
      h[i * batch + b] = std::complex<T>(std::cos(ph), std::sin(ph));
    }
  std::vector<std::complex<double>> orig(h.begin(), h.end());
  parafaft::MI355Backend<T> be(1);
  typename parafaft::MI355Backend<T>::ComplexBuffer d(N * batch);
  h2d(d.data(), h);
  be.create_stage_plan(0, N, batch, d.data(), batch, 1);
  be.execute_stage(0, FFTDirection::Forward, d.data());
  be.sync();
  cvec<T> f(N * batch);
  d2h(f, d.data());
  std::vector<std::complex<double>> ref(orig);
  orc_c2c_batched(reinterpret_cast<double *>(ref.data()), N, batch, batch, 1, -1, 1);
  check(max_err(f, ref) < tol<T>() * N, "strided forward vs oracle", max_err(f, ref));
  be.execute_stage(0, FFTDirection::Backward, d.data());
  be.sync();
  d2h(h, d.data());
  check(max_err(h, orig, 1.0 / N) < tol<T>(), "strided roundtrip", max_err(h, orig, 1.0 / N));
}

// Test 3: several stages on one backend (3D 16x32x64, full forward+backward)
template <typename T> static void test_multiple_stages() {
  const int n0 = 16, n1 = 32, n2 = 64, tot = n0 * n1 * n2;
  std::vector<double> u(2 * tot);
  orc_uniform(u.data(), u.size(), 77);
  cvec<T> h(tot);
  for (int i = 0; i < tot; ++i) h[i] = std::complex<T>(T(u[2 * i]), T(u[2 * i + 1]));
  std::vector<std::complex<double>> orig(h.begin(), h.end());
  parafaft::MI355Backend<T> be(3);
  typename parafaft::MI355Backend<T>::ComplexBuffer d(tot);
  h2d(d.data(), h);
  be.create_stage_plan(0, n2, n0 * n1, d.data(), 1, n2);
  be.create_stage_plan(1, n1, n2, d.data(), n2, 1);
  be.create_stage_plan(2, n0, n1 * n2, d.data(), n1 * n2, 1);
  be.execute_stage(0, FFTDirection::Forward, d.data());
  for (int i = 0; i < n0; ++i) be.execute_stage(1, FFTDirection::Forward, d.data() + i * n1 * n2);
  be.execute_stage(2, FFTDirection::Forward, d.data());
  be.sync();
  cvec<T> f(tot);
  d2h(f, d.data());
  std::vector<std::complex<double>> ref(orig);
  int shape[3] = {n0, n1, n2};
  orc_parafaft_c2c(3, shape, reinterpret_cast<double *>(ref.data()), -1, 0);
  check(max_err(f, ref) < tol<T>() * 300, "3 stages forward vs oracle", max_err(f, ref));
  be.execute_stage(2, FFTDirection::Backward, d.data());
  for (int i = 0; i < n0; ++i) be.execute_stage(1, FFTDirection::Backward, d.data() + i * n1 * n2);
  be.execute_stage(0, FFTDirection::Backward, d.data());
  be.sync();
  d2h(h, d.data());
  check(max_err(h, orig, 1.0 / tot) < tol<T>(), "3 stages roundtrip", max_err(h, orig, 1.0 / tot));
}

// Test 3b: Measure planning (candidate kernels timed on the planning buffer)
static void test_measure_planning() {
  const int n0 = 8, n1 = 1024, n2 = 64, tot = n0 * n1 * n2;
  std::vector<double> u(2 * tot);
  orc_uniform(u.data(), u.size(), 78);
  std::vector<std::complex<float>> h(tot);
  for (int i = 0; i < tot; ++i) h[i] = std::complex<float>(float(u[2 * i]), float(u[2 * i + 1]));
  std::vector<std::complex<double>> ref(h.begin(), h.end());
  parafaft::MI355Backend<float> be(1, parafaft::FFTPlanFlag::Measure);
  typename parafaft::MI355Backend<float>::ComplexBuffer scratch(tot), d(tot);
  const int before = pfx_tune_cache_size();
  be.create_stage_plan(0, n1, n2, scratch.data(), n2, 1);  // middle-axis geometry, tuned
  h2d(d.data(), h);
  for (int i = 0; i < n0; ++i) be.execute_stage(0, FFTDirection::Forward, d.data() + i * n1 * n2);
  be.sync();
  d2h(h, d.data());
  for (int i = 0; i < n0; ++i)
    orc_c2c_batched(reinterpret_cast<double *>(ref.data() + i * n1 * n2), n1, n2, n2, 1, -1, 1);
  double e = max_err(h, ref);
  check(e < 2e-5 * n1 && pfx_tune_cache_size() == before + 1, "Measure planning: tuned + correct", e);
}

// Test 4: one plan applied to different pointers (new-array execute)
template <typename T> static void test_different_pointers() {
  const int N = 64, batch = 2;
  parafaft::MI355Backend<T> be(1);
  typename parafaft::MI355Backend<T>::ComplexBuffer a(N * batch), b(N * batch);
  be.create_stage_plan(0, N, batch, a.data(), 1, N);
  double worst = 0;
  for (int which = 0; which < 2; ++which) {
    std::vector<double> u(2 * N * batch);
    orc_uniform(u.data(), u.size(), 100 + which);
    cvec<T> h(N * batch);
    for (int i = 0; i < N * batch; ++i) h[i] = std::complex<T>(T(u[2 * i]), T(u[2 * i + 1]));
    std::vector<std::complex<double>> ref(h.begin(), h.end());
    auto *p = which ? b.data() : a.data();
    h2d(p, h);
    be.execute_stage(0, FFTDirection::Forward, p);
    be.sync();
    d2h(h, p);
    orc_c2c_batched(reinterpret_cast<double *>(ref.data()), N, batch, 1, N, -1, 1);
    worst = std::max(worst, max_err(h, ref));
  }
  check(worst < tol<T>() * N, "same plan, two buffers", worst);
}

// Test 5: R2C / C2R in place on padded rows
template <typename T> static void test_r2c_c2r(const int N = 64) {
  const int batch = 5, h = N / 2 + 1, pd = 2 * h;
  std::vector<double> u(N * batch);
  orc_uniform(u.data(), u.size(), 5);
  std::vector<T> pad(pd * batch, T(0));
  std::vector<double> ref(pd * batch, 0.0);
  for (int b = 0; b < batch; ++b)
    for (int i = 0; i < N; ++i) {
      pad[b * pd + i] = T(u[b * N + i]);
      ref[b * pd + i] = double(pad[b * pd + i]);
    }
  parafaft::MI355Backend<T> be(1);
  typename parafaft::MI355Backend<T>::Buffer d(pd * batch);
  parafaft::check_pfx(pfx_memcpy_async(d.data(), pad.data(), pad.size() * sizeof(T), nullptr), "h2d");
  be.create_r2c_inplace_plan(N, batch, d.data(), 1, pd);
  be.create_c2r_inplace_plan(N, batch, d.data(), 1, pd);
  pfx_stream_synchronize(nullptr);
  be.execute_r2c_inplace(d.data());
  be.sync();
  std::vector<T> out(pd * batch);
  parafaft::check_pfx(pfx_memcpy_async(out.data(), d.data(), out.size() * sizeof(T), nullptr), "d2h");
  pfx_stream_synchronize(nullptr);
  orc_r2c_inplace(ref.data(), N, batch, 1, pd, 1);
  double e = 0;
  for (int i = 0; i < pd * batch; ++i) e = std::max(e, std::abs(double(out[i]) - ref[i]));
  check(e < tol<T>() * N, "R2C in place vs oracle", e);
  be.execute_c2r_inplace(d.data());
  be.sync();
  parafaft::check_pfx(pfx_memcpy_async(out.data(), d.data(), out.size() * sizeof(T), nullptr), "d2h");
  pfx_stream_synchronize(nullptr);
  e = 0;
  for (int b = 0; b < batch; ++b)
    for (int i = 0; i < N; ++i) e = std::max(e, std::abs(double(out[b * pd + i]) / N - double(pad[b * pd + i])));
  check(e < tol<T>(), "R2C -> C2R roundtrip", e);
}

// Test 5b: lengths without a single-kernel plan go through create_stage_plan
// like any other (hipfftPlanMany accepts any length): the four-step plan for
// lengths that split into two tiled lengths, the chirp convolution else.
template <typename T> static void test_any_length(const int N) {
  const int batch = 3;
  std::vector<double> u(2 * N * batch);
  orc_uniform(u.data(), u.size(), 7 + N);
  cvec<T> h(N * batch);
  for (int i = 0; i < N * batch; ++i) h[i] = std::complex<T>(T(u[2 * i]), T(u[2 * i + 1]));
  std::vector<std::complex<double>> orig(h.begin(), h.end());
  parafaft::MI355Backend<T> be(1);
  typename parafaft::MI355Backend<T>::ComplexBuffer d(N * batch);
  h2d(d.data(), h);
  be.create_stage_plan(0, N, batch, d.data(), 1, N);
  be.execute_stage(0, FFTDirection::Forward, d.data());
  be.sync();
  cvec<T> f(N * batch);
  d2h(f, d.data());
  std::vector<std::complex<double>> ref(orig);
  orc_c2c_batched(reinterpret_cast<double *>(ref.data()), N, batch, 1, N, -1, 1);
  char what[64];
  std::snprintf(what, sizeof(what), "length %d forward vs oracle", N);
  check(max_err(f, ref) < tol<T>() * N, what, max_err(f, ref));
  be.execute_stage(0, FFTDirection::Backward, d.data());
  be.sync();
  d2h(h, d.data());
  std::snprintf(what, sizeof(what), "length %d roundtrip", N);
  check(max_err(h, orig, 1.0 / N) < tol<T>(), what, max_err(h, orig, 1.0 / N));
}

// Test 6: errors surface as std::runtime_error (check_hip style)
static void test_errors() {
  parafaft::MI355Backend<double> be(1);
  bool threw = false;
  try {
    be.create_stage_plan(0, 0, 1, nullptr, 1, 1);
  } catch (const std::runtime_error &e) {
    threw = std::strstr(e.what(), "pfx error in pfx_plan_create_c2c") != nullptr;
  }
  check(threw, "invalid length throws runtime_error", 0);
  threw = false;
  try {
    be.execute_stage(3, FFTDirection::Forward, nullptr);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  check(threw, "bad stage index throws", 0);
}

// Concept surface used by the drivers (compile-time)
static_assert(!parafaft::MI355Backend<double>::use_alltoallw, "packs on device");
static_assert(!parafaft::MI355Backend<double>::handles_distributed, "per-stage driver");
static_assert(std::is_same<parafaft::MI355Backend<float>::FloatType, float>::value, "FloatType");
static_assert(sizeof(parafaft::MI355Backend<double>::Complex) == sizeof(std::complex<double>),
              "Complex layout");
static_assert(std::is_move_constructible<parafaft::MI355Backend<double>>::value, "movable");
static_assert(!std::is_copy_constructible<parafaft::MI355Backend<double>>::value, "non-copyable");

int main() {
  std::printf("MI355Backend<double>\n");
  test_contiguous<double>();
  test_strided<double>();
  test_multiple_stages<double>();
  test_different_pointers<double>();
  test_r2c_c2r<double>();
  std::printf("MI355Backend<float>\n");
  test_contiguous<float>();
  test_strided<float>();
  test_multiple_stages<float>();
  test_different_pointers<float>();
  test_r2c_c2r<float>();
  std::printf("lengths without a single-kernel plan\n");
  for (int n : {144, 1009, 8191, 32768, 100000}) {
    test_any_length<float>(n);
    test_any_length<double>(n);
  }
  std::printf("smooth lengths and rows beyond 4096 in one kernel\n");
  for (int n : {100, 768, 1000, 3072, 8192, 16384}) {
    test_any_length<float>(n);
    test_any_length<double>(n);
  }
  test_r2c_c2r<float>(1001);
  test_r2c_c2r<double>(2 * 1009);
  test_measure_planning();
  test_errors();
  std::printf("%d failure(s)\n", failures);
  return failures;
}
