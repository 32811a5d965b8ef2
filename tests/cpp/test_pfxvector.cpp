// The backend's device buffer type, pfxvector<T> (the hipvector<T> of
// backend/hipfft/fft_backend_hipfft.hpp:58-121), over the cases of the
// reference's test/unit/test_hipvector.hip: allocation, host<->device
// transfers of double and complex data, const access, a large allocation,
// many live allocations, integer and float element types, and resize (old
// contents discarded).  Returns the failure count.
#include <complex>
#include <cstdio>
#include <vector>

#include "parafaft_mi355/fft_backend_mi355.hpp"

using parafaft::pfxvector;

static int failures = 0;
static void report(bool ok, const char *what) {
  std::printf("  %-34s %s\n", what, ok ? "[PASS]" : "[FAIL]");
  if (!ok) ++failures;
}

template <typename T> static void put(pfxvector<T> &v, const std::vector<T> &h) {
  parafaft::check_pfx(pfx_memcpy_async(v.data(), h.data(), h.size() * sizeof(T), nullptr), "h2d");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
}
template <typename T> static std::vector<T> get(const pfxvector<T> &v) {
  std::vector<T> h(v.size());
  parafaft::check_pfx(pfx_memcpy_async(h.data(), v.data(), h.size() * sizeof(T), nullptr), "d2h");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  return h;
}

template <typename T> static bool roundtrip(std::size_t n, T (*gen)(std::size_t)) {
  pfxvector<T> v(n);
  std::vector<T> h(n);
  for (std::size_t i = 0; i < n; ++i) h[i] = gen(i);
  put(v, h);
  return get(v) == h;
}

int main() {
  try {
    {
      pfxvector<double> v(1024);
      report(v.size() == 1024 && v.data() != nullptr, "basic allocation");
      pfxvector<double> e;
      report(e.size() == 0 && e.data() == nullptr, "default-constructed is empty");
    }
    report(roundtrip<double>(4096, [](std::size_t i) { return 0.5 * double(i) - 7.0; }),
           "transfer double");
    report(roundtrip<std::complex<double>>(
               2048, [](std::size_t i) { return std::complex<double>(double(i), -double(i)); }),
           "transfer complex<double>");
    report(roundtrip<int>(1000, [](std::size_t i) { return int(i) * 3 - 500; }), "integer type");
    report(roundtrip<float>(1000, [](std::size_t i) { return float(i) / 7.f; }), "float type");
    {
      pfxvector<double> v(16);
      std::vector<double> h(16, 3.25);
      put(v, h);
      const pfxvector<double> &cv = v;
      const double *p = cv.data();
      report(p == v.data() && get(cv) == h, "const access");
    }
    {
      const std::size_t n = std::size_t(1) << 28;  // 2 GiB of double
      pfxvector<double> v(n);
      std::vector<double> h(n / 1024, 1.5);
      parafaft::check_pfx(pfx_memcpy_async(v.data() + (n - h.size()), h.data(),
                                           h.size() * sizeof(double), nullptr), "h2d");
      parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
      report(v.size() == n && v.data() != nullptr, "large allocation (2 GiB)");
    }
    {
      std::vector<pfxvector<float>> vs;
      for (int k = 0; k < 32; ++k) vs.emplace_back(std::size_t(1000 + k));
      bool ok = true;
      for (int k = 0; k < 32; ++k) ok = ok && vs[k].size() == std::size_t(1000 + k) && vs[k].data();
      for (int k = 1; k < 32; ++k) ok = ok && vs[k].data() != vs[k - 1].data();
      report(ok, "multiple allocations");
    }
    {
      pfxvector<double> v(8);
      v.resize(512);
      bool ok = v.size() == 512 && v.data() != nullptr;
      std::vector<double> h(512);
      for (std::size_t i = 0; i < h.size(); ++i) h[i] = 2.0 * double(i);
      put(v, h);
      ok = ok && get(v) == h;
      v.resize(8);
      ok = ok && v.size() == 8 && v.data() != nullptr;
      v.resize(8);  // same size: a no-op
      report(ok && v.size() == 8, "resize");
    }
    {
      pfxvector<double> a(64);
      double *p = a.data();
      pfxvector<double> b(std::move(a));
      report(b.data() == p && b.size() == 64 && a.data() == nullptr && a.size() == 0, "move");
    }
  } catch (const std::exception &e) {
    std::printf("exception: %s\n", e.what());
    return 100;
  }
  std::printf("%d failure(s)\n", failures);
  return failures;
}
