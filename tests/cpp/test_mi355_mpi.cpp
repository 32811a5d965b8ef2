// The backend as the reference drivers construct and call it, with MPICH's
// <mpi.h> on the include path: MI355Backend<T>(D, MPI_COMM_WORLD,
// FFTPlanFlag::Measure) -- the expression ParaFaFT_C2C / ParaFaFT_R2C use to
// build their backend (parafaft_c2c.hpp:97-99, parafaft_r2c.hpp:119-121).
// No MPI function is called (the backend ignores the communicator), so the
// executable needs MPICH's header only, not its library.
//
// The calls below replay the drivers' call expressions with their argument
// types:
//  - create_stage_plan(int stage, int length, std::size_t batch, Complex*,
//    std::ptrdiff_t stride, std::ptrdiff_t dist) for the last, first and
//    middle axes (parafaft_c2c.hpp:249-287; parafaft_r2c.hpp:1158-1206);
//  - execute_stage on the buffer, and on buffer + i * slice_stride for the
//    middle axis (parafaft_c2c.hpp:866-897);
//  - create_r2c/c2r_inplace_plan(int, std::size_t, FloatType*, 1,
//    std::ptrdiff_t) and execute_r2c/c2r_inplace(FloatType*)
//    (parafaft_r2c.hpp:369, 540, 1175-1178);
//  - per-peer memcpy2d with size_t pitches (parafaft_common.hpp:582-588);
//  - the compile-time statics and the P2P statics' signatures as
//    setup_p2p uses them (parafaft_c2c.hpp:295-303, 906-1046).
// Results are checked against the oracle.  Returns the failure count.
#include <mpi.h>

#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "orc.h"
#include "parafaft_mi355/fft_backend_mi355.hpp"

using parafaft::FFTDirection;
using parafaft::FFTPlanFlag;

static int failures = 0;

static void check(bool ok, const char *what, double err) {
  std::printf("  %-52s err=%.3e %s\n", what, err, ok ? "[PASS]" : "[FAIL]");
  if (!ok) ++failures;
}

template <typename Backend> static void check_statics() {
  static_assert(!Backend::use_alltoallw, "drivers must pack on the device (c2c:576-580)");
  static_assert(!Backend::handles_distributed, "the drivers run the stage loop");
  static_assert(Backend::use_p2p, "setup_p2p sizes its per-transition table");
  static_assert(std::is_same<decltype(Backend::device_id()), int>::value, "device_id");
  static_assert(std::is_same<decltype(Backend::can_access_peer(0, 1)), bool>::value, "peer");
  static_assert(std::is_same<decltype(Backend::peer_link_is_top_tier(0, 1)), bool>::value, "tier");
  static_assert(std::is_same<decltype(Backend::ipc_open_handle(nullptr)), void *>::value, "open");
  using Handle = std::array<char, Backend::ipc_handle_size>;  // parafaft_c2c.hpp:1024
  static_assert(sizeof(Handle) >= 64, "IPC handle size");
  static_assert(std::is_move_constructible<Backend>::value, "move-constructible");
  static_assert(!std::is_copy_constructible<Backend>::value, "non-copyable");
  // what setup_p2p does on one GPU: no peer reports access, so the driver
  // takes the packed exchange and never requests a handle
  const int my_device = Backend::device_id();
  check(!Backend::can_access_peer(my_device, my_device + 1), "no peer access on one GPU", 0);
  Backend::enable_peer_access(my_device + 1);
  check(Backend::peer_link_is_top_tier(my_device, my_device), "self link is top tier", 0);
  Backend::ipc_close_handle(nullptr);
}

// One-rank C2C stage loop of shape n0 x n1 x n2, forward: the stage plans and
// executions of create_backend_plans / perform_fft_on_buffer, with the
// drivers' argument types.  The exchange between stages is the identity on
// one rank, so it is left out.
template <typename T> static void c2c_3d(MPI_Comm comm, FFTPlanFlag flag) {
  using Backend = parafaft::MI355Backend<T>;
  using Complex = typename Backend::Complex;
  constexpr int D = 3;
  const int shape[D] = {12, 20, 16};
  Backend backend(D, comm, flag);
  std::size_t total = 1;
  for (int i = 0; i < D; ++i) total *= static_cast<std::size_t>(shape[i]);
  typename Backend::ComplexBuffer scratch(total), data(total);
  for (int stage = 0; stage < D; ++stage) {
    const int axis = D - 1 - stage;
    const int length = shape[axis];
    if (axis == D - 1) {
      std::size_t batch = 1;
      for (int i = 0; i < D - 1; ++i) batch *= static_cast<std::size_t>(shape[i]);
      backend.create_stage_plan(stage, length, batch, scratch.data(), 1, length);
    } else if (axis == 0) {
      std::size_t batch = 1;
      for (int i = 1; i < D; ++i) batch *= static_cast<std::size_t>(shape[i]);
      std::ptrdiff_t stride = static_cast<std::ptrdiff_t>(batch);
      backend.create_stage_plan(stage, length, batch, scratch.data(), stride, 1);
    } else {
      std::size_t trailing_size = 1;
      for (int i = axis + 1; i < D; ++i) trailing_size *= static_cast<std::size_t>(shape[i]);
      std::ptrdiff_t trailing_dist = static_cast<std::ptrdiff_t>(trailing_size);
      backend.create_stage_plan(stage, length, trailing_size, scratch.data(), trailing_dist, 1);
    }
  }
  std::vector<double> in(2 * total);
  orc_uniform(in.data(), in.size(), 77);
  std::vector<std::complex<T>> h(total);
  for (std::size_t i = 0; i < total; ++i) h[i] = std::complex<T>(T(in[2 * i]), T(in[2 * i + 1]));
  parafaft::check_pfx(pfx_memcpy_async(data.data(), h.data(), total * sizeof(Complex), nullptr),
                      "h2d");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  Complex *buffer = data.data();
  for (int stage = 0; stage < D; ++stage) {
    const int axis = D - 1 - stage;
    if (axis == D - 1 || axis == 0) {
      backend.execute_stage(stage, FFTDirection::Forward, buffer);
    } else {
      std::size_t leading_size = 1, trailing_size = 1;
      for (int i = 0; i < axis; ++i) leading_size *= static_cast<std::size_t>(shape[i]);
      for (int i = axis + 1; i < D; ++i) trailing_size *= static_cast<std::size_t>(shape[i]);
      const std::size_t slice_stride = static_cast<std::size_t>(shape[axis]) * trailing_size;
      for (std::size_t i = 0; i < leading_size; ++i)
        backend.execute_stage(stage, FFTDirection::Forward, buffer + i * slice_stride);
    }
  }
  backend.sync();
  std::vector<std::complex<T>> out(total);
  parafaft::check_pfx(pfx_memcpy_async(out.data(), data.data(), total * sizeof(Complex), nullptr),
                      "d2h");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  std::vector<double> ref(2 * total);
  for (std::size_t i = 0; i < total; ++i) {
    ref[2 * i] = h[i].real();
    ref[2 * i + 1] = h[i].imag();
  }
  orc_parafaft_c2c(D, shape, ref.data(), -1, 1);
  double num = 0, den = 0;
  for (std::size_t i = 0; i < total; ++i) {
    const double dr = out[i].real() - ref[2 * i], di = out[i].imag() - ref[2 * i + 1];
    num += dr * dr + di * di;
    den += ref[2 * i] * ref[2 * i] + ref[2 * i + 1] * ref[2 * i + 1];
  }
  const double e = std::sqrt(num / den);
  const double bound = sizeof(T) == 8 ? 1e-13 : 1e-6 * std::log2(double(total));
  check(e <= bound, sizeof(T) == 8 ? "C2C 12x20x16 double, (D, MPI_Comm, Measure)"
                                   : "C2C 12x20x16 float, (D, MPI_Comm, Measure)", e);
}

// R2C stage 0 of ParaFaFT_R2C on one rank: in-place plans on the padded
// real buffer, planned on reinterpret_cast<FloatType*>(scratch.data()).
template <typename T> static void r2c_rows(MPI_Comm comm) {
  using Backend = parafaft::MI355Backend<T>;
  constexpr int D = 2;
  const int rows = 24, real_length = 40, complex_length = real_length / 2 + 1;
  Backend backend(D, comm, FFTPlanFlag::Estimate);
  std::size_t batch0 = static_cast<std::size_t>(rows);
  std::ptrdiff_t padded_dist = 2 * complex_length;
  typename Backend::ComplexBuffer scratch_b(rows * complex_length), buf(rows * complex_length);
  T *plan_buf = reinterpret_cast<T *>(scratch_b.data());
  backend.create_r2c_inplace_plan(real_length, batch0, plan_buf, 1, padded_dist);
  backend.create_c2r_inplace_plan(real_length, batch0, plan_buf, 1, padded_dist);
  std::vector<double> real(rows * padded_dist, 0.0);
  std::vector<double> u(rows * real_length);
  orc_uniform(u.data(), u.size(), 91);
  for (int r = 0; r < rows; ++r)
    for (int i = 0; i < real_length; ++i) real[r * padded_dist + i] = u[r * real_length + i];
  std::vector<T> h(real.begin(), real.end());
  T *padded_buffer = reinterpret_cast<T *>(buf.data());
  parafaft::check_pfx(pfx_memcpy_async(padded_buffer, h.data(), h.size() * sizeof(T), nullptr),
                      "h2d");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  backend.execute_r2c_inplace(padded_buffer);
  backend.sync();
  std::vector<T> spec(h.size());
  parafaft::check_pfx(pfx_memcpy_async(spec.data(), padded_buffer, h.size() * sizeof(T), nullptr),
                      "d2h");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  std::vector<double> ref(real);
  for (std::size_t i = 0; i < ref.size(); ++i) ref[i] = double(h[i]);
  orc_r2c_inplace(ref.data(), real_length, rows, 1, padded_dist, 1);
  double num = 0, den = 0;
  for (std::size_t i = 0; i < ref.size(); ++i) {
    num += (spec[i] - ref[i]) * (spec[i] - ref[i]);
    den += ref[i] * ref[i];
  }
  const double e = std::sqrt(num / den);
  check(e <= (sizeof(T) == 8 ? 1e-13 : 1e-6 * std::log2(double(real_length))),
        sizeof(T) == 8 ? "R2C rows double, (D, MPI_Comm) ctor" : "R2C rows float, (D, MPI_Comm) ctor",
        e);
  backend.execute_c2r_inplace(padded_buffer);
  backend.sync();
  parafaft::check_pfx(pfx_memcpy_async(spec.data(), padded_buffer, h.size() * sizeof(T), nullptr),
                      "d2h");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  double m = 0;
  for (int r = 0; r < rows; ++r)
    for (int i = 0; i < real_length; ++i)
      m = std::max(m, std::fabs(spec[r * padded_dist + i] / real_length - double(h[r * padded_dist + i])));
  check(m <= (sizeof(T) == 8 ? 1e-13 : 1e-5), "C2R(R2C(x)) / N == x", m);
}

// Per-peer memcpy2d of the pack loop: peer p's block of src_n[p] entries on
// the split axis, src_leading rows, into its packed segment (bytes compared
// with a host copy done the same way).
static void pack_memcpy2d(MPI_Comm comm) {
  using Backend = parafaft::MI355Backend<float>;
  using Complex = Backend::Complex;
  Backend backend(2, comm, FFTPlanFlag::Estimate);
  const int nparts = 4, src_axis_extent = 13;
  const int src_n[nparts] = {4, 3, 3, 3}, src_s[nparts] = {0, 4, 7, 10};
  const std::size_t src_leading = 6, src_trailing = 5, elem_size = sizeof(Complex);
  std::size_t send_displs[nparts], pos = 0;
  for (int p = 0; p < nparts; ++p) {
    send_displs[p] = pos;
    pos += src_leading * src_n[p] * src_trailing;
  }
  const std::size_t total = src_leading * src_axis_extent * src_trailing;
  std::vector<std::complex<float>> h(total);
  for (std::size_t i = 0; i < total; ++i) h[i] = std::complex<float>(float(i), -float(i));
  Backend::ComplexBuffer src(total), pack(total);
  parafaft::check_pfx(pfx_memcpy_async(src.data(), h.data(), total * elem_size, nullptr), "h2d");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  void *src_array = src.data(), *pack_buf = pack.data();
  size_t src_spitch = static_cast<size_t>(src_axis_extent) * src_trailing * elem_size;
  for (int p = 0; p < nparts; ++p) {
    size_t row_bytes = static_cast<size_t>(src_n[p]) * src_trailing * elem_size;
    backend.memcpy2d(reinterpret_cast<char *>(pack_buf) + send_displs[p] * elem_size, row_bytes,
                     reinterpret_cast<char *>(src_array) + static_cast<size_t>(src_s[p]) *
                                                               src_trailing * elem_size,
                     src_spitch, row_bytes, src_leading);
  }
  backend.sync();
  std::vector<std::complex<float>> got(total), want(total);
  parafaft::check_pfx(pfx_memcpy_async(got.data(), pack.data(), total * elem_size, nullptr), "d2h");
  parafaft::check_pfx(pfx_stream_synchronize(nullptr), "sync");
  for (int p = 0; p < nparts; ++p)
    for (std::size_t l = 0; l < src_leading; ++l)
      std::memcpy(&want[send_displs[p] + l * src_n[p] * src_trailing],
                  &h[(l * src_axis_extent + src_s[p]) * src_trailing],
                  src_n[p] * src_trailing * elem_size);
  check(std::memcmp(got.data(), want.data(), total * elem_size) == 0,
        "per-peer memcpy2d pack, bit-exact", 0);
}

int main() {
  std::printf("MI355Backend through the (D, MPI_Comm, FFTPlanFlag) constructor\n");
  check_statics<parafaft::MI355Backend<float>>();
  check_statics<parafaft::MI355Backend<double>>();
  const MPI_Comm comm = MPI_COMM_WORLD;
  c2c_3d<float>(comm, FFTPlanFlag::Measure);
  c2c_3d<double>(comm, FFTPlanFlag::Measure);
  c2c_3d<float>(MPI_COMM_WORLD, FFTPlanFlag::Estimate);
  r2c_rows<float>(comm);
  r2c_rows<double>(comm);
  pack_memcpy2d(comm);
  std::printf("%d failure(s)\n", failures);
  return failures;
}
