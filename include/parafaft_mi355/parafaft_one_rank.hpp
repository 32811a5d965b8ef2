/**
 * @file parafaft_one_rank.hpp
 * @brief One-rank ParaFaFT drivers for any backend of the concept.
 *
 * parafaft_mi355::ParaFaFT_C2C<D, Backend> and ParaFaFT_R2C<D, Backend> restate,
 * for a single rank and without MPI, the stage loops of the reference drivers:
 * plan geometry (parafaft_c2c.hpp:249-287, parafaft_r2c.hpp:1158-1206), the
 * ping-pong parity and per-stage loop (parafaft_c2c.hpp:539-694,
 * parafaft_r2c.hpp:360-542), perform_fft_on_buffer with its per-slice loop on
 * middle axes (parafaft_c2c.hpp:866-897, parafaft_r2c.hpp:1219-1245), and the
 * one-rank exchange, which is an identity flat copy (exchange_local,
 * parafaft_common.hpp:467-543).  Accessor names and meaning follow
 * parafaft_c2c.hpp:348-426 and parafaft_r2c.hpp:552-796.
 *
 * With Mode::Fast the driver runs the pfx one-rank N-D plan instead: one launch
 * per axis, no exchange copies, no per-slice launches (the results are the
 * same: on one rank every exchange is the identity).
 */
#ifndef PARAFAFT_MI355_ONE_RANK_HPP
#define PARAFAFT_MI355_ONE_RANK_HPP

#include <cstddef>
#include <cstdio>
#include <utility>

#include "fft_backend_mi355.hpp"

namespace parafaft_mi355 {

using parafaft::FFTDirection;
using parafaft::FFTPlanFlag;

enum class Mode { Faithful, Fast };

// The backend's own precision code, the split-precision MFMA forms included
// (MI355Backend<float, PFX_F32_SPLIT_F16> runs its N-D plan split too).
template <typename B> constexpr int nd_precision() { return B::precision; }

/// Measure / Patient planning of a one-rank N-D plan: the FFTW_MEASURE
/// analogue for the fast path (kernels, and chunked or stage-by-stage
/// schedule, timed on a temporary array-sized device buffer that is freed
/// again; the user's arrays are never touched).
inline void plan_nd(pfx_ndplan nd, FFTPlanFlag flag, std::size_t bytes, pfx_stream stream) {
  if (flag == FFTPlanFlag::Estimate) return;
  void *tmp = nullptr;
  parafaft::check_pfx(pfx_malloc(&tmp, bytes), "pfx_malloc (planning buffer)");
  const int rc = pfx_ndplan_tune(nd, flag == FFTPlanFlag::Measure ? PFX_PLAN_MEASURE
                                                                 : PFX_PLAN_PATIENT,
                                 tmp, stream);
  pfx_stream_synchronize(stream);
  pfx_free(tmp);
  parafaft::check_pfx(rc, "pfx_ndplan_tune");
}

template <int D, typename Backend = parafaft::MI355Backend<double>>
class ParaFaFT_C2C {
  static_assert(D >= 1, "D >= 1");

public:
  using FloatType = typename Backend::FloatType;
  using Complex = typename Backend::Complex;
  using ComplexBuffer = typename Backend::ComplexBuffer;

  ParaFaFT_C2C(const int global_shape[D], FFTPlanFlag flag = FFTPlanFlag::Estimate,
               Mode mode = Mode::Faithful)
      : backend_(D, flag), mode_(mode) {
    for (int i = 0; i < D; ++i) global_shape_[i] = global_shape[i];
    size_ = 1;
    for (int i = 0; i < D; ++i) size_ *= static_cast<std::size_t>(global_shape_[i]);
    if (mode_ == Mode::Fast) {
      parafaft::check_pfx(pfx_ndplan_create(&nd_, nd_precision<Backend>(), PFX_C2C, D,
                                            global_shape_),
                          "pfx_ndplan_create");
      plan_fast(flag, size_ * sizeof(Complex));
      return;
    }
    scratch_.resize(size_);
    create_backend_plans();
  }

  ~ParaFaFT_C2C() { pfx_ndplan_destroy(nd_); }
  ParaFaFT_C2C(const ParaFaFT_C2C &) = delete;
  ParaFaFT_C2C &operator=(const ParaFaFT_C2C &) = delete;

  std::size_t get_local_size() const { return size_; }
  std::size_t get_required_output_size() const { return size_; }
  void get_local_shape(int shape[D]) const { copy(shape, global_shape_); }
  void get_global_start(int start[D]) const { zero(start); }
  void get_final_shape(int shape[D]) const { copy(shape, global_shape_); }
  void get_final_start(int start[D]) const { zero(start); }
  void get_domain_decomposition(int dec[D]) const {
    for (int i = 0; i < D; ++i) dec[i] = 1;
  }
  /// parafaft_c2c.hpp:439-445: the backend's own buffer when it handles the
  /// distributed transform itself; this backend does not (nullptr).
  Complex *get_buffer() { return nullptr; }
  /// parafaft_c2c.hpp:463, 473-477: a (D-1)-dimensional grid, one rank at 0.
  static constexpr int get_grid_ndims() { return D - 1; }
  void get_grid_coords(int coords[D - 1]) const {
    for (int i = 0; i < D - 1; ++i) coords[i] = 0;
  }
  /// parafaft_c2c.hpp:499-530: every stage's local shape (one rank: global).
  void print_stage_shapes() const {
    for (int stage = 0; stage < D; ++stage) {
      std::printf("Stage %d: [", stage);
      for (int i = 0; i < D; ++i) std::printf(i ? ", %d" : "%d", global_shape_[i]);
      std::printf("]\n");
    }
  }

  /// parafaft_c2c.hpp:539-602 (unnormalised).
  void forward(Complex *data) { run(data, true, 1.0); }

  /// parafaft_c2c.hpp:632-694.  scale != 1 is fused into the last pass
  /// (the reference leaves 1/N to the user).
  void backward(Complex *data, double scale = 1.0) { run(data, false, scale); }

  Backend &backend() { return backend_; }

private:
  void plan_fast(FFTPlanFlag flag, std::size_t bytes) { plan_nd(nd_, flag, bytes, backend_.stream()); }
  static void copy(int *dst, const int *src) {
    for (int i = 0; i < D; ++i) dst[i] = src[i];
  }
  static void zero(int *dst) {
    for (int i = 0; i < D; ++i) dst[i] = 0;
  }
  std::size_t prod(int from, int to) const {
    std::size_t p = 1;
    for (int i = from; i < to; ++i) p *= static_cast<std::size_t>(global_shape_[i]);
    return p;
  }

  void create_backend_plans() {
    for (int stage = 0; stage < D; ++stage) {
      const int axis = D - 1 - stage;
      const int length = global_shape_[axis];
      if (axis == D - 1) {
        backend_.create_stage_plan(stage, length, prod(0, D - 1), scratch_.data(), 1, length);
      } else if (axis == 0) {
        const std::size_t batch = prod(1, D);
        backend_.create_stage_plan(stage, length, batch, scratch_.data(),
                                   static_cast<std::ptrdiff_t>(batch), 1);
      } else {
        const std::size_t trailing = prod(axis + 1, D);
        backend_.create_stage_plan(stage, length, trailing, scratch_.data(),
                                   static_cast<std::ptrdiff_t>(trailing), 1);
      }
    }
  }

  void perform_fft_on_buffer(int stage, int axis, FFTDirection dir, Complex *buf, double scale) {
    auto exec = [&](Complex *p) {
      if (scale == 1.0) backend_.execute_stage(stage, dir, p);
      else backend_.execute_stage_scaled(stage, dir, p, scale);
    };
    if (axis == D - 1 || axis == 0) {
      exec(buf);
    } else {
      const std::size_t leading = prod(0, axis);
      const std::size_t slice = static_cast<std::size_t>(global_shape_[axis]) * prod(axis + 1, D);
      for (std::size_t i = 0; i < leading; ++i) exec(buf + i * slice);
    }
  }

  void run(Complex *data, bool forward, double scale) {
    if (mode_ == Mode::Fast) {
      parafaft::check_pfx(pfx_ndplan_execute(nd_, forward ? PFX_FORWARD : PFX_BACKWARD, data,
                                             scale, backend_.stream()),
                          "pfx_ndplan_execute");
      backend_.sync();
      return;
    }
    const std::size_t bytes = size_ * sizeof(Complex);
    Complex *src, *dst;
    if ((D - 1) % 2 == 0) {
      src = data;
      dst = scratch_.data();
    } else {
      backend_.memcpy(scratch_.data(), data, bytes);
      src = scratch_.data();
      dst = data;
    }
    const FFTDirection dir = forward ? FFTDirection::Forward : FFTDirection::Backward;
    for (int k = 0; k < D; ++k) {
      const int stage = forward ? k : D - 1 - k;
      const int axis = D - 1 - stage;
      perform_fft_on_buffer(stage, axis, dir, src, k == D - 1 ? scale : 1.0);
      if (k < D - 1) {
        backend_.memcpy(dst, src, bytes);  // one-rank exchange_local: identity
        std::swap(src, dst);
      }
    }
    backend_.sync();
  }

  Backend backend_;
  Mode mode_;
  int global_shape_[D];
  std::size_t size_ = 0;
  ComplexBuffer scratch_;
  pfx_ndplan nd_ = nullptr;
};

template <int D, typename Backend = parafaft::MI355Backend<double>>
class ParaFaFT_R2C {
public:
  using FloatType = typename Backend::FloatType;
  using Complex = typename Backend::Complex;
  using ComplexBuffer = typename Backend::ComplexBuffer;

  ParaFaFT_R2C(const int global_shape[D], FFTPlanFlag flag = FFTPlanFlag::Estimate,
               Mode mode = Mode::Faithful)
      : backend_(D, flag), mode_(mode) {
    for (int i = 0; i < D; ++i) real_[i] = cplx_[i] = global_shape[i];
    cplx_[D - 1] = real_[D - 1] / 2 + 1;
    csize_ = 1;
    for (int i = 0; i < D; ++i) csize_ *= static_cast<std::size_t>(cplx_[i]);
    if (mode_ == Mode::Fast) {
      parafaft::check_pfx(pfx_ndplan_create(&nd_, nd_precision<Backend>(), PFX_R2C, D, real_),
                          "pfx_ndplan_create");
      plan_fast(flag, get_required_output_size() * sizeof(FloatType));
      return;
    }
    scratch_.resize(csize_);
    create_backend_plans();
  }

  ~ParaFaFT_R2C() { pfx_ndplan_destroy(nd_); }
  ParaFaFT_R2C(const ParaFaFT_R2C &) = delete;
  ParaFaFT_R2C &operator=(const ParaFaFT_R2C &) = delete;

  /// parafaft_r2c.hpp:573-596, in FloatType scalars.
  std::size_t get_required_output_size() const {
    const std::size_t stage0 = cprod(0, D - 1) * 2 * static_cast<std::size_t>(cplx_[D - 1]);
    return stage0 > 2 * csize_ ? stage0 : 2 * csize_;
  }
  std::size_t get_local_real_size() const { return rprod(0, D); }
  std::size_t get_local_complex_size() const { return csize_; }
  void get_local_real_shape(int s[D]) const {
    for (int i = 0; i < D; ++i) s[i] = real_[i];
  }
  void get_local_complex_shape(int s[D]) const {
    for (int i = 0; i < D; ++i) s[i] = cplx_[i];
  }
  void get_real_global_start(int s[D]) const {
    for (int i = 0; i < D; ++i) s[i] = 0;
  }
  void get_complex_global_start(int s[D]) const {
    for (int i = 0; i < D; ++i) s[i] = 0;
  }
  /// parafaft_r2c.hpp:712-717, 726-742: one rank, a (D-1)-dimensional grid.
  void get_domain_decomposition(int dec[D]) const {
    for (int i = 0; i < D; ++i) dec[i] = 1;
  }
  static constexpr int get_grid_ndims() { return D - 1; }
  void get_grid_coords(int coords[D - 1]) const {
    for (int i = 0; i < D - 1; ++i) coords[i] = 0;
  }
  /// parafaft_r2c.hpp:773-790: backend-owned buffers (none here: nullptr).
  FloatType *get_real_buffer() { return nullptr; }
  Complex *get_buffer() { return nullptr; }
  /// parafaft_r2c.hpp:664-710: global real and complex shapes, and every
  /// stage's input and output shape (one rank: the global ones).
  void print_shapes() const {
    auto line = [](const char *what, const int *v) {
      std::printf("%s: [", what);
      for (int i = 0; i < D; ++i) std::printf(i ? ", %d" : "%d", v[i]);
      std::printf("]\n");
    };
    line("Global real shape", real_);
    line("Global complex shape", cplx_);
    for (int stage = 0; stage < D; ++stage) {
      std::printf("Stage %d ", stage);
      line(stage == 0 ? "input shape (real)" : "input shape", stage == 0 ? real_ : cplx_);
      std::printf("Stage %d ", stage);
      line("output shape (complex)", cplx_);
    }
  }

  /// parafaft_r2c.hpp:360-430.
  void forward_in_place(FloatType *padded) {
    if (mode_ == Mode::Fast) {
      parafaft::check_pfx(pfx_ndplan_execute(nd_, PFX_FORWARD, padded, 1.0, backend_.stream()),
                          "pfx_ndplan_execute");
      backend_.sync();
      return;
    }
    Complex *pc = reinterpret_cast<Complex *>(padded);
    const std::size_t bytes = csize_ * sizeof(Complex);
    backend_.execute_r2c_inplace(padded);
    Complex *src, *dst;
    if ((D - 1) % 2 == 0) {
      src = pc;
      dst = scratch_.data();
    } else {
      backend_.memcpy(scratch_.data(), pc, bytes);
      src = scratch_.data();
      dst = pc;
    }
    for (int stage = 1; stage < D; ++stage) {
      backend_.memcpy(dst, src, bytes);  // one-rank exchange_local: identity
      perform_fft_on_buffer(stage, D - 1 - stage, FFTDirection::Forward, dst);
      std::swap(src, dst);
    }
    backend_.sync();
  }

  /// parafaft_r2c.hpp:473-542 (unnormalised unless scale != 1).
  void backward_in_place(FloatType *padded, double scale = 1.0) {
    if (mode_ == Mode::Fast) {
      parafaft::check_pfx(pfx_ndplan_execute(nd_, PFX_BACKWARD, padded, scale, backend_.stream()),
                          "pfx_ndplan_execute");
      backend_.sync();
      return;
    }
    Complex *pc = reinterpret_cast<Complex *>(padded);
    const std::size_t bytes = csize_ * sizeof(Complex);
    Complex *src, *dst;
    if ((D - 1) % 2 == 0) {
      src = pc;
      dst = scratch_.data();
    } else {
      backend_.memcpy(scratch_.data(), pc, bytes);
      src = scratch_.data();
      dst = pc;
    }
    for (int stage = D - 1; stage >= 1; --stage) {
      perform_fft_on_buffer(stage, D - 1 - stage, FFTDirection::Backward, src);
      backend_.memcpy(dst, src, bytes);
      std::swap(src, dst);
    }
    if (scale == 1.0) backend_.execute_c2r_inplace(padded);
    else backend_.execute_c2r_inplace_scaled(padded, scale);
    backend_.sync();
  }

  /// parafaft_r2c.hpp:804-823: contiguous real rows into padded rows of
  /// 2(N/2+1) scalars, one 2D copy.
  void copy_real_to_padded(const FloatType *real_input, FloatType *padded_output) const {
    const std::size_t n = static_cast<std::size_t>(real_[D - 1]);
    const std::size_t pd = 2 * static_cast<std::size_t>(cplx_[D - 1]);
    backend_.memcpy2d(padded_output, pd * sizeof(FloatType), real_input, n * sizeof(FloatType),
                      n * sizeof(FloatType), rprod(0, D - 1));
  }
  /// parafaft_r2c.hpp:825-847: padded rows back to contiguous real rows.
  void copy_padded_to_real(const FloatType *padded_input, FloatType *real_output) const {
    const std::size_t n = static_cast<std::size_t>(real_[D - 1]);
    const std::size_t pd = 2 * static_cast<std::size_t>(cplx_[D - 1]);
    backend_.memcpy2d(real_output, n * sizeof(FloatType), padded_input, pd * sizeof(FloatType),
                      n * sizeof(FloatType), rprod(0, D - 1));
  }

  /// parafaft_r2c.hpp:333-339: unpadded real input (left untouched) to the
  /// spectrum in `complex_output` (get_required_output_size() scalars, also
  /// the working space).  Fast mode fuses the padding copy into the R2C pass.
  void forward(const FloatType *real_input, Complex *complex_output) {
    if (mode_ == Mode::Fast &&
        fused_oop(PFX_FORWARD, const_cast<FloatType *>(real_input), complex_output, 1.0))
      return;
    copy_real_to_padded(real_input, reinterpret_cast<FloatType *>(complex_output));
    forward_in_place(reinterpret_cast<FloatType *>(complex_output));
  }

  /// parafaft_r2c.hpp:446-453: the spectrum in `complex_input` (working
  /// space, destroyed) to the real output.  Fast mode fuses the padding copy
  /// into the C2R pass.
  void backward(Complex *complex_input, FloatType *real_output, double scale = 1.0) {
    if (mode_ == Mode::Fast && fused_oop(PFX_BACKWARD, real_output, complex_input, scale)) return;
    backward_in_place(reinterpret_cast<FloatType *>(complex_input), scale);
    copy_padded_to_real(reinterpret_cast<FloatType *>(complex_input), real_output);
    backend_.sync();
  }

  Backend &backend() { return backend_; }

private:
  // false when a real-axis plan has no tiled kernel (then copy + in place)
  bool fused_oop(int direction, FloatType *real, Complex *padded, double scale) {
    const int rc =
        pfx_ndplan_execute_r2c_oop(nd_, direction, real, padded, scale, backend_.stream());
    if (rc == PFX_ERR_UNSUPPORTED) return false;
    parafaft::check_pfx(rc, "pfx_ndplan_execute_r2c_oop");
    backend_.sync();
    return true;
  }
  void plan_fast(FFTPlanFlag flag, std::size_t bytes) { plan_nd(nd_, flag, bytes, backend_.stream()); }
  std::size_t cprod(int from, int to) const {
    std::size_t p = 1;
    for (int i = from; i < to; ++i) p *= static_cast<std::size_t>(cplx_[i]);
    return p;
  }
  std::size_t rprod(int from, int to) const {
    std::size_t p = 1;
    for (int i = from; i < to; ++i) p *= static_cast<std::size_t>(real_[i]);
    return p;
  }

  void create_backend_plans() {
    const std::size_t batch0 = rprod(0, D - 1);
    const std::ptrdiff_t pd = 2 * static_cast<std::ptrdiff_t>(cplx_[D - 1]);
    FloatType *plan_buf = reinterpret_cast<FloatType *>(scratch_.data());
    backend_.create_r2c_inplace_plan(real_[D - 1], batch0, plan_buf, 1, pd);
    backend_.create_c2r_inplace_plan(real_[D - 1], batch0, plan_buf, 1, pd);
    for (int stage = 1; stage < D; ++stage) {
      const int axis = D - 1 - stage;
      const int length = cplx_[axis];
      if (axis == 0) {
        const std::size_t batch = cprod(1, D);
        backend_.create_stage_plan(stage, length, batch, scratch_.data(),
                                   static_cast<std::ptrdiff_t>(batch), 1);
      } else {
        const std::size_t trailing = cprod(axis + 1, D);
        backend_.create_stage_plan(stage, length, trailing, scratch_.data(),
                                   static_cast<std::ptrdiff_t>(trailing), 1);
      }
    }
  }

  void perform_fft_on_buffer(int stage, int axis, FFTDirection dir, Complex *buf) {
    if (axis == 0) {
      backend_.execute_stage(stage, dir, buf);
    } else {
      const std::size_t leading = cprod(0, axis);
      const std::size_t slice = static_cast<std::size_t>(cplx_[axis]) * cprod(axis + 1, D);
      for (std::size_t i = 0; i < leading; ++i) backend_.execute_stage(stage, dir, buf + i * slice);
    }
  }

  Backend backend_;
  Mode mode_;
  int real_[D];
  int cplx_[D];
  std::size_t csize_ = 0;
  ComplexBuffer scratch_;
  pfx_ndplan nd_ = nullptr;
};

}  // namespace parafaft_mi355

#endif  // PARAFAFT_MI355_ONE_RANK_HPP
