/**
 * @file fft_backend_mi355.hpp
 * @brief MI355X (gfx950) FFT backend for ParaFaFT over the pfx C ABI.
 *
 * MI355Backend<FloatType> satisfies the backend concept of
 * backend/fft_backend.hpp:45-196 together with the members the drivers use
 * beyond it (memcpy2d, sync, first_touch, use_alltoallw, handles_distributed,
 * use_p2p and the P2P statics), with the signatures HipFFTBackend<T> has
 * (backend/hipfft/fft_backend_hipfft.hpp:178-635).  Every FFT runs in
 * hand-written HIP kernels inside libpfx.so; this header only needs
 * pfx_capi.h and the C++ standard library (no HIP headers), so a translation
 * unit that includes it compiles with any C++17 compiler and links -lpfx.
 *
 * Drop-in: add `#include "mi355/fft_backend_mi355.hpp"` next to the other
 * backend includes at backend/fft_backend.hpp:202-205 and instantiate
 * ParaFaFT_C2C<D, parafaft::MI355Backend<float>> (or <double>).
 */
#ifndef PARAFAFT_BACKEND_MI355_HPP
#define PARAFAFT_BACKEND_MI355_HPP

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../pfx_capi.h"

namespace parafaft {

#ifndef PARAFAFT_BACKEND_HPP
// Standalone use (outside the ParaFaFT tree): the enums of
// backend/fft_backend.hpp:25-43, same names and enumerators.
enum class FFTDirection { Forward, Backward };
enum class FFTPlanFlag { Estimate, Measure, Patient };
#endif

/// Throws std::runtime_error("pfx error in <op>: <detail>") on failure,
/// the style of check_hip (fft_backend_hipfft.hpp:46-51).
inline void check_pfx(int status, const char *operation) {
  if (status != PFX_SUCCESS) {
    throw std::runtime_error(std::string("pfx error in ") + operation + ": " +
                             pfx_error_string(status) + ": " + pfx_last_error_message());
  }
}

/// Owning device vector (the role of hipvector<T>, fft_backend_hipfft.hpp:58-121).
template <typename T> class pfxvector {
public:
  pfxvector() = default;
  explicit pfxvector(std::size_t size) { resize(size); }
  ~pfxvector() {
    if (data_) pfx_free(data_);
  }
  pfxvector(const pfxvector &) = delete;
  pfxvector &operator=(const pfxvector &) = delete;
  pfxvector(pfxvector &&o) noexcept : data_(o.data_), size_(o.size_) {
    o.data_ = nullptr;
    o.size_ = 0;
  }
  pfxvector &operator=(pfxvector &&o) noexcept {
    std::swap(data_, o.data_);
    std::swap(size_, o.size_);
    return *this;
  }
  /// Old contents are discarded (as hipvector::resize).
  void resize(std::size_t n) {
    if (n == size_) return;
    void *p = nullptr;
    check_pfx(pfx_malloc(&p, n * sizeof(T)), "pfx_malloc pfxvector");
    if (data_) pfx_free(data_);
    data_ = static_cast<T *>(p);
    size_ = n;
  }
  T *data() { return data_; }
  const T *data() const { return data_; }
  std::size_t size() const { return size_; }

private:
  T *data_ = nullptr;
  std::size_t size_ = 0;
};

/// complex<T> with the {x, y} layout of hipfftComplex / hipfftDoubleComplex,
/// layout-compatible with std::complex<T>.
template <typename T> struct alignas(2 * sizeof(T)) pfx_complex {
  T x, y;
};

/**
 * @tparam FloatTypeT float or double.
 * @tparam SplitMode  PFX_F32 (plain) or, for float, PFX_F32_SPLIT_F16 /
 *                    PFX_F32_SPLIT_BF16 (radix-16 sub-DFTs on MFMA with hi/lo
 *                    operand splits).
 */
template <typename FloatTypeT = double,
          int SplitMode = std::is_same<FloatTypeT, double>::value ? PFX_F64 : PFX_F32>
class MI355Backend {
  static_assert(std::is_same<FloatTypeT, float>::value || std::is_same<FloatTypeT, double>::value,
                "MI355Backend supports float and double");

public:
  using FloatType = FloatTypeT;
  using Complex = pfx_complex<FloatType>;
  using Buffer = pfxvector<FloatType>;
  using ComplexBuffer = pfxvector<Complex>;

  static constexpr int precision = std::is_same<FloatTypeT, double>::value ? PFX_F64 : SplitMode;

  /// plan_flag: Estimate picks kernels from a static table; Measure / Patient
  /// time the candidate kernels on the planning buffer the driver passes to
  /// create_stage_plan (its scratch buffer, whose contents FFTW_MEASURE also
  /// destroys) and keep the fastest (pfx_plan_tune, cached per geometry).
  explicit MI355Backend(int num_stages, FFTPlanFlag plan_flag = FFTPlanFlag::Estimate)
      : num_stages_(num_stages), plans_(num_stages, nullptr),
        tune_flag_(plan_flag == FFTPlanFlag::Estimate
                       ? PFX_PLAN_ESTIMATE
                       : (plan_flag == FFTPlanFlag::Measure ? PFX_PLAN_MEASURE : PFX_PLAN_PATIENT)) {
    check_pfx(pfx_stream_create(&stream_), "pfx_stream_create");
  }

  /// The (num_stages, MPI_Comm, flag) constructor of the concept
  /// (fft_backend.hpp:66-75); the communicator is not needed on the GPU.
  template <typename CommT,
            typename = typename std::enable_if<!std::is_same<CommT, FFTPlanFlag>::value>::type>
  MI355Backend(int num_stages, CommT comm, FFTPlanFlag plan_flag = FFTPlanFlag::Estimate)
      : MI355Backend(num_stages, plan_flag) {
    (void)comm;
  }

  ~MI355Backend() {
    if (stream_) {
      pfx_stream_synchronize(stream_);
      pfx_stream_destroy(stream_);
    }
    for (pfx_plan p : plans_) pfx_plan_destroy(p);
    pfx_plan_destroy(r2c_);
    pfx_plan_destroy(c2r_);
  }

  MI355Backend(const MI355Backend &) = delete;
  MI355Backend &operator=(const MI355Backend &) = delete;

  MI355Backend(MI355Backend &&o) noexcept
      : num_stages_(o.num_stages_), plans_(std::move(o.plans_)), r2c_(o.r2c_), c2r_(o.c2r_),
        stream_(o.stream_), tune_flag_(o.tune_flag_) {
    o.plans_.assign(o.num_stages_, nullptr);
    o.r2c_ = o.c2r_ = nullptr;
    o.stream_ = nullptr;
  }

  // ---------------------------------------------------------------- C2C
  /// fft_backend_hipfft.hpp:294-314; `data` is a layout reference only.
  void create_stage_plan(int stage, int length, std::size_t batch, Complex *data,
                         std::ptrdiff_t stride, std::ptrdiff_t dist) {
    check_stage(stage);
    pfx_plan p = nullptr;
    check_pfx(pfx_plan_create_c2c(&p, precision, length, checked(batch, "batch"), stride, dist),
              "pfx_plan_create_c2c");
    tune(p, data);
    pfx_plan_destroy(plans_[stage]);
    plans_[stage] = p;
  }

  /// fft_backend_hipfft.hpp:326-339: in place, asynchronous on the stream.
  void execute_stage(int stage, FFTDirection direction, Complex *data) {
    check_stage(stage);
    check_pfx(pfx_execute(plans_[stage], dir(direction), data, stream_), "pfx_execute");
  }

  /// As execute_stage with every output scaled inside the last pass.
  void execute_stage_scaled(int stage, FFTDirection direction, Complex *data, double scale) {
    check_stage(stage);
    check_pfx(pfx_execute_scaled(plans_[stage], dir(direction), data, scale, stream_),
              "pfx_execute_scaled");
  }

  // ----------------------------------------------------------- R2C / C2R
  /// fft_backend_hipfft.hpp:361-384 (dist in FloatType scalars).
  void create_r2c_inplace_plan(int length, std::size_t batch, FloatType *padded_real,
                               std::ptrdiff_t stride, std::ptrdiff_t dist) {
    pfx_plan p = nullptr;
    check_pfx(pfx_plan_create_r2c(&p, precision, length, checked(batch, "batch"), stride, dist),
              "pfx_plan_create_r2c");
    tune(p, padded_real);
    pfx_plan_destroy(r2c_);
    r2c_ = p;
  }

  void execute_r2c_inplace(FloatType *padded_real) {
    check_pfx(pfx_execute(r2c_, PFX_FORWARD, padded_real, stream_), "pfx_execute r2c");
  }

  /// fft_backend_hipfft.hpp:426-447.
  void create_c2r_inplace_plan(int length, std::size_t batch, FloatType *padded_real,
                               std::ptrdiff_t stride, std::ptrdiff_t dist) {
    pfx_plan p = nullptr;
    check_pfx(pfx_plan_create_c2r(&p, precision, length, checked(batch, "batch"), stride, dist),
              "pfx_plan_create_c2r");
    tune(p, padded_real);
    pfx_plan_destroy(c2r_);
    c2r_ = p;
  }

  void execute_c2r_inplace(FloatType *padded_real) {
    check_pfx(pfx_execute(c2r_, PFX_BACKWARD, padded_real, stream_), "pfx_execute c2r");
  }

  void execute_c2r_inplace_scaled(FloatType *padded_real, double scale) {
    check_pfx(pfx_execute_scaled(c2r_, PFX_BACKWARD, padded_real, scale, stream_),
              "pfx_execute_scaled c2r");
  }

  // -------------------------------------------------------------- memory
  void memcpy(void *dest, const void *src, std::size_t bytes) const {
    check_pfx(pfx_memcpy_async(dest, src, bytes, stream_), "pfx_memcpy_async");
  }

  void memcpy2d(void *dst, std::size_t dpitch, const void *src, std::size_t spitch,
                std::size_t width, std::size_t height) const {
    check_pfx(pfx_memcpy2d_async(dst, dpitch, src, spitch, width, height, stream_),
              "pfx_memcpy2d_async");
  }

  void sync() const { check_pfx(pfx_stream_synchronize(stream_), "pfx_stream_synchronize"); }

  /// No host NUMA first-touch on device memory (fft_backend_hipfft.hpp:516).
  void first_touch(void *, std::size_t) const {}

  /// Drivers pack on the device (parafaft_c2c.hpp:219-221, 576-580).
  static constexpr bool use_alltoallw = false;
  static constexpr bool handles_distributed = false;

  // ------------------------------------------------------------ P2P statics
  // Present so the drivers' setup_p2p (parafaft_c2c.hpp:906-1046) sizes its
  // per-transition table; this one-GPU build reports no peer access, so the
  // drivers take the packed exchange and never request IPC handles.
  static constexpr bool use_p2p = true;
  static int device_id() { return 0; }
  static bool can_access_peer(int, int) { return false; }
  static void enable_peer_access(int) {}
  static bool peer_link_is_top_tier(int src, int dst) { return src == dst; }
  static constexpr std::size_t ipc_handle_size = 64;
  static void ipc_get_handle(void *, void *) {
    throw std::runtime_error("MI355Backend: IPC peer buffers are not part of this build");
  }
  static void *ipc_open_handle(const void *) {
    throw std::runtime_error("MI355Backend: IPC peer buffers are not part of this build");
  }
  static void ipc_close_handle(void *) {}

  /// The backend's stream (a hipStream_t).
  pfx_stream stream() const { return stream_; }

private:
  void tune(pfx_plan p, void *planning_buffer) {
    if (tune_flag_ == PFX_PLAN_ESTIMATE || planning_buffer == nullptr) return;
    check_pfx(pfx_plan_tune(p, tune_flag_, planning_buffer, stream_), "pfx_plan_tune");
  }

  static int dir(FFTDirection d) { return d == FFTDirection::Forward ? PFX_FORWARD : PFX_BACKWARD; }

  void check_stage(int stage) const {
    if (stage < 0 || stage >= num_stages_)
      throw std::runtime_error("MI355Backend: stage index out of range");
  }

  static int64_t checked(std::size_t v, const char *name) {
    if (v > static_cast<std::size_t>(std::numeric_limits<int64_t>::max()))
      throw std::runtime_error(std::string("MI355Backend: plan parameter '") + name +
                               "' exceeds int64 range");
    return static_cast<int64_t>(v);
  }

  int num_stages_;
  std::vector<pfx_plan> plans_;
  pfx_plan r2c_ = nullptr;
  pfx_plan c2r_ = nullptr;
  pfx_stream stream_ = nullptr;
  int tune_flag_ = PFX_PLAN_ESTIMATE;
};

}  // namespace parafaft

#endif  // PARAFAFT_BACKEND_MI355_HPP
