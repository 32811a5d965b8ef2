/*
 * pfx_capi.h -- C ABI of the MI355X (gfx950) FFT backend for ParaFaFT.
 *
 * Plain C: pointers, sizes and status codes only.  Every function returns 0
 * (PFX_SUCCESS) or a pfx_status; pfx_error_string() names the code and
 * pfx_last_error_message() gives the detail of the calling thread's last
 * failure.  All device work is enqueued asynchronously on the given stream
 * (a hipStream_t passed as void*; NULL means the null stream).
 *
 * Each entry point names the reference interface it replaces.  Reference
 * paths are relative to the ParaFaFT source tree.
 */
#ifndef PFX_CAPI_H
#define PFX_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfx_plan_t *pfx_plan;
typedef struct pfx_ndplan_t *pfx_ndplan;
typedef void *pfx_stream; /* hipStream_t */

enum pfx_status {
  PFX_SUCCESS = 0,
  PFX_ERR_INVALID_ARG = 1,
  PFX_ERR_UNSUPPORTED = 2,
  PFX_ERR_HIP = 3,
  PFX_ERR_ALLOC = 4,
  PFX_ERR_INTERNAL = 5
};

enum pfx_precision {
  PFX_F32 = 0,            /* complex<float>  / float   */
  PFX_F64 = 1,            /* complex<double> / double  */
  PFX_F32_SPLIT_F16 = 2,  /* complex<float> data, radix-16 sub-DFTs on FP16 MFMA, hi/lo split */
  PFX_F32_SPLIT_BF16 = 3  /* complex<float> data, radix-16 sub-DFTs on BF16 MFMA, hi/lo split */
};

enum pfx_kind { PFX_C2C = 0, PFX_R2C = 1, PFX_C2R = 2 };

enum pfx_direction { PFX_FORWARD = 0, PFX_BACKWARD = 1 };

/* ---------------------------------------------------------------- errors */
const char *pfx_error_string(int status);
const char *pfx_last_error_message(void);
/* Library version string. */
const char *pfx_version(void);

/* ------------------------------------------------ batched strided 1D plans */

/*
 * C2C plan: `batch` transforms of `length` complex elements; element j of
 * transform b is data[b*dist + j*stride] (complex units).  Both directions
 * share one plan.
 * Lengths: one tiled kernel where the length has a radix plan (up to 4096);
 * else a four-step plan when the length splits into two lengths with radix
 * plans (n = n1 * n2, e.g. 144 = 12 * 12, 8192 = 64 * 128, up to 4096^2):
 * two tiled passes, a twiddle pass and a transpose through a scratch buffer
 * of the plan's footprint; any other length up to 2^23 through Bluestein's
 * chirp convolution (zero-padded rows of M >= 2n - 1, two M-point
 * transforms; M the shortest 2^k, 3 2^k, 5 2^k, ... with a one-kernel plan).  These plans own their scratch, so executions of one such
 * plan must not overlap.  Longer lengths that do not split:
 * PFX_ERR_UNSUPPORTED.  (The direct O(n^2) DFT, up to 4096 complex<float> /
 * 2048 complex<double>, remains for pfx_execute_c2c_noalloc.)
 * Replaces: HipFFTBackend::create_stage_plan -> hipfftPlanMany(Z2Z/C2C)
 *   (backend/hipfft/fft_backend_hipfft.hpp:294-314); geometry as in
 *   FFTWBackend::create_stage_plan (backend/fftw3/fft_backend_fftw.hpp:387-424).
 */
int pfx_plan_create_c2c(pfx_plan *plan, int precision, int64_t length, int64_t batch,
                        int64_t stride, int64_t dist);

/*
 * Two-level C2C plan: transform b = o*inner_batch + i starts at
 * data[o*outer_dist + i*dist].  One launch covers what the reference issues
 * as outer_batch separate execute_stage calls on a middle axis.
 * Replaces: the leading_size loop of ParaFaFT_C2C::perform_fft_on_buffer
 *   (parafaft_c2c.hpp:875-896; parafaft_r2c.hpp:1230-1243).
 */
int pfx_plan_create_c2c_many(pfx_plan *plan, int precision, int64_t length,
                             int64_t inner_batch, int64_t stride, int64_t dist,
                             int64_t outer_batch, int64_t outer_dist);

/*
 * In-place padded R2C plan: row b starts at scalar b*dist_scalars; its
 * `length` reals are replaced by length/2+1 complex values (complex row
 * distance dist_scalars/2).  stride must be 1 (the only value ParaFaFT uses).
 * Lengths: even lengths whose half has a radix plan run one kernel; other
 * even lengths the half-length C2C plan (as above) of the rows read as
 * complex pairs, with a post-pass (C2R: a pre-pass); odd lengths a length-n
 * C2C plan on scratch rows (the reals with zero imaginary parts; C2R: the
 * Hermitian extension).
 * Replaces: HipFFTBackend::create_r2c_inplace_plan (fft_backend_hipfft.hpp:361-384),
 *   FFTWBackend::create_r2c_inplace_plan (fft_backend_fftw.hpp:439-469).
 */
int pfx_plan_create_r2c(pfx_plan *plan, int precision, int64_t length, int64_t batch,
                        int64_t stride, int64_t dist_scalars);

/*
 * In-place padded C2R plan (unnormalised), mirror of the R2C plan.
 * Replaces: HipFFTBackend::create_c2r_inplace_plan (fft_backend_hipfft.hpp:426-447),
 *   FFTWBackend::create_c2r_inplace_plan (fft_backend_fftw.hpp:503-530).
 */
int pfx_plan_create_c2r(pfx_plan *plan, int precision, int64_t length, int64_t batch,
                        int64_t stride, int64_t dist_scalars);

/*
 * Execute a plan in place on `data` (any element-aligned device pointer with
 * the plan's layout: the new-array execute semantics of
 * backend/fft_backend.hpp:107-111).  direction: PFX_FORWARD = exp(-2 pi i jk/n),
 * PFX_BACKWARD = exp(+2 pi i jk/n); both unnormalised.  R2C plans accept only
 * PFX_FORWARD, C2R plans only PFX_BACKWARD.
 * Replaces: HipFFTBackend::execute_stage / execute_r2c_inplace /
 *   execute_c2r_inplace (fft_backend_hipfft.hpp:326-339, 396-405, 459-467).
 */
int pfx_execute(pfx_plan plan, int direction, void *data, pfx_stream stream);

/*
 * As pfx_execute, with every output multiplied by `scale` inside the last
 * pass (the user-side 1/N normalisation loop of example/c2c.cpp:27-31 fused
 * into the transform).
 */
int pfx_execute_scaled(pfx_plan plan, int direction, void *data, double scale,
                       pfx_stream stream);

/*
 * Out-of-place forms of R2C / C2R plans: the real rows are unpadded,
 * `real_dist` scalars apart (even, >= the length); the complex side is the
 * plan's padded layout.  The R2C's first pass reads the real rows and the
 * C2R's last pass writes them, so the padding copy of ParaFaFT_R2C::forward
 * (copy_real_to_padded, parafaft_r2c.hpp:333-339, 804-823) and ::backward
 * (copy_padded_to_real, :446-453, 825-847) is fused into the transform.
 * The C2R form leaves `padded_in` unchanged.  Plans without a tiled kernel
 * return PFX_ERR_UNSUPPORTED (copy, then transform in place).
 */
int pfx_execute_r2c_oop(pfx_plan plan, const void *real_in, int64_t real_dist,
                        void *padded_out, double scale, pfx_stream stream);
int pfx_execute_c2r_oop(pfx_plan plan, void *padded_in, void *real_out, int64_t real_dist,
                        double scale, pfx_stream stream);

/*
 * One axis of a row-major block split into per-peer segments, in the order
 * the pencil exchange packs it (parafaft_common.hpp:576-590): element
 * (lead, e, trail), e on the split axis, is complex element
 * (lead * count[p] + e - start[p]) * trailing + trail of segments[p], p being
 * the peer whose [start[p], start[p] + count[p]) holds e.  The blocks must
 * tile the transform length in ascending order.  Arrays are host arrays of
 * nparts (1..16) entries; segments[p] are device addresses.
 */
typedef struct {
  int nparts;
  void *const *segments;
  const int *start;
  const int *count;
  uint64_t trailing;
} pfx_segmap;

/*
 * pfx_execute with the pencil exchange fused into the transform: when `out`
 * is given the last pass stores straight into the send segments (the pack of
 * parafaft_common.hpp:576-590), when `in` is given the first pass loads from
 * the receive segments (the unpack of :654-668).  Transform b of the plan is
 * (lead, trail) = (b / trailing, b % trailing) of the map.  Results are
 * bit-identical to execute-then-pack / unpack-then-execute.  Tiled C2C and
 * R2C (`out` only) plans of plain precision.
 */
int pfx_execute_segmented(pfx_plan plan, int direction, void *data, const pfx_segmap *in,
                          const pfx_segmap *out, double scale, pfx_stream stream);

int pfx_plan_destroy(pfx_plan plan);

/*
 * One C2C transform with no plan and no allocation: for a length with a
 * tiled kernel, the plan pfx_plan_create_c2c(precision, length, batch,
 * stride, dist) would build (its default kernel, not a tuned one) is set up
 * on the stack, its twiddle tables are generated into a fixed device scratch
 * array of the library (a kernel, on `stream`), and the transform runs as
 * pfx_execute_scaled would: same kernel, same tables, so the same bits.  A
 * length without a tiled kernel runs the direct O(n^2) DFT (the planned
 * path's four-step and chirp plans need scratch of their own).  Nothing is allocated on the
 * device or the host and no call synchronises, so callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops) can use it.
 * Limits: PFX_F32 / PFX_F64, length <= 4096 (larger lengths are refused with
 * PFX_ERR_INVALID_ARG; the scratch holds 2 x 4096 complex<double>).  Every
 * call on one GPU uses that GPU's one scratch array, so calls that overlap
 * on one GPU (different streams, or threads) race; calls in one stream are
 * ordered.  No counterpart in the reference: it is for callers that cannot
 * allocate.
 */
int pfx_execute_c2c_noalloc(int precision, int64_t length, int64_t batch, int64_t stride,
                            int64_t dist, int direction, void *data, double scale,
                            pfx_stream stream);

/*
 * Planning effort (the FFTPlanFlag of backend/fft_backend.hpp:39-43).
 * PFX_PLAN_ESTIMATE: a static priority table picks the kernel.
 * PFX_PLAN_MEASURE / PFX_PLAN_PATIENT: pfx_plan_tune times every candidate
 * kernel of the plan (PATIENT with more repetitions) and keeps the fastest;
 * results go to a process-wide tuning cache keyed by the plan geometry, which
 * later plans of the same key reuse without timing.
 */
enum pfx_plan_flag { PFX_PLAN_ESTIMATE = 0, PFX_PLAN_MEASURE = 1, PFX_PLAN_PATIENT = 2 };

/*
 * Times the candidate kernels of `plan` on `scratch` (a device buffer with the
 * plan's layout; its contents are destroyed, as FFTW_MEASURE destroys the
 * planning array) and switches the plan to the fastest.  A cache hit returns
 * without timing.  flag PFX_PLAN_ESTIMATE is a no-op.
 * Replaces: FFTW_MEASURE / FFTW_PATIENT planning in FFTWBackend::create_stage_plan
 *   (backend/fftw3/fft_backend_fftw.hpp:387-424, 743-754).
 */
int pfx_plan_tune(pfx_plan plan, int flag, void *scratch, pfx_stream stream);

/* Number of candidate kernels for `plan` (the set pfx_plan_tune times). */
int pfx_plan_num_candidates(pfx_plan plan, int *count);
/* Switches `plan` to candidate `index` (0 <= index < count). */
int pfx_plan_use_candidate(pfx_plan plan, int index);

/*
 * Saves / loads / clears the tuning cache (text, one entry per line).
 * Replaces: FFTW wisdom export / import (backend/fftw3/fft_backend_fftw.hpp:666-694).
 */
int pfx_tune_cache_save(const char *path);
/* Lines with a key not in the cache's key format or a kernel this build does
 * not have are dropped; the load still succeeds with the rest, and
 * pfx_last_error_message() then names the file and the dropped count. */
int pfx_tune_cache_load(const char *path);
/* Lines the last pfx_tune_cache_load dropped. */
int pfx_tune_cache_dropped(void);
int pfx_tune_cache_clear(void);
/* Number of entries in the tuning cache. */
int pfx_tune_cache_size(void);

/* Kernel chosen for a plan: name buffer, transforms per workgroup, threads
 * per workgroup.  Diagnostics only. */
int pfx_plan_describe(pfx_plan plan, char *name, size_t name_len, int *per_block,
                      int *threads);

/* --------------------------------------------------------------- memory */

/* Replaces: hipvector<T>::resize / ~hipvector (fft_backend_hipfft.hpp:58-121). */
int pfx_malloc(void **ptr, size_t bytes);
int pfx_free(void *ptr);

/* Replaces: HipFFTBackend::memcpy (fft_backend_hipfft.hpp:481-484). */
int pfx_memcpy_async(void *dst, const void *src, size_t bytes, pfx_stream stream);

/*
 * Strided 2D copy: `height` rows of `width` bytes.  Device-to-device copies
 * run as a kernel (no 4 GiB limit on the total extent); copies touching host
 * memory go through the HIP runtime.
 * Replaces: HipFFTBackend::memcpy2d (fft_backend_hipfft.hpp:496-501).
 */
int pfx_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch,
                       size_t width, size_t height, pfx_stream stream);

/* Replaces: the HIP stream owned by HipFFTBackend (fft_backend_hipfft.hpp:206, 231-235, 506-508). */
int pfx_stream_create(pfx_stream *stream);
int pfx_stream_destroy(pfx_stream stream);
int pfx_stream_synchronize(pfx_stream stream);

/* ------------------------------------------------------------ pack/unpack */

/*
 * Pack for every peer p: src_leading rows of src_n[p]*src_trailing elements
 * of elem_size bytes, read from src + src_s[p]*src_trailing (row pitch
 * src_extent*src_trailing elements) and written contiguously at
 * pack + send_displs[p].  Bit-exact C-order subarray serialisation.  One
 * launch for all peers.  Arrays are host arrays of nparts entries.
 * Replaces: the pack loops of exchange_packed / exchange_hybrid
 *   (parafaft_common.hpp:576-590, 711-730).
 */
int pfx_pack(void *pack, const void *src, int nparts, const int *src_n, const int *src_s,
             const uint64_t *send_displs, uint64_t src_leading, uint64_t src_trailing,
             uint64_t src_extent, uint64_t elem_size, pfx_stream stream);

/*
 * Unpack, the mirror of pfx_pack: for every peer p, dst_leading rows of
 * dst_n[p]*dst_trailing elements from recv + recv_displs[p] into
 * dst + dst_s[p]*dst_trailing with row pitch dst_extent*dst_trailing.
 * Replaces: the unpack loops (parafaft_common.hpp:654-668, 844-857).
 */
int pfx_unpack(void *dst, const void *recv, int nparts, const int *dst_n, const int *dst_s,
               const uint64_t *recv_displs, uint64_t dst_leading, uint64_t dst_trailing,
               uint64_t dst_extent, uint64_t elem_size, pfx_stream stream);

/* ------------------------------------------------------------------ scale */

/*
 * data[i] *= s for `count` scalars of the given precision (PFX_F32 / PFX_F64).
 * Replaces: the user-side normalisation loop after backward
 *   (test/c2c/test_fftw_c2c_roundtrip_nd.cpp:73-78, example/c2c.cpp:27-31).
 */
int pfx_scale(void *data, int64_t count, int precision, double s, pfx_stream stream);

/* --------------------------------------------- one-rank N-D fast path */

/*
 * Whole-array in-place transform of a row-major array of `ndim` axes on one
 * GPU: one launch per axis, no scratch copies.  kind PFX_C2C: complex array of
 * `shape`.  kind PFX_R2C: padded real array, last axis stored as
 * 2*(shape[ndim-1]/2+1) scalars; forward = R2C, backward = C2R.
 * Replaces, on one rank: ParaFaFT_C2C::forward/backward (parafaft_c2c.hpp:539-694)
 *   and ParaFaFT_R2C::forward_in_place/backward_in_place (parafaft_r2c.hpp:360-542)
 *   without the exchange_local copies (parafaft_common.hpp:528-535).
 */
int pfx_ndplan_create(pfx_ndplan *plan, int precision, int kind, int ndim, const int *shape);

/* scale: multiplies the result inside the last pass (1.0 = unnormalised). */
int pfx_ndplan_execute(pfx_ndplan plan, int direction, void *data, double scale,
                       pfx_stream stream);

/*
 * R2C N-D plans, out of place (ParaFaFT_R2C::forward(real, complex) and
 * ::backward(complex, real), parafaft_r2c.hpp:333-339, 446-453): forward reads
 * the unpadded real array `real` and leaves the spectrum in `padded`;
 * backward transforms `padded` in place (destroying it) and writes the real
 * result to `real`.  The padding copies are fused into the first R2C and the
 * last C2R pass.
 */
int pfx_ndplan_execute_r2c_oop(pfx_ndplan plan, int direction, void *real, void *padded,
                               double scale, pfx_stream stream);
int pfx_ndplan_destroy(pfx_ndplan plan);

/* Tunes every stage plan (pfx_plan_tune) on `scratch`, a device buffer of the
 * whole array (contents destroyed), and keeps the chunked schedule (below)
 * only if it times faster. */
int pfx_ndplan_tune(pfx_ndplan plan, int flag, void *scratch, pfx_stream stream);

/*
 * Chunked schedule: for ndim >= 3 arrays larger than twice the Infinity Cache
 * budget (PFX_ND_CHUNK_MB, 256 MB by default), every stage but axis 0 runs on
 * `planes` axis-0 planes at a time, so a chunk's later stages find its data
 * in the cache; axis 0 then runs over the whole array.  Same kernels, same
 * results.  `planes` divides n0; 0 = stage by stage.  No counterpart in the
 * reference (its stage loop is parafaft_c2c.hpp:539-602).
 */
int pfx_ndplan_chunk(pfx_ndplan plan, int *planes);
int pfx_ndplan_set_chunk(pfx_ndplan plan, int planes);
/* Chunk plan of stage s (0..ndim-2; ndim-1 is the C2R of an R2C plan), applied
 * at data + c * planes * (plane elements), for per-stage timing. */
int pfx_ndplan_chunk_stage(pfx_ndplan plan, int stage, pfx_plan *out);

/* Stage plan of axis-stage s (0 = last axis), for per-stage timing. */
int pfx_ndplan_stage(pfx_ndplan plan, int stage, pfx_plan *out);

/*
 * Split axis 0 (C2C FP32/FP64, chunked schedule): n0 = q G, with
 * x0 = x1 + q x2 and k0 = k2 + G k1.  Per chunk, one kernel takes the
 * G-point DFTs along axis 0 of planes x1 + q x2, the twiddle W_n0^(x1 k2)
 * and the FFT along the last axis, and stores plane G x1 + k2 of a scratch
 * array (so a chunk's planes stay consecutive); the middle axes run there;
 * then the q-point DFTs along axis 0 (stride G planes) go from the scratch
 * into the array.  Axis 0's column runs get q / n0 of the rows to cross
 * instead of n0, at the cost of one array-sized scratch the plan holds.
 * g = 0 turns it off; pfx_ndplan_tune (Measure/Patient) keeps it only if it
 * times at least 2% faster than the chunked schedule.  Changing the chunk
 * turns it off.  The scratch belongs to the plan: executions of one split
 * plan must not overlap (on different streams).
 * No counterpart in the reference (its stage loop is
 * parafaft_c2c.hpp:539-602).
 */
int pfx_ndplan_split(pfx_ndplan plan, int *g);
int pfx_ndplan_set_split(pfx_ndplan plan, int g);
/* The q-point plan of the split schedule (applied from the scratch into the
 * array, forward), for per-stage timing and kernel names. */
int pfx_ndplan_split_stage(pfx_ndplan plan, pfx_plan *out);

/* ------------------------------------------- many uneven transforms */

/*
 * One transform of a batch: `ndim` (1..4) axes of `shape`, in place at the
 * device pointer `data`.  kind PFX_C2C: complex array; kind PFX_R2C: padded
 * real array (last axis stored as 2*(shape[ndim-1]/2+1) scalars), forward =
 * R2C, backward = C2R.
 */
typedef struct {
  int kind;
  int ndim;
  int shape[4];
  void *data;
} pfx_batch_desc;

typedef struct pfx_batch_t *pfx_batch;

/*
 * Plans `count` independent transforms of different shapes and kinds.  Their
 * stage work is grouped by kernel variant and stage level, so one launch runs
 * the same kernel over every transform that needs it (a device job table maps
 * each workgroup to its transform and tile).  The data pointers are part of
 * the plan.
 * Replaces: one ParaFaFT_C2C / ParaFaFT_R2C object and plan set per shape,
 *   each issuing its own launches (parafaft_c2c.hpp:539-694).
 */
int pfx_batch_create(pfx_batch *batch, int precision, int count, const pfx_batch_desc *descs);

/* All transforms, grouped launches (stage level by stage level). */
/*
 * Plan every transform of the batch with `flag` (pfx_plan_tune on each stage,
 * each distinct geometry timed once thanks to the tuning cache), using each
 * transform's own buffer as the planning scratch, then rebuild the grouped
 * schedule with the chosen kernels.  The FFTW_MEASURE analogue of
 * one-plan-per-shape (fft_backend_fftw.hpp:743-754); buffer contents are
 * lost, as with FFTW_MEASURE.  No counterpart in the reference's batch API,
 * which has none.
 */
int pfx_batch_tune(pfx_batch batch, int flag, pfx_stream stream);

int pfx_batch_execute(pfx_batch batch, int direction, pfx_stream stream);

/* All transforms, one launch per transform stage (the ungrouped baseline). */
int pfx_batch_execute_each(pfx_batch batch, int direction, pfx_stream stream);

/*
 * All transforms in ONE persistent launch (plus a small kernel that clears
 * its counters): a grid of one 1024-thread workgroup per CU claims tickets
 * (one tile of one transform stage) from per-XCD queues -- every transform
 * runs on one XCD, so its stages share an L2 -- and stage s of a transform
 * starts only when every tile of its stage s - 1 has completed (device-side
 * counters).  Complex<float> batches whose stage lengths all
 * have radix plans; otherwise PFX_ERR_UNSUPPORTED (use pfx_batch_execute).
 * Where a stage's kernel and the persistent body share their radices (the
 * default plans), the result is bit-identical to pfx_batch_execute.  A wait
 * that exceeds the spin limit ends the launch early and is reported by
 * pfx_batch_persistent_status.  No counterpart in the reference (its
 * nearest analogue is one ParaFaFT object per shape issuing its own
 * launches, parafaft_c2c.hpp:539-602).
 */
int pfx_batch_execute_persistent(pfx_batch batch, int direction, pfx_stream stream);

/* Synchronises `stream`, then sets *failed to 0 if the last persistent
 * launch of `batch` completed, 1 if it stopped at its spin limit, or 2 if a
 * ticket queue was left with work (no workgroup ran on that XCD); after a
 * nonzero status the results are incomplete. */
int pfx_batch_persistent_status(pfx_batch batch, pfx_stream stream, int *failed);

/* Polls of a completion counter that may find it short before a persistent
 * launch gives up (default 2^21, seconds of waiting; a legitimate wait takes
 * microseconds).  0: give up at the first poll that finds it short. */
int pfx_batch_set_spin_limit(pfx_batch batch, int polls);

/* Jobs (transform stages), tickets (tiles) and workgroups of the persistent
 * schedule for one direction; PFX_ERR_UNSUPPORTED if the batch has none. */
int pfx_batch_persistent_info(pfx_batch batch, int direction, int *jobs, int *tickets,
                              int *workgroups);

/* Launch counts of the two schedules for one direction. */
int pfx_batch_launches(pfx_batch batch, int direction, int *grouped, int *each);

int pfx_batch_destroy(pfx_batch batch);

/* --------------------------------------------------------- HIP graphs */

typedef struct pfx_graph_t *pfx_graph;

/*
 * Capture every pfx call enqueued on `stream` (a non-null stream) between
 * begin and end into an instantiated HIP graph; replay it with
 * pfx_graph_launch.  Captured calls must not synchronise.  No library call
 * records a memset node.
 */
int pfx_graph_begin_capture(pfx_stream stream);
int pfx_graph_end_capture(pfx_stream stream, pfx_graph *graph);
int pfx_graph_launch(pfx_graph graph, pfx_stream stream);
int pfx_graph_destroy(pfx_graph graph);

#ifdef __cplusplus
}
#endif

#endif /* PFX_CAPI_H */
