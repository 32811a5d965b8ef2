// pfx_persist.hip -- the persistent kernel of the uneven batch (see
// pfx_persist.hpp for the scheme) and its body table.
#include <hip/hip_runtime.h>

// bodies read the thread index through an opaque statement (pfx_stockham.hpp)
#define PFX_OPAQUE_TID 1
#include "pfx_persist.hpp"
#include "pfx_persist_list.inc"

namespace pfx {

namespace {

// Host-side descriptions, in the order of the device dispatch below.
#define PFX_PB_ROW(TPT, C, ...) \
  {KIND_C2C, prod<__VA_ARGS__>(), 0, TPT, C, radices<__VA_ARGS__>(), "row:" #__VA_ARGS__},
#define PFX_PB_WROW(TPT, C, ...) \
  {KIND_C2C, prod<__VA_ARGS__>(), 0, TPT, C, radices<__VA_ARGS__>(), "rowwide:" #__VA_ARGS__},
#define PFX_PB_COL(TPT, C, ...) \
  {KIND_C2C, prod<__VA_ARGS__>(), 1, TPT, C, radices<__VA_ARGS__>(), "col:" #__VA_ARGS__},
#define PFX_PB_WCOL(TPT, C, ...) \
  {KIND_C2C, prod<__VA_ARGS__>(), 1, TPT, C, radices<__VA_ARGS__>(), "wide:" #__VA_ARGS__},
#define PFX_PB_WCOLG(TPT, C, ...) \
  {KIND_C2C, prod<__VA_ARGS__>(), 1, TPT, C, radices<__VA_ARGS__>(), "widegtw:" #__VA_ARGS__},
#define PFX_PB_REAL(TPT, C, ...)                                                              \
  {KIND_R2C, prod<__VA_ARGS__>(), 0, TPT, C, radices<__VA_ARGS__>(), "r2c:" #__VA_ARGS__},   \
  {KIND_C2R, prod<__VA_ARGS__>(), 0, TPT, C, radices<__VA_ARGS__>(), "c2r:" #__VA_ARGS__},
const PersistBody kBodies[] = {PFX_PERSIST_LIST};
#undef PFX_PB_ROW
#undef PFX_PB_WROW
#undef PFX_PB_COL
#undef PFX_PB_WCOL
#undef PFX_PB_WCOLG
#undef PFX_PB_REAL

using T = float;

// Every body fills exactly one workgroup and fits the block.
#define PFX_PB_CHECK(TPT, C, LDS)                                            \
  static_assert((TPT) * (C) == kPersistThreads, "tile of one workgroup");    \
  static_assert((LDS) <= kPersistLds, "tile within the workgroup's LDS");

// The tile of job `a`: `body` picks the instantiation.  Bodies end with their
// last stores issued; the caller waits for them.
__device__ __forceinline__ void run_body(int body, const StockhamArgs &a, int64_t tile,
                                         int64_t ntiles, unsigned char *smem) {
  int id = 0;
#define PFX_PB_ROW(TPT, C, ...)                                                          \
  if (body == id++) {                                                                    \
    PFX_PB_CHECK(TPT, C, (C2CLds<T, TPT, C, false, false, __VA_ARGS__>::bytes))         \
    c2c_body_in<T, TPT, C, false, false, __VA_ARGS__>(a, tile, ntiles, smem);            \
    return;                                                                              \
  }
#define PFX_PB_COL(TPT, C, ...)                                                          \
  if (body == id++) {                                                                    \
    PFX_PB_CHECK(TPT, C, (C2CLds<T, TPT, C, true, false, __VA_ARGS__>::bytes))          \
    c2c_body_in<T, TPT, C, true, false, __VA_ARGS__>(a, tile, ntiles, smem);             \
    return;                                                                              \
  }
#define PFX_PB_WROW(TPT, C, ...)                                                               \
  if (body == id++) {                                                                          \
    PFX_PB_CHECK(TPT, C, (WideLds<T, TPT, C, true, false, true, __VA_ARGS__>::bytes))         \
    c2c_wide_body_in<T, TPT, C, true, false, false, true, __VA_ARGS__>(a, tile, ntiles, smem); \
    return;                                                                                    \
  }
#define PFX_PB_WCOL(TPT, C, ...)                                                                \
  if (body == id++) {                                                                           \
    PFX_PB_CHECK(TPT, C, (WideLds<T, TPT, C, false, false, false, __VA_ARGS__>::bytes))        \
    c2c_wide_body_in<T, TPT, C, false, false, false, false, __VA_ARGS__>(a, tile, ntiles, smem); \
    return;                                                                                     \
  }
#define PFX_PB_WCOLG(TPT, C, ...)                                                              \
  if (body == id++) {                                                                          \
    PFX_PB_CHECK(TPT, C, (WideLds<T, TPT, C, false, false, true, __VA_ARGS__>::bytes))        \
    c2c_wide_body_in<T, TPT, C, false, false, false, true, __VA_ARGS__>(a, tile, ntiles, smem); \
    return;                                                                                    \
  }
#define PFX_PB_REAL(TPT, C, ...)                                 \
  if (body == id++) {                                            \
    PFX_PB_CHECK(TPT, C, (RealLds<T, TPT, C, __VA_ARGS__>::bytes)) \
    r2c_body_in<T, TPT, C, false, __VA_ARGS__>(a, tile, smem);   \
    return;                                                      \
  }                                                              \
  if (body == id++) {                                            \
    c2r_body_in<T, TPT, C, false, __VA_ARGS__>(a, tile, smem);   \
    return;                                                      \
  }
  PFX_PERSIST_LIST
#undef PFX_PB_ROW
#undef PFX_PB_WROW
#undef PFX_PB_COL
#undef PFX_PB_WCOL
#undef PFX_PB_WCOLG
#undef PFX_PB_REAL
}

// s_waitcnt vmcnt(0): this wave's outstanding vector memory operations done
// (expcnt and lgkmcnt fields at their no-wait maxima).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Global-address-space views of the kernel's tables (they arrive as kernel
// arguments, but a helper that takes them from a struct in memory would see
// generic pointers and use flat accesses).
using gint = __attribute__((address_space(1))) int;
using cgint = const __attribute__((address_space(1))) int;

__device__ __forceinline__ int load_agent(cgint *p) {
  return __hip_atomic_load((const int *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int add_agent(gint *p, int v) {
  return __hip_atomic_fetch_add((int *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Job holding position p of a queue (qrow: that queue's njobs + 1 offsets).
__device__ __forceinline__ int job_at(cgint *qrow, int njobs, int p) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (qrow[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The job holding a run of positions of the queue, kept by the claiming lane
// so that consecutive tickets of one job need no table search.
struct QRun {
  int job = 0, lo = 0, hi = 0;  // positions [lo, hi)
};

// Waits until counter `done` reaches `need`.  False if the launch is aborted
// or this wait exceeds the spin limit (which aborts the launch).
__device__ __forceinline__ bool wait_done(cgint *done, int need, gint *ctl, int spin_limit) {
  int spins = 0;
  while (load_agent(done) < need) {
    if (load_agent(ctl + kCtlAbort)) return false;
    if (spins++ >= spin_limit) {
      __hip_atomic_store((int *)(ctl + kCtlError), kPersistSpinLimit, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store((int *)(ctl + kCtlAbort), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

// The XCD this workgroup runs on (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcc_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15;
}

// Makes what other CUs of this XCD stored visible to this workgroup's loads:
// the agent-scope acquire (buffer_inv sc1: this CU's L1 and the XCD's L2
// drop their clean lines; the producers' stores, which this XCD's L2 holds,
// stay).  An L1-only invalidate (buffer_inv sc0) was measured not to be
// enough: consumers then read stale values.
__device__ __forceinline__ void acquire_xcd() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__global__ void __launch_bounds__(kPersistThreads) persist_kernel(PersistArgs pa) {
  __shared__ __align__(16) unsigned char smem[kPersistLds];
  // the ticket is handed to the workgroup through the first words of the
  // block; a barrier separates reading them from the body's first LDS write
  int *slot = reinterpret_cast<int *>(smem);
  cgint *jobs = (cgint *)pa.jobs;  // PJob = {body, ntiles, dep, need}
  gint *ctl = (gint *)pa.ctl;
  gint *done = (gint *)pa.done;
  const int nj = pa.njobs;
  const int myq = xcc_id() % pa.nqueues;
  cgint *qrow = (cgint *)pa.qstart + (size_t)myq * (nj + 1);
  const int qlen = qrow[nj];
  gint *head = ctl + myq * kCtlStride;
  // uniform state: every thread tracks it alike from the broadcast tickets
  int sig_job = -1, sig_cnt = 0;  // completed tiles of sig_job not yet counted
  int acq = -1;                   // the dependency this workgroup last waited for
  // claiming lane (thread 0): the next claimed position and the job run
  QRun run;
  int p = 0;
  if (threadIdx.x == 0) p = add_agent(head, 1);
  for (;;) {
    if (threadIdx.x == 0) {
      int job = -1, tile = 0;
      if (p < qlen) {
        if (p < run.lo || p >= run.hi) {
          run.job = job_at(qrow, nj, p);
          run.lo = qrow[run.job];
          run.hi = qrow[run.job + 1];
        }
        job = run.job;
        tile = p - run.lo;
      }
      slot[0] = job;
      slot[1] = tile;
    }
    __syncthreads();
    const int job = slot[0];
    const int tile = slot[1];
    __syncthreads();
    const int dep = job >= 0 ? jobs[(size_t)job * 4 + 2] : -1;
    const bool need_wait = dep >= 0 && dep != acq;
    // Counting is batched per job: the tiles of one job done back to back
    // are counted together, when the workgroup moves to another job, before
    // it waits for anything, and before it exits -- so it never waits while
    // it holds uncounted tiles, and every tile is counted.
    if (sig_cnt > 0 && (job < 0 || job != sig_job || need_wait)) {
      wait_vmem();  // every wave's stores of the counted tiles are in the L2
      __syncthreads();
      if (threadIdx.x == 0) add_agent(done + sig_job, sig_cnt);
      sig_cnt = 0;
    }
    if (job < 0) break;
    if (need_wait) {
      if (threadIdx.x == 0) {
        const bool ok = wait_done(done + dep, jobs[(size_t)job * 4 + 3], ctl, pa.spin_limit);
        acquire_xcd();
        slot[2] = ok ? 1 : 0;
      }
      __syncthreads();
      const bool ok = slot[2] != 0;
      __syncthreads();
      if (!ok) break;
      acq = dep;  // its data, once seen, stays visible (lines refilled from the L2)
    }
    // the next ticket is claimed while this one runs (its atomic's latency
    // hides behind the body's first loads)
    if (threadIdx.x == 0) p = add_agent(head, 1);
    run_body(jobs[(size_t)job * 4], pa.args[job], tile, jobs[(size_t)job * 4 + 1], smem);
    __syncthreads();
    sig_job = job;  // a different job's count was flushed before this one ran
    ++sig_cnt;
  }
  // the last workgroup to finish checks that every queue was drained (a
  // queue of an XCD that ran no workgroup would leave its transforms undone)
  if (threadIdx.x == 0) {
    if (add_agent(ctl + kCtlExit, 1) == (int)gridDim.x - 1) {
      for (int q = 0; q < pa.nqueues; ++q)
        if (load_agent(ctl + q * kCtlStride) < pa.qstart[(size_t)q * (nj + 1) + nj] &&
            load_agent(ctl + kCtlError) == 0)
          __hip_atomic_store((int *)(ctl + kCtlError), kPersistUndrained, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void persist_reset_kernel(int *ctl, int *done, int njobs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kCtlInts) ctl[i] = 0;
  if (i < njobs) done[i] = 0;
}

}  // namespace

const PersistBody *persist_bodies(int *count) {
  *count = (int)(sizeof(kBodies) / sizeof(kBodies[0]));
  return kBodies;
}

int persist_queues() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 1;
  const int q = cus / 32;
  return q < 1 ? 1 : (q > kPersistQueues ? kPersistQueues : q);
}

int persist_grid() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persist_kernel, kPersistThreads, 0) !=
      hipSuccess)
    return 0;
  return cus * per_cu;
}

hipError_t launch_persist(const PersistArgs &a, int grid, hipStream_t s) {
  const int n = a.njobs > kCtlInts ? a.njobs : kCtlInts;
  persist_reset_kernel<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(a.ctl, a.done, a.njobs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  persist_kernel<<<dim3(grid), dim3(kPersistThreads), 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace pfx
