// pfx_persist.hpp -- one persistent launch for a batch of uneven transforms.
//
// The batch's stage work is cut into tickets: one ticket is one workgroup
// tile of one job, a job being one stage of one transform.  A fixed grid of
// 1024-thread workgroups (every one resident: one per CU) claims tickets
// until none are left; a switch over a fixed list of kernel bodies, every one
// instantiated for exactly 1024 threads and at most 160 KiB of LDS, runs the
// tile.  Stage s of a transform waits until every tile of its stage s - 1 has
// completed (per-job counters in global memory).
//
// Every transform belongs to one XCD: its tickets sit in that XCD's queue,
// which only the workgroups running on that XCD (hardware register XCC_ID)
// claim.  All stages of a transform therefore exchange data through one L2:
// a producer's stores are complete in that L2 when its waves' vector memory
// counters drain (s_waitcnt vmcnt(0)), after which the job counter is
// incremented; a consumer that sees the count issues an agent-scope acquire
// (L1 and L2 drop clean lines; the L2's dirty lines -- the producers' data
// -- stay) before its workgroup reads.  No L2 write-back is needed, which,
// issued per tile on the whole XCD's L2, cost more than the tiles; and the
// acquire is issued once per dependency a workgroup meets, not per tile.
// Transforms are spread over the XCDs largest first, each to the least
// loaded.
//
// Progress: a queue lists its jobs so that a job's dependency comes earlier;
// a workgroup finishes the ticket it runs before it starts the next, and
// holds at most one claimed ticket that has not started; a job's tiles are
// counted before the workgroup waits on anything.  The earliest unfinished
// job of a queue therefore has its dependency done and its tickets either
// running or claimable.  Every wait is still bounded: a spin limit and an
// abort word end the launch with an error flag instead of a hang; and a
// queue left undrained (no workgroup ran on its XCD) is reported too.
#pragma once
#include "pfx_internal.hpp"

namespace pfx {

constexpr int kPersistThreads = 1024;
constexpr int kPersistLds = 163840;  // the whole LDS of a CU
constexpr int kPersistQueues = 8;    // one ticket queue per XCD
constexpr int kCtlStride = 32;       // ints between control words (128 B)
constexpr int kCtlAbort = kPersistQueues * kCtlStride;  // set: every workgroup stops
constexpr int kCtlError = kCtlAbort + kCtlStride;       // nonzero: the launch is incomplete
constexpr int kCtlExit = kCtlError + kCtlStride;        // workgroups that have finished
constexpr int kCtlInts = kCtlExit + kCtlStride;
// kCtlError values
constexpr int kPersistSpinLimit = 1;   // a wait exceeded the spin limit
constexpr int kPersistUndrained = 2;   // a queue had tickets left at the end

// A kernel body of the persistent kernel (host-side description).
struct PersistBody {
  int kind;       // KIND_C2C / KIND_R2C / KIND_C2R
  int n;          // complex FFT length in LDS (R2C/C2R: half the real length)
  int col;        // 1: batch-contiguous layout, 0: row layout
  int tpt;        // threads per transform
  int per_block;  // transforms per tile
  Radices rad;
  const char *name;
};
const PersistBody *persist_bodies(int *count);

struct PJob {
  int body;    // index into persist_bodies()
  int ntiles;  // tiles of this job
  int dep;     // job that must complete first (-1: none)
  int need;    // tiles of `dep`
};
static_assert(sizeof(PJob) == 4 * sizeof(int), "the kernel reads jobs as 4 ints");

struct PersistArgs {
  const PJob *jobs;
  const StockhamArgs *args;  // one per job (tile order 0: tiles are taken in order)
  const int *qstart;         // kPersistQueues rows of njobs + 1 ticket offsets
  int njobs;
  int *ctl;                  // kCtlInts control words (queue heads, abort, error)
  int *done;                 // njobs completed-tile counters
  int spin_limit;            // polls of a counter before a wait gives up
  int nqueues;               // queues in use: the XCDs of the device
};

// Workgroups of a persistent launch: every CU's resident count (all of them
// must be resident at once).
int persist_grid();
// Ticket queues: one per XCD of the current device (32 CUs per XCD).
int persist_queues();
// Zeroes the control words and the job counters (a kernel, so the call can
// be captured into a graph), then launches the persistent kernel.
hipError_t launch_persist(const PersistArgs &a, int grid, hipStream_t s);

}  // namespace pfx
