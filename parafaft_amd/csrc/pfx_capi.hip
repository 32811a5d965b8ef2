// pfx_capi.hip -- implementation of the C ABI declared in include/pfx_capi.h.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pfx_capi.h"
#include "pfx_internal.hpp"
#include "pfx_persist.hpp"

using namespace pfx;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(PFX_ERR_HIP, "HIP error in %s: %s", what, hipGetErrorString(e));
}

#define PFX_HIP(call, what)                    \
  do {                                         \
    hipError_t e_ = (call);                    \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// The kernel registry: every table's entries, with the grouped and segmented
// launchers (compiled in units of their own, listed in the same order)
// merged in by index.
struct Registry {
  std::vector<KernelEntry> storage;
  std::vector<const KernelEntry *> entries;
  Registry() {
    struct Src {
      const KernelEntry *(*table)(int *);
      const GroupLaunchFn *(*group)(int *);
      const LaunchFn *(*seg)(int *);
    };
    const Src srcs[] = {{table_f32_row, group_f32_row, seg_f32_row},
                        {table_f32_col, group_f32_col, seg_f32_col},
                        {table_f32_real, group_f32_real, seg_f32_real},
                        {table_f64_row, group_f64_row, seg_f64_row},
                        {table_f64_col, group_f64_col, seg_f64_col},
                        {table_f64_real, group_f64_real, seg_f64_real},
                        {table_f32_w3232, group_f32_w3232, seg_f32_w3232},
                        {table_split, nullptr, nullptr},
                        {table_split32, nullptr, nullptr},
                        {table_nt, group_nt, nullptr},
                        {table_f32_w3232nt, nullptr, nullptr}};
    for (const Src &src : srcs) {
      int n = 0, ng = 0, ns = 0;
      const KernelEntry *t = src.table(&n);
      const GroupLaunchFn *g = src.group ? src.group(&ng) : nullptr;
      const LaunchFn *sg = src.seg ? src.seg(&ns) : nullptr;
      if ((g && ng != n) || (sg && ns != n)) {
        fprintf(stderr, "pfx: kernel lists out of step (%d entries, %d grouped, %d segmented)\n",
                n, ng, ns);
        abort();
      }
      for (int i = 0; i < n; ++i) {
        KernelEntry e = t[i];
        e.launch_group = g ? g[i] : nullptr;
        e.launch_seg = sg ? sg[i] : nullptr;
        storage.push_back(e);
      }
    }
    for (const KernelEntry &e : storage) entries.push_back(&e);
  }
};

const Registry &registry() {
  static Registry r;
  return r;
}

// XCD-aware tile order for batch-contiguous kernels (PFX_XCD_REMAP=0 turns it
// off, for measurement).
// Both are read when a plan is created.
// PFX_XCD_REMAP=0: launch order; 1 (default): one contiguous run of tiles
// per XCD; K > 1: XCDs take turns on chunks of K tiles.
int xcd_remap_mode() {
  const char *s = getenv("PFX_XCD_REMAP");
  return s ? atoi(s) : 1;
}

// Override of the batch-contiguous tile: PFX_COL_TILE=<transforms per workgroup>.
int col_tile_override() {
  const char *s = getenv("PFX_COL_TILE");
  return s ? atoi(s) : 0;
}

// Lengths without a tiled kernel run the long plans (four-step or chirp),
// which measured 1.1-57x faster than the direct DFT from length 7 to 4093
// (256 MB of rows or columns); PFX_LONG_MIN=<length> (measurement) keeps
// the direct DFT below that length, up to its limit.
int64_t long_min_length() {
  const char *s = getenv("PFX_LONG_MIN");
  return s ? atoll(s) : 2;
}

// Four-step plans make two passes where both lengths have the fused
// kernels; PFX_FOURSTEP_PASSES=4 (measurement) keeps the four-pass form
// (tiled sub-plans, twiddle, transpose).
bool fourstep_two_pass() {
  const char *s = getenv("PFX_FOURSTEP_PASSES");
  return !(s && atoi(s) == 4);
}

// Chirp plans whose convolution length has a tiled radix plan run one fused
// kernel; PFX_CHIRP_FUSED=0 (measurement) keeps the five-pass form.
bool chirp_fused_enabled() {
  const char *s = getenv("PFX_CHIRP_FUSED");
  return !(s && atoi(s) == 0);
}

// Kernel-choice override for measurement: PFX_KERNEL_MATCH=<substring of the
// radix description>, e.g. "wide:32, 32".
const char *kernel_match_override() { return getenv("PFX_KERNEL_MATCH"); }

const KernelEntry *find_kernel(int prec, int kind, int64_t n, bool col) {
  const KernelEntry *best = nullptr;
  // Highest `preferred` wins unless PFX_COL_TILE names an available tile.
  const int want = col ? col_tile_override() : 0;
  for (const KernelEntry *e : registry().entries) {
    if (e->prec != prec || e->kind != kind || e->n != n) continue;
    if (kind == KIND_C2C && e->col != (col ? 1 : 0)) continue;
    const char *m = kernel_match_override();
    if (m && *m && strstr(e->radices, m)) return e;
    if (want > 0 && e->per_block == want) return e;
    if (!best || e->preferred > best->preferred) best = e;
  }
  return best;
}

int base_prec(int precision) {
  return precision == PFX_F64 ? 1 : 0;
}

}  // namespace

// A length with no single-kernel plan beyond the direct-DFT limit
// (pfx_fourstep.hip has the element-wise passes).
//  LONG_FOURSTEP (C2C, n = n1 * n2 with tiled n1 and n2): `a` transforms
//    along j2 (length n2, element stride n1 s) from the caller's array into
//    `scratch` at the same offsets, the twiddle W_n^(j1 k2) follows, `c`
//    transforms along j1 (length n1) in place on the scratch, and the
//    transpose writes n2 k1 + k2 back.
//  LONG_CHIRP (C2C lengths that do not split, Bluestein): the chirp w[j] =
//    exp(-pi i j^2 / n) premultiplies into scratch rows of M >= 2n - 1
//    (zero-padded), `c` (length M) forward, times the filter's spectrum,
//    `c` backward, and w postmultiplies (with 1/M and the caller's scale).
//  LONG_REAL_HALF (even real length n = 2h): `c` is the length-h C2C plan
//    of the padded rows read as complex pairs, with the R2C post-pass / C2R
//    pre-pass around it.
//  LONG_REAL_FULL (odd real length): `c` is a length-n C2C plan on scratch
//    rows of n complex values, the real rows loaded into / stored from them.
enum { LONG_FOURSTEP = 0, LONG_CHIRP = 1, LONG_REAL_HALF = 2, LONG_REAL_FULL = 3 };
// Scratch rows of the chirp and odd-real modes cover at most this many bytes
// (the batch runs in slices of that many rows): bounded memory, and a slice
// stays in the Infinity Cache between its passes.
constexpr int64_t kLongRowsBytes = int64_t(256) << 20;
constexpr int64_t kChirpMaxLength = int64_t(1) << 23;  // M <= 2^24 = 4096^2 splits
struct pfx_plan_t;
struct LongPlan {
  int mode = LONG_FOURSTEP;
  int64_t n1 = 0, n2 = 0;
  int64_t M = 0;            // LONG_CHIRP: convolution length
  int64_t rows = 0;         // LONG_CHIRP / LONG_REAL_FULL: scratch rows (a slice of the batch)
  pfx_plan_t *a = nullptr;
  pfx_plan_t *c = nullptr;
  // four-step, batch-contiguous caller: the fused first pass (its per-pass
  // table `ptw1`) and the second pass `c2`, instead of a, twiddle, c and
  // the transpose
  pfx_plan_t *c2 = nullptr;
  void *ptw1 = nullptr;
  // four-step, rows (any caller that is not batch-contiguous): both passes
  // fused (pfx_fourstep.hip), `ptw1` the length-n2 pass's table, `ptw2` the
  // length-n1 pass's; no sub-plans
  const ChirpFused *cf = nullptr;  // LONG_CHIRP in one kernel (`ptw1` its pass table)
  bool rows2 = false;
  void *ptw2 = nullptr;
  void *tw2 = nullptr;  // rows2: W_n^(c k2) at 16 k2 + c

  void *scratch = nullptr;  // the mode's work rows (four-step: one outer
                            // slice's footprint of the caller's layout)
  void *tw = nullptr;       // four-step W_n^i, i < n; real halves W_n^k; chirp w
  void *filt = nullptr;     // LONG_CHIRP: the filters' spectra, forward then backward
};

struct pfx_plan_t {
  int precision = PFX_F32;
  int kind = PFX_C2C;
  int64_t length = 0;     // transform length (real length for R2C/C2R)
  int64_t inner = 0, outer = 1;
  int64_t stride = 1, dist = 0, dist_outer = 0;
  const KernelEntry *kernel = nullptr;  // null: direct-DFT fallback or `lp`
  int xcd_remap = 1;
  void *tw = nullptr;
  void *ptw = nullptr;  // per-pass table of the tiled kernel
  int64_t tw_len = 0;
  LongPlan *lp = nullptr;  // long lengths (kernel is null)
};

namespace {

// Tuning cache: plan geometry -> chosen kernel (precision code, radix
// description and workgroup tile identify it uniquely in the registry).
std::mutex g_tune_mu;
std::map<std::string, std::string> g_tune;
int g_tune_dropped = 0;  // lines the last pfx_tune_cache_load dropped

std::string entry_id(const KernelEntry *k) {
  return std::to_string(k->prec) + "|" + std::to_string(k->kind) + "|" + std::to_string(k->n) +
         "|" + std::to_string(k->col) + "|" + std::to_string(k->per_block) + "|" + k->radices;
}

const KernelEntry *entry_by_id(const std::string &id) {
  for (const KernelEntry *e : registry().entries)
    if (entry_id(e) == id) return e;
  return nullptr;
}

// A cache value is "<entry id>|order=<tile order>".
bool parse_tuned(const std::string &v, const KernelEntry **k, int *order) {
  const size_t pos = v.rfind("|order=");
  const std::string id = pos == std::string::npos ? v : v.substr(0, pos);
  *order = pos == std::string::npos ? 1 : atoi(v.c_str() + pos + 7);
  *k = entry_by_id(id);
  return *k != nullptr;
}

std::string plan_key(const pfx_plan_t *p) {
  char buf[256];
  snprintf(buf, sizeof(buf), "p%d k%d n%lld i%lld o%lld s%lld d%lld do%lld", p->precision, p->kind,
           (long long)p->length, (long long)p->inner, (long long)p->outer, (long long)p->stride,
           (long long)p->dist, (long long)p->dist_outer);
  return buf;
}

// True when the kernel `e` computes the transform the plan's default
// kernel `k0` computes (same kind, length, layout and arithmetic) and one
// launch of it fits the grid limit.  A tuning-cache value is input (a
// loaded file may pair a key with another length's or layout's kernel), so
// every use of one goes through this check.
bool kernel_fits_plan(const pfx_plan_t *p, const KernelEntry *k0, const KernelEntry *e,
                      int order) {
  if (!k0 || !e || order < 0) return false;
  if (e->kind != k0->kind || e->n != k0->n || e->col != k0->col || e->prec != k0->prec)
    return false;
  return (p->inner * p->outer + e->per_block - 1) / e->per_block <= 0x7fffffff;
}

// The cache entry of the plan's key, if there is one and it passes
// kernel_fits_plan against `k0`.  Caller holds g_tune_mu.
bool cached_choice(const pfx_plan_t *p, const KernelEntry *k0, const KernelEntry **e,
                   int *order) {
  auto it = g_tune.find(plan_key(p));
  if (it == g_tune.end() || !parse_tuned(it->second, e, order)) return false;
  return kernel_fits_plan(p, k0, *e, *order);
}

// Valid-looking plan key: the fields plan_key writes, in its order, and
// optionally the chunk suffix of an in-context choice (" chunkP ofN").
bool is_plan_key(const char *s) {
  int pr, kd, P, n0;
  long long n, i, o, st, d, dout;
  char tail;
  int fields = sscanf(s, "p%d k%d n%lld i%lld o%lld s%lld d%lld do%lld%c", &pr, &kd, &n, &i, &o,
                      &st, &d, &dout, &tail);
  if (fields == 8) return true;
  if (sscanf(s, "p%d k%d n%lld i%lld o%lld s%lld d%lld do%lld chunk%d of%d%c", &pr, &kd, &n, &i,
             &o, &st, &d, &dout, &P, &n0, &tail) == 10)
    return true;
  // a backward chunk plan's in-context choice
  char bwd[4];
  return sscanf(s, "p%d k%d n%lld i%lld o%lld s%lld d%lld do%lld chunk%d of%d %3s%c", &pr, &kd, &n,
                &i, &o, &st, &d, &dout, &P, &n0, bwd, &tail) == 11 &&
         !strcmp(bwd, "bwd");
}

}  // namespace

struct pfx_ndplan_t {
  int precision = PFX_F32;
  int kind = PFX_C2C;
  int ndim = 0;
  std::vector<int> shape;
  std::vector<pfx_plan> stages;  // stages[s] transforms axis ndim-1-s (+ C2R for R2C)
  // Chunked schedule: every stage except axis 0 runs on `chunk` axis-0 planes
  // at a time (cstages: the same plans restricted to a chunk, + C2R for R2C),
  // so a chunk's later stages find its data in the Infinity Cache.  0 = off.
  int chunk = 0;
  int64_t plane_elems = 0;        // complex elements per axis-0 plane
  std::vector<pfx_plan> cstages;
  // The same chunk plans for backward transforms, whose chunks run the middle
  // axes before the last (reading HBM first, the cache second): tuned in the
  // backward schedule, so their kernels may differ from the forward ones.
  std::vector<pfx_plan> bstages;
  // Split axis 0 (C2C, chunked, n0 = q G; pfx_planerow.hip): per chunk, the
  // G-point DFTs along axis 0 fused with the last axis (into the scratch),
  // then the middle axes on the scratch; then the q-point DFTs along axis 0
  // from the scratch into the array.  0 = off.
  int split = 0;
  void *split_buf = nullptr;      // scratch: one whole array
  void *split_tw = nullptr;       // W_n0^i, i < n0
  void *split_ptw = nullptr;      // per-pass table of the fused kernel's last-axis radices
  pfx_plan split_col = nullptr;   // the q-point pass
};

static int split_set(pfx_ndplan p, int g);
static void split_free(pfx_ndplan p);

extern "C" {

const char *pfx_error_string(int status) {
  switch (status) {
    case PFX_SUCCESS: return "success";
    case PFX_ERR_INVALID_ARG: return "invalid argument";
    case PFX_ERR_UNSUPPORTED: return "unsupported configuration";
    case PFX_ERR_HIP: return "HIP runtime error";
    case PFX_ERR_ALLOC: return "allocation failure";
    case PFX_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

const char *pfx_last_error_message(void) { return g_last_error.c_str(); }

const char *pfx_version(void) { return "parafaft-mi355 0.1 (gfx950)"; }

static int plan_create(pfx_plan *out, int precision, int kind, int64_t length, int64_t inner,
                       int64_t stride, int64_t dist, int64_t outer, int64_t dist_outer);

// Releases a long plan's parts (the plan itself stays).
static void free_long(pfx_plan p) {
  if (!p->lp) return;
  pfx_plan_destroy(p->lp->a);
  pfx_plan_destroy(p->lp->c);
  pfx_plan_destroy(p->lp->c2);
  if (p->lp->ptw1) hipFree(p->lp->ptw1);
  if (p->lp->ptw2) hipFree(p->lp->ptw2);
  if (p->lp->tw2) hipFree(p->lp->tw2);
  if (p->lp->scratch) hipFree(p->lp->scratch);
  if (p->lp->tw) hipFree(p->lp->tw);
  if (p->lp->filt) hipFree(p->lp->filt);
  delete p->lp;
  p->lp = nullptr;
}

// Sub-plans, scratch and twiddles of a long plan (see LongPlan).
// Convolution length of a chirp plan over n points: the shortest M >= 2n - 1
// of the forms b 2^k (b = 1, 3, 5, 15, 25, 125) that runs as one row
// kernel; else the shortest such M that splits into two tiled lengths (the
// length-M/f pass of contiguous rows runs batch-contiguous, the length-f
// pass in rows).  A one-kernel M at most twice the shortest costs less than
// any four-step M.  2^k always qualifies up to 4096^2.
static int64_t chirp_length(int bp, int64_t n) {
  static const int64_t kBase[] = {1, 3, 5, 15, 25, 125};
  const int64_t need = 2 * n - 1;
  int64_t one = 0, split = 0;
  for (const int64_t b : kBase) {
    int64_t m = b;
    while (m < need) m <<= 1;
    if (find_kernel(bp, KIND_C2C, m, false)) {
      if (!one || m < one) one = m;
      continue;
    }
    if (split && m >= split) continue;
    for (int64_t f = 2; f * f <= m; ++f)
      if (m % f == 0 && find_kernel(bp, KIND_C2C, m / f, true) &&
          find_kernel(bp, KIND_C2C, f, false)) {
        split = m;
        break;
      }
  }
  return one ? one : split;
}

static int make_long(pfx_plan p) {
  const int bp = base_prec(p->precision);
  const size_t esz = bp ? 16 : 8;
  LongPlan *lp = new (std::nothrow) LongPlan;
  if (!lp) return fail(PFX_ERR_ALLOC, "out of host memory");
  p->lp = lp;
  const int64_t n = p->length, B = p->inner, st = p->stride, d = p->dist;
  const int64_t total = p->inner * p->outer;
  int rc = PFX_SUCCESS;
  if (p->kind == PFX_C2C) {
    // The passes' transforms: for a batch-contiguous caller (dist 1) the
    // caller's batch is their inner index, so both passes are
    // batch-contiguous; otherwise the inner index is j1 (length-n2 pass:
    // stride n1 st, dist st, batch-contiguous when st == 1) or k2
    // (length-n1 pass: stride st, dist n1 st, a row pass).
    const bool bc = d == 1 && st > 1;
    // the most even split into two tiled lengths (n1 <= n2) whose passes
    // have kernels in the layouts they run in (some lengths have row
    // kernels only)
    for (int64_t f = 2; f * f <= n; ++f)
      if (n % f == 0 && find_kernel(bp, KIND_C2C, n / f, st == 1 || bc) &&
          find_kernel(bp, KIND_C2C, f, bc))
        lp->n1 = f;
    if (lp->n1) {
      lp->mode = LONG_FOURSTEP;
      lp->n2 = n / lp->n1;
      const int64_t n1 = lp->n1, n2 = lp->n2;
      Radices rad1, radn1;
      const bool plain = p->precision == PFX_F32 || p->precision == PFX_F64;
      if (!bc && plain && fourstep_two_pass() && fourstep_rows_radices(n2, &rad1) &&
          fourstep_rows_radices(n1, &radn1)) {
        // two passes: the length-n2 DFTs with the twiddle into the scratch,
        // then the length-n1 DFTs stored transposed into the caller's array
        lp->rows2 = true;
        PFX_HIP(hipMalloc(&lp->ptw1, (size_t)n2 * esz), "hipMalloc four-step pass table");
        PFX_HIP(make_pass_twiddles(bp, rad1, lp->ptw1, nullptr), "four-step pass table");
        PFX_HIP(hipMalloc(&lp->ptw2, (size_t)n1 * esz), "hipMalloc four-step pass table");
        PFX_HIP(make_pass_twiddles(bp, radn1, lp->ptw2, nullptr), "four-step pass table");
      } else if (bc && plain && fourstep_two_pass() && fourstep_cols_pass1_radices(n2, &rad1)) {
        // two passes: the fused first pass, then the length-n1 DFTs from
        // the scratch ((n2 j1 + k2) st + b) into the caller's array
        rc = plan_create(&lp->c2, p->precision, PFX_C2C, n1, B, n2 * st, 1, n2, st);
        if (rc != PFX_SUCCESS) return rc;
        PFX_HIP(hipMalloc(&lp->ptw1, (size_t)n2 * esz), "hipMalloc four-step pass table");
        PFX_HIP(make_pass_twiddles(bp, rad1, lp->ptw1, nullptr), "four-step pass table");
      } else if (bc) {
        rc = plan_create(&lp->a, p->precision, PFX_C2C, n2, B, n1 * st, 1, n1, st);
        if (rc == PFX_SUCCESS)
          rc = plan_create(&lp->c, p->precision, PFX_C2C, n1, B, st, 1, n2, n1 * st);
      } else {
        rc = plan_create(&lp->a, p->precision, PFX_C2C, n2, n1, n1 * st, st, B, d);
        if (rc == PFX_SUCCESS)
          rc = plan_create(&lp->c, p->precision, PFX_C2C, n1, n2, st, n1 * st, B, d);
      }
      if (rc != PFX_SUCCESS) return rc;
      if (B > 0) {
        const int64_t span = (B - 1) * d + (n - 1) * st + 1;
        PFX_HIP(hipMalloc(&lp->scratch, (size_t)span * esz), "hipMalloc long-plan scratch");
      }
      PFX_HIP(hipMalloc(&lp->tw, (size_t)n * esz), "hipMalloc long-plan twiddles");
      PFX_HIP(make_twiddles(bp, n, lp->tw, nullptr), "long-plan twiddles");
      if (lp->rows2) {
        PFX_HIP(hipMalloc(&lp->tw2, (size_t)(16 * n2) * esz), "hipMalloc four-step row table");
        PFX_HIP(make_fourstep_rows_table(bp, lp->tw2, lp->tw, n2, nullptr),
                "four-step row table");
      }
    } else {
      if (n > kChirpMaxLength)
        return fail(PFX_ERR_UNSUPPORTED, "no split into two tiled lengths, and beyond the "
                                         "chirp-convolution limit of %lld",
                    (long long)kChirpMaxLength);
      lp->mode = LONG_CHIRP;
      lp->M = chirp_length(bp, n);
      if ((p->precision == PFX_F32 || p->precision == PFX_F64) && chirp_fused_enabled())
        lp->cf = bp ? chirp_fused_f64(lp->M) : chirp_fused_f32(lp->M);
      if (lp->cf) {
        // one kernel per execution, in place: no scratch rows, no sub-plan
        PFX_HIP(hipMalloc(&lp->ptw1, (size_t)lp->M * esz), "hipMalloc chirp pass table");
        PFX_HIP(make_pass_twiddles(bp, lp->cf->rad, lp->ptw1, nullptr), "chirp pass table");
      } else {
        lp->rows = std::max<int64_t>(
            1, std::min<int64_t>(total, kLongRowsBytes / (lp->M * (int64_t)esz)));
        rc = plan_create(&lp->c, p->precision, PFX_C2C, lp->M, lp->rows, 1, lp->M, 1, 0);
        if (rc != PFX_SUCCESS) return rc;
        if (total > 0)
          PFX_HIP(hipMalloc(&lp->scratch, (size_t)(lp->rows * lp->M) * esz),
                  "hipMalloc chirp rows");
      }
      PFX_HIP(hipMalloc(&lp->tw, (size_t)n * esz), "hipMalloc chirp");
      PFX_HIP(hipMalloc(&lp->filt, (size_t)(2 * lp->M) * esz), "hipMalloc chirp filters");
      PFX_HIP(make_chirp(bp, lp->tw, n, nullptr), "chirp");
      PFX_HIP(make_chirp_filter(bp, lp->filt, lp->tw, n, lp->M, 0, nullptr), "chirp filter");
      PFX_HIP(make_chirp_filter(bp, static_cast<char *>(lp->filt) + (size_t)lp->M * esz, lp->tw,
                                n, lp->M, 1, nullptr), "chirp filter");
      pfx_plan f = nullptr;
      rc = plan_create(&f, p->precision, PFX_C2C, lp->M, 2, 1, lp->M, 1, 0);
      if (rc == PFX_SUCCESS) rc = pfx_execute(f, PFX_FORWARD, lp->filt, nullptr);
      const hipError_t e = hipStreamSynchronize(nullptr);
      pfx_plan_destroy(f);
      if (rc != PFX_SUCCESS) return rc;
      PFX_HIP(e, "chirp filter spectra");
    }
  } else if (n % 2 == 0) {
    lp->mode = LONG_REAL_HALF;
    rc = plan_create(&lp->c, p->precision, PFX_C2C, n / 2, B, 1, d / 2, p->outer,
                     p->dist_outer / 2);
    if (rc != PFX_SUCCESS) return rc;
    PFX_HIP(hipMalloc(&lp->tw, (size_t)n * esz), "hipMalloc long-plan twiddles");
    PFX_HIP(make_twiddles(bp, n, lp->tw, nullptr), "long-plan twiddles");
  } else {
    lp->mode = LONG_REAL_FULL;
    lp->rows = std::max<int64_t>(1, std::min<int64_t>(total, kLongRowsBytes / (n * (int64_t)esz)));
    rc = plan_create(&lp->c, p->precision, PFX_C2C, n, lp->rows, 1, n, 1, 0);
    if (rc != PFX_SUCCESS) return rc;
    if (total > 0)
      PFX_HIP(hipMalloc(&lp->scratch, (size_t)(lp->rows * n) * esz), "hipMalloc real rows");
  }
  PFX_HIP(hipStreamSynchronize(nullptr), "long-plan tables");
  return PFX_SUCCESS;
}

static int plan_create(pfx_plan *out, int precision, int kind, int64_t length, int64_t inner,
                       int64_t stride, int64_t dist, int64_t outer, int64_t dist_outer) {
  if (!out) return fail(PFX_ERR_INVALID_ARG, "plan pointer is null");
  *out = nullptr;
  if (precision != PFX_F32 && precision != PFX_F64 && precision != PFX_F32_SPLIT_F16 &&
      precision != PFX_F32_SPLIT_BF16)
    return fail(PFX_ERR_UNSUPPORTED, "unknown precision %d", precision);
  if (length <= 0) return fail(PFX_ERR_INVALID_ARG, "length must be positive (got %lld)", (long long)length);
  if (inner < 0 || outer < 0) return fail(PFX_ERR_INVALID_ARG, "batch must be non-negative");
  if (inner > 0 && outer > INT64_MAX / inner)
    return fail(PFX_ERR_INVALID_ARG, "batch %lld x %lld overflows", (long long)inner,
                (long long)outer);
  if (stride <= 0) return fail(PFX_ERR_INVALID_ARG, "stride must be positive (got %lld)", (long long)stride);
  if ((kind == PFX_R2C || kind == PFX_C2R) && stride != 1)
    return fail(PFX_ERR_UNSUPPORTED, "R2C/C2R plans support stride 1 only (got %lld)", (long long)stride);
  if ((kind == PFX_R2C || kind == PFX_C2R) && dist < 2 * (length / 2 + 1) && inner > 1)
    return fail(PFX_ERR_INVALID_ARG, "padded dist %lld is below 2*(n/2+1)", (long long)dist);
  if ((kind == PFX_R2C || kind == PFX_C2R) && (dist % 2) != 0 && inner > 1)
    return fail(PFX_ERR_INVALID_ARG, "padded dist must be even (complex-aligned rows)");
  const int64_t total = inner * outer;  // the grid is sized per tile and checked below

  pfx_plan p = new (std::nothrow) pfx_plan_t;
  if (!p) return fail(PFX_ERR_ALLOC, "out of host memory");
  p->precision = precision;
  p->kind = kind;
  p->length = length;
  p->inner = inner;
  p->outer = outer;
  p->stride = stride;
  p->dist = dist;
  p->dist_outer = dist_outer;
  p->xcd_remap = xcd_remap_mode();

  const int bp = base_prec(precision);
  const size_t esz = bp ? 16 : 8;
  int64_t lds_n = length;
  int kk = KIND_C2C;
  if (kind == PFX_R2C || kind == PFX_C2R) {
    kk = kind == PFX_R2C ? KIND_R2C : KIND_C2R;
    lds_n = (length % 2 == 0) ? length / 2 : -1;
  }
  // Batch-contiguous layout when consecutive transforms are adjacent.
  const bool col = (kind == PFX_C2C) && dist == 1 && stride > 1;
  // Split precisions use their MFMA kernels where the plan has radix-16
  // passes and the plain FP32 kernels elsewhere.
  const bool split = precision == PFX_F32_SPLIT_F16 || precision == PFX_F32_SPLIT_BF16;
  if (lds_n > 0 && split && kk == KIND_C2C) p->kernel = find_kernel(precision, kk, lds_n, col);
  if (lds_n > 0 && !p->kernel) p->kernel = find_kernel(bp, kk, lds_n, col);
  if (p->kernel) {  // a tuned choice for this geometry wins over the table
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const KernelEntry *e = nullptr;
    int order = 1;
    if (cached_choice(p, p->kernel, &e, &order)) {
      p->kernel = e;
      p->xcd_remap = order;
    }
  }
  if (!p->kernel && (length > naive_max_length(bp) || length >= long_min_length())) {
    const int rc = make_long(p);
    if (rc == PFX_SUCCESS) {
      *out = p;
      return PFX_SUCCESS;
    }
    const std::string msg = g_last_error;
    if (length > naive_max_length(bp)) {
      pfx_plan_destroy(p);
      return fail(rc, "length %lld has no tiled kernel and exceeds the direct-DFT fallback "
                      "limit of %d: %s", (long long)length, naive_max_length(bp), msg.c_str());
    }
    free_long(p);  // within the direct DFT's reach: that instead
  }
  if (p->kernel) {
    const int64_t blocks = (total + p->kernel->per_block - 1) / p->kernel->per_block;
    if (blocks > 0x7fffffff) {
      delete p;
      return fail(PFX_ERR_UNSUPPORTED, "batch too large for one launch");
    }
  } else if (total > 0x7fffffff) {
    delete p;
    return fail(PFX_ERR_UNSUPPORTED, "batch too large for one launch");
  }
  // Twiddle table: C2C W_n (n entries); R2C/C2R tiled W_{2h}=W_n (n entries);
  // fallback: W_n.
  p->tw_len = length;
  hipError_t e = hipMalloc(&p->tw, (size_t)p->tw_len * esz);
  if (e != hipSuccess) {
    delete p;
    return hip_fail(e, "hipMalloc twiddles");
  }
  e = make_twiddles(bp, p->tw_len, p->tw, nullptr);
  if (e == hipSuccess && p->kernel) {
    e = hipMalloc(&p->ptw, (size_t)p->kernel->n * esz);
    if (e == hipSuccess) e = make_pass_twiddles(bp, p->kernel->rad, p->ptw, nullptr);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    hipFree(p->tw);
    if (p->ptw) hipFree(p->ptw);
    delete p;
    return hip_fail(e, "twiddle generation");
  }
  *out = p;
  return PFX_SUCCESS;
}

// Candidates for a plan: every registry entry that computes the same
// transform (same kind and LDS length, same data precision family, same layout).
static std::vector<const KernelEntry *> candidates(const pfx_plan_t *p) {
  std::vector<const KernelEntry *> out;
  if (!p->kernel) return out;
  const KernelEntry *k0 = p->kernel;
  for (const KernelEntry *e : registry().entries) {
    if (e->kind != k0->kind || e->n != k0->n || e->col != k0->col) continue;
    if (e->prec != k0->prec) continue;  // same arithmetic (plain or the same split form)
    // one launch must fit the grid limit with this tile
    if ((p->inner * p->outer + e->per_block - 1) / e->per_block > 0x7fffffff) continue;
    out.push_back(e);
  }
  return out;
}

static hipError_t set_kernel(pfx_plan_t *p, const KernelEntry *k) {
  if (k == p->kernel) return hipSuccess;
  const size_t esz = base_prec(p->precision) ? 16 : 8;
  void *ptw = nullptr;
  hipError_t e = hipMalloc(&ptw, (size_t)k->n * esz);
  if (e == hipSuccess) e = make_pass_twiddles(base_prec(p->precision), k->rad, ptw, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    if (ptw) hipFree(ptw);
    return e;
  }
  if (p->ptw) hipFree(p->ptw);
  p->ptw = ptw;
  p->kernel = k;
  return hipSuccess;
}

int pfx_plan_create_c2c(pfx_plan *plan, int precision, int64_t length, int64_t batch,
                        int64_t stride, int64_t dist) {
  return plan_create(plan, precision, PFX_C2C, length, batch, stride, dist, 1, 0);
}

int pfx_plan_create_c2c_many(pfx_plan *plan, int precision, int64_t length, int64_t inner_batch,
                             int64_t stride, int64_t dist, int64_t outer_batch,
                             int64_t outer_dist) {
  return plan_create(plan, precision, PFX_C2C, length, inner_batch, stride, dist, outer_batch,
                     outer_dist);
}

int pfx_plan_create_r2c(pfx_plan *plan, int precision, int64_t length, int64_t batch,
                        int64_t stride, int64_t dist_scalars) {
  return plan_create(plan, precision, PFX_R2C, length, batch, stride, dist_scalars, 1, 0);
}

int pfx_plan_create_c2r(pfx_plan *plan, int precision, int64_t length, int64_t batch,
                        int64_t stride, int64_t dist_scalars) {
  return plan_create(plan, precision, PFX_C2R, length, batch, stride, dist_scalars, 1, 0);
}

// Launch a plan's kernel; `real_oop` / `real_dist` (scalars) make an R2C
// plan load its rows from an unpadded real array, or a C2R plan store them
// into one, instead of the padded rows of `data`.
static int execute_impl(pfx_plan p, int direction, void *data, double scale, pfx_stream stream,
                        void *real_oop, int64_t real_dist);

// A C2C tiled plan from `src` into `dst` at the same offsets (src null: in
// place on dst), outputs times `scale`.
static int launch_tiled(pfx_plan p, int direction, void *dst, const void *src, hipStream_t s,
                        double scale = 1.0) {
  if (!p->kernel) return fail(PFX_ERR_INTERNAL, "sub-plan without a tiled kernel");
  StockhamArgs a;
  a.data = dst;
  a.src = src;
  a.src_dist = p->dist;
  a.src_dist_outer = p->dist_outer;
  a.tw = p->tw;
  a.ptw = p->ptw;
  a.total = p->inner * p->outer;
  a.inner = p->inner;
  a.stride = p->stride;
  a.dist = p->dist;
  a.dist_outer = p->dist_outer;
  a.scale = scale;
  a.use_scale = scale != 1.0;
  a.backward = direction == PFX_BACKWARD;
  a.xcd_remap = p->xcd_remap;
  a.in_seg = a.out_seg = 0;
  if (a.total == 0) return PFX_SUCCESS;
  p->kernel->launch(a, s);
  PFX_HIP(hipGetLastError(), "kernel launch");
  return PFX_SUCCESS;
}

// The steps of a long plan (see LongPlan), on one stream.  The scratch is
// the plan's own: executions of one long plan must not overlap.
static int run_long(pfx_plan p, int direction, void *data, double scale, hipStream_t s) {
  LongPlan *lp = p->lp;
  const int bp = base_prec(p->precision);
  const size_t esz = bp ? 16 : 8;
  // The chirp and odd-real modes run the batch in slices of lp->rows
  // transforms; the sub-plan always transforms all scratch rows (on the
  // last, short slice the rows past it are stale and unused).
  if (lp->mode == LONG_CHIRP && lp->cf) {
    ChirpArgs ca;
    ca.n = p->length;
    ca.M = lp->M;
    ca.t0 = 0;
    ca.total = p->inner * p->outer;
    ca.inner = p->inner;
    ca.dist = p->dist;
    ca.dist_outer = p->dist_outer;
    ca.stride = p->stride;
    ca.backward = direction == PFX_BACKWARD;
    ca.scale = scale;
    if (ca.total == 0) return PFX_SUCCESS;
    // batch-contiguous tiles where neighbouring transforms are adjacent
    const bool col = p->dist == 1 && p->stride > 1;
    (col ? lp->cf->col : lp->cf->row)(data, lp->tw, lp->filt, lp->ptw1, ca, s);
    PFX_HIP(hipGetLastError(), "chirp kernel");
    return PFX_SUCCESS;
  }
  if (lp->mode == LONG_CHIRP) {
    const int64_t total = p->inner * p->outer;
    ChirpArgs ca;
    ca.n = p->length;
    ca.M = lp->M;
    ca.inner = p->inner;
    ca.dist = p->dist;
    ca.dist_outer = p->dist_outer;
    ca.stride = p->stride;
    ca.backward = direction == PFX_BACKWARD;
    ca.scale = scale;
    const char *B = static_cast<const char *>(lp->filt) + (ca.backward ? (size_t)lp->M * esz : 0);
    for (int64_t t0 = 0; t0 < total; t0 += lp->rows) {
      ca.t0 = t0;
      ca.total = std::min(lp->rows, total - t0);
      PFX_HIP(launch_chirp_pre(bp, data, lp->scratch, lp->tw, ca, s), "chirp premultiply");
      int rc = pfx_execute_scaled(lp->c, PFX_FORWARD, lp->scratch, 1.0, s);
      if (rc != PFX_SUCCESS) return rc;
      PFX_HIP(launch_chirp_mul(bp, lp->scratch, B, ca, s), "chirp filter");
      rc = pfx_execute_scaled(lp->c, PFX_BACKWARD, lp->scratch, 1.0, s);
      if (rc != PFX_SUCCESS) return rc;
      PFX_HIP(launch_chirp_post(bp, lp->scratch, data, lp->tw, ca, s), "chirp postmultiply");
    }
    return PFX_SUCCESS;
  }
  if (lp->mode == LONG_REAL_FULL) {
    RealFullArgs fa;
    fa.n = p->length;
    fa.dist = p->dist;
    fa.use_scale = scale != 1.0;
    fa.scale = scale;
    const bool r2c = p->kind == PFX_R2C;
    for (int64_t b0 = 0; b0 < p->inner; b0 += lp->rows) {
      fa.batch = std::min(lp->rows, p->inner - b0);
      char *rows = static_cast<char *>(data) + (size_t)(b0 * p->dist) * (esz / 2);
      PFX_HIP(launch_real_full(bp, r2c ? 0 : 2, rows, lp->scratch, fa, s), "real rows load");
      int rc = pfx_execute_scaled(lp->c, r2c ? PFX_FORWARD : PFX_BACKWARD, lp->scratch, 1.0, s);
      if (rc != PFX_SUCCESS) return rc;
      PFX_HIP(launch_real_full(bp, r2c ? 1 : 3, rows, lp->scratch, fa, s), "real rows store");
    }
    return PFX_SUCCESS;
  }
  if (p->kind == PFX_C2C) {
    FourStepArgs fa;
    fa.n1 = lp->n1;
    fa.n2 = lp->n2;
    fa.stride = p->stride;
    fa.batch = p->inner;
    fa.dist = p->dist;
    fa.backward = direction == PFX_BACKWARD;
    for (int64_t o = 0; o < p->outer; ++o) {
      char *base = static_cast<char *>(data) + (size_t)(o * p->dist_outer) * esz;
      if (lp->rows2) {
        FsRowArgs ra;
        ra.src = base;
        ra.dst = lp->scratch;
        ra.ptw = lp->ptw1;
        ra.tw = lp->tw;
        ra.twc = lp->tw2;
        ra.n1 = lp->n1;
        ra.n2 = lp->n2;
        ra.st = p->stride;
        ra.dist = p->dist;
        ra.rows = p->inner;
        ra.tiles = 0;
        ra.backward = direction == PFX_BACKWARD;
        ra.scale = 1.0;
        PFX_HIP(launch_fourstep_rows_pass(bp, 1, ra, s), "four-step rows first pass");
        ra.src = lp->scratch;
        ra.dst = base;
        ra.ptw = lp->ptw2;
        ra.scale = scale;
        PFX_HIP(launch_fourstep_rows_pass(bp, 2, ra, s), "four-step rows second pass");
        continue;
      }
      if (lp->c2) {
        FsColArgs ca;
        ca.src = base;
        ca.dst = lp->scratch;
        ca.ptw = lp->ptw1;
        ca.tw = lp->tw;
        ca.n1 = lp->n1;
        ca.n2 = lp->n2;
        ca.st = p->stride;
        ca.batch = p->inner;
        ca.tiles_per_j1 = 0;
        ca.backward = direction == PFX_BACKWARD;
        PFX_HIP(launch_fourstep_cols_pass1(bp, ca, s), "four-step first pass");
        const int rc = launch_tiled(lp->c2, direction, base, lp->scratch, s, scale);
        if (rc != PFX_SUCCESS) return rc;
        continue;
      }
      int rc = launch_tiled(lp->a, direction, lp->scratch, base, s);
      if (rc != PFX_SUCCESS) return rc;
      fa.use_scale = 0;
      fa.scale = 1.0;
      PFX_HIP(launch_fourstep_twiddle(bp, lp->scratch, lp->tw, fa, s), "four-step twiddle");
      rc = launch_tiled(lp->c, direction, lp->scratch, nullptr, s);
      if (rc != PFX_SUCCESS) return rc;
      fa.use_scale = scale != 1.0;
      fa.scale = scale;
      PFX_HIP(launch_fourstep_transpose(bp, lp->scratch, base, fa, s), "four-step transpose");
    }
    return PFX_SUCCESS;
  }
  RealPassArgs ra;
  ra.h = p->length / 2;
  ra.batch = p->inner;
  ra.dist = p->dist / 2;
  ra.use_scale = scale != 1.0;
  ra.scale = scale;
  if (p->kind == PFX_R2C) {
    int rc = pfx_execute_scaled(lp->c, PFX_FORWARD, data, 1.0, s);
    if (rc != PFX_SUCCESS) return rc;
    PFX_HIP(launch_real_post(bp, data, lp->tw, ra, s), "long R2C post-pass");
    return PFX_SUCCESS;
  }
  PFX_HIP(launch_real_pre(bp, data, lp->tw, ra, s), "long C2R pre-pass");
  return pfx_execute_scaled(lp->c, PFX_BACKWARD, data, 1.0, s);
}

int pfx_execute_scaled(pfx_plan p, int direction, void *data, double scale, pfx_stream stream) {
  return execute_impl(p, direction, data, scale, stream, nullptr, 0);
}

int pfx_execute_r2c_oop(pfx_plan p, const void *real_in, int64_t real_dist, void *padded_out,
                        double scale, pfx_stream stream) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (p->kind != PFX_R2C) return fail(PFX_ERR_INVALID_ARG, "not an R2C plan");
  return execute_impl(p, PFX_FORWARD, padded_out, scale, stream, const_cast<void *>(real_in),
                      real_dist);
}

int pfx_execute_c2r_oop(pfx_plan p, void *padded_in, void *real_out, int64_t real_dist,
                        double scale, pfx_stream stream) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (p->kind != PFX_C2R) return fail(PFX_ERR_INVALID_ARG, "not a C2R plan");
  return execute_impl(p, PFX_BACKWARD, padded_in, scale, stream, real_out, real_dist);
}

static int execute_impl(pfx_plan p, int direction, void *data, double scale, pfx_stream stream,
                        void *real_oop, int64_t real_dist) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (direction != PFX_FORWARD && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "direction must be PFX_FORWARD or PFX_BACKWARD");
  if (p->kind == PFX_R2C && direction != PFX_FORWARD)
    return fail(PFX_ERR_INVALID_ARG, "an R2C plan only runs forward");
  if (p->kind == PFX_C2R && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "a C2R plan only runs backward");
  const int64_t total = p->inner * p->outer;
  if (total == 0) return PFX_SUCCESS;
  if (!data) return fail(PFX_ERR_INVALID_ARG, "data is null");
  if (real_oop) {
    if (!p->kernel)
      return fail(PFX_ERR_UNSUPPORTED,
                  "out-of-place R2C/C2R needs a tiled kernel (even length with a radix plan)");
    if (real_dist < p->length || real_dist % 2 != 0 || (p->outer > 1))
      return fail(PFX_ERR_INVALID_ARG,
                  "real rows need an even distance of at least the length (got %lld)",
                  (long long)real_dist);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p->lp) return run_long(p, direction, data, scale, s);
  const bool real = p->kind != PFX_C2C;
  if (p->kernel) {
    StockhamArgs a;
    if (real_oop && p->kind == PFX_R2C) {
      a.src = real_oop;               // complex pairs of the unpadded real rows
      a.src_dist = real_dist / 2;
    } else if (real_oop) {
      a.dst = real_oop;
      a.dst_dist = real_dist / 2;
    }
    a.data = data;
    a.tw = p->tw;
    a.ptw = p->ptw;
    a.total = total;
    a.inner = p->inner;
    a.stride = p->stride;
    // R2C/C2R tiled kernels address the row as complex values.
    a.dist = real ? p->dist / 2 : p->dist;
    a.dist_outer = real ? p->dist_outer / 2 : p->dist_outer;
    a.scale = scale;
    a.use_scale = scale != 1.0;
    a.backward = direction == PFX_BACKWARD;
    a.xcd_remap = p->xcd_remap;
    a.in_seg = a.out_seg = 0;
    p->kernel->launch(a, s);
  } else {
    NaiveArgs a;
    a.data = data;
    a.tw = p->tw;
    a.total = total;
    a.inner = p->inner;
    a.stride = p->stride;
    a.dist = p->dist;
    a.dist_outer = p->dist_outer;
    a.n = (int)p->length;
    a.kind = p->kind == PFX_C2C ? KIND_C2C : (p->kind == PFX_R2C ? KIND_R2C : KIND_C2R);
    a.backward = direction == PFX_BACKWARD;
    a.scale = scale;
    a.use_scale = scale != 1.0;
    launch_naive(base_prec(p->precision), a, s);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return PFX_SUCCESS;
}

static int to_segmap(const pfx_segmap *m, int length, int64_t total, SegMap *out) {
  if (m->nparts < 1 || m->nparts > kMaxSegPeers || !m->segments || !m->start || !m->count ||
      m->trailing == 0)
    return fail(PFX_ERR_INVALID_ARG, "bad segment map (nparts must be 1..%d)", kMaxSegPeers);
  out->nparts = m->nparts;
  int end = 0;
  for (int q = 0; q < m->nparts; ++q) {
    if (m->start[q] != end || m->count[q] < 0 || m->count[q] > length - end)
      return fail(PFX_ERR_INVALID_ARG, "segment blocks must tile the axis in order");
    if (m->count[q] > 0 && !m->segments[q])
      return fail(PFX_ERR_INVALID_ARG, "segment %d has no address", q);
    end += m->count[q];
    out->s[q] = m->start[q];
    out->n[q] = m->count[q];
    out->seg[q] = m->segments[q];
  }
  if (end != length)
    return fail(PFX_ERR_INVALID_ARG, "segment blocks cover %d of %d elements", end, length);
  // the kernels index a segment's rows (lead * count + e) in 32 bits
  int maxn = 0;
  for (int q = 0; q < m->nparts; ++q) maxn = std::max(maxn, m->count[q]);
  const int64_t leads = (total + (int64_t)m->trailing - 1) / (int64_t)m->trailing;
  if (m->trailing >= (1ull << 32) || (leads + 1) * (int64_t)maxn >= (1ll << 31))
    return fail(PFX_ERR_UNSUPPORTED, "segment of %lld rows exceeds 32-bit row indexing",
                (long long)((leads + 1) * maxn));
  out->trailing = (int64_t)m->trailing;
  return PFX_SUCCESS;
}

int pfx_execute_segmented(pfx_plan p, int direction, void *data, const pfx_segmap *in,
                          const pfx_segmap *out, double scale, pfx_stream stream) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (!in && !out) return pfx_execute_scaled(p, direction, data, scale, stream);
  if (!p->kernel || p->kernel->prec > 1 || !p->kernel->launch_seg)
    return fail(PFX_ERR_UNSUPPORTED, "segmented execution needs a tiled plain-precision kernel");
  if (p->kind == PFX_C2R || (p->kind == PFX_R2C && in))
    return fail(PFX_ERR_UNSUPPORTED, "segmented loads are for C2C plans; C2R is not fused");
  if ((p->kind == PFX_R2C && direction != PFX_FORWARD) ||
      (direction != PFX_FORWARD && direction != PFX_BACKWARD))
    return fail(PFX_ERR_INVALID_ARG, "bad direction");
  const int64_t total = p->inner * p->outer;
  if (total == 0) return PFX_SUCCESS;
  if ((!in || !out) && !data) return fail(PFX_ERR_INVALID_ARG, "data is null");
  const bool real = p->kind != PFX_C2C;
  StockhamArgs a;
  a.src = nullptr;
  a.data = data;
  a.tw = p->tw;
  a.ptw = p->ptw;
  a.total = total;
  a.inner = p->inner;
  a.stride = p->stride;
  a.dist = real ? p->dist / 2 : p->dist;
  a.dist_outer = real ? p->dist_outer / 2 : p->dist_outer;
  a.scale = scale;
  a.use_scale = scale != 1.0;
  a.backward = direction == PFX_BACKWARD;
  a.xcd_remap = p->xcd_remap;
  a.in_seg = in ? 1 : 0;
  a.out_seg = out ? 1 : 0;
  if (in) {
    int rc = to_segmap(in, (int)p->length, total, &a.in_map);
    if (rc) return rc;
    // Segments back to back with one lead per block (the receive side of an
    // exchange that gathers the outermost axis) are the plan's own strided
    // layout: load plainly from the first segment, out of place.
    const size_t esz = p->precision == PFX_F64 ? 16 : 8;
    bool plain = total <= (int64_t)in->trailing && p->stride == (int64_t)in->trailing &&
                 p->dist == 1 && (p->outer == 1 || p->dist_outer == p->inner);
    for (int q = 0; plain && q < in->nparts; ++q)
      plain = static_cast<const char *>(in->segments[q]) ==
              static_cast<const char *>(in->segments[0]) + (size_t)in->start[q] * in->trailing * esz;
    if (plain) {
      a.src = in->segments[0];
      a.src_dist = a.dist;
      a.src_dist_outer = a.dist_outer;
      a.in_seg = 0;
    }
  }
  if (out) {
    int rc = to_segmap(out, real ? (int)p->length / 2 + 1 : (int)p->length, total, &a.out_map);
    if (rc) return rc;
  }
  (a.in_seg || a.out_seg ? p->kernel->launch_seg : p->kernel->launch)(
      a, static_cast<hipStream_t>(stream));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return PFX_SUCCESS;
}

int pfx_execute(pfx_plan plan, int direction, void *data, pfx_stream stream) {
  return pfx_execute_scaled(plan, direction, data, 1.0, stream);
}

int pfx_plan_tune(pfx_plan p, int flag, void *scratch, pfx_stream stream) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (p->lp && flag != PFX_PLAN_ESTIMATE) {
    // the sub-plans: `a` in place on the caller's scratch (its footprint is
    // the plan's), `c` on the plan's own scratch (C2C) or the caller's rows
    if (!scratch) return fail(PFX_ERR_INVALID_ARG, "tuning needs a scratch buffer");
    if (p->lp->rows2 || p->lp->cf) return PFX_SUCCESS;  // fused kernels: nothing to choose
    if (p->lp->a) {
      int rc = pfx_plan_tune(p->lp->a, flag, scratch, stream);
      if (rc != PFX_SUCCESS) return rc;
    }
    if (p->lp->c2)  // in place on the caller's scratch: its footprint is the plan's
      return pfx_plan_tune(p->lp->c2, flag, scratch, stream);
    // `c` works on the plan's own rows except for the real halves (the
    // caller's rows read as complex pairs)
    if (p->lp->mode != LONG_REAL_HALF && !p->lp->scratch) return PFX_SUCCESS;  // empty batch
    return pfx_plan_tune(p->lp->c, flag,
                         p->lp->mode == LONG_REAL_HALF ? scratch : p->lp->scratch, stream);
  }
  if (flag == PFX_PLAN_ESTIMATE || !p->kernel) return PFX_SUCCESS;
  if (flag != PFX_PLAN_MEASURE && flag != PFX_PLAN_PATIENT)
    return fail(PFX_ERR_INVALID_ARG, "unknown planning flag %d", flag);
  if (!scratch) return fail(PFX_ERR_INVALID_ARG, "tuning needs a scratch buffer");
  const std::string key = plan_key(p);
  {
    // a cached choice is taken only if it computes this plan's transform
    // (the plan's current kernel has the same kind, length, layout and
    // arithmetic as every candidate); otherwise it is measured afresh
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const KernelEntry *e = nullptr;
    int order = 1;
    if (cached_choice(p, p->kernel, &e, &order)) {
      PFX_HIP(set_kernel(p, e), "tuned kernel");
      p->xcd_remap = order;
      return PFX_SUCCESS;
    }
  }
  const int reps = flag == PFX_PLAN_PATIENT ? 10 : 3;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t a, b;
  PFX_HIP(hipEventCreate(&a), "hipEventCreate");
  PFX_HIP(hipEventCreate(&b), "hipEventCreate");
  const KernelEntry *best = p->kernel;
  float best_ms = -1.f;
  const int dir = p->kind == PFX_C2R ? PFX_BACKWARD : PFX_FORWARD;
  int best_order = p->xcd_remap;
  for (const KernelEntry *k : candidates(p)) {
    hipError_t e = set_kernel(p, k);
    if (e != hipSuccess) continue;
    // batch-contiguous kernels also try XCD tile orders (contiguous runs per
    // XCD, or XCDs taking turns on chunks of 8 / 32 tiles)
    std::vector<int> orders = {p->xcd_remap};
    if (k->col) orders = {1, 8, 32};
    for (int order : orders) {
      p->xcd_remap = order;
      int rc = pfx_execute(p, dir, scratch, stream);  // warm-up
      if (rc != PFX_SUCCESS) continue;
      hipEventRecord(a, s);
      for (int r = 0; r < reps && rc == PFX_SUCCESS; ++r) rc = pfx_execute(p, dir, scratch, stream);
      hipEventRecord(b, s);
      if (hipEventSynchronize(b) != hipSuccess || rc != PFX_SUCCESS) continue;
      float ms = 0.f;
      hipEventElapsedTime(&ms, a, b);
      if (best_ms < 0.f || ms < best_ms) {
        best_ms = ms;
        best = k;
        best_order = order;
      }
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  PFX_HIP(set_kernel(p, best), "tuned kernel");
  p->xcd_remap = best_order;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune[key] = entry_id(best) + "|order=" + std::to_string(best_order);
  return PFX_SUCCESS;
}

int pfx_plan_num_candidates(pfx_plan p, int *count) {
  if (!p || !count) return fail(PFX_ERR_INVALID_ARG, "null argument");
  *count = (int)candidates(p).size();
  return PFX_SUCCESS;
}

int pfx_plan_use_candidate(pfx_plan p, int index) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  std::vector<const KernelEntry *> c = candidates(p);
  if (index < 0 || index >= (int)c.size())
    return fail(PFX_ERR_INVALID_ARG, "candidate %d out of range (%zu)", index, c.size());
  PFX_HIP(set_kernel(p, c[index]), "candidate kernel");
  return PFX_SUCCESS;
}

int pfx_tune_cache_save(const char *path) {
  if (!path) return fail(PFX_ERR_INVALID_ARG, "path is null");
  FILE *f = fopen(path, "w");
  if (!f) return fail(PFX_ERR_INVALID_ARG, "cannot write %s", path);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  fprintf(f, "# pfx tuning cache v1 (gfx950)\n");
  for (auto &kv : g_tune) fprintf(f, "%s\t%s\n", kv.first.c_str(), kv.second.c_str());
  fclose(f);
  return PFX_SUCCESS;
}

int pfx_tune_cache_load(const char *path) {
  if (!path) return fail(PFX_ERR_INVALID_ARG, "path is null");
  FILE *f = fopen(path, "r");
  if (!f) return fail(PFX_ERR_INVALID_ARG, "cannot read %s", path);
  char line[1024];
  std::lock_guard<std::mutex> lock(g_tune_mu);
  int bad = 0;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char *tab = strchr(line, '\t');
    if (!tab) { ++bad; continue; }
    *tab = 0;
    std::string val(tab + 1);
    while (!val.empty() && (val.back() == '\n' || val.back() == '\r')) val.pop_back();
    const KernelEntry *ke = nullptr;
    int order = 1;
    // a key not in plan_key's format, or a kernel not in this build
    if (!is_plan_key(line) || !parse_tuned(val, &ke, &order) || order < 0) { ++bad; continue; }
    g_tune[line] = val;
  }
  fclose(f);
  g_tune_dropped = bad;
  if (bad > 0) {
    // the load succeeds with the valid lines; the message says what was dropped
    fail(PFX_SUCCESS, "tuning cache %s: %d line(s) dropped (bad key or kernel not in this build)",
         path, bad);
  }
  return PFX_SUCCESS;
}

int pfx_tune_cache_dropped(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return g_tune_dropped;
}

int pfx_tune_cache_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
  return PFX_SUCCESS;
}

int pfx_tune_cache_size(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return (int)g_tune.size();
}

int pfx_plan_destroy(pfx_plan p) {
  if (!p) return PFX_SUCCESS;
  if (p->tw) hipFree(p->tw);
  if (p->ptw) hipFree(p->ptw);
  free_long(p);
  delete p;
  return PFX_SUCCESS;
}

int pfx_plan_describe(pfx_plan p, char *name, size_t name_len, int *per_block, int *threads) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  char buf[256];
  if (p->kernel) {
    const KernelEntry *k = p->kernel;
    static const char *pn[] = {"f32", "f64", "f32split_f16", "f32split_bf16"};
    snprintf(buf, sizeof(buf), "%s_%s_%s_n%d_r[%s]_c%d_t%d", pn[k->prec & 3],
             k->kind == KIND_C2C ? "c2c" : (k->kind == KIND_R2C ? "r2c" : "c2r"),
             k->col ? "col" : "row", k->n, k->radices, k->per_block, k->tpt);
    if (per_block) *per_block = k->per_block;
    if (threads) *threads = k->tpt * k->per_block;
  } else if (p->lp && p->lp->mode == LONG_CHIRP) {
    snprintf(buf, sizeof(buf), "%s_c2c_bluestein_n%lld_m%lld",
             p->precision == PFX_F64 ? "f64" : "f32", (long long)p->length,
             (long long)p->lp->M);
    if (per_block) *per_block = 1;
    if (threads) *threads = 256;
  } else if (p->lp && p->kind == PFX_C2C) {
    snprintf(buf, sizeof(buf), "%s_c2c_fourstep_n%lld_%lldx%lld",
             p->precision == PFX_F64 ? "f64" : "f32", (long long)p->length,
             (long long)p->lp->n1, (long long)p->lp->n2);
    if (per_block) *per_block = 1;
    if (threads) *threads = 256;
  } else if (p->lp) {
    snprintf(buf, sizeof(buf), "%s_%s_long_n%lld", p->precision == PFX_F64 ? "f64" : "f32",
             p->kind == PFX_R2C ? "r2c" : "c2r", (long long)p->length);
    if (per_block) *per_block = 1;
    if (threads) *threads = 256;
  } else {
    snprintf(buf, sizeof(buf), "%s_direct_dft_n%lld", p->precision == PFX_F64 ? "f64" : "f32",
             (long long)p->length);
    if (per_block) *per_block = 1;
    if (threads) *threads = 256;
  }
  if (name && name_len) {
    strncpy(name, buf, name_len - 1);
    name[name_len - 1] = 0;
  }
  return PFX_SUCCESS;
}

// ------------------------------------------------------------- no-alloc C2C

}  // extern "C"

namespace {

// Device scratch of pfx_execute_c2c_noalloc: the length-n table W_n^i and
// the per-pass table of the kernel, n <= kNoAllocMaxLength complex<double>
// (128 KiB).  Each GPU the library runs on holds its own copy until the
// process ends.  In host code `scratch` is a stand-in for the symbol, not a
// device address: the device address comes from hipGetSymbolAddress, once
// per device, and only that address is passed to kernels.
constexpr int kNoAllocMaxLength = 4096;
constexpr size_t kNoAllocScratchBytes = 2 * (size_t)kNoAllocMaxLength * 16;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device (null: not looked up)

// The kernel registry is built when the library is loaded, so no call of
// the no-alloc entry point allocates host memory for it.
const Registry &g_registry_at_load = registry();

int scratch_of_current_device(void **out) {
  int dev = 0;
  PFX_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(PFX_ERR_UNSUPPORTED, "device %d beyond the %d the scratch table holds", dev,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    PFX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(PFX_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PFX_SUCCESS;
}

}  // namespace

extern "C" {

int pfx_execute_c2c_noalloc(int precision, int64_t length, int64_t batch, int64_t stride,
                            int64_t dist, int direction, void *data, double scale,
                            pfx_stream stream) {
  if (precision != PFX_F32 && precision != PFX_F64)
    return fail(PFX_ERR_UNSUPPORTED, "the no-alloc transform takes PFX_F32 or PFX_F64 (got %d)",
                precision);
  if (length <= 0) return fail(PFX_ERR_INVALID_ARG, "length must be positive (got %lld)", (long long)length);
  if (length > kNoAllocMaxLength)
    return fail(PFX_ERR_INVALID_ARG,
                "length %lld exceeds the no-alloc scratch maximum of %d (use a plan)",
                (long long)length, kNoAllocMaxLength);
  if (batch < 0) return fail(PFX_ERR_INVALID_ARG, "batch must be non-negative");
  if (stride <= 0) return fail(PFX_ERR_INVALID_ARG, "stride must be positive (got %lld)", (long long)stride);
  if (direction != PFX_FORWARD && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "direction must be PFX_FORWARD or PFX_BACKWARD");
  if (batch == 0) return PFX_SUCCESS;
  if (!data) return fail(PFX_ERR_INVALID_ARG, "data is null");
  // the plan pfx_plan_create_c2c would build (its default kernel), on the
  // stack, with both twiddle tables in the device scratch
  pfx_plan_t p;
  p.precision = precision;
  p.kind = PFX_C2C;
  p.length = length;
  p.inner = batch;
  p.outer = 1;
  p.stride = stride;
  p.dist = dist;
  p.dist_outer = 0;
  p.xcd_remap = xcd_remap_mode();
  const int bp = base_prec(precision);
  const bool col = dist == 1 && stride > 1;
  p.kernel = find_kernel(bp, KIND_C2C, length, col);
  if (!p.kernel && length > naive_max_length(bp))
    return fail(PFX_ERR_UNSUPPORTED,
                "length %lld has no tiled kernel and exceeds the direct-DFT fallback limit of %d",
                (long long)length, naive_max_length(bp));
  const int64_t per = p.kernel ? p.kernel->per_block : 1;
  if ((batch + per - 1) / per > 0x7fffffff)
    return fail(PFX_ERR_UNSUPPORTED, "batch too large for one launch");
  void *base = nullptr;
  int rc = scratch_of_current_device(&base);
  if (rc != PFX_SUCCESS) return rc;
  const size_t esz = bp ? 16 : 8;
  p.tw = base;
  p.ptw = static_cast<unsigned char *>(base) + (size_t)length * esz;
  p.tw_len = length;
  hipStream_t s = static_cast<hipStream_t>(stream);
  PFX_HIP(make_twiddles(bp, length, p.tw, s), "twiddle table");
  if (p.kernel) PFX_HIP(make_pass_twiddles(bp, p.kernel->rad, p.ptw, s), "pass table");
  return execute_impl(&p, direction, data, scale, stream, nullptr, 0);
}

// ------------------------------------------------------------------ memory

int pfx_malloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(PFX_ERR_INVALID_ARG, "ptr is null");
  *ptr = nullptr;
  if (bytes == 0) return PFX_SUCCESS;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail(PFX_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return PFX_SUCCESS;
}

int pfx_free(void *ptr) {
  if (!ptr) return PFX_SUCCESS;
  PFX_HIP(hipFree(ptr), "hipFree");
  return PFX_SUCCESS;
}

int pfx_memcpy_async(void *dst, const void *src, size_t bytes, pfx_stream stream) {
  if (bytes == 0 || dst == src) return PFX_SUCCESS;
  PFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, static_cast<hipStream_t>(stream)),
          "hipMemcpyAsync");
  return PFX_SUCCESS;
}

static bool is_device_ptr(const void *p) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice;
}

int pfx_memcpy2d_async(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                       size_t height, pfx_stream stream) {
  if (width == 0 || height == 0) return PFX_SUCCESS;
  if (dpitch < width || spitch < width)
    return fail(PFX_ERR_INVALID_ARG, "pitch smaller than width");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (is_device_ptr(dst) && is_device_ptr(src)) {
    PFX_HIP(launch_copy2d(dst, dpitch, src, spitch, width, height, s), "copy2d kernel");
  } else {
    PFX_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDefault, s),
            "hipMemcpy2DAsync");
  }
  return PFX_SUCCESS;
}

int pfx_stream_create(pfx_stream *stream) {
  if (!stream) return fail(PFX_ERR_INVALID_ARG, "stream is null");
  hipStream_t s;
  PFX_HIP(hipStreamCreate(&s), "hipStreamCreate");
  *stream = s;
  return PFX_SUCCESS;
}

int pfx_stream_destroy(pfx_stream stream) {
  if (!stream) return PFX_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipStreamSynchronize(s);
  PFX_HIP(hipStreamDestroy(s), "hipStreamDestroy");
  return PFX_SUCCESS;
}

int pfx_stream_synchronize(pfx_stream stream) {
  PFX_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
  return PFX_SUCCESS;
}

// ------------------------------------------------------------ pack/unpack

static int pack_common(bool pack, void *dst, const void *src, int nparts, const int *n,
                       const int *st, const uint64_t *displs, uint64_t leading, uint64_t trailing,
                       uint64_t extent, uint64_t esz, pfx_stream stream) {
  if (nparts <= 0) return fail(PFX_ERR_INVALID_ARG, "nparts must be positive");
  if (!n || !st || !displs) return fail(PFX_ERR_INVALID_ARG, "geometry arrays are null");
  if (esz == 0) return fail(PFX_ERR_INVALID_ARG, "elem_size is zero");
  std::vector<PackSeg> segs;
  segs.reserve(nparts);
  for (int p = 0; p < nparts; ++p) {
    if (n[p] < 0 || st[p] < 0 || (uint64_t)st[p] + (uint64_t)n[p] > extent)
      return fail(PFX_ERR_INVALID_ARG, "peer %d block [%d, %d) exceeds axis extent %llu", p,
                  st[p], st[p] + n[p], (unsigned long long)extent);
    PackSeg sg;
    sg.row_bytes = (uint64_t)n[p] * trailing * esz;
    const uint64_t strided = (uint64_t)st[p] * trailing * esz;
    const uint64_t packed = displs[p] * esz;
    sg.dst_off = pack ? packed : strided;
    sg.src_off = pack ? strided : packed;
    if (sg.row_bytes) segs.push_back(sg);
  }
  if (segs.empty() || leading == 0) return PFX_SUCCESS;
  const uint64_t pitch = extent * trailing * esz;
  hipError_t e = launch_pack(dst, src, segs.data(), (int)segs.size(), leading, 0, 0,
                             pack ? 0 : pitch, pack ? pitch : 0, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, pack ? "pack kernel" : "unpack kernel");
  return PFX_SUCCESS;
}

int pfx_pack(void *pack, const void *src, int nparts, const int *src_n, const int *src_s,
             const uint64_t *send_displs, uint64_t src_leading, uint64_t src_trailing,
             uint64_t src_extent, uint64_t elem_size, pfx_stream stream) {
  return pack_common(true, pack, src, nparts, src_n, src_s, send_displs, src_leading,
                     src_trailing, src_extent, elem_size, stream);
}

int pfx_unpack(void *dst, const void *recv, int nparts, const int *dst_n, const int *dst_s,
               const uint64_t *recv_displs, uint64_t dst_leading, uint64_t dst_trailing,
               uint64_t dst_extent, uint64_t elem_size, pfx_stream stream) {
  return pack_common(false, dst, recv, nparts, dst_n, dst_s, recv_displs, dst_leading,
                     dst_trailing, dst_extent, elem_size, stream);
}

int pfx_scale(void *data, int64_t count, int precision, double s, pfx_stream stream) {
  if (count < 0) return fail(PFX_ERR_INVALID_ARG, "count must be non-negative");
  if (precision != PFX_F32 && precision != PFX_F64)
    return fail(PFX_ERR_INVALID_ARG, "precision must be PFX_F32 or PFX_F64");
  if (count && !data) return fail(PFX_ERR_INVALID_ARG, "data is null");
  PFX_HIP(launch_scale(data, count, base_prec(precision), s, static_cast<hipStream_t>(stream)),
          "scale kernel");
  return PFX_SUCCESS;
}

// ------------------------------------------------------ one-rank N-D path

}  // extern "C"

namespace {

int64_t env_mb_nd(const char *name, int64_t dflt) {
  const char *v = getenv(name);
  return v && *v ? (int64_t)atoll(v) : dflt;
}

// Plans of stages 0..D-2 restricted to P axis-0 planes (+ the C2R on P
// planes), into `out`.
int make_chunk_list(pfx_ndplan p, const std::vector<int64_t> &cs, int P,
                    std::vector<pfx_plan> &out) {
  const int ndim = p->ndim;
  const int prec = p->precision;
  for (int s = 0; s < ndim - 1; ++s) {
    const int axis = ndim - 1 - s;
    int64_t lead = P, trail = 1;
    for (int i = 1; i < axis; ++i) lead *= cs[i];
    for (int i = axis + 1; i < ndim; ++i) trail *= cs[i];
    pfx_plan pl = nullptr;
    int rc;
    if (axis == ndim - 1)
      rc = p->kind == PFX_R2C
               ? pfx_plan_create_r2c(&pl, prec, p->shape[axis], lead, 1, 2 * cs[axis])
               : pfx_plan_create_c2c(&pl, prec, cs[axis], lead, 1, cs[axis]);
    else
      rc = pfx_plan_create_c2c_many(&pl, prec, cs[axis], trail, trail, 1, lead, cs[axis] * trail);
    if (rc != PFX_SUCCESS) return rc;
    out.push_back(pl);
  }
  if (p->kind == PFX_R2C) {
    int64_t lead = P;
    for (int i = 1; i < ndim - 1; ++i) lead *= cs[i];
    pfx_plan pl = nullptr;
    int rc = pfx_plan_create_c2r(&pl, prec, p->shape[ndim - 1], lead, 1, 2 * cs[ndim - 1]);
    if (rc != PFX_SUCCESS) return rc;
    out.push_back(pl);
  }
  return PFX_SUCCESS;
}

// The forward chunk plans and their backward twins.
int make_chunk_plans(pfx_ndplan p, const std::vector<int64_t> &cs, int P) {
  int rc = make_chunk_list(p, cs, P, p->cstages);
  if (rc == PFX_SUCCESS) rc = make_chunk_list(p, cs, P, p->bstages);
  if (rc != PFX_SUCCESS) return rc;
  p->chunk = P;
  return PFX_SUCCESS;
}

}  // namespace

extern "C" {

int pfx_ndplan_create(pfx_ndplan *out, int precision, int kind, int ndim, const int *shape) {
  if (!out) return fail(PFX_ERR_INVALID_ARG, "plan pointer is null");
  *out = nullptr;
  if (ndim < 1 || ndim > 16 || !shape) return fail(PFX_ERR_INVALID_ARG, "ndim must be in [1, 16]");
  if (kind != PFX_C2C && kind != PFX_R2C)
    return fail(PFX_ERR_INVALID_ARG, "kind must be PFX_C2C or PFX_R2C");
  for (int i = 0; i < ndim; ++i)
    if (shape[i] <= 0) return fail(PFX_ERR_INVALID_ARG, "shape[%d] must be positive", i);
  pfx_ndplan p = new (std::nothrow) pfx_ndplan_t;
  if (!p) return fail(PFX_ERR_ALLOC, "out of host memory");
  p->precision = precision;
  p->kind = kind;
  p->ndim = ndim;
  p->shape.assign(shape, shape + ndim);
  std::vector<int64_t> cs(shape, shape + ndim);  // complex shape
  if (kind == PFX_R2C) cs[ndim - 1] = shape[ndim - 1] / 2 + 1;
  p->stages.assign(ndim, nullptr);
  for (int s = 0; s < ndim; ++s) {
    const int axis = ndim - 1 - s;
    int64_t lead = 1, trail = 1;
    for (int i = 0; i < axis; ++i) lead *= cs[i];
    for (int i = axis + 1; i < ndim; ++i) trail *= cs[i];
    int rc;
    if (axis == ndim - 1) {
      if (kind == PFX_R2C) {
        // forward plan created here; the C2R plan is kept as a separate stage slot below
        rc = pfx_plan_create_r2c(&p->stages[s], precision, shape[axis], lead, 1, 2 * cs[axis]);
      } else {
        rc = pfx_plan_create_c2c(&p->stages[s], precision, cs[axis], lead, 1, cs[axis]);
      }
    } else if (axis == 0) {
      rc = pfx_plan_create_c2c(&p->stages[s], precision, cs[axis], trail, trail, 1);
    } else {
      rc = pfx_plan_create_c2c_many(&p->stages[s], precision, cs[axis], trail, trail, 1, lead,
                                    cs[axis] * trail);
    }
    if (rc != PFX_SUCCESS) {
      std::string msg = g_last_error;
      pfx_ndplan_destroy(p);
      return fail(rc, "stage %d (axis %d): %s", s, axis, msg.c_str());
    }
  }
  if (kind == PFX_R2C) {
    pfx_plan c2r = nullptr;
    int64_t lead = 1;
    for (int i = 0; i < ndim - 1; ++i) lead *= cs[i];
    int rc = pfx_plan_create_c2r(&c2r, precision, shape[ndim - 1], lead, 1, 2 * cs[ndim - 1]);
    if (rc != PFX_SUCCESS) {
      std::string msg = g_last_error;
      pfx_ndplan_destroy(p);
      return fail(rc, "C2R stage: %s", msg.c_str());
    }
    p->stages.push_back(c2r);
  }
  // Chunked schedule for arrays well beyond the Infinity Cache: the largest
  // divisor P of n0 whose P planes fit the budget.
  p->plane_elems = 1;
  for (int i = 1; i < ndim; ++i) p->plane_elems *= cs[i];
  const int64_t esz = base_prec(precision) ? 16 : 8;
  const int64_t budget = env_mb_nd("PFX_ND_CHUNK_MB", 256) << 20;
  const int64_t plane_bytes = p->plane_elems * esz;
  int P = 0;
  if (ndim >= 3 && budget > 0 && (int64_t)shape[0] * plane_bytes > 2 * budget)
    for (int d = shape[0] - 1; d >= 1; --d)
      if (shape[0] % d == 0 && d * plane_bytes <= budget) {
        P = d;
        break;
      }
  if (P > 0) {
    int rc = make_chunk_plans(p, cs, P);
    if (rc != PFX_SUCCESS) {
      std::string msg = g_last_error;
      pfx_ndplan_destroy(p);
      return fail(rc, "chunk plans: %s", msg.c_str());
    }
  }
  *out = p;
  return PFX_SUCCESS;
}

int pfx_ndplan_chunk(pfx_ndplan p, int *planes) {
  if (!p || !planes) return fail(PFX_ERR_INVALID_ARG, "null argument");
  *planes = p->chunk;
  return PFX_SUCCESS;
}

int pfx_ndplan_chunk_stage(pfx_ndplan p, int stage, pfx_plan *out) {
  if (!p || !out) return fail(PFX_ERR_INVALID_ARG, "null argument");
  if (p->chunk == 0) return fail(PFX_ERR_INVALID_ARG, "the plan runs stage by stage");
  if (stage < 0 || stage >= (int)p->cstages.size()) return fail(PFX_ERR_INVALID_ARG, "bad stage");
  *out = p->cstages[stage];
  return PFX_SUCCESS;
}

int pfx_ndplan_split(pfx_ndplan p, int *g) {
  if (!p || !g) return fail(PFX_ERR_INVALID_ARG, "null argument");
  *g = p->split;
  return PFX_SUCCESS;
}

int pfx_ndplan_set_split(pfx_ndplan p, int g) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (g < 0) return fail(PFX_ERR_INVALID_ARG, "split factor must be >= 0");
  return split_set(p, g);
}

int pfx_ndplan_split_stage(pfx_ndplan p, pfx_plan *out) {
  if (!p || !out) return fail(PFX_ERR_INVALID_ARG, "null argument");
  if (p->split == 0) return fail(PFX_ERR_INVALID_ARG, "axis 0 is not split");
  *out = p->split_col;
  return PFX_SUCCESS;
}

int pfx_ndplan_set_chunk(pfx_ndplan p, int planes) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (planes == p->chunk) return PFX_SUCCESS;
  split_free(p);  // the split schedule is made for one chunk size
  if (planes != 0 && (p->ndim < 3 || planes < 1 || planes >= p->shape[0] ||
                      p->shape[0] % planes != 0))
    return fail(PFX_ERR_INVALID_ARG, "chunk must divide n0 = %d and be below it (ndim >= 3)",
                p->ndim ? p->shape[0] : 0);
  for (pfx_plan c : p->cstages) pfx_plan_destroy(c);
  for (pfx_plan c : p->bstages) pfx_plan_destroy(c);
  p->cstages.clear();
  p->bstages.clear();
  p->chunk = 0;
  if (planes == 0) return PFX_SUCCESS;
  std::vector<int64_t> cs(p->shape.begin(), p->shape.end());
  if (p->kind == PFX_R2C) cs[p->ndim - 1] = p->shape[p->ndim - 1] / 2 + 1;
  return make_chunk_plans(p, cs, planes);
}

// The split-axis-0 schedule (see pfx_ndplan_t::split): per chunk of P planes
// (P / G values of x1), the fused pass from the array into planes
// [c P, (c + 1) P) of the scratch and the middle axes there; then the q-point
// pass from the scratch into the array, with the scale.  A backward runs the
// same steps conjugated (each kernel's re/im swap), which inverts the whole
// product in the same order.
static int split_run(pfx_ndplan p, int direction, void *data, double scale, hipStream_t s) {
  const int D = p->ndim;
  const int bp = base_prec(p->precision);
  const int64_t esz = bp ? 16 : 8;
  const int G = p->split;
  const int64_t q = p->shape[0] / G;
  const int nchunks = p->shape[0] / p->chunk;
  const size_t chunk_bytes = (size_t)(p->chunk * p->plane_elems * esz);
  const bool fwd = direction == PFX_FORWARD;
  PlaneRowArgs a;
  a.src = data;
  a.dst = p->split_buf;
  a.tw = p->split_tw;
  a.ptw = p->split_ptw;
  a.plane = p->plane_elems;
  a.rows = p->plane_elems / p->shape[D - 1];
  a.nx1 = p->chunk / G;
  a.in_a = 1;
  a.in_b = q;
  a.out_a = G;
  a.out_b = 1;
  a.backward = fwd ? 0 : 1;
  for (int c = 0; c < nchunks; ++c) {
    char *base = static_cast<char *>(p->split_buf) + (size_t)c * chunk_bytes;
    a.x1_begin = (int64_t)c * a.nx1;
    PFX_HIP(launch_plane_row(bp, p->shape[D - 1], G, a, s), "split-axis pass");
    for (int k = 1; k < D - 1; ++k) {
      int rc = pfx_execute_scaled(fwd ? p->cstages[k] : p->bstages[k], direction, base, 1.0, s);
      if (rc != PFX_SUCCESS) return rc;
    }
  }
  return launch_tiled(p->split_col, direction, data, p->split_buf, s, scale);
}

static void split_free(pfx_ndplan p) {
  if (p->split_col) pfx_plan_destroy(p->split_col);
  if (p->split_buf) hipFree(p->split_buf);
  if (p->split_tw) hipFree(p->split_tw);
  if (p->split_ptw) hipFree(p->split_ptw);
  p->split_col = nullptr;
  p->split_buf = p->split_tw = p->split_ptw = nullptr;
  p->split = 0;
}

// Turns the split-axis-0 schedule on with G planes per fused tile (the
// scratch, tables and q-point plan are made here), or off (g = 0).
static int split_set(pfx_ndplan p, int g) {
  if (g == p->split) return PFX_SUCCESS;
  split_free(p);
  if (g == 0) return PFX_SUCCESS;
  const int D = p->ndim;
  const int bp = base_prec(p->precision);
  Radices rad;
  if (p->kind != PFX_C2C || (p->precision != PFX_F32 && p->precision != PFX_F64))
    return fail(PFX_ERR_UNSUPPORTED, "split axis 0: C2C FP32/FP64 plans only");
  if (D < 3 || p->chunk == 0)
    return fail(PFX_ERR_UNSUPPORTED, "split axis 0 needs the chunked schedule (ndim >= 3)");
  if (g < 2 || p->shape[0] % g != 0 || p->chunk % g != 0)
    return fail(PFX_ERR_INVALID_ARG, "split factor %d must divide n0 = %d and the chunk (%d planes)",
                g, p->shape[0], p->chunk);
  if (!plane_row_radices(bp, p->shape[D - 1], g, &rad))
    return fail(PFX_ERR_UNSUPPORTED, "no fused split-axis kernel for last-axis length %d, %d planes",
                p->shape[D - 1], g);
  const int64_t esz = bp ? 16 : 8;
  const int64_t q = p->shape[0] / g;
  int rc = pfx_plan_create_c2c_many(&p->split_col, p->precision, q, p->plane_elems,
                                    (int64_t)g * p->plane_elems, 1, g, p->plane_elems);
  if (rc != PFX_SUCCESS) return rc;
  if (!p->split_col->kernel || p->split_col->lp) {
    split_free(p);
    return fail(PFX_ERR_UNSUPPORTED, "split axis 0: no tiled kernel for the %lld-point pass",
                (long long)q);
  }
  const size_t bytes = (size_t)p->shape[0] * (size_t)p->plane_elems * (size_t)esz;
  hipError_t e = hipMalloc(&p->split_buf, bytes);
  if (e == hipSuccess) e = hipMalloc(&p->split_tw, (size_t)p->shape[0] * esz);
  if (e == hipSuccess) e = hipMalloc(&p->split_ptw, (size_t)p->shape[D - 1] * esz);
  if (e == hipSuccess) e = make_twiddles(bp, p->shape[0], p->split_tw, nullptr);
  if (e == hipSuccess) e = make_pass_twiddles(bp, rad, p->split_ptw, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    split_free(p);
    // an allocation that did not fit stays the runtime's last error; clear it
    // so that the next launch check does not report it as its own
    (void)hipGetLastError();
    return hip_fail(e, "split axis 0 scratch");
  }
  p->split = g;
  return PFX_SUCCESS;
}

// One N-D transform.  `real` (R2C plans only): the unpadded real array, read
// by the first R2C pass (forward) or written by the last C2R pass (backward)
// in place of the padded rows of `data`.
static int nd_run(pfx_ndplan p, int direction, void *data, double scale, pfx_stream stream,
                  void *real) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (direction != PFX_FORWARD && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "bad direction");
  const int D = p->ndim;
  const int64_t nlast = p->shape[D - 1];
  // the R2C (forward) or C2R (backward) stage of the real axis
  auto exec = [&](pfx_plan pl, char *base, double sc, char *real_base) -> int {
    if (real_base && pl->kind != PFX_C2C)
      return execute_impl(pl, direction, base, sc, stream, real_base, nlast);
    return pfx_execute_scaled(pl, direction, base, sc, stream);
  };
  if (p->split > 0 && p->chunk > 0) return split_run(p, direction, data, scale,
                                                     static_cast<hipStream_t>(stream));
  // Axis order.  The transforms along different axes commute, and a C2C
  // backward is every forward pass conjugated (each kernel's re/im swap), so
  // a C2C backward runs the forward's order: last axis first, which lets a
  // chunk's middle axes read what its row pass just wrote from the Infinity
  // Cache.  R2C plans run backward in reverse (the C2R pass takes the
  // complex half-spectrum and must come last).
  const bool fwd_order = direction == PFX_FORWARD || p->kind == PFX_C2C;
  if (p->chunk > 0) {
    const int64_t esz = base_prec(p->precision) ? 16 : 8;
    const int nchunks = p->shape[0] / p->chunk;
    const size_t chunk_bytes = (size_t)(p->chunk * p->plane_elems * esz);
    int64_t real_plane = 1;
    for (int i = 1; i < D; ++i) real_plane *= p->shape[i];
    const size_t chunk_real_bytes = (size_t)(p->chunk * real_plane * (esz / 2));
    auto run_chunks = [&](double last_scale) -> int {
      for (int c = 0; c < nchunks; ++c) {
        char *base = static_cast<char *>(data) + (size_t)c * chunk_bytes;
        char *rbase = real ? static_cast<char *>(real) + (size_t)c * chunk_real_bytes : nullptr;
        for (int k = 0; k < D - 1; ++k) {
          // forward order: stages 0..D-2; reverse: D-2..0 (C2R last for R2C)
          const int st = fwd_order ? k : D - 2 - k;
          const std::vector<pfx_plan> &cl = direction == PFX_FORWARD ? p->cstages : p->bstages;
          pfx_plan pl = cl[st];
          if (p->kind == PFX_R2C && st == 0 && direction == PFX_BACKWARD) pl = cl[D - 1];
          const double sc = (k == D - 2) ? last_scale : 1.0;
          int rc = exec(pl, base, sc, rbase);
          if (rc != PFX_SUCCESS) return rc;
        }
      }
      return PFX_SUCCESS;
    };
    if (fwd_order) {
      int rc = run_chunks(1.0);
      if (rc != PFX_SUCCESS) return rc;
      return pfx_execute_scaled(p->stages[D - 1], direction, data, scale, stream);
    }
    int rc = pfx_execute_scaled(p->stages[D - 1], direction, data, 1.0, stream);
    if (rc != PFX_SUCCESS) return rc;
    return run_chunks(scale);
  }
  for (int k = 0; k < D; ++k) {
    const int s = fwd_order ? k : D - 1 - k;
    pfx_plan pl = p->stages[s];
    if (p->kind == PFX_R2C && s == 0 && direction == PFX_BACKWARD) pl = p->stages[D];
    const double sc = (k == D - 1) ? scale : 1.0;
    int rc = exec(pl, static_cast<char *>(data), sc, static_cast<char *>(real));
    if (rc != PFX_SUCCESS) return rc;
  }
  return PFX_SUCCESS;
}

int pfx_ndplan_execute(pfx_ndplan p, int direction, void *data, double scale, pfx_stream stream) {
  return nd_run(p, direction, data, scale, stream, nullptr);
}

int pfx_ndplan_execute_r2c_oop(pfx_ndplan p, int direction, void *real, void *padded,
                               double scale, pfx_stream stream) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  if (p->kind != PFX_R2C) return fail(PFX_ERR_INVALID_ARG, "not an R2C plan");
  if (!real || !padded) return fail(PFX_ERR_INVALID_ARG, "null buffer");
  if (direction != PFX_FORWARD && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "bad direction");
  // Decided before any stage runs: a backward that found out at its last
  // (C2R) stage would leave the earlier stages applied to `padded`.
  const int D = p->ndim;
  const pfx_plan real_stage = direction == PFX_FORWARD
                                  ? (p->chunk > 0 ? p->cstages[0] : p->stages[0])
                                  : (p->chunk > 0 ? p->bstages[D - 1] : p->stages[D]);
  if (!real_stage->kernel)
    return fail(PFX_ERR_UNSUPPORTED,
                "out-of-place R2C/C2R needs a tiled kernel (even length with a radix plan)");
  return nd_run(p, direction, padded, scale, stream, real);
}

int pfx_ndplan_destroy(pfx_ndplan p) {
  if (!p) return PFX_SUCCESS;
  split_free(p);
  for (pfx_plan s : p->stages) pfx_plan_destroy(s);
  for (pfx_plan s : p->cstages) pfx_plan_destroy(s);
  for (pfx_plan s : p->bstages) pfx_plan_destroy(s);
  delete p;
  return PFX_SUCCESS;
}

// Times the chunked forward against the split-axis-0 schedule for every
// factor G with a fused kernel (each with its q-point pass tuned) and keeps
// a split only if it is at least 2% faster (so timing noise does not trade
// one schedule for the other); `force`: the best split in any case.
// PFX_ND_SPLIT=G restricts the candidates to one factor.  A scratch that
// cannot be allocated leaves the chunked schedule.
static int split_tune(pfx_ndplan p, int flag, void *scratch, pfx_stream stream, bool force) {
  if (p->kind != PFX_C2C || p->chunk == 0 || p->ndim < 3 ||
      (p->precision != PFX_F32 && p->precision != PFX_F64))
    return PFX_SUCCESS;
  const int D = p->ndim;
  const int bp = base_prec(p->precision);
  std::vector<int> gs;
  const char *only = getenv("PFX_ND_SPLIT");
  for (int g : {4, 8, 16}) {
    if (only && *only && atoi(only) != g) continue;
    if (p->shape[0] % g == 0 && p->chunk % g == 0 && plane_row_radices(bp, p->shape[D - 1], g, nullptr))
      gs.push_back(g);
  }
  if (gs.empty()) return PFX_SUCCESS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t a, b;
  PFX_HIP(hipEventCreate(&a), "hipEventCreate");
  PFX_HIP(hipEventCreate(&b), "hipEventCreate");
  const int reps = flag == PFX_PLAN_PATIENT ? 6 : 3;
  auto time_fwd = [&](float *ms) -> int {
    int rc = pfx_ndplan_execute(p, PFX_FORWARD, scratch, 1.0, stream);  // warm-up
    hipEventRecord(a, s);
    for (int r = 0; r < reps && rc == PFX_SUCCESS; ++r)
      rc = pfx_ndplan_execute(p, PFX_FORWARD, scratch, 1.0, stream);
    hipEventRecord(b, s);
    if (hipEventSynchronize(b) != hipSuccess) return PFX_ERR_HIP;
    hipEventElapsedTime(ms, a, b);
    return rc;
  };
  split_free(p);
  float best = -1.f;
  int best_g = 0, rc = PFX_SUCCESS;
  if (!force) rc = time_fwd(&best);
  for (int g : gs) {
    if (rc != PFX_SUCCESS) break;
    if (split_set(p, g) != PFX_SUCCESS) continue;  // no scratch: stay chunked
    rc = pfx_plan_tune(p->split_col, flag, scratch, stream);
    float ms = 0.f;
    if (rc == PFX_SUCCESS) rc = time_fwd(&ms);
    const float need = best_g == 0 && !force ? 0.98f * best : best;
    if (rc == PFX_SUCCESS && (best < 0.f || ms < need)) {
      best = ms;
      best_g = g;
    }
    split_free(p);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != PFX_SUCCESS) return rc;
  if (best_g == 0) return PFX_SUCCESS;
  rc = split_set(p, best_g);
  if (rc == PFX_SUCCESS) rc = pfx_plan_tune(p->split_col, flag, scratch, stream);
  return rc;
}

int pfx_ndplan_tune(pfx_ndplan p, int flag, void *scratch, pfx_stream stream) {
  if (!p) return fail(PFX_ERR_INVALID_ARG, "plan is null");
  for (pfx_plan s : p->stages) {
    int rc = pfx_plan_tune(s, flag, scratch, stream);
    if (rc != PFX_SUCCESS) return rc;
  }
  // Chunk plans start from the kernel and tile order the full stage was
  // tuned to; then each chunked forward stage is re-tuned in context, timing
  // the whole chunked schedule per candidate (a chunk's short launches alone
  // are too noisy, and the best kernel for a pass that reads from the
  // Infinity Cache differs from the best for one that reads HBM).
  for (std::vector<pfx_plan> *list : {&p->cstages, &p->bstages}) {
    for (size_t c = 0; c < list->size(); ++c) {
      const size_t src = (p->kind == PFX_R2C && c + 1 == list->size()) ? p->ndim : c;
      pfx_plan full = p->stages[src], part = (*list)[c];
      if (flag != PFX_PLAN_ESTIMATE && full->kernel && part->kernel) {
        PFX_HIP(set_kernel(part, full->kernel), "chunk kernel");
        part->xcd_remap = full->xcd_remap;
      }
    }
  }
  if (p->chunk > 0 && flag != PFX_PLAN_ESTIMATE) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t a, b;
    PFX_HIP(hipEventCreate(&a), "hipEventCreate");
    PFX_HIP(hipEventCreate(&b), "hipEventCreate");
    const int reps = flag == PFX_PLAN_PATIENT ? 4 : 2;
    auto time_schedule = [&](int direction, float *ms) -> int {
      int rc = pfx_ndplan_execute(p, direction, scratch, 1.0, stream);  // warm-up
      hipEventRecord(a, s);
      for (int r = 0; r < reps && rc == PFX_SUCCESS; ++r)
        rc = pfx_ndplan_execute(p, direction, scratch, 1.0, stream);
      hipEventRecord(b, s);
      if (hipEventSynchronize(b) != hipSuccess) return PFX_ERR_HIP;
      hipEventElapsedTime(ms, a, b);
      return rc;
    };
    // Each chunk stage a direction runs, re-tuned in that direction's whole
    // chunked schedule; the in-context choice is cached under the chunk
    // plan's key with a suffix (a cache file replays it, as for single plans).
    auto tune_in_context = [&](pfx_plan part, int direction) -> int {
      if (!part->kernel) return PFX_SUCCESS;
      const std::string ctx_key = plan_key(part) + " chunk" + std::to_string(p->chunk) + " of" +
                                  std::to_string(p->shape[0]) +
                                  (direction == PFX_BACKWARD ? " bwd" : "");
      {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune.find(ctx_key);
        const KernelEntry *e = nullptr;
        int order = 1;
        if (it != g_tune.end() && parse_tuned(it->second, &e, &order) &&
            kernel_fits_plan(part, part->kernel, e, order)) {
          PFX_HIP(set_kernel(part, e), "chunk kernel");
          part->xcd_remap = order;
          return PFX_SUCCESS;
        }
      }
      const KernelEntry *best = part->kernel;
      int best_order = part->xcd_remap;
      float best_ms = -1.f;
      for (const KernelEntry *k : candidates(part)) {
        if (set_kernel(part, k) != hipSuccess) continue;
        std::vector<int> orders = {part->xcd_remap};
        if (k->col) orders = {1, 8, 32};
        for (int order : orders) {
          part->xcd_remap = order;
          float ms = 0.f;
          if (time_schedule(direction, &ms) != PFX_SUCCESS) continue;
          if (best_ms < 0.f || ms < best_ms) {
            best_ms = ms;
            best = k;
            best_order = order;
          }
        }
      }
      PFX_HIP(set_kernel(part, best), "chunk kernel");
      part->xcd_remap = best_order;
      std::lock_guard<std::mutex> lock(g_tune_mu);
      g_tune[ctx_key] = entry_id(best) + "|order=" + std::to_string(best_order);
      return PFX_SUCCESS;
    };
    int rc = PFX_SUCCESS;
    for (int st = 0; st + 1 < p->ndim && rc == PFX_SUCCESS; ++st)
      rc = tune_in_context(p->cstages[st], PFX_FORWARD);
    // backward: the same stages, except that an R2C plan's last axis runs
    // its C2R plan
    for (int st = 0; st + 1 < p->ndim && rc == PFX_SUCCESS; ++st)
      rc = tune_in_context(p->bstages[p->kind == PFX_R2C && st == 0 ? p->ndim - 1 : st],
                           PFX_BACKWARD);
    if (rc != PFX_SUCCESS) {
      hipEventDestroy(a);
      hipEventDestroy(b);
      return rc;
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
  }
  // PFX_ND_SCHEDULE=chunked|stagewise|split overrides the comparisons below
  // (for measurement, e.g. counter runs that serialise dispatches): chunked
  // keeps the chunks without a split axis 0, split keeps the chunks and the
  // fastest split
  const char *forced = getenv("PFX_ND_SCHEDULE");
  if (forced && !strcmp(forced, "stagewise")) return pfx_ndplan_set_chunk(p, 0);
  const bool keep_chunks = forced && (!strcmp(forced, "chunked") || !strcmp(forced, "split"));
  if (p->chunk > 0 && flag != PFX_PLAN_ESTIMATE && !keep_chunks) {
    // keep the chunked schedule only if it times faster than stage-by-stage
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t a, b;
    PFX_HIP(hipEventCreate(&a), "hipEventCreate");
    PFX_HIP(hipEventCreate(&b), "hipEventCreate");
    const int reps = flag == PFX_PLAN_PATIENT ? 6 : 2;
    float ms[2] = {0.f, 0.f};
    const int chunk = p->chunk;
    int rc = PFX_SUCCESS;
    for (int mode = 0; mode < 2 && rc == PFX_SUCCESS; ++mode) {
      p->chunk = mode ? chunk : 0;
      rc = pfx_ndplan_execute(p, PFX_FORWARD, scratch, 1.0, stream);  // warm-up
      hipEventRecord(a, s);
      for (int r = 0; r < reps && rc == PFX_SUCCESS; ++r)
        rc = pfx_ndplan_execute(p, PFX_FORWARD, scratch, 1.0, stream);
      hipEventRecord(b, s);
      if (hipEventSynchronize(b) == hipSuccess) hipEventElapsedTime(&ms[mode], a, b);
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    p->chunk = chunk;
    if (rc != PFX_SUCCESS) return rc;
    if (ms[0] <= ms[1]) return pfx_ndplan_set_chunk(p, 0);
  }
  if (flag != PFX_PLAN_ESTIMATE && !(forced && !strcmp(forced, "chunked")))
    return split_tune(p, flag, scratch, stream, forced && !strcmp(forced, "split"));
  return PFX_SUCCESS;
}

int pfx_ndplan_stage(pfx_ndplan p, int stage, pfx_plan *out) {
  if (!p || !out) return fail(PFX_ERR_INVALID_ARG, "null argument");
  if (stage < 0 || stage >= (int)p->stages.size()) return fail(PFX_ERR_INVALID_ARG, "bad stage");
  *out = p->stages[stage];
  return PFX_SUCCESS;
}

// ------------------------------------------------- many uneven transforms

}  // extern "C"

namespace {

struct Group {
  const KernelEntry *kernel;
  std::vector<StockhamArgs> jobs;
  std::vector<int64_t> start;  // workgroup start of every job (+ total)
  StockhamArgs *d_jobs = nullptr;
  int64_t *d_start = nullptr;
};

struct Level {
  std::vector<Group> groups;
  std::vector<std::pair<pfx_plan, void *>> direct;  // stages without a tiled kernel
};

StockhamArgs job_args(pfx_plan p, void *data, int direction) {
  const bool real = p->kind != PFX_C2C;
  StockhamArgs a;
  a.src = nullptr;
  a.data = data;
  a.tw = p->tw;
  a.ptw = p->ptw;
  a.total = p->inner * p->outer;
  a.inner = p->inner;
  a.stride = p->stride;
  a.dist = real ? p->dist / 2 : p->dist;
  a.dist_outer = real ? p->dist_outer / 2 : p->dist_outer;
  a.scale = 1.0;
  a.use_scale = 0;
  a.backward = direction == PFX_BACKWARD;
  a.xcd_remap = p->xcd_remap;
  a.in_seg = a.out_seg = 0;
  return a;
}

}  // namespace

namespace {

// The persistent schedule of one direction: one job per transform stage, in
// an order where every job's dependency (the same transform's previous
// stage) comes first; tiles of each job split over the per-XCD queues.
struct Persist {
  bool ok = false;
  std::string why;  // why there is no persistent schedule
  std::vector<PJob> jobs;
  std::vector<StockhamArgs> args;
  std::vector<int> qstart;  // kPersistQueues rows of njobs + 1 offsets
  int tickets = 0;
  PJob *d_jobs = nullptr;
  StockhamArgs *d_args = nullptr;
  int *d_qstart = nullptr;
};

}  // namespace

struct pfx_batch_t {
  int precision = PFX_F32;
  std::vector<pfx_ndplan> plans;
  std::vector<void *> data;
  std::vector<Level> levels[2];  // [direction]: every chunk's stage levels, in order
  Persist persist[2];            // [direction]
  int *d_ctl = nullptr;          // control words + per-job counters (persistent launch)
  std::map<std::string, void *> persist_tw;  // pass tables for bodies with other radices
  int spin_limit = 1 << 21;
  int grid = 0;
  int nqueues = 1;
};

struct pfx_graph_t {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

}  // extern "C"

namespace {

void free_levels(pfx_batch b) {
  for (auto &dir : b->levels) {
    for (auto &lv : dir)
      for (auto &g : lv.groups) {
        if (g.d_jobs) (void)hipFree(g.d_jobs);
        if (g.d_start) (void)hipFree(g.d_start);
      }
    dir.clear();
  }
}

// Bytes a transform occupies in HBM (its complex view).
int64_t footprint(pfx_ndplan nd) {
  int64_t elems = 1;
  for (int k = 0; k + 1 < nd->ndim; ++k) elems *= nd->shape[k];
  const int last = nd->shape[nd->ndim - 1];
  elems *= nd->kind == PFX_C2C ? last : last / 2 + 1;
  return elems * (nd->precision == PFX_F64 ? 16 : 8);
}

int64_t env_mb(const char *name, int64_t dflt) {
  const char *v = getenv(name);
  return v && *v ? (int64_t)atoll(v) : dflt;
}

// The schedule: transforms of at least kSmall bytes are packed, largest
// first, into chunks of at most kChunk bytes, and each chunk runs all its
// stages before the next starts, so the later passes of a chunk find its
// data still in the 256 MB Infinity Cache (MALL).  Everything smaller forms
// one chunk whose stages are grouped across all its transforms (those are
// too small to fill the GPU one at a time).  Within a chunk, each stage level
// is one grouped launch per kernel variant, with device job tables (each job
// starting at a multiple of 8 workgroups for the XCD-aware order).
int build_levels(pfx_batch b) {
  free_levels(b);
  const int count = (int)b->plans.size();
  const int64_t chunk_bytes = env_mb("PFX_BATCH_CHUNK_MB", 192) << 20;
  const int64_t small_bytes = env_mb("PFX_BATCH_SMALL_MB", 16) << 20;
  std::vector<std::vector<int>> chunks;
  std::vector<int> small, big;
  for (int i = 0; i < count; ++i)
    (footprint(b->plans[i]) >= small_bytes && chunk_bytes > 0 ? big : small).push_back(i);
  std::stable_sort(big.begin(), big.end(), [&](int x, int y) {
    return footprint(b->plans[x]) > footprint(b->plans[y]);
  });
  int64_t used = 0;
  for (int i : big) {
    const int64_t f = footprint(b->plans[i]);
    if (chunks.empty() || used + f > chunk_bytes) {
      chunks.emplace_back();
      used = 0;
    }
    chunks.back().push_back(i);
    used += f;
  }
  if (!small.empty()) chunks.push_back(small);
  for (int dir = 0; dir < 2; ++dir) {
    for (const std::vector<int> &chunk : chunks) {
      int maxd = 0;
      for (int i : chunk) maxd = std::max(maxd, b->plans[i]->ndim);
      const size_t base = b->levels[dir].size();
      b->levels[dir].resize(base + maxd);
      for (int i : chunk) {
        pfx_ndplan nd = b->plans[i];
        const int D = nd->ndim;
        for (int k = 0; k < D; ++k) {
          const int st = dir == PFX_FORWARD ? k : D - 1 - k;
          pfx_plan pl = nd->stages[st];
          if (nd->kind == PFX_R2C && st == 0 && dir == PFX_BACKWARD) pl = nd->stages[D];
          Level &lv = b->levels[dir][base + k];
          if (!pl->kernel || !pl->kernel->launch_group) {
            lv.direct.push_back({pl, b->data[i]});
            continue;
          }
          Group *g = nullptr;
          for (auto &gg : lv.groups)
            if (gg.kernel == pl->kernel) g = &gg;
          if (!g) {
            lv.groups.push_back(Group{pl->kernel, {}, {}, nullptr, nullptr});
            g = &lv.groups.back();
          }
          g->jobs.push_back(job_args(pl, b->data[i], dir));
        }
      }
    }
    for (auto &lv : b->levels[dir])
      for (auto &g : lv.groups) {
        int64_t pos = 0;
        for (const StockhamArgs &a : g.jobs) {
          g.start.push_back(pos);
          const int64_t tiles = (a.total + g.kernel->per_block - 1) / g.kernel->per_block;
          pos += (tiles + 7) / 8 * 8;
        }
        g.start.push_back(pos);
        if (pos > 0x7fffffff) return fail(PFX_ERR_UNSUPPORTED, "grouped launch exceeds the grid limit");
        hipError_t e = hipMalloc(&g.d_jobs, g.jobs.size() * sizeof(StockhamArgs));
        if (e == hipSuccess) e = hipMalloc(&g.d_start, g.start.size() * sizeof(int64_t));
        if (e == hipSuccess)
          e = hipMemcpy(g.d_jobs, g.jobs.data(), g.jobs.size() * sizeof(StockhamArgs),
                        hipMemcpyHostToDevice);
        if (e == hipSuccess)
          e = hipMemcpy(g.d_start, g.start.data(), g.start.size() * sizeof(int64_t),
                        hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "batch job tables");
      }
  }
  return PFX_SUCCESS;
}

void free_persist(pfx_batch b) {
  for (Persist &P : b->persist) {
    if (P.d_jobs) (void)hipFree(P.d_jobs);
    if (P.d_args) (void)hipFree(P.d_args);
    if (P.d_qstart) (void)hipFree(P.d_qstart);
    P = Persist();
  }
  if (b->d_ctl) (void)hipFree(b->d_ctl);
  b->d_ctl = nullptr;
}

bool same_radices(const Radices &x, const Radices &y) {
  if (x.n != y.n) return false;
  for (int i = 0; i < x.n; ++i)
    if (x.r[i] != y.r[i]) return false;
  return true;
}

// The persistent body computing stage plan `pl`: same kind, length and
// layout, with the radices of the plan's kernel when a body has them (then
// the arithmetic, and the result, are the plan's bit for bit).  -1: none.
int persist_body_for(const pfx_plan_t *pl) {
  if (!pl->kernel) return -1;  // lengths without a radix plan (direct DFT)
  int nb = 0;
  const PersistBody *B = persist_bodies(&nb);
  const int kk = pl->kind == PFX_C2C ? KIND_C2C : (pl->kind == PFX_R2C ? KIND_R2C : KIND_C2R);
  int col;
  if (pl->kind != PFX_C2C || pl->stride == 1) col = 0;
  else if (pl->dist == 1) col = 1;
  else return -1;
  if (pl->kind != PFX_C2C && pl->stride != 1) return -1;
  int any = -1;
  for (int i = 0; i < nb; ++i) {
    if (B[i].kind != kk || B[i].n != pl->kernel->n || B[i].col != col) continue;
    if (same_radices(B[i].rad, pl->kernel->rad)) return i;
    if (any < 0) any = i;
  }
  return any;
}

// Per-pass twiddle table of a body's radices (the plan's own when they match).
int body_pass_table(pfx_batch b, const pfx_plan_t *pl, const PersistBody &body,
                    const void **out) {
  if (same_radices(body.rad, pl->kernel->rad)) {
    *out = pl->ptw;
    return PFX_SUCCESS;
  }
  std::string key = std::to_string(body.n) + ":" + body.name;
  auto it = b->persist_tw.find(key);
  if (it == b->persist_tw.end()) {
    void *t = nullptr;
    hipError_t e = hipMalloc(&t, (size_t)body.n * 8);
    if (e == hipSuccess) e = make_pass_twiddles(0, body.rad, t, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      if (t) (void)hipFree(t);
      return hip_fail(e, "persistent pass table");
    }
    it = b->persist_tw.emplace(key, t).first;
  }
  *out = it->second;
  return PFX_SUCCESS;
}

// Builds both directions' persistent schedules.  Every transform goes to
// one queue (one XCD), largest first to the queue with the fewest bytes so
// far.  Within a queue the transforms below PFX_BATCH_SMALL_MB (16) MiB come
// first, stage level by stage level; the larger ones follow pipelined,
// largest first: stage k of the r-th of them sits in wave r + k L
// (L = PFX_PERSIST_LAG, 1 by default), so a stage's tickets come a wave
// after those of the stage it waits for, while the next transform's first
// stage keeps the XCD busy.
int build_persist(pfx_batch b) {
  free_persist(b);
  const int nq = persist_queues();
  b->nqueues = nq;
  for (int dir = 0; dir < 2; ++dir) {
    Persist &P = b->persist[dir];
    P.ok = false;
    if (b->precision != PFX_F32) {
      P.why = "the persistent kernel is built for complex<float> batches";
      continue;
    }
    b->grid = persist_grid();
    if (b->grid <= 0) {
      P.why = "no device occupancy for the persistent kernel";
      continue;
    }
    const int count = (int)b->plans.size();
    const int64_t small_bytes = env_mb("PFX_BATCH_SMALL_MB", 16) << 20;
    const int lag = (int)env_mb("PFX_PERSIST_LAG", 1);
    // transform -> queue, largest first to the least loaded queue
    std::vector<int> by_size(count);
    for (int i = 0; i < count; ++i) by_size[i] = i;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int x, int y) {
      return footprint(b->plans[x]) > footprint(b->plans[y]);
    });
    std::vector<int64_t> load(nq, 0);
    std::vector<std::vector<int>> small(nq), big(nq);  // in size order
    for (int i : by_size) {
      const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[q] += footprint(b->plans[i]);
      (footprint(b->plans[i]) >= small_bytes ? big : small)[q].push_back(i);
    }
    std::vector<std::pair<int, int>> order;  // (transform, stage step k), queue by queue
    std::vector<int> order_q;
    for (int q = 0; q < nq; ++q) {
      int maxd = 0;
      for (int i : small[q]) maxd = std::max(maxd, b->plans[i]->ndim);
      for (int k = 0; k < maxd; ++k)
        for (int i : small[q])
          if (k < b->plans[i]->ndim) {
            order.push_back({i, k});
            order_q.push_back(q);
          }
      struct W { int w, k, r; };
      std::vector<W> waves;
      for (int r = 0; r < (int)big[q].size(); ++r)
        for (int k = 0; k < b->plans[big[q][r]]->ndim; ++k) waves.push_back({r + k * lag, k, r});
      std::stable_sort(waves.begin(), waves.end(), [](const W &x, const W &y) {
        return x.w != y.w ? x.w < y.w : x.k < y.k;
      });
      for (const W &w : waves) {
        order.push_back({big[q][w.r], w.k});
        order_q.push_back(q);
      }
    }
    int nb = 0;
    const PersistBody *B = persist_bodies(&nb);
    std::vector<int> last(count, -1);  // job of each transform's latest stage
    std::vector<int> job_q;            // queue of every job
    bool ok = true;
    for (size_t o = 0; o < order.size(); ++o) {
      const int i = order[o].first, k = order[o].second;
      pfx_ndplan nd = b->plans[i];
      const int D = nd->ndim;
      const int st = dir == PFX_FORWARD ? k : D - 1 - k;
      pfx_plan pl = nd->stages[st];
      if (nd->kind == PFX_R2C && st == 0 && dir == PFX_BACKWARD) pl = nd->stages[D];
      const int body = persist_body_for(pl);
      if (body < 0) {
        P.why = "transform " + std::to_string(i) + " has a stage (length " +
                std::to_string(pl->length) + ") without a persistent kernel body";
        ok = false;
        break;
      }
      StockhamArgs a = job_args(pl, b->data[i], dir);
      a.xcd_remap = 0;  // a queue's tiles run in order on one XCD
      int rc = body_pass_table(b, pl, B[body], &a.ptw);
      if (rc != PFX_SUCCESS) return rc;
      if (a.total == 0) continue;
      const int64_t tiles = (a.total + B[body].per_block - 1) / B[body].per_block;
      if (tiles > (1 << 30)) {
        P.why = "a stage has too many tiles for the ticket queues";
        ok = false;
        break;
      }
      PJob J;
      J.body = body;
      J.ntiles = (int)tiles;
      J.dep = last[i];
      J.need = J.dep >= 0 ? P.jobs[J.dep].ntiles : 0;
      last[i] = (int)P.jobs.size();
      P.jobs.push_back(J);
      P.args.push_back(a);
      job_q.push_back(order_q[o]);
    }
    if (!ok) {
      P.jobs.clear();
      P.args.clear();
      continue;
    }
    const int nj = (int)P.jobs.size();
    // every wait is on an earlier job of the same queue, for all of its
    // tiles (the progress argument of the persistent kernel rests on this)
    for (int j = 0; j < nj; ++j) {
      const PJob &J = P.jobs[j];
      if (J.body < 0 || J.body >= nb || J.ntiles < 1 || J.dep >= j ||
          (J.dep >= 0 && (J.need != P.jobs[J.dep].ntiles || job_q[J.dep] != job_q[j])))
        return fail(PFX_ERR_INTERNAL, "persistent schedule: job %d breaks the ticket order", j);
    }
    P.qstart.assign((size_t)kPersistQueues * (nj + 1), 0);
    int64_t tickets = 0;
    for (int x = 0; x < kPersistQueues; ++x) {
      int64_t pos = 0;
      for (int j = 0; j < nj; ++j) {
        P.qstart[(size_t)x * (nj + 1) + j] = (int)pos;
        if (job_q[j] == x) pos += P.jobs[j].ntiles;
      }
      P.qstart[(size_t)x * (nj + 1) + nj] = (int)pos;
      tickets += pos;
      if (pos > 0x7fffffff) return fail(PFX_ERR_UNSUPPORTED, "ticket queue exceeds 2^31 entries");
    }
    P.tickets = (int)tickets;
    if (nj > 0) {
      hipError_t e = hipMalloc(&P.d_jobs, nj * sizeof(PJob));
      if (e == hipSuccess) e = hipMalloc(&P.d_args, nj * sizeof(StockhamArgs));
      if (e == hipSuccess) e = hipMalloc(&P.d_qstart, P.qstart.size() * sizeof(int));
      if (e == hipSuccess)
        e = hipMemcpy(P.d_jobs, P.jobs.data(), nj * sizeof(PJob), hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(P.d_args, P.args.data(), nj * sizeof(StockhamArgs), hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(P.d_qstart, P.qstart.data(), P.qstart.size() * sizeof(int),
                      hipMemcpyHostToDevice);
      if (e != hipSuccess) return hip_fail(e, "persistent job tables");
    }
    P.ok = true;
  }
  size_t ctl = kCtlInts;
  ctl += std::max(b->persist[0].jobs.size(), b->persist[1].jobs.size());
  hipError_t e = hipMalloc(&b->d_ctl, ctl * sizeof(int));
  if (e != hipSuccess) return hip_fail(e, "persistent control words");
  return PFX_SUCCESS;
}

int build_schedules(pfx_batch b) {
  int rc = build_levels(b);
  if (rc == PFX_SUCCESS) rc = build_persist(b);
  return rc;
}

}  // namespace

extern "C" {

int pfx_batch_destroy(pfx_batch b) {
  if (!b) return PFX_SUCCESS;
  free_levels(b);
  free_persist(b);
  for (auto &kv : b->persist_tw) (void)hipFree(kv.second);
  for (pfx_ndplan p : b->plans) pfx_ndplan_destroy(p);
  delete b;
  return PFX_SUCCESS;
}

int pfx_batch_create(pfx_batch *out, int precision, int count, const pfx_batch_desc *descs) {
  if (!out) return fail(PFX_ERR_INVALID_ARG, "batch pointer is null");
  *out = nullptr;
  if (count < 0 || (count > 0 && !descs)) return fail(PFX_ERR_INVALID_ARG, "bad descriptor list");
  pfx_batch b = new (std::nothrow) pfx_batch_t;
  if (!b) return fail(PFX_ERR_ALLOC, "out of host memory");
  b->precision = precision;
  for (int i = 0; i < count; ++i) {
    const pfx_batch_desc &d = descs[i];
    if (d.ndim < 1 || d.ndim > 4) {
      pfx_batch_destroy(b);
      return fail(PFX_ERR_INVALID_ARG, "transform %d: ndim must be 1..4", i);
    }
    pfx_ndplan nd = nullptr;
    int rc = pfx_ndplan_create(&nd, precision, d.kind, d.ndim, d.shape);
    if (rc != PFX_SUCCESS) {
      std::string msg = g_last_error;
      pfx_batch_destroy(b);
      return fail(rc, "transform %d: %s", i, msg.c_str());
    }
    b->plans.push_back(nd);
    b->data.push_back(d.data);
  }
  int rc = build_schedules(b);
  if (rc != PFX_SUCCESS) {
    std::string msg = g_last_error;
    pfx_batch_destroy(b);
    return fail(rc, "%s", msg.c_str());
  }
  *out = b;
  return PFX_SUCCESS;
}

int pfx_batch_tune(pfx_batch b, int flag, pfx_stream stream) {
  if (!b) return fail(PFX_ERR_INVALID_ARG, "batch is null");
  for (size_t i = 0; i < b->plans.size(); ++i) {
    // the transform's own buffer is the planning scratch (its contents are lost)
    int rc = pfx_ndplan_tune(b->plans[i], flag, b->data[i], stream);
    if (rc != PFX_SUCCESS) return rc;
  }
  hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "planning");
  return build_schedules(b);
}

int pfx_batch_execute(pfx_batch b, int direction, pfx_stream stream) {
  if (!b) return fail(PFX_ERR_INVALID_ARG, "batch is null");
  if (direction != PFX_FORWARD && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "bad direction");
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (auto &lv : b->levels[direction]) {
    for (auto &g : lv.groups) {
      GroupArgs ga{g.d_jobs, g.d_start, (int)g.jobs.size()};
      g.kernel->launch_group(ga, g.start.back(), s);
    }
    for (auto &dp : lv.direct) {
      int rc = pfx_execute(dp.first, direction, dp.second, stream);
      if (rc != PFX_SUCCESS) return rc;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "grouped launch");
  return PFX_SUCCESS;
}

int pfx_batch_execute_each(pfx_batch b, int direction, pfx_stream stream) {
  if (!b) return fail(PFX_ERR_INVALID_ARG, "batch is null");
  for (size_t i = 0; i < b->plans.size(); ++i) {
    int rc = pfx_ndplan_execute(b->plans[i], direction, b->data[i], 1.0, stream);
    if (rc != PFX_SUCCESS) return rc;
  }
  return PFX_SUCCESS;
}

int pfx_batch_execute_persistent(pfx_batch b, int direction, pfx_stream stream) {
  if (!b) return fail(PFX_ERR_INVALID_ARG, "batch is null");
  if (direction != PFX_FORWARD && direction != PFX_BACKWARD)
    return fail(PFX_ERR_INVALID_ARG, "bad direction");
  Persist &P = b->persist[direction];
  if (!P.ok) return fail(PFX_ERR_UNSUPPORTED, "no persistent schedule: %s", P.why.c_str());
  if (P.jobs.empty()) return PFX_SUCCESS;
  PersistArgs a;
  a.jobs = P.d_jobs;
  a.args = P.d_args;
  a.qstart = P.d_qstart;
  a.njobs = (int)P.jobs.size();
  a.ctl = b->d_ctl;
  a.done = b->d_ctl + kCtlInts;
  a.spin_limit = b->spin_limit;
  a.nqueues = b->nqueues;
  PFX_HIP(launch_persist(a, b->grid, static_cast<hipStream_t>(stream)), "persistent launch");
  return PFX_SUCCESS;
}

int pfx_batch_persistent_status(pfx_batch b, pfx_stream stream, int *failed) {
  if (!b || !failed) return fail(PFX_ERR_INVALID_ARG, "null argument");
  *failed = 0;
  if (!b->d_ctl) return PFX_SUCCESS;
  PFX_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
  int err = 0;
  PFX_HIP(hipMemcpy(&err, b->d_ctl + kCtlError, sizeof(int), hipMemcpyDeviceToHost),
          "persistent status");
  *failed = err;
  return PFX_SUCCESS;
}

int pfx_batch_set_spin_limit(pfx_batch b, int polls) {
  if (!b) return fail(PFX_ERR_INVALID_ARG, "batch is null");
  if (polls < 0) return fail(PFX_ERR_INVALID_ARG, "spin limit must not be negative");
  b->spin_limit = polls;
  return PFX_SUCCESS;
}

int pfx_batch_persistent_info(pfx_batch b, int direction, int *jobs, int *tickets,
                              int *workgroups) {
  if (!b || direction < 0 || direction > 1) return fail(PFX_ERR_INVALID_ARG, "bad argument");
  const Persist &P = b->persist[direction];
  if (!P.ok) return fail(PFX_ERR_UNSUPPORTED, "no persistent schedule: %s", P.why.c_str());
  if (jobs) *jobs = (int)P.jobs.size();
  if (tickets) *tickets = P.tickets;
  if (workgroups) *workgroups = b->grid;
  return PFX_SUCCESS;
}

int pfx_batch_launches(pfx_batch b, int direction, int *grouped, int *each) {
  if (!b || direction < 0 || direction > 1) return fail(PFX_ERR_INVALID_ARG, "bad argument");
  int g = 0, e = 0;
  for (auto &lv : b->levels[direction]) g += (int)(lv.groups.size() + lv.direct.size());
  for (pfx_ndplan p : b->plans) e += p->ndim;
  if (grouped) *grouped = g;
  if (each) *each = e;
  return PFX_SUCCESS;
}

// ---------------------------------------------------------------- graphs

int pfx_graph_begin_capture(pfx_stream stream) {
  if (!stream) return fail(PFX_ERR_INVALID_ARG, "capture needs a non-null stream");
  PFX_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal),
          "hipStreamBeginCapture");
  return PFX_SUCCESS;
}

int pfx_graph_end_capture(pfx_stream stream, pfx_graph *out) {
  if (!stream || !out) return fail(PFX_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  pfx_graph g = new (std::nothrow) pfx_graph_t;
  if (!g) return fail(PFX_ERR_ALLOC, "out of host memory");
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return hip_fail(e, "graph capture");
  }
  *out = g;
  return PFX_SUCCESS;
}

int pfx_graph_launch(pfx_graph g, pfx_stream stream) {
  if (!g) return fail(PFX_ERR_INVALID_ARG, "graph is null");
  PFX_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)), "hipGraphLaunch");
  return PFX_SUCCESS;
}

int pfx_graph_destroy(pfx_graph g) {
  if (!g) return PFX_SUCCESS;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return PFX_SUCCESS;
}

}  // extern "C"
