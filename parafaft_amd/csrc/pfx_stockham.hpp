// pfx_stockham.hpp -- batched, strided 1D FFT kernels for gfx950 (CDNA4).
//
// One workgroup transforms C independent 1D sequences of length N that are
// fully resident in LDS.  The transform is a mixed-radix Stockham (autosort)
// decimation-in-time: pass p with radix R and span Ns (product of the earlier
// radices) maps butterfly jj in [0, N/R) to inputs jj + m N/R and outputs
// (jj / Ns) Ns R + (jj % Ns) + m Ns, after multiplying input m by
// W_{Ns R}^{(jj % Ns) m}.  The first pass reads straight from HBM and the last
// pass writes straight to HBM; passes in between exchange through LDS.
//
// Two thread layouts cover the strided axes of a pencil:
//   ROW  (stride == 1): the T threads of one transform are adjacent lanes, so
//        a first-pass load of element j + m N/R is a contiguous run per wave.
//   COL  (dist == 1, the batch index is the contiguous one): the C transforms
//        of a workgroup are adjacent lanes, so every load/store of one element
//        index touches C consecutive complex values -- coalesced across the
//        batch, which is how the axis-0 and middle-axis passes of a row-major
//        pencil reach HBM bandwidth.
// Addresses are general: element e of transform b lives at
//   (b / inner) * dist_outer + (b % inner) * dist + e * stride
// so the middle axis of an N-D array (the reference's loop of leading_size
// execute_stage calls) runs as a single launch.
//
// Kinds: C2C (both directions), R2C (forward, half-length complex FFT + fused
// post-twiddle), C2R (backward, fused pre-twiddle + half-length inverse FFT).
// An optional scale is fused into the final HBM store.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "pfx_butterfly.hpp"

namespace pfx {

enum Kind { KIND_C2C = 0, KIND_R2C = 1, KIND_C2R = 2 };

template <int R, int... Rest> constexpr int first_of() { return R; }

// Element i of a table in device (global) memory, read through a
// global-address-space pointer (a pointer that came from a job table in
// memory would otherwise be read with a generic, flat, load).
template <typename T> __device__ __forceinline__ V2<T> gload(const V2<T>* p, int64_t i) {
  typedef T E2 __attribute__((ext_vector_type(2)));
  const E2 x = *((const __attribute__((address_space(1))) E2*)(p + i));
  return mk<T>(x.x, x.y);
}

// Thread index of the calling lane in its workgroup, as the kernel bodies
// read it.  A kernel that runs many bodies in a loop (the persistent kernel)
// defines PFX_OPAQUE_TID before including this header: the index then comes
// out of an empty asm statement at every use, so the compiler cannot hoist
// the bodies' index and LDS-address arithmetic out of the loop, where all of
// it would stay live across every body and spill.
__device__ __forceinline__ int lane_tid() {
#ifdef PFX_OPAQUE_TID
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
#else
  return threadIdx.x;
#endif
}

// One axis of a row-major block split into per-peer segments, in the order
// the pencil exchange packs it (parafaft_common.hpp:576-590): element
// (lead, e, trail) of the block, with e on the split axis, lives at
//   seg[p][(lead * n[p] + (e - s[p])) * trailing + trail]
// p being the peer whose block [s[p], s[p] + n[p]) holds e.  Every segment
// has its own address, so the rank's own block can sit in its receive
// segment while the others sit in the send buffer.  A stage's last pass can
// store straight into this layout (the pack) and the next stage's first
// pass can load from it (the unpack).
constexpr int kMaxSegPeers = 16;
struct SegMap {
  int nparts;
  int s[kMaxSegPeers];
  int n[kMaxSegPeers];
  void* seg[kMaxSegPeers];
  int64_t trailing;
};

// Address of element e of the transform at (lead, trail); a transform's
// (lead, trail) is split from its index once, at set-up.  The lanes of a
// wave that load or store together hold the consecutive elements e0 + j of
// a block of `blk` elements (j < blk their position in the transform), with
// e0 the same for the whole wave.  So the peer of e0 is found with scalar
// work (loop and loads from the kernel-argument segment), and per lane only
// the peer boundaries inside the block need a select; usually there are
// none.  No per-lane indexing into the map (that would move the kernel
// arguments to scratch).  The host guarantees lead * n[p] + n[p] < 2^31 and
// trailing < 2^32.
using KSegMap = const __attribute__((address_space(4))) SegMap;

template <typename T>
__device__ __forceinline__ V2<T>* seg_addr(KSegMap* m, uint32_t lead, int64_t trail, int e,
                                           int e0, int blk) {
  int p0 = 0;
  for (int q = 1; q < m->nparts; ++q)
    if (e0 >= m->s[q]) p0 = q;
  V2<T>* base = static_cast<V2<T>*>(m->seg[p0]);
  int nn = m->n[p0], ss = m->s[p0];
  for (int q = p0 + 1; q < m->nparts && m->s[q] < e0 + blk; ++q) {
    const bool here = e >= m->s[q];
    base = here ? static_cast<V2<T>*>(m->seg[q]) : base;
    nn = here ? m->n[q] : nn;
    ss = here ? m->s[q] : ss;
  }
  const uint32_t row = lead * (uint32_t)nn + (uint32_t)(e - ss);
  return base + ((uint64_t)row * (uint64_t)m->trailing + trail);
}

struct StockhamArgs {
  void* data;            // complex view of the buffer (R2C/C2R: padded rows)
  const void* src = nullptr;  // out of place: first-pass loads come from here
                              // (nullptr: in place, from data), transform b at
                              // (b / inner) src_dist_outer + (b % inner) src_dist
  int64_t src_dist = 0, src_dist_outer = 0;
  void* dst = nullptr;        // out of place: last-pass stores go here (nullptr:
                              // into data), at dst_dist / dst_dist_outer likewise
  int64_t dst_dist = 0, dst_dist_outer = 0;
  const void* tw;        // W_L^i, i in [0, L): L = N (C2C) or 2N (R2C/C2R)
  const void* ptw;       // per-pass table of the LDS FFT: entry m*Ns + kk - 1
                         // holds W_{Ns R}^{kk m} for the pass with span Ns
  int64_t total;         // number of transforms = inner * outer
  int64_t inner;         // transforms per outer slice
  int64_t stride;        // element stride inside a transform (complex units)
  int64_t dist;          // distance between transforms inside a slice
  int64_t dist_outer;    // distance between outer slices
  double scale;          // multiplies every output (1.0 = none)
  int backward;          // C2C only: 1 = exp(+i...), via re/im swap
  int use_scale;
  int xcd_remap;         // tile order: 0 launch order; 1 one contiguous run of
                         // tiles per XCD; K > 1: XCDs take turns on chunks of K
  int in_seg;            // 1: first pass loads from in_map (fused unpack)
  int out_seg;           // 1: last pass stores into out_map (fused pack)
  SegMap in_map, out_map;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8), and each XCD has its own L2.  Batch-contiguous tiles that share
// a cache line must run on the same XCD, so tile ids are remapped to give XCD
// x the contiguous range [x q + min(x, r), ...) of the nb = 8 q + r tiles.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int64_t xcd_tile(int64_t i, int64_t nb) {
  const int64_t q = nb / kNumXcd, r = nb % kNumXcd;
  const int64_t x = i % kNumXcd, k = i / kNumXcd;
  return x * q + (x < r ? x : r) + k;
}
// XCD x takes chunks x, x + 8, x + 16, ... of K consecutive tiles, so all
// XCDs work inside one window of 8 K tiles at a time.  Used when 8 K divides
// the tile count; otherwise the contiguous-run order.
__device__ __forceinline__ int64_t xcd_tile_chunked(int64_t i, int64_t nb, int64_t K) {
  if (K <= 1 || nb % (kNumXcd * K) != 0) return xcd_tile(i, nb);
  const int64_t x = i % kNumXcd, k = i / kNumXcd;
  return ((k / K) * kNumXcd + x) * K + (k % K);
}
__device__ __forceinline__ int64_t tile_order(int64_t i, int64_t nb, int mode) {
  if (mode == 0) return i;
  if (mode == 1) return xcd_tile(i, nb);
  return xcd_tile_chunked(i, nb, mode);
}

template <typename T, int N, int C, bool COL> struct LdsLayout {
  // ROW: one padded row of N per transform; one pad slot every 16 entries
  // keeps the first pass's stride-R writes off a single bank.
  static constexpr int NP = N + N / 16 + 1;
  static constexpr int size = COL ? N * C : NP * C;
  __device__ static __forceinline__ int at(int idx, int c) {
    if constexpr (COL) return idx * C + c;
    else return c * NP + idx + (idx >> 4);
  }
};

// One Stockham pass executed by one thread for its Q butterflies.
// SRC/DST: 0 = LDS, 1 = HBM.
template <typename T, int N, int R, int Ns, int TPT, class LoadF, class StoreF>
__device__ __forceinline__ void stockham_pass(int j, const V2<T>* __restrict__ stw, LoadF load, StoreF store) {
  constexpr int NB = N / R;                  // butterflies in this pass
  constexpr int Q = (NB + TPT - 1) / TPT;    // butterflies per thread
  constexpr bool EXACT = (NB % TPT) == 0;
  V2<T> v[Q][R];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int jj = j + q * TPT;
    if (EXACT || jj < NB) {
#pragma unroll
      for (int m = 0; m < R; ++m) v[q][m] = load(jj + m * NB);
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int jj = j + q * TPT;
    if (EXACT || jj < NB) {
      if constexpr (Ns > 1) {
        const int kk = jj % Ns;
#pragma unroll
        for (int m = 1; m < R; ++m) v[q][m] = cmul<T>(v[q][m], stw[m * Ns + kk - 1]);
      }
      bfly<T, R>(v[q]);
    }
  }
  store.before();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int jj = j + q * TPT;
    if (EXACT || jj < NB) {
      const int kk = jj % Ns;
      const int ob = (jj / Ns) * Ns * R + kk;
#pragma unroll
      for (int m = 0; m < R; ++m) store(ob + m * Ns, v[q][m]);
    }
  }
  store.after();
}

// HBM side of one transform.  SEG = false is the plain strided row/column;
// SEG = true adds the segment maps of the fused exchange (a separate kernel
// instance, so the plain kernels carry none of its addressing).
// Segment lookup tables in LDS, built once per workgroup by the segmented
// C2C kernels that have room for them: entry e of the split axis holds
// (p << 12) | (e - s[p]) (N <= 4096), and each part its segment address and
// block count.  An access is then two LDS reads and one multiply-add, with
// no per-access scalar search between a lane's loads.
struct SegPart {
  void* seg;
  int n;
  int pad;
};
template <int N> struct SegTabs {
  static_assert(N <= 4096, "table offsets are 12 bits");
  static constexpr int part_bytes = kMaxSegPeers * (int)sizeof(SegPart);
  static constexpr int bytes = 2 * part_bytes + 2 * ((N * 2 + 15) / 16 * 16);
  __device__ static __forceinline__ SegPart* parts(unsigned char* m, int side) {
    return reinterpret_cast<SegPart*>(m + side * part_bytes);
  }
  __device__ static __forceinline__ uint16_t* ent(unsigned char* m, int side) {
    return reinterpret_cast<uint16_t*>(m + 2 * part_bytes + side * ((N * 2 + 15) / 16 * 16));
  }
  // all threads of the workgroup take part; a barrier must follow
  __device__ static __forceinline__ void fill(unsigned char* m, int side, KSegMap* map) {
    SegPart* pp = parts(m, side);
    for (int q = 0; q < map->nparts; ++q)
      if ((int)threadIdx.x == q) {
        SegPart sp;
        sp.seg = map->seg[q];
        sp.n = map->n[q];
        sp.pad = 0;
        pp[q] = sp;
      }
    uint16_t* en = ent(m, side);
    for (int e = threadIdx.x; e < N; e += blockDim.x) {
      int p = 0, s0 = 0;
      for (int q = 1; q < map->nparts; ++q) {
        const int sq = map->s[q];
        if (e >= sq) {
          p = q;
          s0 = sq;
        }
      }
      en[e] = (uint16_t)((p << 12) | (e - s0));
    }
  }
};

// Part tables of the cursor form, one per map side: entry q holds segment
// q's address, block count and first element; entry nparts is a sentinel
// whose first element exceeds every index.  A lane whose accesses along the
// split axis come in increasing order keeps a cursor on its current part
// and moves it forward past a boundary (rare: blocks are long), so an
// access is one 32-bit add and one multiply-add, with no table read.
struct SegPartS {
  void* seg;
  int n;
  int s;
};
static_assert(sizeof(SegPartS) == 16, "part table entries are 16 bytes");
struct SegCursorTabs {
  static constexpr int side_bytes = (kMaxSegPeers + 1) * (int)sizeof(SegPartS);
  static constexpr int bytes = 2 * side_bytes;
  __device__ static __forceinline__ SegPartS* parts(unsigned char* m, int side) {
    return reinterpret_cast<SegPartS*>(m + side * side_bytes);
  }
  // all threads of the workgroup take part; a barrier must follow
  __device__ static __forceinline__ void fill(unsigned char* m, int side, KSegMap* map) {
    SegPartS* pp = parts(m, side);
    const int np = map->nparts;
    for (int q = 0; q < np; ++q)
      if ((int)threadIdx.x == q) {
        SegPartS sp;
        sp.seg = map->seg[q];
        sp.n = map->n[q];
        sp.s = map->s[q];
        pp[q] = sp;
      }
    if ((int)threadIdx.x == np) {
      SegPartS sp;
      sp.seg = nullptr;
      sp.n = 0;
      sp.s = 0x7fffffff;
      pp[np] = sp;
    }
  }
};

// NT: non-temporal HBM accesses (streaming, no cache allocation) for a
// pass whose data is not read again soon; measured 4% faster on the
// 1024^3 axis-0 column pattern when both loads and stores stream.
template <bool B> struct BoolC {
  static constexpr bool value = B;
};

template <typename T, bool SEG = false, bool TAB = false, bool NT = false> struct GlobalPtr {
  static constexpr bool kSeg = false;
  V2<T>* p;
  const V2<T>* pin;          // load side (== p in place)
  int64_t stride;
  bool swap_io;
  bool valid;
  T scale;
  bool use_scale;
  // The data is device (global) memory.  Its pointers reach a kernel either
  // as kernel arguments (the compiler then knows they are global) or from a
  // job table in memory (grouped and persistent kernels), where it cannot
  // tell and would issue generic (flat) accesses: those count against both
  // the vector-memory and the LDS wait counters, so every LDS access between
  // them forces a full wait.  Accesses go through global-address-space
  // two-element vectors instead.
  typedef T E2 __attribute__((ext_vector_type(2)));
  using GE2 = __attribute__((address_space(1))) E2;
  __device__ __forceinline__ V2<T> load(int e) const {
    const GE2* q = (const GE2*)(pin + (int64_t)e * stride);
    if constexpr (NT) {
      const E2 x = __builtin_nontemporal_load(q);
      return mk<T>(x.x, x.y);
    } else {
      const E2 x = *q;
      return mk<T>(x.x, x.y);
    }
  }
  __device__ __forceinline__ void store(int e, V2<T> v) const {
    GE2* q = (GE2*)(p + (int64_t)e * stride);
    E2 x;
    x.x = v.x;
    x.y = v.y;
    if constexpr (NT) {
      __builtin_nontemporal_store(x, q);
    } else {
      *q = x;
    }
  }
  // the segmented form's accessors with the map branch already taken
  template <bool M> __device__ __forceinline__ V2<T> load_as(int e) const { return load(e); }
  template <bool M> __device__ __forceinline__ void store_as(int e, V2<T> v) const { store(e, v); }
};

template <typename T, bool TAB, bool NT> struct GlobalPtr<T, true, TAB, NT> {
  static constexpr bool kSeg = true;
  V2<T>* p;
  const V2<T>* pin;
  int64_t stride;
  bool swap_io;
  bool valid;
  T scale;
  bool use_scale;
  KSegMap* in_map;           // non-null: loads come from a segment layout
  KSegMap* out_map;          // non-null: stores go to a segment layout
  uint32_t in_lead, out_lead;  // transform's place in each map
  uint32_t in_trail, out_trail;  // (< trailing < 2^32)
  int lane_j;                // this lane's position among its transform's threads
  int blk;                   // threads per transform
  // TAB: the maps' LDS tables (set by the kernel body after building them)
  const uint16_t* in_ent;
  const uint16_t* out_ent;
  const SegPart* in_parts;
  const SegPart* out_parts;
  uint32_t in_trailing, out_trailing;
  // first element of the wave's block (e - lane_j is the same on every lane)
  __device__ __forceinline__ int block_start(int e) const {
    return __builtin_amdgcn_readfirstlane(e - lane_j);
  }
  // segment addresses are device memory: accessed as global (not generic,
  // flat) two-element vectors, one load or store instruction per element
  typedef T E2 __attribute__((ext_vector_type(2)));
  using GV = __attribute__((address_space(1))) E2;
  // (trailing < 2^32 and trail < trailing: one 32 x 32 + 64-bit multiply-add)
  __device__ __forceinline__ static GV* tab_addr(const uint16_t* ent, const SegPart* parts,
                                                 uint32_t lead, uint32_t trail,
                                                 uint32_t trailing, int e) {
    const uint32_t t = ent[e];
    const SegPart sp = parts[t >> 12];
    const uint32_t row = lead * (uint32_t)sp.n + (t & 0xFFFu);
    return (GV*)sp.seg + ((uint64_t)row * trailing + trail);
  }
  // M: the map's side of the branch (true: the segment layout)
  template <bool M> __device__ __forceinline__ V2<T> load_as(int e) const {
    if constexpr (M) {
      if constexpr (TAB) {
        const E2 x = *tab_addr(in_ent, in_parts, in_lead, in_trail, in_trailing, e);
        return mk<T>(x.x, x.y);
      } else {
        return *seg_addr<T>(in_map, in_lead, in_trail, e, block_start(e), blk);
      }
    } else {
      return pin[(int64_t)e * stride];
    }
  }
  template <bool M> __device__ __forceinline__ void store_as(int e, V2<T> v) const {
    if constexpr (M) {
      if constexpr (TAB) {
        E2 x;
        x.x = v.x;
        x.y = v.y;
        *tab_addr(out_ent, out_parts, out_lead, out_trail, out_trailing, e) = x;
      } else {
        *seg_addr<T>(out_map, out_lead, out_trail, e, block_start(e), blk) = v;
      }
    } else {
      p[(int64_t)e * stride] = v;
    }
  }
  // cursor form (the kernel body sets the tables after filling them)
  const SegPartS* in_pp;
  const SegPartS* out_pp;
  struct Cursor {
    GV* base;    // the current part's segment
    int rowoff;  // lead * n - s of the current part: row of element e is rowoff + e
    int snext;   // first element of the next part
    int p;
  };
  __device__ __forceinline__ static void cur_set(Cursor& c, const SegPartS* pp, uint32_t lead) {
    const SegPartS sp = pp[c.p];
    c.base = (GV*)sp.seg;
    c.rowoff = (int)(lead * (uint32_t)sp.n) - sp.s;
    c.snext = pp[c.p + 1].s;
  }
  // a cursor on the part holding e0, the smallest element the lane accesses
  __device__ __forceinline__ static Cursor cur_begin(const SegPartS* pp, uint32_t lead, int e0) {
    Cursor c;
    c.p = 0;
    while (e0 >= pp[c.p + 1].s) ++c.p;
    cur_set(c, pp, lead);
    return c;
  }
  // address of element e >= every element this cursor has addressed before
  __device__ __forceinline__ static GV* cur_addr(Cursor& c, const SegPartS* pp, uint32_t lead,
                                                 uint32_t trail, uint32_t trailing, int e) {
    if (e >= c.snext) {
      do {
        ++c.p;
      } while (e >= pp[c.p + 1].s);
      cur_set(c, pp, lead);
    }
    return c.base + ((uint64_t)(uint32_t)(c.rowoff + e) * trailing + trail);
  }
  __device__ __forceinline__ Cursor in_cursor(int e0) const {
    return cur_begin(in_pp, in_lead, e0);
  }
  __device__ __forceinline__ Cursor out_cursor(int e0) const {
    return cur_begin(out_pp, out_lead, e0);
  }
  __device__ __forceinline__ V2<T> load_cur(Cursor& c, int e) const {
    const E2 x = *cur_addr(c, in_pp, in_lead, in_trail, in_trailing, e);
    return mk<T>(x.x, x.y);
  }
  __device__ __forceinline__ void store_cur(Cursor& c, int e, V2<T> v) const {
    E2 x;
    x.x = v.x;
    x.y = v.y;
    *cur_addr(c, out_pp, out_lead, out_trail, out_trailing, e) = x;
  }
  __device__ __forceinline__ V2<T> load(int e) const {
    return in_map ? load_as<true>(e) : load_as<false>(e);
  }
  __device__ __forceinline__ void store(int e, V2<T> v) const {
    if (out_map) store_as<true>(e, v);
    else store_as<false>(e, v);
  }
};

// f(BoolC<M>) with M the side of g's load (store) map branch, taken once for
// a whole loop of accesses rather than per access: the plain and segmented
// loops are then straight-line code each.
template <class G, class F> __device__ __forceinline__ void with_in_map(const G& g, F&& f) {
  if constexpr (G::kSeg) {
    if (g.in_map) f(BoolC<true>());
    else f(BoolC<false>());
  } else {
    f(BoolC<false>());
  }
}
template <class G, class F> __device__ __forceinline__ void with_out_map(const G& g, F&& f) {
  if constexpr (G::kSeg) {
    if (g.out_map) f(BoolC<true>());
    else f(BoolC<false>());
  } else {
    f(BoolC<false>());
  }
}

// Build the LDS tables of the maps a segmented launch uses and point g at
// them.  Every thread of the workgroup calls this (it ends in a barrier).
template <typename T, int N>
__device__ __forceinline__ void seg_tables(GlobalPtr<T, true, true, false>& g, unsigned char* mem) {
  using ST = SegTabs<N>;
  if (g.in_map) ST::fill(mem, 0, g.in_map);
  if (g.out_map) ST::fill(mem, 1, g.out_map);
  __syncthreads();
  g.in_ent = ST::ent(mem, 0);
  g.out_ent = ST::ent(mem, 1);
  g.in_parts = ST::parts(mem, 0);
  g.out_parts = ST::parts(mem, 1);
  g.in_trailing = g.in_map ? (uint32_t)g.in_map->trailing : 0u;
  g.out_trailing = g.out_map ? (uint32_t)g.out_map->trailing : 0u;
}

// The cursor form's part tables (every thread calls this; it ends in a
// barrier).
template <typename T, bool TAB, bool NT>
__device__ __forceinline__ void seg_cursor_tables(GlobalPtr<T, true, TAB, NT>& g,
                                                  unsigned char* mem) {
  if (g.in_map) SegCursorTabs::fill(mem, 0, g.in_map);
  if (g.out_map) SegCursorTabs::fill(mem, 1, g.out_map);
  __syncthreads();
  g.in_pp = SegCursorTabs::parts(mem, 0);
  g.out_pp = SegCursorTabs::parts(mem, 1);
  g.in_trailing = g.in_map ? (uint32_t)g.in_map->trailing : 0u;
  g.out_trailing = g.out_map ? (uint32_t)g.out_map->trailing : 0u;
}

// LDS budget of one workgroup; a segmented kernel builds the tables when its
// own block leaves room for them.
constexpr int kLdsMax = 163840;
template <bool SEG, int N, int BYTES> constexpr bool seg_tab_fits() {
  if constexpr (!SEG || N > 4096) return false;
  else return BYTES + SegTabs<N>::bytes <= kLdsMax;
}
template <bool SEG, int BYTES> constexpr bool seg_cursor_fits() {
  return SEG && BYTES + SegCursorTabs::bytes <= kLdsMax;
}

__device__ __forceinline__ int64_t xform_offset(int64_t b, const StockhamArgs& a) {
  const int64_t o = b / a.inner;
  const int64_t i = b - o * a.inner;
  return o * a.dist_outer + i * a.dist;
}
__device__ __forceinline__ int64_t xform_offset(int64_t b, const StockhamArgs& a, int64_t dist,
                                                int64_t dist_outer) {
  const int64_t o = b / a.inner;
  const int64_t i = b - o * a.inner;
  return o * dist_outer + i * dist;
}

template <typename T, bool SEG, bool TAB, bool NT>
__device__ __forceinline__ void init_global(GlobalPtr<T, SEG, TAB, NT>& g, const StockhamArgs& a,
                                            int64_t b) {
  g.valid = b < a.total;
  if constexpr (!SEG) {
    const int64_t off = g.valid ? xform_offset(b, a) : 0;
    g.p = a.dst ? static_cast<V2<T>*>(a.dst) +
                      (g.valid ? xform_offset(b, a, a.dst_dist, a.dst_dist_outer) : 0)
                : static_cast<V2<T>*>(a.data) + off;
    g.pin = a.src ? static_cast<const V2<T>*>(a.src) +
                        (g.valid ? xform_offset(b, a, a.src_dist, a.src_dist_outer) : 0)
                  : static_cast<const V2<T>*>(a.data) + off;
  } else {
    const int64_t off = g.valid && !(a.in_seg && a.out_seg) ? xform_offset(b, a) : 0;
    const bool v = g.valid && !a.in_seg;
    g.p = static_cast<V2<T>*>(a.data) + off;
    g.pin = a.src ? static_cast<const V2<T>*>(a.src) +
                        (v ? xform_offset(b, a, a.src_dist, a.src_dist_outer) : 0)
                  : static_cast<const V2<T>*>(a.data) + off;
    // Only single-plan kernels run segmented, and their one argument (the
    // StockhamArgs) sits at offset 0 of the kernel-argument segment; taking
    // the address of the parameter itself would copy it to scratch.
    const auto* kern =
        (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
    g.in_map = a.in_seg ? reinterpret_cast<KSegMap*>(kern + offsetof(StockhamArgs, in_map)) : nullptr;
    g.out_map = a.out_seg ? reinterpret_cast<KSegMap*>(kern + offsetof(StockhamArgs, out_map)) : nullptr;
    g.in_lead = g.in_trail = g.out_lead = g.out_trail = 0;
    if (a.in_seg) {
      const int64_t l = b / a.in_map.trailing;
      g.in_lead = (uint32_t)l;
      g.in_trail = b - l * a.in_map.trailing;
    }
    if (a.out_seg) {
      const int64_t l = b / a.out_map.trailing;
      g.out_lead = (uint32_t)l;
      g.out_trail = b - l * a.out_map.trailing;
    }
  }
}

struct NoHook {
  __device__ void operator()() const {}
};

// Run passes PI.. of the radix list.  The first pass may read HBM (FIRST_G),
// the last may write HBM (LAST_G); everything else goes through LDS.
template <typename T, int N, int TPT, int C, bool COL, int TWS, bool FIRST_G, bool LAST_G,
          int PI, int Ns, int... Rs>
struct Passes;

template <typename T, int N, int TPT, int C, bool COL, int TWS, bool FIRST_G, bool LAST_G, int PI,
          int Ns>
struct Passes<T, N, TPT, C, COL, TWS, FIRST_G, LAST_G, PI, Ns> {
  template <class G, class H = NoHook>
  __device__ static __forceinline__ void run(int, int, V2<T>*, const V2<T>*, const G&,
                                             const H& = H()) {}
};

template <typename T, int N, int TPT, int C, bool COL, int TWS, bool FIRST_G, bool LAST_G, int PI,
          int Ns, int R, int... Rest>
struct Passes<T, N, TPT, C, COL, TWS, FIRST_G, LAST_G, PI, Ns, R, Rest...> {
  using L = LdsLayout<T, N, C, COL>;
  // hook() runs once in the first pass, after its loads and butterflies and
  // before the barrier that precedes its LDS stores (a kernel commits its
  // LDS twiddle copy there, which the second pass is the first to read)
  template <class G, class H = NoHook>
  __device__ static __forceinline__ void run(int j, int c, V2<T>* lds, const V2<T>* tw, const G& g,
                                             const H& hook = H()) {
    constexpr bool IS_FIRST = (PI == 0);
    constexpr bool IS_LAST = (sizeof...(Rest) == 0);
    auto load = [&](int e) -> V2<T> {
      if constexpr (IS_FIRST && FIRST_G) {
        V2<T> x = mk<T>(T(0), T(0));
        if (g.valid) x = g.load(e);
        return g.swap_io ? swp<T>(x) : x;
      } else {
        return lds[L::at(e, c)];
      }
    };
    struct StoreLds {
      V2<T>* lds; int c; const H* hook;
      __device__ void before() const {
        if constexpr (IS_FIRST) (*hook)();
        __syncthreads();
      }
      __device__ void after() const { __syncthreads(); }
      __device__ void operator()(int e, V2<T> v) const { lds[L::at(e, c)] = v; }
    };
    struct StoreG {
      const G* g;
      // A single pass that both reads and writes HBM in place must see every
      // lane's loads done before any lane stores (only matters when TPT > 1).
      __device__ void before() const {
        if constexpr (IS_FIRST && FIRST_G && TPT > 1) __syncthreads();
      }
      __device__ void after() const {}
      __device__ void operator()(int e, V2<T> v) const {
        if (g->use_scale) v = scl<T>(v, g->scale);
        if (g->swap_io) v = swp<T>(v);
        if (g->valid) g->store(e, v);
      }
    };
    if constexpr (IS_LAST && LAST_G) {
      stockham_pass<T, N, R, Ns, TPT>(j, tw, load, StoreG{&g});
    } else {
      stockham_pass<T, N, R, Ns, TPT>(j, tw, load, StoreLds{lds, c, &hook});
    }
    Passes<T, N, TPT, C, COL, TWS, FIRST_G, LAST_G, PI + 1, Ns * R, Rest...>::run(j, c, lds, tw, g);
  }
};

// Copy of the per-pass twiddle table (N - 1 entries) into LDS by a
// workgroup of NTHR threads, in two steps: fetch() issues this thread's
// loads into registers, commit() stores them to LDS and ends in a barrier.
// A kernel issues its first-pass HBM loads in between, so the twiddle loads
// and the data loads are in flight together instead of one after the other
// with a barrier between.  Only plans with more than one pass multiply by
// twiddles.
template <typename T, int N, int NPASS, int NTHR> struct TwiddleStage {
  static constexpr int IT = NPASS > 1 ? (N - 1 + NTHR - 1) / NTHR : 0;
  V2<T> t[IT > 0 ? IT : 1];
  __device__ __forceinline__ void fetch(const void* ptw) {
    if constexpr (IT > 0) {
      const V2<T>* src = static_cast<const V2<T>*>(ptw);
#pragma unroll
      for (int i = 0; i < IT; ++i) {
        const int e = lane_tid() + i * NTHR;
        if (e < N - 1) t[i] = gload<T>(src, e);
      }
    }
  }
  // the LDS stores alone (the caller's next barrier publishes them)
  __device__ __forceinline__ void put(V2<T>* stw) const {
    if constexpr (IT > 0) {
#pragma unroll
      for (int i = 0; i < IT; ++i) {
        const int e = lane_tid() + i * NTHR;
        if (e < N - 1) stw[e] = t[i];
      }
    }
  }
  __device__ __forceinline__ void commit(V2<T>* stw) {
    if constexpr (IT > 0) {
      put(stw);
      __syncthreads();
    }
  }
};

// Both steps at once (kernels whose first pass is not issued separately).
template <typename T, int N, int NPASS, int NTHR>
__device__ __forceinline__ const V2<T>* stage_twiddles(V2<T>* stw, const void* ptw) {
  TwiddleStage<T, N, NPASS, NTHR> ts;
  ts.fetch(ptw);
  ts.commit(stw);
  return stw;
}

// One LDS block per kernel body, the per-pass twiddle table (TW entries of
// V2<T>) first and the data tile (DATA entries of D) after it.  A single
// block fixes the layout: separately declared arrays are placed by the
// compiler, and with the small twiddle table behind a 128 KB tile its
// addresses no longer fit the 16-bit offset of ds_read, which costs one
// VGPR per twiddle address and spilled the 1024-thread kernels.
template <typename T, typename D, int TW, int DATA> struct LdsBlock {
  static constexpr int tw_bytes = (TW * (int)sizeof(V2<T>) + 15) / 16 * 16;
  static constexpr int bytes = tw_bytes + DATA * (int)sizeof(D);
  __device__ static __forceinline__ V2<T>* tw(unsigned char* b) {
    return reinterpret_cast<V2<T>*>(b);
  }
  __device__ static __forceinline__ D* data(unsigned char* b) {
    return reinterpret_cast<D*>(b + tw_bytes);
  }
};

template <int TPT, int C, bool COL>
__device__ __forceinline__ void thread_coords(int& j, int& c) {
  const int tid = lane_tid();
  if constexpr (COL) { c = tid % C; j = tid / C; }
  else { j = tid % TPT; c = tid / TPT; }
}

// ---------------------------------------------------------------------------
// C2C
// ---------------------------------------------------------------------------
// Kernel bodies take the workgroup's tile id and the tile count of its
// launch (or of its job in a grouped launch), so the single-plan kernels and
// the grouped multi-transform kernels share them.
// LDS bytes of one c2c_body workgroup (tile, twiddle table, segment tables).
template <typename T, int TPT, int C, bool COL, bool SEG, int... Rs> struct C2CLds {
  static constexpr int N = (Rs * ... * 1);
  using B = LdsBlock<T, V2<T>, (sizeof...(Rs) > 1 ? N : 1), LdsLayout<T, N, C, COL>::size>;
  static constexpr bool TAB = seg_tab_fits<SEG, N, B::bytes>();
  static constexpr int bytes = B::bytes + (TAB ? SegTabs<(TAB ? N : 1)>::bytes : 0);
};

// The body on an LDS block of C2CLds<...>::bytes the caller provides (a
// kernel's own static block, or the shared block of the persistent kernel).
template <typename T, int TPT, int C, bool COL, bool SEG, int... Rs>
__device__ __forceinline__ void c2c_body_in(const StockhamArgs& a, int64_t tile, int64_t ntiles,
                                            unsigned char* smem) {
  constexpr int N = (Rs * ... * 1);
  using L = LdsLayout<T, N, C, COL>;
  using B = LdsBlock<T, V2<T>, (sizeof...(Rs) > 1 ? N : 1), L::size>;
  constexpr bool TAB = seg_tab_fits<SEG, N, B::bytes>();
  V2<T>* lds = B::data(smem);
  V2<T>* stw = B::tw(smem);
  // the LDS twiddle copy: fetched now, stored by the first pass (which uses
  // none) before its barrier, so its loads overlap the first-pass loads
  TwiddleStage<T, N, sizeof...(Rs), TPT * C> ts;
  ts.fetch(a.ptw);
  const V2<T>* tw = stw;
  int j, c;
  thread_coords<TPT, C, COL>(j, c);
  if (COL) tile = tile_order(tile, ntiles, a.xcd_remap);
  const int64_t b = tile * C + c;
  GlobalPtr<T, SEG, TAB> g;
  init_global(g, a, b);
  if constexpr (SEG) {
    g.lane_j = j;
    g.blk = TPT;
  }
  if constexpr (TAB) seg_tables<T, N>(g, smem + B::bytes);
  g.stride = a.stride;
  g.swap_io = a.backward != 0;
  g.scale = T(a.scale);
  g.use_scale = a.use_scale != 0;
  Passes<T, N, TPT, C, COL, 1, true, true, 0, 1, Rs...>::run(j, c, lds, tw, g,
                                                             [&] { ts.put(stw); });
}

template <typename T, int TPT, int C, bool COL, bool SEG, int... Rs>
__device__ __forceinline__ void c2c_body(const StockhamArgs& a, int64_t tile, int64_t ntiles) {
  __shared__ __align__(16) unsigned char smem[C2CLds<T, TPT, C, COL, SEG, Rs...>::bytes];
  c2c_body_in<T, TPT, C, COL, SEG, Rs...>(a, tile, ntiles, smem);
}

template <typename T, int TPT, int C, bool COL, bool SEG, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_kernel(StockhamArgs a) {
  c2c_body<T, TPT, C, COL, SEG, Rs...>(a, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// R2C (forward): row of 2N reals = N complex z[n] = x[2n] + i x[2n+1];
// Z = FFT_N(z); X[k] = (Z[k] + conj Z[N-k])/2 - (i/2) W_2N^k (Z[k] - conj Z[N-k]),
// k = 0..N, written as N+1 complex values over the same padded row.
// Row layout only (stride 1).  a.dist is in complex units (= padded dist / 2).
// ---------------------------------------------------------------------------
// LDS bytes of one r2c_body / c2r_body workgroup.
template <typename T, int TPT, int C, int... Rs> struct RealLds {
  static constexpr int N = (Rs * ... * 1);
  static constexpr int bytes =
      LdsBlock<T, V2<T>, (sizeof...(Rs) > 1 ? N : 1), LdsLayout<T, N, C, false>::size>::bytes;
};

template <typename T, int TPT, int C, bool SEG, int... Rs>
__device__ __forceinline__ void r2c_body_in(const StockhamArgs& a, int64_t tile,
                                            unsigned char* smem) {
  constexpr int N = (Rs * ... * 1);
  using L = LdsLayout<T, N, C, false>;
  using B = LdsBlock<T, V2<T>, (sizeof...(Rs) > 1 ? N : 1), L::size>;
  V2<T>* lds = B::data(smem);
  V2<T>* stw = B::tw(smem);
  TwiddleStage<T, N, sizeof...(Rs), TPT * C> ts;  // stored by the first pass
  ts.fetch(a.ptw);
  const V2<T>* ptw = stw;
  int j, c;
  thread_coords<TPT, C, false>(j, c);
  const int64_t b = tile * C + c;
  GlobalPtr<T, SEG> g;
  init_global(g, a, b);
  if constexpr (SEG) {
    g.lane_j = j;
    g.blk = TPT;
  }
  g.stride = 1;
  g.swap_io = false;
  g.scale = T(a.scale);
  g.use_scale = a.use_scale != 0;
  const V2<T>* tw = static_cast<const V2<T>*>(a.tw);  // W_{2N}^i, i < 2N
  Passes<T, N, TPT, C, false, 2, true, false, 0, 1, Rs...>::run(j, c, lds, ptw, g,
                                                                [&] { ts.put(stw); });
  // Segmented output: the post-twiddle writes X back into the LDS row (each
  // thread owns its pairs (k, N-k), so in place is safe; X[N] takes the
  // row's spare slot N), then one uniform loop over the peers' blocks stores
  // each block as a run, with no per-element peer search.
  bool to_lds = false;
  if constexpr (SEG) to_lds = g.out_map != nullptr;
  // post-twiddle: pairs (k, N-k) for k = j + i TPT <= N/2; the twiddle
  // loads of all of them are issued before the first is used
  const T half = T(0.5);
  constexpr int KI = (N / 2 + TPT) / TPT;
  V2<T> twk[KI], twm[KI];
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const int k = j + i * TPT;
    if (k <= N / 2) {
      twk[i] = gload<T>(tw, k);
      twm[i] = gload<T>(tw, N - k);
    }
  }
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const int k = j + i * TPT;
    if (k > N / 2) continue;
    const int km = (N - k) % N;
    V2<T> A = lds[L::at(k, c)];
    V2<T> Bm = lds[L::at(km, c)];
    // X[k]
    {
      V2<T> B = conj<T>(Bm);
      V2<T> s = add<T>(A, B), d = sub<T>(A, B);
      V2<T> wd = cmul<T>(twk[i], d);
      V2<T> x = mk<T>(half * (s.x + wd.y), half * (s.y - wd.x));
      if (g.use_scale) x = scl<T>(x, g.scale);
      if (to_lds) lds[L::at(k, c)] = x;
      else if (g.valid) g.store(k, x);
    }
    // X[N-k] unless k is its own mirror (2k == N); k = 0 gives X[N]
    if (2 * k != N) {
      const int kk = N - k;
      V2<T> A2 = Bm, B2 = conj<T>(A);
      V2<T> s = add<T>(A2, B2), d = sub<T>(A2, B2);
      V2<T> wd = cmul<T>(twm[i], d);
      V2<T> x = mk<T>(half * (s.x + wd.y), half * (s.y - wd.x));
      if (g.use_scale) x = scl<T>(x, g.scale);
      if (to_lds) lds[L::at(kk, c)] = x;
      else if (g.valid) g.store(kk, x);
    }
  }
  if constexpr (SEG) {
    if (to_lds) {
      __syncthreads();
      KSegMap* m = g.out_map;
      const int64_t T_ = m->trailing;
      for (int q = 0; q < m->nparts; ++q) {
        const int s0 = m->s[q], n0 = m->n[q];
        V2<T>* base = static_cast<V2<T>*>(m->seg[q]) + g.out_trail;
        const uint32_t row0 = g.out_lead * (uint32_t)n0;
        for (int e = s0 + j; e < s0 + n0; e += TPT)
          if (g.valid) base[(int64_t)(row0 + (uint32_t)(e - s0)) * T_] = lds[L::at(e, c)];
      }
    }
  }
}

template <typename T, int TPT, int C, bool SEG, int... Rs>
__device__ __forceinline__ void r2c_body(const StockhamArgs& a, int64_t tile) {
  __shared__ __align__(16) unsigned char smem[RealLds<T, TPT, C, Rs...>::bytes];
  r2c_body_in<T, TPT, C, SEG, Rs...>(a, tile, smem);
}

template <typename T, int TPT, int C, bool SEG, int... Rs>
__global__ void __launch_bounds__(TPT * C) r2c_kernel(StockhamArgs a) {
  r2c_body<T, TPT, C, SEG, Rs...>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// C2R (backward): X[0..N] -> Z[k] = (X[k] + conj X[N-k]) + i W_2N^-k (X[k] - conj X[N-k]),
// Im X[0] and Im X[N] ignored (Hermitian input); z = IFFT_N(Z) (unnormalised)
// gives z[n] = x[2n] + i x[2n+1], stored as the N complex slots of the row.
// ---------------------------------------------------------------------------
template <typename T, int TPT, int C, bool SEG, int... Rs>
__device__ __forceinline__ void c2r_body_in(const StockhamArgs& a, int64_t tile,
                                            unsigned char* smem) {
  constexpr int N = (Rs * ... * 1);
  using L = LdsLayout<T, N, C, false>;
  using B = LdsBlock<T, V2<T>, (sizeof...(Rs) > 1 ? N : 1), L::size>;
  V2<T>* lds = B::data(smem);
  V2<T>* stw = B::tw(smem);
  TwiddleStage<T, N, sizeof...(Rs), TPT * C> ts;  // stored with the pre-twiddled row
  ts.fetch(a.ptw);
  const V2<T>* ptw = stw;
  int j, c;
  thread_coords<TPT, C, false>(j, c);
  const int64_t b = tile * C + c;
  GlobalPtr<T, SEG> g;
  init_global(g, a, b);
  if constexpr (SEG) {
    g.lane_j = j;
    g.blk = TPT;
  }
  g.stride = 1;
  g.swap_io = true;  // inverse FFT via re/im swap; pre-twiddled data is stored swapped
  g.scale = T(a.scale);
  g.use_scale = a.use_scale != 0;
  const V2<T>* tw = static_cast<const V2<T>*>(a.tw);
  // every load of the pairs (k, N - k), k = j + i TPT <= N/2, is issued
  // before the first is used, so they are in flight together
  constexpr int KI = (N / 2 + TPT) / TPT;
  V2<T> xk[KI], xm[KI];
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const int k = j + i * TPT;
    xk[i] = mk<T>(T(0), T(0));
    xm[i] = xk[i];
    if (g.valid && k <= N / 2) {
      xk[i] = g.load(k);
      xm[i] = g.load(N - k);
    }
  }
#pragma unroll
  for (int i = 0; i < KI; ++i) {
    const int k = j + i * TPT;
    if (k > N / 2) continue;
    const int km = N - k;
    V2<T> Xk = xk[i], Xm = xm[i];
    if (k == 0) { Xk.y = T(0); Xm.y = T(0); }
    // Z[k]
    {
      V2<T> B = conj<T>(Xm);
      V2<T> s = add<T>(Xk, B), d = sub<T>(Xk, B);
      V2<T> wd = cmulc<T>(d, gload<T>(tw, k));  // W^-k d
      V2<T> z = mk<T>(s.x - wd.y, s.y + wd.x);  // s + i wd
      lds[L::at(k, c)] = swp<T>(z);
    }
    if (k != 0 && 2 * k != N) {
      V2<T> B = conj<T>(Xk);
      V2<T> s = add<T>(Xm, B), d = sub<T>(Xm, B);
      V2<T> wd = cmulc<T>(d, gload<T>(tw, km));
      V2<T> z = mk<T>(s.x - wd.y, s.y + wd.x);
      lds[L::at(km, c)] = swp<T>(z);
    }
  }
  ts.put(stw);
  __syncthreads();
  Passes<T, N, TPT, C, false, 2, false, true, 0, 1, Rs...>::run(j, c, lds, ptw, g);
}

template <typename T, int TPT, int C, bool SEG, int... Rs>
__device__ __forceinline__ void c2r_body(const StockhamArgs& a, int64_t tile) {
  __shared__ __align__(16) unsigned char smem[RealLds<T, TPT, C, Rs...>::bytes];
  c2r_body_in<T, TPT, C, SEG, Rs...>(a, tile, smem);
}

template <typename T, int TPT, int C, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2r_kernel(StockhamArgs a) {
  c2r_body<T, TPT, C, false, Rs...>(a, blockIdx.x);
}

}  // namespace pfx

namespace pfx {

// ---------------------------------------------------------------------------
// Wide C2C: each thread holds E = N / TPT elements of its transform in
// registers across all passes, and the LDS exchange between passes moves one
// component at a time (re, then im), so a tile needs N * C scalars of LDS
// rather than N * C complex values.  Two layouts:
//  - batch-contiguous (ROWL = false): the C transforms of a workgroup are
//    adjacent lanes, so every element index is one C * 8-byte run in HBM
//    (256 B for C = 32 complex<float>); per-pass twiddles staged in LDS;
//  - row (ROWL = true): the TPT threads of a transform are adjacent lanes,
//    so first-pass loads and last-pass stores are contiguous runs of the
//    row; twiddles are read from the (cached) global table, which keeps the
//    workgroup's LDS at N * C scalars and several workgroups per CU.
// Requires every radix to divide E (all passes are exact).
// ---------------------------------------------------------------------------
template <int N, int C, bool ROWL, bool GTW = ROWL> struct WideLayout {
  // twiddles from the global table (the row layout's, and batch-contiguous
  // tiles that leave the LDS to the data); otherwise staged in LDS
  static constexpr bool global_tw = GTW;
  // row: one pad slot every 16 entries keeps the stride-R writes of a pass
  // with Ns = 1 off a single bank
  static constexpr int NP = N + N / 16 + 1;
  static constexpr int size = ROWL ? NP * C : N * C;
  __device__ static __forceinline__ int at(int e, int c) {
    if constexpr (ROWL) return c * NP + e + (e >> 4);
    else return e * C + c;
  }
};

template <typename T, int N, int TPT, class WL, int E, int PI, int Ns, int... Rs>
struct WidePasses;

template <typename T, int N, int TPT, class WL, int E, int PI, int Ns>
struct WidePasses<T, N, TPT, WL, E, PI, Ns> {
  template <bool CUR = false, class G>
  __device__ static __forceinline__ void run(int, int, V2<T>*, T*, const V2<T>*, const G&) {}
};

template <typename T, int N, int TPT, class WL, int E, int PI, int Ns, int R, int... Rest>
struct WidePasses<T, N, TPT, WL, E, PI, Ns, R, Rest...> {
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  // CUR: the segmented stores of the last pass go through a cursor
  template <bool CUR = false, class G>
  __device__ static __forceinline__ void run(int j, int c, V2<T>* v, T* lds, const V2<T>* tw, const G& g) {
    constexpr int Q = E / R;
    // twiddle + butterflies on the values this thread holds
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int jj = j + q * TPT;
      if constexpr (Ns > 1) {
        const int kk = jj % Ns;
#pragma unroll
        for (int m = 1; m < R; ++m) {
          const int ti = m * Ns + kk - 1;
          V2<T> w;
          if constexpr (WL::global_tw) w = gload<T>(tw, ti);
          else w = tw[ti];
          v[q * R + m] = cmul<T>(v[q * R + m], w);
        }
      }
      bfly<T, R>(v + q * R);
    }
    if constexpr (sizeof...(Rest) == 0) {
      // last pass: outputs straight to HBM (segmented: the map branch is
      // taken once for the whole loop, which keeps the plain and segment
      // store loops straight-line code; the plain kernels keep their own)
      auto stores = [&](auto M) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int jj = j + q * TPT;
          const int ob = (jj / Ns) * Ns * R + jj % Ns;
#pragma unroll
          for (int m = 0; m < R; ++m) {
            V2<T> x = v[q * R + m];
            if (g.use_scale) x = scl<T>(x, g.scale);
            if (g.swap_io) x = swp<T>(x);
            if (g.valid) g.template store_as<decltype(M)::value>(ob + m * Ns, x);
          }
        }
      };
      if constexpr (CUR) {
        // last pass: Ns R = N, so element j + TPT (q + Q m) in (m, q) order
        // is increasing and one cursor walks the segments
        static_assert(Ns * R == N, "the last pass spans the transform");
        with_out_map(g, [&](auto M) __attribute__((always_inline)) {
          if constexpr (decltype(M)::value) {
            auto cur = g.out_cursor(j);
#pragma unroll
            for (int m = 0; m < R; ++m) {
#pragma unroll
              for (int q = 0; q < Q; ++q) {
                V2<T> x = v[q * R + m];
                if (g.use_scale) x = scl<T>(x, g.scale);
                if (g.swap_io) x = swp<T>(x);
                if (g.valid) g.store_cur(cur, j + q * TPT + m * Ns, x);
              }
            }
          } else {
            stores(BoolC<false>());
          }
        });
      } else if constexpr (G::kSeg) {
        with_out_map(g, stores);
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int jj = j + q * TPT;
          const int ob = (jj / Ns) * Ns * R + jj % Ns;
#pragma unroll
          for (int m = 0; m < R; ++m) {
            V2<T> x = v[q * R + m];
            if (g.use_scale) x = scl<T>(x, g.scale);
            if (g.swap_io) x = swp<T>(x);
            if (g.valid) g.store(ob + m * Ns, x);
          }
        }
      }
    } else {
      constexpr int R2 = first_of<Rest...>();
      constexpr int Q2 = E / R2;
      constexpr int NB2 = N / R2;
      // One component at a time: every lane writes its v[].x (or .y) to LDS,
      // and after the barrier that component of v is dead, so the next
      // pass's values are read straight back into it (no second array).
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int jj = j + q * TPT;
          const int ob = (jj / Ns) * Ns * R + jj % Ns;
#pragma unroll
          for (int m = 0; m < R; ++m)
            lds[WL::at(ob + m * Ns, c)] = comp ? v[q * R + m].y : v[q * R + m].x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q2; ++q) {
          const int jj = j + q * TPT;
#pragma unroll
          for (int m = 0; m < R2; ++m) {
            const T s = lds[WL::at(jj + m * NB2, c)];
            if (comp) v[q * R2 + m].y = s;
            else v[q * R2 + m].x = s;
          }
        }
        __syncthreads();
      }
      WidePasses<T, N, TPT, WL, E, PI + 1, Ns * R, Rest...>::template run<CUR>(j, c, v, lds, tw, g);
    }
  }
};

// LDS bytes of one wide-body workgroup.
template <typename T, int TPT, int C, bool ROWL, bool SEG, bool GTW, int... Rs> struct WideLds {
  static constexpr int N = (Rs * ... * 1);
  using B = LdsBlock<T, T, (!GTW && sizeof...(Rs) > 1 ? N : 0), WideLayout<N, C, ROWL, GTW>::size>;
  // segmented: cursors on a part table where it fits, else the kernel
  // arguments' map searched per wave
  static constexpr bool CUR = seg_cursor_fits<SEG, B::bytes>();
  static constexpr int bytes = B::bytes + (CUR ? SegCursorTabs::bytes : 0);
};

// GTW: twiddles read from the global table (always for the row layout; a
// batch-contiguous tile can use it to leave the whole LDS block to its data).
template <typename T, int TPT, int C, bool ROWL, bool SEG, bool NT, bool GTW, int... Rs>
__device__ __forceinline__ void c2c_wide_body_in(const StockhamArgs& a, int64_t tile,
                                                 int64_t ntiles, unsigned char* smem) {
  static_assert(GTW || !ROWL, "the row layout reads twiddles from the global table");
  constexpr int N = (Rs * ... * 1);
  constexpr int E = N / TPT;
  constexpr int R0 = first_of<Rs...>();
  constexpr int Q0 = E / R0;
  constexpr int NB0 = N / R0;
  using WL = WideLayout<N, C, ROWL, GTW>;
  using B = LdsBlock<T, T, (!GTW && sizeof...(Rs) > 1 ? N : 0), WL::size>;
  constexpr bool CUR = seg_cursor_fits<SEG, B::bytes>();
  T* lds = B::data(smem);
  const V2<T>* tw;
  int c, j;
  // Column layout: the LDS twiddle copy is fetched first and committed after
  // the first-pass loads are issued (the first pass uses no twiddles), except
  // in 1024-thread tiles of 32 elements per thread: at their 128-VGPR limit
  // the fetched entries spill, and those measured faster committing first.
  constexpr bool EARLY = !(TPT * C == 1024 && E >= 32);
  TwiddleStage<T, N, (GTW ? 1 : (int)sizeof...(Rs)), TPT * C> tstage;
  if constexpr (GTW) {
    tw = static_cast<const V2<T>*>(a.ptw);
  } else {
    tstage.fetch(a.ptw);
    if constexpr (!EARLY) tstage.commit(B::tw(smem));
    tw = B::tw(smem);
  }
  if constexpr (ROWL) {
    j = lane_tid() % TPT;
    c = lane_tid() / TPT;
  } else {
    c = lane_tid() % C;
    j = lane_tid() / C;
    tile = tile_order(tile, ntiles, a.xcd_remap);
  }
  const int64_t b = tile * C + c;
  GlobalPtr<T, SEG, false, NT && !SEG> g;
  init_global(g, a, b);
  if constexpr (SEG) {
    g.lane_j = j;
    g.blk = TPT;
  }
  if constexpr (CUR) seg_cursor_tables(g, smem + B::bytes);
  g.stride = a.stride;
  g.swap_io = a.backward != 0;
  g.scale = T(a.scale);
  g.use_scale = a.use_scale != 0;
  V2<T> v[E];
  if constexpr (CUR) {
    // Element j + TPT (q + Q0 m) in (m, q) order: increasing, so one cursor
    // walks the segments (the map branch is taken once for the loop).
    with_in_map(g, [&](auto M) __attribute__((always_inline)) {
      if constexpr (decltype(M)::value) {
        auto cur = g.in_cursor(j);
#pragma unroll
        for (int m = 0; m < R0; ++m) {
#pragma unroll
          for (int q = 0; q < Q0; ++q) {
            V2<T> x = mk<T>(T(0), T(0));
            if (g.valid) x = g.load_cur(cur, j + q * TPT + m * NB0);
            v[q * R0 + m] = g.swap_io ? swp<T>(x) : x;
          }
        }
      } else {
#pragma unroll
        for (int q = 0; q < Q0; ++q) {
          const int jj = j + q * TPT;
#pragma unroll
          for (int m = 0; m < R0; ++m) {
            V2<T> x = mk<T>(T(0), T(0));
            if (g.valid) x = g.template load_as<false>(jj + m * NB0);
            v[q * R0 + m] = g.swap_io ? swp<T>(x) : x;
          }
        }
      }
    });
  } else if constexpr (SEG) {
    // the map branch taken once for the whole first-pass load loop
    with_in_map(g, [&](auto M) __attribute__((always_inline)) {
#pragma unroll
      for (int q = 0; q < Q0; ++q) {
        const int jj = j + q * TPT;
#pragma unroll
        for (int m = 0; m < R0; ++m) {
          V2<T> x = mk<T>(T(0), T(0));
          if (g.valid) x = g.template load_as<decltype(M)::value>(jj + m * NB0);
          v[q * R0 + m] = g.swap_io ? swp<T>(x) : x;
        }
      }
    });
  } else {
#pragma unroll
    for (int q = 0; q < Q0; ++q) {
      const int jj = j + q * TPT;
#pragma unroll
      for (int m = 0; m < R0; ++m) {
        V2<T> x = mk<T>(T(0), T(0));
        if (g.valid) x = g.load(jj + m * NB0);
        v[q * R0 + m] = g.swap_io ? swp<T>(x) : x;
      }
    }
  }
  if constexpr (!GTW && EARLY) tstage.commit(B::tw(smem));
  WidePasses<T, N, TPT, WL, E, 0, 1, Rs...>::template run<CUR>(j, c, v, lds, tw, g);
}

template <typename T, int TPT, int C, bool ROWL, bool SEG, bool NT, int... Rs>
__device__ __forceinline__ void c2c_wide_body_x(const StockhamArgs& a, int64_t tile, int64_t ntiles) {
  __shared__ __align__(16) unsigned char smem[WideLds<T, TPT, C, ROWL, SEG, ROWL, Rs...>::bytes];
  c2c_wide_body_in<T, TPT, C, ROWL, SEG, NT, ROWL, Rs...>(a, tile, ntiles, smem);
}

template <typename T, int TPT, int C, bool ROWL, bool SEG, int... Rs>
__device__ __forceinline__ void c2c_wide_body(const StockhamArgs& a, int64_t tile, int64_t ntiles) {
  c2c_wide_body_x<T, TPT, C, ROWL, SEG, false, Rs...>(a, tile, ntiles);
}

template <typename T, int TPT, int C, bool ROWL, bool SEG, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_wide_kernel(StockhamArgs a) {
  c2c_wide_body<T, TPT, C, ROWL, SEG, Rs...>(a, blockIdx.x, gridDim.x);
}

// The same kernel with non-temporal HBM loads and stores (plain form only).
template <typename T, int TPT, int C, bool ROWL, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_wide_nt_kernel(StockhamArgs a) {
  c2c_wide_body_x<T, TPT, C, ROWL, false, true, Rs...>(a, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Grouped launches: one launch runs the same kernel variant over many jobs
// (independent transforms, or stages of them, with different geometry).  Job
// j owns the workgroups [start[j], start[j+1]); every job's range begins at a
// multiple of 8 so the XCD-aware tile order holds inside each job.
// ---------------------------------------------------------------------------
struct GroupArgs {
  const StockhamArgs* jobs;
  const int64_t* start;  // njobs + 1 entries
  int njobs;
};

__device__ __forceinline__ int find_job(const GroupArgs& ga, int64_t blk) {
  int lo = 0, hi = ga.njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ga.start[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <typename T, int TPT, int C, bool COL, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_group_kernel(GroupArgs ga) {
  const int jb = find_job(ga, blockIdx.x);
  const int64_t s0 = ga.start[jb];
  c2c_body<T, TPT, C, COL, false, Rs...>(ga.jobs[jb], blockIdx.x - s0, ga.start[jb + 1] - s0);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_wide_group_kernel(GroupArgs ga) {
  const int jb = find_job(ga, blockIdx.x);
  const int64_t s0 = ga.start[jb];
  c2c_wide_body<T, TPT, C, ROWL, false, Rs...>(ga.jobs[jb], blockIdx.x - s0,
                                              ga.start[jb + 1] - s0);
}
template <typename T, int TPT, int C, bool ROWL, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2c_wide_nt_group_kernel(GroupArgs ga) {
  const int jb = find_job(ga, blockIdx.x);
  const int64_t s0 = ga.start[jb];
  c2c_wide_body_x<T, TPT, C, ROWL, false, true, Rs...>(ga.jobs[jb], blockIdx.x - s0,
                                                      ga.start[jb + 1] - s0);
}
template <typename T, int TPT, int C, int... Rs>
__global__ void __launch_bounds__(TPT * C) r2c_group_kernel(GroupArgs ga) {
  const int jb = find_job(ga, blockIdx.x);
  r2c_body<T, TPT, C, false, Rs...>(ga.jobs[jb], blockIdx.x - ga.start[jb]);
}
template <typename T, int TPT, int C, int... Rs>
__global__ void __launch_bounds__(TPT * C) c2r_group_kernel(GroupArgs ga) {
  const int jb = find_job(ga, blockIdx.x);
  c2r_body<T, TPT, C, false, Rs...>(ga.jobs[jb], blockIdx.x - ga.start[jb]);
}

}  // namespace pfx
