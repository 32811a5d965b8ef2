// pfx_fsrun.hpp -- the Stockham passes of one LDS-resident tile for kernels
// that address HBM through their own functors (the fused four-step passes,
// the fused chirp convolution): the first pass loads from HBM straight into
// the butterflies' registers, the last stores from them, and the passes
// between run LDS to LDS.
#pragma once
#include "pfx_stockham.hpp"

namespace pfx {

template <typename T, int N, int TPT, int C, bool COL, int Ns, int... Rs> struct FsPasses;
template <typename T, int N, int TPT, int C, bool COL, int Ns>
struct FsPasses<T, N, TPT, C, COL, Ns> {
  __device__ static __forceinline__ void run(int, int, V2<T>*, const V2<T>*) {}
};
template <typename T, int N, int TPT, int C, bool COL, int Ns, int R, int... Rest>
struct FsPasses<T, N, TPT, C, COL, Ns, R, Rest...> {
  using L = LdsLayout<T, N, C, COL>;
  struct StoreLds {
    V2<T>* lds;
    int c;
    __device__ void before() const { __syncthreads(); }
    __device__ void after() const { __syncthreads(); }
    __device__ void operator()(int e, V2<T> v) const { lds[L::at(e, c)] = v; }
  };
  __device__ static __forceinline__ void run(int j, int c, V2<T>* lds, const V2<T>* stw) {
    auto load = [&](int e) -> V2<T> { return lds[L::at(e, c)]; };
    stockham_pass<T, N, R, Ns, TPT>(j, stw, load, StoreLds{lds, c});
    FsPasses<T, N, TPT, C, COL, Ns * R, Rest...>::run(j, c, lds, stw);
  }
};

// The passes of one tile with the first pass loading from HBM through `ld`
// and the last storing to HBM through `st` (element and output indices of
// the transform), the passes between them in LDS.
template <typename T, int N, int TPT, int C, bool COL, int Ns, int... Rs> struct FsRun;
template <typename T, int N, int TPT, int C, bool COL, int Ns>
struct FsRun<T, N, TPT, C, COL, Ns> {
  template <class LD, class ST>
  __device__ static __forceinline__ void run(int, int, V2<T>*, const V2<T>*, const LD&,
                                             const ST&) {}
};
template <typename T, int N, int TPT, int C, bool COL, int Ns, int R, int... Rest>
struct FsRun<T, N, TPT, C, COL, Ns, R, Rest...> {
  using P = FsPasses<T, N, TPT, C, COL, Ns, R, Rest...>;
  template <class LD, class ST>
  __device__ static __forceinline__ void run(int j, int c, V2<T>* lds, const V2<T>* stw,
                                             const LD& ld, const ST& st) {
    using L = LdsLayout<T, N, C, COL>;
    auto from_lds = [&](int e) -> V2<T> { return lds[L::at(e, c)]; };
    typename P::StoreLds to_lds{lds, c};
    constexpr bool first = Ns == 1, last = sizeof...(Rest) == 0;
    if constexpr (first && last) stockham_pass<T, N, R, Ns, TPT>(j, stw, ld, st);
    else if constexpr (first) stockham_pass<T, N, R, Ns, TPT>(j, stw, ld, to_lds);
    else if constexpr (last) stockham_pass<T, N, R, Ns, TPT>(j, stw, from_lds, st);
    else stockham_pass<T, N, R, Ns, TPT>(j, stw, from_lds, to_lds);
    FsRun<T, N, TPT, C, COL, Ns * R, Rest...>::run(j, c, lds, stw, ld, st);
  }
};

}  // namespace pfx
