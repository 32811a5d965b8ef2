// pfx_butterfly.hpp -- small forward DFT butterflies (exp(-2 pi i jk/R)) on
// complex values held in registers.  The backward transform is obtained by
// swapping re/im on the way in and out (DFT_+(x) = swap(DFT_-(swap(x)))), so
// only forward butterflies exist.  Constants are the exact mathematical values
// rounded to the working precision.
#pragma once
#include <hip/hip_runtime.h>

namespace pfx {

template <typename T> struct V2T;
template <> struct V2T<float> { using type = float2; };
template <> struct V2T<double> { using type = double2; };
template <typename T> using V2 = typename V2T<T>::type;

template <typename T> __host__ __device__ __forceinline__ V2<T> mk(T x, T y) { V2<T> r; r.x = x; r.y = y; return r; }
template <typename T> __host__ __device__ __forceinline__ V2<T> add(V2<T> a, V2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ __forceinline__ V2<T> sub(V2<T> a, V2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ __forceinline__ V2<T> cmul(V2<T> a, V2<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T> __host__ __device__ __forceinline__ V2<T> cmulc(V2<T> a, V2<T> b) {
  return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __host__ __device__ __forceinline__ V2<T> conj(V2<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __host__ __device__ __forceinline__ V2<T> swp(V2<T> a) { return mk<T>(a.y, a.x); }
// -i * a
template <typename T> __host__ __device__ __forceinline__ V2<T> mul_mi(V2<T> a) { return mk<T>(a.y, -a.x); }
// +i * a
template <typename T> __host__ __device__ __forceinline__ V2<T> mul_pi(V2<T> a) { return mk<T>(-a.y, a.x); }
template <typename T> __host__ __device__ __forceinline__ V2<T> scl(V2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }

namespace k {
constexpr double SQRT1_2 = 0.70710678118654752440084436210484904;
constexpr double COS_PI8 = 0.92387953251128675612818318939678829;
constexpr double SIN_PI8 = 0.38268343236508977172845998403039887;
constexpr double SQRT3_2 = 0.86602540378443864676372317075293618;
constexpr double C5_1 = 0.30901699437494742410229341718281906;   // cos(2pi/5)
constexpr double C5_2 = -0.80901699437494742410229341718281906;  // cos(4pi/5)
constexpr double S5_1 = 0.95105651629515357211643933337938214;   // sin(2pi/5)
constexpr double S5_2 = 0.58778525229247312916870595463907277;   // sin(4pi/5)
constexpr double COS_PI16 = 0.98078528040323044912618223613424369;
constexpr double SIN_PI16 = 0.19509032201612826784828486847702224;
constexpr double COS_3PI16 = 0.83146961230254523707878837761790576;
constexpr double SIN_3PI16 = 0.55557023301960222474283081394853287;
}  // namespace k

template <typename T> __host__ __device__ __forceinline__ void bfly2(V2<T>* v) {
  V2<T> a = v[0], b = v[1];
  v[0] = add<T>(a, b);
  v[1] = sub<T>(a, b);
}

// in-place radix-4 on v[0], v[s], v[2s], v[3s] (natural order in and out)
template <typename T, int S = 1> __host__ __device__ __forceinline__ void bfly4s(V2<T>* v) {
  V2<T> a0 = v[0], a1 = v[S], a2 = v[2 * S], a3 = v[3 * S];
  V2<T> t0 = add<T>(a0, a2), t1 = sub<T>(a0, a2);
  V2<T> t2 = add<T>(a1, a3), t3 = mul_mi<T>(sub<T>(a1, a3));
  v[0] = add<T>(t0, t2);
  v[2 * S] = sub<T>(t0, t2);
  v[S] = add<T>(t1, t3);
  v[3 * S] = sub<T>(t1, t3);
}
template <typename T> __host__ __device__ __forceinline__ void bfly4(V2<T>* v) { bfly4s<T, 1>(v); }

template <typename T> __host__ __device__ __forceinline__ void bfly3(V2<T>* v) {
  const T h = T(k::SQRT3_2);
  V2<T> a = v[0], b = v[1], c = v[2];
  V2<T> t1 = add<T>(b, c);
  V2<T> t2 = mk<T>(a.x - T(0.5) * t1.x, a.y - T(0.5) * t1.y);
  V2<T> t3 = scl<T>(sub<T>(b, c), h);
  v[0] = add<T>(a, t1);
  v[1] = add<T>(t2, mul_mi<T>(t3));
  v[2] = add<T>(t2, mul_pi<T>(t3));
}

template <typename T> __host__ __device__ __forceinline__ void bfly5(V2<T>* v) {
  const T c1 = T(k::C5_1), c2 = T(k::C5_2), s1 = T(k::S5_1), s2 = T(k::S5_2);
  V2<T> a0 = v[0];
  V2<T> b1 = add<T>(v[1], v[4]), b2 = add<T>(v[2], v[3]);
  V2<T> d1 = sub<T>(v[1], v[4]), d2 = sub<T>(v[2], v[3]);
  V2<T> p1 = mk<T>(a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y);
  V2<T> p2 = mk<T>(a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y);
  V2<T> q1 = mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
  V2<T> q2 = mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
  v[0] = mk<T>(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
  v[1] = add<T>(p1, mul_mi<T>(q1));
  v[4] = add<T>(p1, mul_pi<T>(q1));
  v[2] = add<T>(p2, mul_mi<T>(q2));
  v[3] = add<T>(p2, mul_pi<T>(q2));
}

// radix-8 = 2 x radix-4: evens / odds, twiddle W8^k on the odds, combine
template <typename T> __host__ __device__ __forceinline__ void bfly8(V2<T>* v) {
  const T r = T(k::SQRT1_2);
  V2<T> e[4] = {v[0], v[2], v[4], v[6]};
  V2<T> o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<T>(e);
  bfly4<T>(o);
  // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = -(1 + i)/sqrt2
  o[1] = mk<T>((o[1].x + o[1].y) * r, (o[1].y - o[1].x) * r);
  o[2] = mul_mi<T>(o[2]);
  o[3] = mk<T>((o[3].y - o[3].x) * r, -(o[3].x + o[3].y) * r);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = add<T>(e[q], o[q]);
    v[q + 4] = sub<T>(e[q], o[q]);
  }
}

// radix-16 as 4 x 4 (decimation in time, n = 4 n1 + n2, k = k1 + 4 k2)
template <typename T> __host__ __device__ __forceinline__ void bfly16(V2<T>* v) {
  const T c = T(k::COS_PI8), s = T(k::SIN_PI8), r = T(k::SQRT1_2);
  // step 1: for each n2, DFT4 over n1 of v[4 n1 + n2]  -> F[n2][k1] stored at v[4 k1 + n2]
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) bfly4s<T, 4>(v + n2);
  // step 2: multiply F[n2][k1] (at v[4 k1 + n2]) by W16^(n2 k1)
  // k1 = 1: W^1, W^2, W^3 on n2 = 1, 2, 3
  v[5] = mk<T>(v[5].x * c + v[5].y * s, v[5].y * c - v[5].x * s);   // W16^1 = c - i s
  v[6] = mk<T>((v[6].x + v[6].y) * r, (v[6].y - v[6].x) * r);        // W16^2
  v[7] = mk<T>(v[7].x * s + v[7].y * c, v[7].y * s - v[7].x * c);   // W16^3 = s - i c
  // k1 = 2: W^2, W^4, W^6
  v[9] = mk<T>((v[9].x + v[9].y) * r, (v[9].y - v[9].x) * r);        // W16^2
  v[10] = mul_mi<T>(v[10]);                                          // W16^4 = -i
  v[11] = mk<T>((v[11].y - v[11].x) * r, -(v[11].x + v[11].y) * r);  // W16^6 = -(1+i)/sqrt2
  // k1 = 3: W^3, W^6, W^9
  v[13] = mk<T>(v[13].x * s + v[13].y * c, v[13].y * s - v[13].x * c);  // W16^3
  v[14] = mk<T>((v[14].y - v[14].x) * r, -(v[14].x + v[14].y) * r);     // W16^6
  v[15] = mk<T>(-v[15].x * c - v[15].y * s, -v[15].y * c + v[15].x * s); // W16^9 = -c + i s
  // step 3: for each k1, DFT4 over n2 of v[4 k1 + n2] -> X[k1 + 4 k2]
  V2<T> t[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    bfly4s<T, 1>(v + 4 * k1);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = v[4 * k1 + k2];
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = t[i];
}

// radix-32 = 2 x radix-16 (decimation in time): X[k] = E[k] + W32^k O[k],
// X[k + 16] = E[k] - W32^k O[k], W32^k = cos(pi k/16) - i sin(pi k/16).
template <typename T> __host__ __device__ __forceinline__ void bfly32(V2<T>* v) {
  const T c1 = T(k::COS_PI16), s1 = T(k::SIN_PI16);
  const T c2 = T(k::COS_PI8), s2 = T(k::SIN_PI8);
  const T c3 = T(k::COS_3PI16), s3 = T(k::SIN_3PI16);
  const T r = T(k::SQRT1_2);
  V2<T> e[16], o[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    e[i] = v[2 * i];
    o[i] = v[2 * i + 1];
  }
  bfly16<T>(e);
  bfly16<T>(o);
  // (cos, sin) of pi k / 16 for k = 1..15; the twiddle is cos - i sin
  const T cs[16][2] = {{T(1), T(0)}, {c1, s1}, {c2, s2}, {c3, s3}, {r, r},    {s3, c3},
                       {s2, c2},     {s1, c1}, {T(0), T(1)}, {-s1, c1}, {-s2, c2}, {-s3, c3},
                       {-r, r},      {-c3, s3}, {-c2, s2}, {-c1, s1}};
#pragma unroll
  for (int q = 1; q < 16; ++q) {
    if (q == 8) {
      o[q] = mul_mi<T>(o[q]);
    } else {
      const T c = cs[q][0], s = cs[q][1];
      o[q] = mk<T>(o[q].x * c + o[q].y * s, o[q].y * c - o[q].x * s);
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    v[q] = add<T>(e[q], o[q]);
    v[q + 16] = sub<T>(e[q], o[q]);
  }
}

template <typename T, int R> __host__ __device__ __forceinline__ void bfly(V2<T>* v) {
  if constexpr (R == 2) bfly2<T>(v);
  else if constexpr (R == 3) bfly3<T>(v);
  else if constexpr (R == 4) bfly4<T>(v);
  else if constexpr (R == 5) bfly5<T>(v);
  else if constexpr (R == 8) bfly8<T>(v);
  else if constexpr (R == 16) bfly16<T>(v);
  else if constexpr (R == 32) bfly32<T>(v);
  else static_assert(R == 2, "unsupported radix");
}

}  // namespace pfx
