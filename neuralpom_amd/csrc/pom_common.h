// Common device infrastructure for the MI355X POM step kernels.
//
// Layout: every 3-D field of one member is (im, jm, KB) with k fastest, exactly
// the reference's tensor layout (src/neuralpom/cores/explicit/base.py:162-174),
// so torch tensors are used in place.  An ensemble packs members back to back:
// member m's 3-D field starts at off3[m] elements and its 2-D field at off2[m].
//
// Value types: P<T> (passive), D<T> (value + one tangent) and DS<T, M> (value
// + one tangent lane per input field in the mask M).  Every stencil operator
// is written once as a point function over an accessor.  The forward kernel
// instantiates it with an accessor that returns P<T>; the adjoint gather
// instantiates it once per stencil offset o with an accessor that returns DS
// for every input field read at exactly o.  Activity is a property of the
// type, so inactive arithmetic disappears at compile time and each gather term
// is the exact partial derivative d out(q - o) / d in(q), evaluated at the
// saved primal state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace pom {

constexpr int KB = 16;        // sigma levels (reference base.py:27)
constexpr int KBM1 = KB - 1;
constexpr int KBM2 = KB - 2;
constexpr int TPB = 256;      // threads per block for point kernels

// ------------------------------------------------------------------ values
template <class T> struct P { T v; };
template <class T> struct D { T v, d; };

template <class X> struct is_active : std::false_type {};
template <class T> struct is_active<D<T>> : std::true_type {};
template <class X> inline constexpr bool active_v = is_active<X>::value;

template <class T> __device__ __forceinline__ T val(P<T> a) { return a.v; }
template <class T> __device__ __forceinline__ T val(D<T> a) { return a.v; }
template <class T> __device__ __forceinline__ T tan_(P<T>) { return T(0); }
template <class T> __device__ __forceinline__ T tan_(D<T> a) { return a.d; }

template <class S> inline constexpr bool arith_v = std::is_arithmetic<S>::value;

// P op P
template <class T> __device__ __forceinline__ P<T> operator+(P<T> a, P<T> b) { return {a.v + b.v}; }
template <class T> __device__ __forceinline__ P<T> operator-(P<T> a, P<T> b) { return {a.v - b.v}; }
template <class T> __device__ __forceinline__ P<T> operator*(P<T> a, P<T> b) { return {a.v * b.v}; }
template <class T> __device__ __forceinline__ P<T> operator/(P<T> a, P<T> b) { return {a.v / b.v}; }
template <class T> __device__ __forceinline__ P<T> operator-(P<T> a) { return {-a.v}; }
// D op D
template <class T> __device__ __forceinline__ D<T> operator+(D<T> a, D<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __device__ __forceinline__ D<T> operator-(D<T> a, D<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __device__ __forceinline__ D<T> operator*(D<T> a, D<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class T> __device__ __forceinline__ D<T> operator/(D<T> a, D<T> b) {
  T q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
template <class T> __device__ __forceinline__ D<T> operator-(D<T> a) { return {-a.v, -a.d}; }
// D op P / P op D
template <class T> __device__ __forceinline__ D<T> operator+(D<T> a, P<T> b) { return {a.v + b.v, a.d}; }
template <class T> __device__ __forceinline__ D<T> operator+(P<T> a, D<T> b) { return {a.v + b.v, b.d}; }
template <class T> __device__ __forceinline__ D<T> operator-(D<T> a, P<T> b) { return {a.v - b.v, a.d}; }
template <class T> __device__ __forceinline__ D<T> operator-(P<T> a, D<T> b) { return {a.v - b.v, -b.d}; }
template <class T> __device__ __forceinline__ D<T> operator*(D<T> a, P<T> b) { return {a.v * b.v, a.d * b.v}; }
template <class T> __device__ __forceinline__ D<T> operator*(P<T> a, D<T> b) { return {a.v * b.v, a.v * b.d}; }
template <class T> __device__ __forceinline__ D<T> operator/(D<T> a, P<T> b) { return {a.v / b.v, a.d / b.v}; }
template <class T> __device__ __forceinline__ D<T> operator/(P<T> a, D<T> b) {
  T q = a.v / b.v;
  return {q, -q * b.d / b.v};
}
// scalar literals (cast to T, as a python float is cast to the tensor dtype)
#define POM_SCALAR_OPS(OP)                                                                   \
  template <class T, class S, std::enable_if_t<arith_v<S>, int> = 0>                         \
  __device__ __forceinline__ auto operator OP(S s, P<T> b) { return P<T>{T(s)} OP b; }       \
  template <class T, class S, std::enable_if_t<arith_v<S>, int> = 0>                         \
  __device__ __forceinline__ auto operator OP(P<T> a, S s) { return a OP P<T>{T(s)}; }       \
  template <class T, class S, std::enable_if_t<arith_v<S>, int> = 0>                         \
  __device__ __forceinline__ auto operator OP(S s, D<T> b) { return P<T>{T(s)} OP b; }       \
  template <class T, class S, std::enable_if_t<arith_v<S>, int> = 0>                         \
  __device__ __forceinline__ auto operator OP(D<T> a, S s) { return a OP P<T>{T(s)}; }
POM_SCALAR_OPS(+)
POM_SCALAR_OPS(-)
POM_SCALAR_OPS(*)
POM_SCALAR_OPS(/)
#undef POM_SCALAR_OPS

template <class T> __device__ __forceinline__ P<T> psqrt(P<T> a) { return {sqrt(a.v)}; }
template <class T> __device__ __forceinline__ D<T> psqrt(D<T> a) {
  T s = sqrt(a.v);
  return {s, a.d / (T(2) * s)};
}
// |x| with d|x|/dx = sign(x) (0 at 0), as torch.abs
template <class T> __device__ __forceinline__ P<T> pabs(P<T> a) { return {fabs(a.v)}; }
template <class T> __device__ __forceinline__ D<T> pabs(D<T> a) {
  T s = a.v > T(0) ? T(1) : (a.v < T(0) ? T(-1) : T(0));
  return {fabs(a.v), s * a.d};
}
template <class A> __device__ __forceinline__ auto sq(A a) { return a * a; }

// select by a primal condition; activity promotes
template <class T> __device__ __forceinline__ P<T> sel(bool c, P<T> a, P<T> b) { return c ? a : b; }
template <class T> __device__ __forceinline__ D<T> sel(bool c, D<T> a, D<T> b) { return c ? a : b; }
template <class T> __device__ __forceinline__ D<T> sel(bool c, D<T> a, P<T> b) { return c ? a : D<T>{b.v, T(0)}; }
template <class T> __device__ __forceinline__ D<T> sel(bool c, P<T> a, D<T> b) { return c ? D<T>{a.v, T(0)} : b; }

// torch.clamp(x, min=lo): gradient passes where x >= lo
template <class X, class T> __device__ __forceinline__ X clamp_min(X a, T lo) {
  if (val(a) >= lo) return a;
  X r{};
  r.v = lo;
  return r;
}
// torch.clamp(x, max=hi): gradient passes where x <= hi
template <class X, class T> __device__ __forceinline__ X clamp_max(X a, T hi) {
  if (val(a) <= hi) return a;
  X r{};
  r.v = hi;
  return r;
}
// elementwise maximum selected by the primal (tie -> first argument)
template <class X, class Y> __device__ __forceinline__ auto pmax(X a, Y b);

// common activity type of two values, and conversion into it
template <class X> __device__ X mkval();   // declared only: for decltype
template <class X, class Y> using common_t = decltype(mkval<X>() + mkval<Y>());
template <class R, class X> __device__ __forceinline__ R cvt(X x);
template <class X, class Y> __device__ __forceinline__ auto pmax(X a, Y b) {
  using R = common_t<X, Y>;
  return val(a) >= val(b) ? cvt<R>(a) : cvt<R>(b);
}

template <class X> struct scalar_of;
template <class T> struct scalar_of<P<T>> { using type = T; };
template <class T> struct scalar_of<D<T>> { using type = T; };

// ------------------------------------------------------------------ multi-seed tangents
// DS<T, M>: a value and its partial derivatives with respect to the input
// fields whose bits are set in the compile-time mask M (one lane per field,
// stored densely in bit order).  Used by the fused adjoint gathers: one
// evaluation of a point function at a stencil offset gives the partials with
// respect to every input field at that offset.  Arithmetic per lane follows D.
constexpr int popc(uint32_t x) { return x ? int(x & 1u) + popc(x >> 1) : 0; }
constexpr int slot(uint32_t m, int b) { return popc(m & ((1u << b) - 1u)); }

template <class T, uint32_t M> struct DS {
  T v;
  T d[popc(M) > 0 ? popc(M) : 1];
};
template <class T, uint32_t M> struct is_active<DS<T, M>> : std::true_type {};
template <class T, uint32_t M> struct scalar_of<DS<T, M>> { using type = T; };
template <class T, uint32_t M> __device__ __forceinline__ T val(const DS<T, M>& a) { return a.v; }

template <class X> struct mask_of { static constexpr uint32_t value = 0; };
template <class T, uint32_t M> struct mask_of<DS<T, M>> { static constexpr uint32_t value = M; };

// f(integral_constant<int, b>) for every set bit b of M, in increasing order
template <uint32_t M, int B = 0, class F> __device__ __forceinline__ void for_bits(F&& f) {
  if constexpr (B < 32) {
    if constexpr ((M >> B) & 1u) f(std::integral_constant<int, B>{});
    for_bits<M, B + 1>(f);
  }
}
// lane b of x (0 when x does not carry b)
template <int B, class T, uint32_t M> __device__ __forceinline__ T lane(const DS<T, M>& x) {
  if constexpr ((M >> B) & 1u) return x.d[slot(M, B)];
  else return T(0);
}

// DS op DS
template <class T, uint32_t A, uint32_t B>
__device__ __forceinline__ DS<T, A | B> operator+(const DS<T, A>& a, const DS<T, B>& b) {
  DS<T, A | B> r;
  r.v = a.v + b.v;
  for_bits<A | B>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr bool ha = (A >> i) & 1u, hb = (B >> i) & 1u;
    if constexpr (ha && hb) r.d[slot(A | B, i)] = a.d[slot(A, i)] + b.d[slot(B, i)];
    else if constexpr (ha) r.d[slot(A | B, i)] = a.d[slot(A, i)];
    else r.d[slot(A | B, i)] = b.d[slot(B, i)];
  });
  return r;
}
template <class T, uint32_t A, uint32_t B>
__device__ __forceinline__ DS<T, A | B> operator-(const DS<T, A>& a, const DS<T, B>& b) {
  DS<T, A | B> r;
  r.v = a.v - b.v;
  for_bits<A | B>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr bool ha = (A >> i) & 1u, hb = (B >> i) & 1u;
    if constexpr (ha && hb) r.d[slot(A | B, i)] = a.d[slot(A, i)] - b.d[slot(B, i)];
    else if constexpr (ha) r.d[slot(A | B, i)] = a.d[slot(A, i)];
    else r.d[slot(A | B, i)] = -b.d[slot(B, i)];
  });
  return r;
}
template <class T, uint32_t A, uint32_t B>
__device__ __forceinline__ DS<T, A | B> operator*(const DS<T, A>& a, const DS<T, B>& b) {
  DS<T, A | B> r;
  r.v = a.v * b.v;
  for_bits<A | B>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr bool ha = (A >> i) & 1u, hb = (B >> i) & 1u;
    if constexpr (ha && hb) r.d[slot(A | B, i)] = a.d[slot(A, i)] * b.v + a.v * b.d[slot(B, i)];
    else if constexpr (ha) r.d[slot(A | B, i)] = a.d[slot(A, i)] * b.v;
    else r.d[slot(A | B, i)] = a.v * b.d[slot(B, i)];
  });
  return r;
}
// Tangent lanes of a quotient are scaled by one reciprocal of the divisor
// (the hardware reciprocal in float, 1 ulp; an IEEE division in double)
// instead of one division per lane.  The primal value stays the exact
// quotient, so every primal comparison of an adjoint gather sees the values
// the forward map computed.
__device__ __forceinline__ float rcp_lane(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_lane(double x) { return 1.0 / x; }
// a / b given rb = rcp_lane(b): in float the product with the hardware
// reciprocal (within 2 ulp, instead of the ten-instruction IEEE division),
// in double the exact quotient.  For values that only feed tangent lanes.
__device__ __forceinline__ float div_lane(float a, float b, float rb) { return a * rb; }
__device__ __forceinline__ double div_lane(double a, double b, double) { return a / b; }

template <class T, uint32_t A, uint32_t B>
__device__ __forceinline__ DS<T, A | B> operator/(const DS<T, A>& a, const DS<T, B>& b) {
  DS<T, A | B> r;
  const T rb = rcp_lane(b.v);
  const T q = div_lane(a.v, b.v, rb);
  r.v = q;
  for_bits<A | B>([&](auto I) {
    constexpr int i = decltype(I)::value;
    constexpr bool ha = (A >> i) & 1u, hb = (B >> i) & 1u;
    if constexpr (ha && hb) r.d[slot(A | B, i)] = (a.d[slot(A, i)] - q * b.d[slot(B, i)]) * rb;
    else if constexpr (ha) r.d[slot(A | B, i)] = a.d[slot(A, i)] * rb;
    else r.d[slot(A | B, i)] = -q * b.d[slot(B, i)] * rb;
  });
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator-(const DS<T, A>& a) {
  DS<T, A> r;
  r.v = -a.v;
  for (int i = 0; i < popc(A); ++i) r.d[i] = -a.d[i];
  return r;
}
// DS op P / P op DS
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator+(const DS<T, A>& a, P<T> b) {
  DS<T, A> r = a;
  r.v = a.v + b.v;
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator+(P<T> a, const DS<T, A>& b) {
  DS<T, A> r = b;
  r.v = a.v + b.v;
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator-(const DS<T, A>& a, P<T> b) {
  DS<T, A> r = a;
  r.v = a.v - b.v;
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator-(P<T> a, const DS<T, A>& b) {
  DS<T, A> r;
  r.v = a.v - b.v;
  for (int i = 0; i < popc(A); ++i) r.d[i] = -b.d[i];
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator*(const DS<T, A>& a, P<T> b) {
  DS<T, A> r;
  r.v = a.v * b.v;
  for (int i = 0; i < popc(A); ++i) r.d[i] = a.d[i] * b.v;
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator*(P<T> a, const DS<T, A>& b) {
  DS<T, A> r;
  r.v = a.v * b.v;
  for (int i = 0; i < popc(A); ++i) r.d[i] = a.v * b.d[i];
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator/(const DS<T, A>& a, P<T> b) {
  DS<T, A> r;
  const T rb = rcp_lane(b.v);
  r.v = div_lane(a.v, b.v, rb);
  for (int i = 0; i < popc(A); ++i) r.d[i] = a.d[i] * rb;
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> operator/(P<T> a, const DS<T, A>& b) {
  DS<T, A> r;
  const T rb = rcp_lane(b.v);
  const T q = div_lane(a.v, b.v, rb);
  r.v = q;
  const T qr = -q * rb;
  for (int i = 0; i < popc(A); ++i) r.d[i] = qr * b.d[i];
  return r;
}
#define POM_DS_SCALAR_OPS(OP)                                                                        \
  template <class T, uint32_t A, class S, std::enable_if_t<arith_v<S>, int> = 0>                     \
  __device__ __forceinline__ auto operator OP(S s, const DS<T, A>& b) { return P<T>{T(s)} OP b; }    \
  template <class T, uint32_t A, class S, std::enable_if_t<arith_v<S>, int> = 0>                     \
  __device__ __forceinline__ auto operator OP(const DS<T, A>& a, S s) { return a OP P<T>{T(s)}; }
POM_DS_SCALAR_OPS(+)
POM_DS_SCALAR_OPS(-)
POM_DS_SCALAR_OPS(*)
POM_DS_SCALAR_OPS(/)
#undef POM_DS_SCALAR_OPS

template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> psqrt(const DS<T, A>& a) {
  DS<T, A> r;
  const T s = sqrt(a.v);
  r.v = s;
  const T h = rcp_lane(T(2) * s);
  for (int i = 0; i < popc(A); ++i) r.d[i] = a.d[i] * h;
  return r;
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> pabs(const DS<T, A>& a) {
  DS<T, A> r;
  const T sg = a.v > T(0) ? T(1) : (a.v < T(0) ? T(-1) : T(0));
  r.v = fabs(a.v);
  for (int i = 0; i < popc(A); ++i) r.d[i] = sg * a.d[i];
  return r;
}
// every lane of x multiplied by s (value unchanged)
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> lanes_scaled(const DS<T, A>& x, T s) {
  DS<T, A> r;
  r.v = x.v;
  for (int i = 0; i < popc(A); ++i) r.d[i] = x.d[i] * s;
  return r;
}

// conversion into a common activity type (P -> D / DS, DS<A> -> DS<A|B>)
template <class R, class X> __device__ __forceinline__ R cvt(X x) {
  if constexpr (std::is_same_v<R, X>) {
    return x;
  } else if constexpr (mask_of<R>::value != 0) {
    constexpr uint32_t MR = mask_of<R>::value, MX = mask_of<X>::value;
    static_assert((MX & ~MR) == 0u, "cvt would drop tangent lanes");
    R r{};
    r.v = x.v;
    if constexpr (MX != 0u) for_bits<MX>([&](auto I) {
      constexpr int i = decltype(I)::value;
      r.d[slot(MR, i)] = x.d[slot(MX, i)];
    });
    return r;
  } else {
    return R{x.v, decltype(x.v)(0)};
  }
}
template <class T, uint32_t A, uint32_t B>
__device__ __forceinline__ DS<T, A | B> sel(bool c, const DS<T, A>& a, const DS<T, B>& b) {
  using R = DS<T, A | B>;
  return c ? cvt<R>(a) : cvt<R>(b);
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> sel(bool c, const DS<T, A>& a, P<T> b) {
  return c ? a : cvt<DS<T, A>>(b);
}
template <class T, uint32_t A> __device__ __forceinline__ DS<T, A> sel(bool c, P<T> a, const DS<T, A>& b) {
  return c ? cvt<DS<T, A>>(a) : b;
}

// ------------------------------------------------------------------ batch
struct Member {
  int32_t im, jm;
  int64_t off2, off3;
};

// Water columns near a member's edge, for the operators that read
// neighbours only there (EDGE_BAND): entry (m, i * jm + j) for every column
// of member m whose distance d = min(i, im-1-i, j, jm-1-j) to the nearest
// edge is below BAND_W, ordered by d, so that the first nband[w] entries are
// the columns with d < w.
constexpr int BAND_W = 4;

struct Batch {
  const Member* mem;   // [M]
  const int32_t* blk3; // [M+1] prefix of 3-D point blocks
  const int32_t* blk2; // [M+1] prefix of 2-D point blocks
  int32_t M;
  const int2* band = nullptr;          // the edge-band columns (null: none built)
  int32_t nband[BAND_W + 1] = {};
};

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup x runs on
// XCD x % 8), and each XCD has its own L2.  The logical block index gives XCD
// k one contiguous range of blocks, so the neighbouring rows (i +- 1, i +- 2)
// a stencil reads were mostly fetched into the same L2 by blocks running at
// about the same time.  A bijection of [0, gridDim.x) for any grid size.
#ifndef POM_XCD_REMAP
#define POM_XCD_REMAP 1
#endif
constexpr int NXCD = 8;
__device__ __forceinline__ int logical_block() {
  const int x = blockIdx.x;
  if constexpr (POM_XCD_REMAP) {
    const int n = gridDim.x, q = n / NXCD, r = n % NXCD, k = x % NXCD;
    return k * q + (k < r ? k : r) + x / NXCD;
  } else {
    return x;
  }
}

// member owning block b (all lanes agree -> scalar)
__device__ __forceinline__ int member_of(const int32_t* blk, int M, int b) {
  int lo = 0, hi = M;   // invariant blk[lo] <= b < blk[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (blk[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// position of the current thread in a 3-D launch; false if past the member
struct Pos {
  int m, im, jm, i, j, k;
  int64_t o2, o3;   // member base offsets
};

// The member's fields are set before the per-lane range check, so that they
// hold the block's one value on every path: the compiler then keeps them (and
// every field base formed from them) in scalar registers, and a load is that
// base plus a 32-bit lane offset.  Set after the check, they would be merged
// with the undefined values of the lanes past the member's end and counted as
// per-lane, and every load would form a 64-bit address per lane.  Local point
// indices fit in 32 bits (a member's 3-D field is below 2 GiB).
__device__ __forceinline__ bool pos3_at(const Batch& B, int b, Pos& p) {
  if (b >= B.blk3[B.M]) return false;
  const int m = member_of(B.blk3, B.M, b);
  const Member mm = B.mem[m];
  p.m = m; p.im = mm.im; p.jm = mm.jm; p.o2 = mm.off2; p.o3 = mm.off3;
  const uint32_t local = (uint32_t)(b - B.blk3[m]) * TPB + threadIdx.x;
  const uint32_t n = (uint32_t)mm.im * (uint32_t)mm.jm * KB;
  if (local >= n) return false;
  p.k = (int)(local & (KB - 1));
  const uint32_t col = local >> 4;
  p.i = (int)(col / (uint32_t)mm.jm);
  p.j = (int)(col - (uint32_t)p.i * (uint32_t)mm.jm);
  return true;
}

__device__ __forceinline__ bool pos3(const Batch& B, Pos& p) { return pos3_at(B, logical_block(), p); }

__device__ __forceinline__ bool pos2_at(const Batch& B, int b, Pos& p) {
  if (b >= B.blk2[B.M]) return false;
  const int m = member_of(B.blk2, B.M, b);
  const Member mm = B.mem[m];
  p.m = m; p.im = mm.im; p.jm = mm.jm; p.o2 = mm.off2; p.o3 = mm.off3;   // before the per-lane check (pos3_at)
  const uint32_t local = (uint32_t)(b - B.blk2[m]) * TPB + threadIdx.x;
  const uint32_t n = (uint32_t)mm.im * (uint32_t)mm.jm;
  if (local >= n) return false;
  p.k = 0;
  p.i = (int)(local / (uint32_t)mm.jm);
  p.j = (int)(local - (uint32_t)p.i * (uint32_t)mm.jm);
  return true;
}

__device__ __forceinline__ bool pos2(const Batch& B, Pos& p) { return pos2_at(B, logical_block(), p); }

__device__ __forceinline__ int clampi(int x, int n) { return x < 0 ? 0 : (x >= n ? n - 1 : x); }

__device__ __forceinline__ bool interior_pos(const Pos& p) {
  return p.i >= 1 && p.i <= p.im - 2 && p.j >= 1 && p.j <= p.jm - 2;
}

__device__ __forceinline__ int64_t ix3(const Pos& p, int i, int j, int k) {
  return p.o3 + ((int64_t)i * p.jm + j) * KB + k;
}
__device__ __forceinline__ int64_t ix2(const Pos& p, int i, int j) {
  return p.o2 + (int64_t)i * p.jm + j;
}

// Element idx of a member's field, given the member's base pointer.  The
// byte offset is formed in 32 bits (a member's field is below 2 GiB, checked
// when the batch is created).
template <class T> __device__ __forceinline__ T ld_local(const T* base, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (idx * (uint32_t)sizeof(T)));
}
// Element idx of a member's field f, o being the member's first element (o2
// or o3).  The whole byte offset (o + idx) * sizeof(T) is formed in 32 bits
// (a field of the whole batch is below 4 GiB, checked when the batch is
// created), so a load takes the field pointer itself as its scalar base and
// one 32-bit lane offset per point: no 64-bit base is formed per field, and a
// kernel that reads many fields keeps only their pointers in scalar registers
// (with a base per field, forty fields exceed the scalar register file and
// the bases spill into vector lanes).
template <class T> __device__ __forceinline__ uint32_t boff(int64_t o, uint32_t idx) {
  return ((uint32_t)o + idx) * (uint32_t)sizeof(T);
}
template <class T> __device__ __forceinline__ T ld_at(const T* f, int64_t o, uint32_t idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(f) + boff<T>(o, idx));
}
template <class T> __device__ __forceinline__ T& ref_at(T* f, int64_t o, uint32_t idx) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(f) + boff<T>(o, idx));
}
__device__ __forceinline__ uint32_t lix3(const Pos& p, int i, int j, int k) {
  return ((uint32_t)i * (uint32_t)p.jm + (uint32_t)j) * (uint32_t)KB + (uint32_t)k;
}
__device__ __forceinline__ uint32_t lix2(const Pos& p, int i, int j) {
  return (uint32_t)i * (uint32_t)p.jm + (uint32_t)j;
}

// ------------------------------------------------------------------ params
// Scalars shared by the batch (double; cast to T at use) plus per-member
// arrays: ramp[m] and the vertical grid vg[m][4][KB] = (z, zz, dz, dzz).
struct Params {
  double dti, dte, dti2, dte2, ispi, isp2i;
  double alpha, smoth, horcon, grav, rhoref, kappa, umol, tprni, small;
  double tbias, sbias, hmax, rfe, rfw, rfn, rfs;
  double gamma_baro;                               // post-solve Shapiro weight (semi-implicit core)
  double gamma_ts, gamma_uv;                       // tracer / internal-velocity Shapiro weights (semi-implicit core)
  double aam_min;                                  // Smagorinsky viscosity floor (semi-implicit core; 0: none)
  int32_t isplit, ispadv, ntp, nbct, nbcs, iext;   // iext: current substep
  int32_t core;                                    // 0 explicit, 1 semi-implicit
};

// Levels are clamped into [0, KB): operators evaluate both sides of their
// level conditions (IFZ / BR select) and the discarded side may ask for k - 1
// at the surface or k + 1 at the bottom.
template <class T> struct Vert {
  const T* vg;   // [M][4][KB]
  __device__ __forceinline__ static int kc(int k) { return k < 0 ? 0 : (k > KB - 1 ? KB - 1 : k); }
  __device__ __forceinline__ T z(int m, int k) const { return vg[(m * 4 + 0) * KB + kc(k)]; }
  __device__ __forceinline__ T zz(int m, int k) const { return vg[(m * 4 + 1) * KB + kc(k)]; }
  __device__ __forceinline__ T dz(int m, int k) const { return vg[(m * 4 + 2) * KB + kc(k)]; }
  __device__ __forceinline__ T dzz(int m, int k) const { return vg[(m * 4 + 3) * KB + kc(k)]; }
};

// ------------------------------------------------------------------ accessors
// Fields of an operator are addressed by small integer ids.  ids below N3 are
// 3-D, ids in [N3, N3+N2) are 2-D.  Geometry (passive) has its own table.
constexpr int MAXF = 24;
constexpr int MAXG = 16;

template <class T> struct FieldTab {
  const T* f[MAXF];    // state inputs (3-D then 2-D)
  const T* g[MAXG];    // geometry / passive 2-D
};

// SF: seeded field id (-1: none); seed offset (SI, SJ, SK) relative to the
// evaluation point.
template <class T, int SF = -1, int SI = 0, int SJ = 0, int SK = 0>
struct Acc {
  const FieldTab<T>* tab;
  Pos p;          // evaluation point (p.i, p.j, p.k)
  Vert<T> vert;
  const T* ramp;

  // Indices are clamped into the member so that a branch which is computed
  // but not selected can never read outside the member's storage.
  __device__ __forceinline__ int ci(int i) const { return i < 0 ? 0 : (i >= p.im ? p.im - 1 : i); }
  __device__ __forceinline__ int cj(int j) const { return j < 0 ? 0 : (j >= p.jm ? p.jm - 1 : j); }
  __device__ __forceinline__ int ck(int k) const { return k < 0 ? 0 : (k >= KB ? KB - 1 : k); }

  template <int F, int DI, int DJ, int DK>
  __device__ __forceinline__ auto at() const {
    T x = ld_at(tab->f[F], p.o3, lix3(p, ci(p.i + DI), cj(p.j + DJ), ck(p.k + DK)));
    if constexpr (F == SF && DI == SI && DJ == SJ && DK == SK) return D<T>{x, T(1)};
    else return P<T>{x};
  }
  template <int F, int DI, int DJ>
  __device__ __forceinline__ auto at2() const {
    T x = ld_at(tab->f[F], p.o2, lix2(p, ci(p.i + DI), cj(p.j + DJ)));
    if constexpr (F == SF && DI == SI && DJ == SJ) return D<T>{x, T(1)};
    else return P<T>{x};
  }
  template <int G, int DI, int DJ>
  __device__ __forceinline__ P<T> g() const {
    return P<T>{ld_at(tab->g[G], p.o2, lix2(p, ci(p.i + DI), cj(p.j + DJ)))};
  }
  __device__ __forceinline__ T z(int k) const { return vert.z(p.m, k); }
  __device__ __forceinline__ T zz(int k) const { return vert.zz(p.m, k); }
  __device__ __forceinline__ T dz(int k) const { return vert.dz(p.m, k); }
  __device__ __forceinline__ T dzz(int k) const { return vert.dzz(p.m, k); }
  __device__ __forceinline__ T rampv() const { return ramp[p.m]; }
};

// Forward accessor whose fields F0 and F1 at the evaluation point itself are
// given values (the outputs of an operator evaluated just before at the same
// point); every other access reads memory as Acc does.
template <class T, int F0, int F1>
struct AccSub : Acc<T> {
  T v0, v1;
  template <int F, int DI, int DJ>
  __device__ __forceinline__ P<T> at2() const {
    if constexpr (F == F0 || F == F1) {
      static_assert(DI == 0 && DJ == 0, "a substituted field is read at the evaluation point only");
      return P<T>{F == F0 ? v0 : v1};
    } else {
      return Acc<T>::template at2<F, DI, DJ>();
    }
  }
};

// Forward accessor whose field FE at the offsets (0, 0), (-1, 0) and (0, -1)
// and fields F0, F1 (when not -1) at the point itself are given values (computed
// just before at the same point); every other access reads memory as Acc does.
template <class T, int FE, int F0 = -1, int F1 = -1>
struct AccElf : Acc<T> {
  T e0, ew, es;   // FE at (0, 0), (-1, 0), (0, -1)
  T v0, v1;       // F0, F1 at (0, 0)
  template <int F, int DI, int DJ>
  __device__ __forceinline__ P<T> at2() const {
    if constexpr (F == FE) {
      static_assert((DI == 0 && DJ == 0) || (DI == -1 && DJ == 0) || (DI == 0 && DJ == -1),
                    "the given field is read at (0, 0), (-1, 0) and (0, -1) only");
      return P<T>{DI == -1 ? ew : (DJ == -1 ? es : e0)};
    } else if constexpr (F == F0 || F == F1) {
      static_assert(DI == 0 && DJ == 0, "a substituted field is read at the evaluation point only");
      return P<T>{F == F0 ? v0 : v1};
    } else {
      return Acc<T>::template at2<F, DI, DJ>();
    }
  }
};

// Forward accessor that reads every 2-D field and geometry array at a fixed
// offset (SI, SJ) from the wrapped accessor's point: an operator's formula
// evaluated at a neighbour of the evaluation point.
template <class AC, int SI, int SJ>
struct AccShift {
  const AC& a;
  template <int F, int DI, int DJ>
  __device__ __forceinline__ auto at2() const { return a.template at2<F, DI + SI, DJ + SJ>(); }
  template <int G, int DI, int DJ>
  __device__ __forceinline__ auto g() const { return a.template g<G, DI + SI, DJ + SJ>(); }
};

// Clamped neighbour indices around a gather thread's point q: every access of
// a gather (evaluation point q - o, access offset a) lands on q + (a - o),
// |a - o| <= 4 horizontally and <= 2 vertically, so one table per thread
// replaces the per-access clamping and index arithmetic.
struct NbrIdx {
  int32_t ri[9];   // clamp(q.i + d) * jm, d = -4..4
  int32_t cj[9];   // clamp(q.j + d)
  int32_t ck[5];   // clamp(q.k + d), d = -2..2
};
__device__ __forceinline__ void make_nbr(const Pos& q, NbrIdx& t) {
#pragma unroll
  for (int d = -4; d <= 4; ++d) {
    const int i = q.i + d, j = q.j + d;
    t.ri[d + 4] = (i < 0 ? 0 : (i >= q.im ? q.im - 1 : i)) * q.jm;
    t.cj[d + 4] = j < 0 ? 0 : (j >= q.jm ? q.jm - 1 : j);
  }
#pragma unroll
  for (int d = -2; d <= 2; ++d) {
    const int k = q.k + d;
    t.ck[d + 2] = k < 0 ? 0 : (k >= KB ? KB - 1 : k);
  }
}

// Gather accessor: evaluation point p = q - (OI, OJ, OK); every field of MASK
// is seeded at the offset (OI, OJ, OK) from p, i.e. at q (2-D fields when
// the caller keeps them in MASK, which it does only for OK = 0).
template <class T, uint32_t MASK, int OI, int OJ, int OK>
struct AccT {
  const FieldTab<T>* tab;
  Pos p;
  Vert<T> vert;
  const T* ramp;
  const NbrIdx* nb;
  template <int DI, int DJ> __device__ __forceinline__ int32_t loc() const {
    static_assert(DI - OI >= -4 && DI - OI <= 4 && DJ - OJ >= -4 && DJ - OJ <= 4, "stencil beyond table");
    return nb->ri[DI - OI + 4] + nb->cj[DJ - OJ + 4];
  }
  template <int F, int DI, int DJ, int DK>
  __device__ __forceinline__ auto at() const {
    static_assert(DK - OK >= -2 && DK - OK <= 2, "stencil beyond table");
    T x = ld_at(tab->f[F], p.o3, (uint32_t)(loc<DI, DJ>() * KB + nb->ck[DK - OK + 2]));
    if constexpr (((MASK >> F) & 1u) && DI == OI && DJ == OJ && DK == OK) return DS<T, (1u << F)>{x, {T(1)}};
    else return P<T>{x};
  }
  template <int F, int DI, int DJ>
  __device__ __forceinline__ auto at2() const {
    T x = ld_at(tab->f[F], p.o2, (uint32_t)loc<DI, DJ>());
    if constexpr (((MASK >> F) & 1u) && DI == OI && DJ == OJ) return DS<T, (1u << F)>{x, {T(1)}};
    else return P<T>{x};
  }
  template <int G, int DI, int DJ>
  __device__ __forceinline__ P<T> g() const {
    return P<T>{ld_at(tab->g[G], p.o2, (uint32_t)loc<DI, DJ>())};
  }
  __device__ __forceinline__ T z(int k) const { return vert.z(p.m, k); }
  __device__ __forceinline__ T zz(int k) const { return vert.zz(p.m, k); }
  __device__ __forceinline__ T dz(int k) const { return vert.dz(p.m, k); }
  __device__ __forceinline__ T dzz(int k) const { return vert.dzz(p.m, k); }
  __device__ __forceinline__ T rampv() const { return ramp[p.m]; }
};

// The gather accessor of an evaluation point known, at compile time, to lie
// away from every edge (farther than the operator's EDGE_BAND): operators
// with an edge band then take their interior formula alone (deep_v), so the
// dual-number types carry only the lanes that formula has.
template <class T, uint32_t MASK, int OI, int OJ, int OK>
struct AccTDeep : AccT<T, MASK, OI, OJ, OK> {
  static constexpr bool deep = true;
};
template <class AC, class = void> struct deep_acc : std::false_type {};
template <class AC> struct deep_acc<AC, std::void_t<decltype(AC::deep)>> : std::bool_constant<AC::deep> {};
template <class AC> inline constexpr bool deep_v = deep_acc<AC>::value;

// Accessor seeding every field of MASK at the offset (SI, SJ, SK): 3-D fields
// at (SI, SJ, SK), 2-D fields at (SI, SJ) (the caller leaves 2-D bits out of
// MASK when SK != 0).
template <class T, uint32_t MASK, int SI, int SJ, int SK>
struct AccM {
  const FieldTab<T>* tab;
  Pos p;
  Vert<T> vert;
  const T* ramp;
  __device__ __forceinline__ int ci(int i) const { return i < 0 ? 0 : (i >= p.im ? p.im - 1 : i); }
  __device__ __forceinline__ int cj(int j) const { return j < 0 ? 0 : (j >= p.jm ? p.jm - 1 : j); }
  __device__ __forceinline__ int ck(int k) const { return k < 0 ? 0 : (k >= KB ? KB - 1 : k); }
  template <int F, int DI, int DJ, int DK>
  __device__ __forceinline__ auto at() const {
    T x = ld_at(tab->f[F], p.o3, lix3(p, ci(p.i + DI), cj(p.j + DJ), ck(p.k + DK)));
    if constexpr (((MASK >> F) & 1u) && DI == SI && DJ == SJ && DK == SK) return DS<T, (1u << F)>{x, {T(1)}};
    else return P<T>{x};
  }
  template <int F, int DI, int DJ>
  __device__ __forceinline__ auto at2() const {
    T x = ld_at(tab->f[F], p.o2, lix2(p, ci(p.i + DI), cj(p.j + DJ)));
    if constexpr (((MASK >> F) & 1u) && DI == SI && DJ == SJ) return DS<T, (1u << F)>{x, {T(1)}};
    else return P<T>{x};
  }
  template <int G, int DI, int DJ>
  __device__ __forceinline__ P<T> g() const {
    return P<T>{ld_at(tab->g[G], p.o2, lix2(p, ci(p.i + DI), cj(p.j + DJ)))};
  }
  __device__ __forceinline__ T z(int k) const { return vert.z(p.m, k); }
  __device__ __forceinline__ T zz(int k) const { return vert.zz(p.m, k); }
  __device__ __forceinline__ T dz(int k) const { return vert.dz(p.m, k); }
  __device__ __forceinline__ T dzz(int k) const { return vert.dzz(p.m, k); }
  __device__ __forceinline__ T rampv() const { return ramp[p.m]; }
};

// compile-time loop
template <int A, int B, class F> __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (A <= B) {
    f(std::integral_constant<int, A>{});
    static_for<A + 1, B>(f);
  }
}

template <int A, int B, class F> inline void static_for_host(F&& f) {
  if constexpr (A <= B) {
    f(std::integral_constant<int, A>{});
    static_for_host<A + 1, B>(f);
  }
}

// shorthand inside point functions (A is the accessor object)
// value of f() if cond, else zero of the same activity type.  For passive
// values (the forward maps) both sides are evaluated and the result
// selected: every accessor clamps its indices, so the unselected side only
// reads valid memory (its value, possibly inf or NaN at a boundary, is
// discarded), and without a branch the compiler issues all loads of an
// operator together.  Active values (the adjoint gathers) branch, since
// their tangent lanes make the unselected side expensive.
template <class F> __device__ __forceinline__ auto ifz(bool cond, F f) {
  using R = decltype(f());
  if constexpr (active_v<R>) {
    R r{};
    if (cond) r = f();
    return r;
  } else {
    const R v = f();
    return cond ? v : R{};
  }
}
template <class FA, class FB> __device__ __forceinline__ auto br(bool cond, FA fa, FB fb) {
  using R = common_t<decltype(fa()), decltype(fb())>;
  if constexpr (active_v<R>) {
    if (cond) return cvt<R>(fa());
    return cvt<R>(fb());
  } else {
    const R a = cvt<R>(fa()), b = cvt<R>(fb());
    return cond ? a : b;
  }
}
#define IFZ(cond, ...) ::pom::ifz((cond), [&]() { return (__VA_ARGS__); })
// arguments holding template commas must be parenthesised
#define BR(cond, ea, eb) ::pom::br((cond), [&]() { return (ea); }, [&]() { return (eb); })
#define AT(F, di, dj, dk) (A.template at<F, di, dj, dk>())
#define AT2(F, di, dj) (A.template at2<F, di, dj>())
#define G2(G, di, dj) (A.template g<G, di, dj>())

}  // namespace pom
