// Point functions of the explicit POM step (stencil operators).
//
// Each operator struct defines
//   ids of its state inputs (3-D ids first, then 2-D ids) and geometry ids,
//   NOUT outputs, OUT3 (outputs are 3-D) and the adjoint stencil radius,
//   template <int O> eval(A, prm): output O at the accessor's point.
// eval() defines the output at EVERY point of the member (boundary rules and
// untouched entries included), so the forward kernel is a plain map and the
// adjoint gather needs no separate "valid" rule.  Arithmetic order follows the
// reference line by line (cited per operator) so float64 runs agree with the
// reference to round-off.
#pragma once
#include "pom_common.h"

namespace pom {

template <class AC> __device__ __forceinline__ bool interior(const AC& A) {
  return A.p.i >= 1 && A.p.i <= A.p.im - 2 && A.p.j >= 1 && A.p.j <= A.p.jm - 2;
}

// =====================================================================
// advct: horizontal momentum advection + diffusion + curvature
// reference: src/neuralpom/cores/explicit/baroclinic.py:14-260
// =====================================================================
struct OpAdvct {
  enum { U, V, UB, VB, AAM, N3, DT = N3, NIN };
  enum { GDX, GDY, GARU, GARV };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 2, RJ = 2, RK = 0;

  // n / d for a geometry-only denominator d: n times the lane reciprocal of d
  // (in float the hardware reciprocal, within 2 ulp of the quotient, instead
  // of the ten-instruction IEEE division; in double the exact reciprocal)
  template <class N, class T> __device__ static auto qdiv(const N& n, P<T> d) { return n * P<T>{rcp_lane(d.v)}; }

  // curvature term at (i+OI, j): baroclinic.py:57-69
  template <int OI, int OJ, class AC> __device__ static auto curv(const AC& A) {
    auto t1 = (AT(V, OI, OJ + 1, 0) + AT(V, OI, OJ, 0)) * (G2(GDY, OI + 1, OJ) - G2(GDY, OI - 1, OJ));
    auto t2 = (AT(U, OI + 1, OJ, 0) + AT(U, OI, OJ, 0)) * (G2(GDX, OI, OJ + 1) - G2(GDX, OI, OJ - 1));
    return qdiv(0.25 * (t1 - t2), G2(GDX, OI, OJ) * G2(GDY, OI, OJ) + 1e-20);
  }
  // u-equation x-flux at (i+OI, j), valid for i+OI in [1, im-2]: :76-84, :114-119, :139
  template <int OI, class AC> __device__ static auto xfu(const AC& A) {
    auto tu1 = (AT2(DT, OI + 1, 0) + AT2(DT, OI, 0)) * AT(U, OI + 1, 0, 0);
    auto tu2 = (AT2(DT, OI, 0) + AT2(DT, OI - 1, 0)) * AT(U, OI, 0, 0);
    auto adv = 0.125 * (tu1 + tu2) * (AT(U, OI + 1, 0, 0) + AT(U, OI, 0, 0));
    auto gu = qdiv(AT(UB, OI + 1, 0, 0) - AT(UB, OI, 0, 0), G2(GDX, OI, 0) + 1e-20);
    auto visc = AT2(DT, OI, 0) * AT(AAM, OI, 0, 0) * 2.0 * gu;
    return (adv - visc) * G2(GDY, OI, 0);
  }
  // dt-weighted corner viscosity at (i+OI, j+OJ): :121-134 / :212-223
  template <int OI, int OJ, class AC> __device__ static auto corner_visc(const AC& A) {
    auto dtavg = 0.25 * (AT2(DT, OI, OJ) + AT2(DT, OI - 1, OJ) + AT2(DT, OI, OJ - 1) + AT2(DT, OI - 1, OJ - 1));
    auto asum = AT(AAM, OI, OJ, 0) + AT(AAM, OI - 1, OJ, 0) + AT(AAM, OI, OJ - 1, 0) + AT(AAM, OI - 1, OJ - 1, 0);
    auto dtaam = dtavg * asum;
    auto ddy = G2(GDY, OI, OJ) + G2(GDY, OI - 1, OJ) + G2(GDY, OI, OJ - 1) + G2(GDY, OI - 1, OJ - 1);
    auto ddx = G2(GDX, OI, OJ) + G2(GDX, OI - 1, OJ) + G2(GDX, OI, OJ - 1) + G2(GDX, OI - 1, OJ - 1);
    auto du_dy = qdiv(AT(UB, OI, OJ, 0) - AT(UB, OI, OJ - 1, 0), ddy + 1e-20);
    auto dv_dx = qdiv(AT(VB, OI, OJ, 0) - AT(VB, OI - 1, OJ, 0), ddx + 1e-20);
    return dtaam * (du_dy + dv_dx);
  }
  // u-equation y-flux at (i, j+OJ), j+OJ in [1, jm-1], i in [1, im-2]: :86-105, :127-141
  template <int OJ, class AC> __device__ static auto yfu(const AC& A) {
    auto tv1 = (AT2(DT, 0, OJ) + AT2(DT, 0, OJ - 1)) * AT(V, 0, OJ, 0);
    auto tv2 = (AT2(DT, -1, OJ) + AT2(DT, -1, OJ - 1)) * AT(V, -1, OJ, 0);
    auto adv = 0.125 * (tv1 + tv2) * (AT(U, 0, OJ, 0) + AT(U, 0, OJ - 1, 0));
    auto dxs = G2(GDX, 0, OJ) + G2(GDX, -1, OJ) + G2(GDX, 0, OJ - 1) + G2(GDX, -1, OJ - 1);
    return (adv - corner_visc<0, OJ>(A)) * (0.25 * dxs);
  }
  // v-equation x-flux at (i+OI, j), i+OI in [1, im-1], j in [1, jm-2]: :174-192, :205-232
  template <int OI, class AC> __device__ static auto xfv(const AC& A) {
    auto tu1 = (AT2(DT, OI, 0) + AT2(DT, OI - 1, 0)) * AT(U, OI, 0, 0);
    auto tu2 = (AT2(DT, OI, -1) + AT2(DT, OI - 1, -1)) * AT(U, OI, -1, 0);
    auto adv = 0.125 * (tu1 + tu2) * (AT(V, OI, 0, 0) + AT(V, OI - 1, 0, 0));
    auto dys = G2(GDY, OI, 0) + G2(GDY, OI - 1, 0) + G2(GDY, OI, -1) + G2(GDY, OI - 1, -1);
    return (adv - corner_visc<OI, 0>(A)) * (0.25 * dys);
  }
  // v-equation y-flux at (i, j+OJ), j+OJ in [1, jm-2]: :194-203, :225-233
  template <int OJ, class AC> __device__ static auto yfv(const AC& A) {
    auto tv1 = (AT2(DT, 0, OJ + 1) + AT2(DT, 0, OJ)) * AT(V, 0, OJ + 1, 0);
    auto tv2 = (AT2(DT, 0, OJ) + AT2(DT, 0, OJ - 1)) * AT(V, 0, OJ, 0);
    auto adv = 0.125 * (tv1 + tv2) * (AT(V, 0, OJ + 1, 0) + AT(V, 0, OJ, 0));
    auto gv = qdiv(AT(VB, 0, OJ + 1, 0) - AT(VB, 0, OJ, 0), G2(GDY, 0, OJ) + 1e-20);
    auto visc = AT2(DT, 0, OJ) * AT(AAM, 0, OJ, 0) * 2.0 * gv;
    return (adv - visc) * G2(GDX, 0, OJ);
  }

  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    const int i = A.p.i, j = A.p.j, k = A.p.k;
    const bool in = interior(A) && k < KBM1;
    if constexpr (O == 0) {
      // advx: :143-163
      return IFZ(in, ([&]() {
        auto r = xfu<0>(A) - IFZ(i >= 2, xfu<-1>(A)) + yfu<1>(A) - yfu<0>(A);
        auto c1 = curv<0, 0>(A) * AT2(DT, 0, 0) * (AT(V, 0, 1, 0) + AT(V, 0, 0, 0));
        auto c2 = curv<-1, 0>(A) * AT2(DT, -1, 0) * (AT(V, -1, 1, 0) + AT(V, -1, 0, 0));
        auto corr = G2(GARU, 0, 0) * 0.25 * (c1 + c2);
        return BR(i >= 2, r - corr, r);
      }()));
    } else {
      // advy: :235-256
      return IFZ(in, ([&]() {
        auto r = xfv<1>(A) - xfv<0>(A) + yfv<0>(A) - IFZ(j >= 2, yfv<-1>(A));
        auto c1 = curv<0, 0>(A) * AT2(DT, 0, 0) * (AT(U, 1, 0, 0) + AT(U, 0, 0, 0));
        auto c2 = curv<0, -1>(A) * AT2(DT, 0, -1) * (AT(U, 1, -1, 0) + AT(U, 0, -1, 0));
        auto corr = G2(GARV, 0, 0) * 0.25 * (c1 + c2);
        return BR(j >= 2, r + corr, r);
      }()));
    }
  }
};

// =====================================================================
// Smagorinsky horizontal viscosity
// reference: src/neuralpom/cores/explicit/time_loop.py:10-16, 192-233
// =====================================================================
template <class T> __device__ __forceinline__ P<T> norm3(P<T> a, P<T> b, P<T> c) {
  return {sqrt(a.v * a.v + b.v * b.v + c.v * c.v)};
}
template <class X, class Y, class Z> __device__ __forceinline__ auto norm3(X a, Y b, Z c) {
  // torch.linalg.vector_norm: d|x| = x.dx/|x|, zero gradient at |x| = 0
  using T = typename scalar_of<X>::type;
  T av = val(a), bv = val(b), cv = val(c);
  T n = sqrt(av * av + bv * bv + cv * cv);
  if constexpr ((mask_of<X>::value | mask_of<Y>::value | mask_of<Z>::value) != 0u) {
    auto num = a * P<T>{av} + b * P<T>{bv} + c * P<T>{cv};
    auto r = num;
    r.v = n;
    for (int i = 0; i < popc(mask_of<decltype(num)>::value); ++i) r.d[i] = n > T(0) ? num.d[i] / n : T(0);
    return r;
  } else {
    T d = n > T(0) ? (av * tan_(a) + bv * tan_(b) + cv * tan_(c)) / n : T(0);
    return D<T>{n, d};
  }
}

struct OpSmag {
  enum { U, V, AAM, N3, NIN = N3 };
  enum { GDX, GDY };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const bool in = interior(A) && A.p.k < KBM1;
    auto f = [&]() {
      auto dx = G2(GDX, 0, 0), dy = G2(GDY, 0, 0);
      auto du_dx = (AT(U, 1, 0, 0) - AT(U, 0, 0, 0)) / dx;
      auto dv_dy = (AT(V, 0, 1, 0) - AT(V, 0, 0, 0)) / dy;
      auto ut = 0.25 * (AT(U, 0, 1, 0) + AT(U, 1, 1, 0) - AT(U, 0, -1, 0) - AT(U, 1, -1, 0)) / dy;
      auto vt = 0.25 * (AT(V, 1, 0, 0) + AT(V, 1, 1, 0) - AT(V, -1, 0, 0) - AT(V, -1, 1, 0)) / dx;
      auto sh = (ut + vt) * 0.70710678118654757;
      return prm.horcon * dx * dy * norm3(du_dx, dv_dy, sh);
    };
    return BR(in, f(), AT(AAM, 0, 0, 0));
  }
};

// =====================================================================
// advq: turbulence advection + diffusion
// reference: src/neuralpom/cores/explicit/tracers.py:14-156
// =====================================================================
struct OpAdvq {
  enum { QB, Q, U, V, W, AAM, N3, DT = N3, ETB, ETF, NIN };
  enum { GDX, GDY, GH, GART, GDUM, GDVM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 1;
  template <int OI, class AC> __device__ static auto xf(const AC& A) {   // :41-47, :62-75, :94-95
    auto adv = 0.125 * (AT(Q, OI, 0, 0) + AT(Q, OI - 1, 0, 0)) * (AT2(DT, OI, 0) + AT2(DT, OI - 1, 0)) *
               (AT(U, OI, 0, 0) + AT(U, OI, 0, -1));
    auto asum = AT(AAM, OI, 0, 0) + AT(AAM, OI - 1, 0, 0) + AT(AAM, OI, 0, -1) + AT(AAM, OI - 1, 0, -1);
    auto dif = 0.25 * asum * (G2(GH, OI, 0) + G2(GH, OI - 1, 0)) * (AT(QB, OI, 0, 0) - AT(QB, OI - 1, 0, 0)) *
               G2(GDUM, OI, 0) / (G2(GDX, OI, 0) + G2(GDX, OI - 1, 0));
    return (adv - dif) * (0.5 * (G2(GDY, OI, 0) + G2(GDY, OI - 1, 0)));
  }
  template <int OJ, class AC> __device__ static auto yf(const AC& A) {   // :49-56, :77-99
    auto adv = 0.125 * (AT(Q, 0, OJ, 0) + AT(Q, 0, OJ - 1, 0)) * (AT2(DT, 0, OJ) + AT2(DT, 0, OJ - 1)) *
               (AT(V, 0, OJ, 0) + AT(V, 0, OJ, -1));
    auto asum = AT(AAM, 0, OJ, 0) + AT(AAM, 0, OJ - 1, 0) + AT(AAM, 0, OJ, -1) + AT(AAM, 0, OJ - 1, -1);
    auto dif = 0.25 * asum * (G2(GH, 0, OJ) + G2(GH, 0, OJ - 1)) * (AT(QB, 0, OJ, 0) - AT(QB, 0, OJ - 1, 0)) *
               G2(GDVM, 0, OJ) / (G2(GDY, 0, OJ) + G2(GDY, 0, OJ - 1));
    return (adv - dif) * (0.5 * (G2(GDX, 0, OJ) + G2(GDX, 0, OJ - 1)));
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int k = A.p.k;
    const bool in = interior(A) && k >= 1 && k < KBM1;
    return IFZ(in, ([&]() {   // :124-154
      auto art = G2(GART, 0, 0);
      auto vert = (AT(W, 0, 0, -1) * AT(Q, 0, 0, -1) - AT(W, 0, 0, 1) * AT(Q, 0, 0, 1)) * art /
                  (A.dz(k) + A.dz(k - 1));
      auto hdiv = xf<1>(A) - xf<0>(A) + yf<1>(A) - yf<0>(A);
      auto tb = (G2(GH, 0, 0) + AT2(ETB, 0, 0)) * art;
      auto tf = (G2(GH, 0, 0) + AT2(ETF, 0, 0)) * art;
      return (tb * AT(QB, 0, 0, 0) - prm.dti2 * (vert + hdiv)) / (tf + 1e-20);
    }()));
  }
};

// =====================================================================
// advt1: centred tracer advection + anomaly diffusion
// reference: src/neuralpom/cores/explicit/tracers.py:433-569
// =====================================================================
struct OpAdvt1 {
  enum { FB, F, FCLIM, U, V, W, AAM, N3, DT = N3, ETB, ETF, NIN };
  enum { GDX, GDY, GH, GART, GDUM, GDVM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 1;
  template <int OI, class AC> __device__ static auto xf(const AC& A, double tprni) {  // :479-508
    auto adv = 0.25 * (AT2(DT, OI, 0) + AT2(DT, OI - 1, 0)) * (AT(F, OI, 0, 0) + AT(F, OI - 1, 0, 0)) *
               AT(U, OI, 0, 0);
    auto fa0 = AT(FB, OI, 0, 0) - AT(FCLIM, OI, 0, 0);
    auto fa1 = AT(FB, OI - 1, 0, 0) - AT(FCLIM, OI - 1, 0, 0);
    auto dif = 0.5 * (AT(AAM, OI, 0, 0) + AT(AAM, OI - 1, 0, 0)) * (G2(GH, OI, 0) + G2(GH, OI - 1, 0)) * tprni *
               (fa0 - fa1) * G2(GDUM, OI, 0) / (G2(GDX, OI, 0) + G2(GDX, OI - 1, 0));
    return (adv - dif) * (0.5 * (G2(GDY, OI, 0) + G2(GDY, OI - 1, 0)));
  }
  template <int OJ, class AC> __device__ static auto yf(const AC& A, double tprni) {  // :484-522
    auto adv = 0.25 * (AT2(DT, 0, OJ) + AT2(DT, 0, OJ - 1)) * (AT(F, 0, OJ, 0) + AT(F, 0, OJ - 1, 0)) *
               AT(V, 0, OJ, 0);
    auto fa0 = AT(FB, 0, OJ, 0) - AT(FCLIM, 0, OJ, 0);
    auto fa1 = AT(FB, 0, OJ - 1, 0) - AT(FCLIM, 0, OJ - 1, 0);
    auto dif = 0.5 * (AT(AAM, 0, OJ, 0) + AT(AAM, 0, OJ - 1, 0)) * (G2(GH, 0, OJ) + G2(GH, 0, OJ - 1)) * tprni *
               (fa0 - fa1) * G2(GDVM, 0, OJ) / (G2(GDY, 0, OJ) + G2(GDY, 0, OJ - 1));
    return (adv - dif) * (0.5 * (G2(GDX, 0, OJ) + G2(GDX, 0, OJ - 1)));
  }
  // vertical flux at level k+OK: :530-541
  template <int OK, class AC> __device__ static auto zf(const AC& A) {
    const int kk = A.p.k + OK;
    auto art = G2(GART, 0, 0);
    auto top = AT(F, 0, 0, OK) * AT(W, 0, 0, OK) * art;
    auto mid = 0.5 * (AT(F, 0, 0, OK - 1) + AT(F, 0, 0, OK)) * AT(W, 0, 0, OK) * art;
    return IFZ(kk < KBM1, BR(kk == 0, top, mid));
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int k = A.p.k;
    const bool in = interior(A) && k < KBM1;
    return IFZ(in, ([&]() {   // :546-567
      auto hdiv = xf<1>(A, prm.tprni) - xf<0>(A, prm.tprni) + yf<1>(A, prm.tprni) - yf<0>(A, prm.tprni);
      auto vdiv = (zf<0>(A) - zf<1>(A)) / A.dz(k);
      auto art = G2(GART, 0, 0);
      auto tb = (G2(GH, 0, 0) + AT2(ETB, 0, 0)) * art;
      auto tf = (G2(GH, 0, 0) + AT2(ETF, 0, 0)) * art;
      return (tb * AT(FB, 0, 0, 0) - prm.dti2 * (hdiv + vdiv)) / (tf + 1e-20);
    }()));
  }
};

// =====================================================================
// advu / advv: internal momentum tendencies and leapfrog
// reference: src/neuralpom/cores/explicit/momentum.py:14-248
// =====================================================================
struct OpAdvu {
  enum { ADVX, DRHOX, U, V, W, UB, N3, DT = N3, EGF, EGB, ETF, ETB, NIN };
  enum { GARU, GCOR, GDY, GH, GEATM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 1;
  template <int OK, class AC> __device__ static auto vfl(const AC& A) {   // :45-51
    const int kk = A.p.k + OK;
    return IFZ(kk >= 1 && kk < KBM1,
               0.25 * (AT(W, 0, 0, OK) + AT(W, -1, 0, OK)) * (AT(U, 0, 0, OK) + AT(U, 0, 0, OK - 1)));
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int k = A.p.k;
    const bool in = interior(A) && k < KBM1;
    return IFZ(in, ([&]() {   // :67-129
      auto aru = G2(GARU, 0, 0);
      auto vdiv = (vfl<0>(A) - vfl<1>(A)) * aru / A.dz(k);
      auto c1 = (G2(GCOR, 0, 0) * AT2(DT, 0, 0)) * (AT(V, 0, 1, 0) + AT(V, 0, 0, 0));
      auto c2 = (G2(GCOR, -1, 0) * AT2(DT, -1, 0)) * (AT(V, -1, 1, 0) + AT(V, -1, 0, 0));
      auto corio = -aru * 0.25 * (c1 + c2);
      auto geg = (AT2(EGF, 0, 0) - AT2(EGF, -1, 0)) + (AT2(EGB, 0, 0) - AT2(EGB, -1, 0));
      auto gpa = 2.0 * (G2(GEATM, 0, 0) - G2(GEATM, -1, 0));
      auto pg = (prm.grav * 0.125) * ((AT2(DT, 0, 0) + AT2(DT, -1, 0)) * (geg + gpa) * (G2(GDY, 0, 0) + G2(GDY, -1, 0)));
      auto rhs = AT(ADVX, 0, 0, 0) + vdiv + corio + pg + AT(DRHOX, 0, 0, 0);
      auto hs = G2(GH, 0, 0) + G2(GH, -1, 0);
      auto dnew = (hs + AT2(ETF, 0, 0) + AT2(ETF, -1, 0)) * aru;
      auto dold = (hs + AT2(ETB, 0, 0) + AT2(ETB, -1, 0)) * aru;
      return (dold * AT(UB, 0, 0, 0) - (2.0 * prm.dti2) * rhs) / (dnew + 1e-20);
    }()));
  }
};

struct OpAdvv {
  enum { ADVY, DRHOY, U, V, W, VB, N3, DT = N3, EGF, EGB, ETF, ETB, NIN };
  enum { GARV, GCOR, GDX, GH, GEATM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 1;
  template <int OK, class AC> __device__ static auto vfl(const AC& A) {   // :164-169
    const int kk = A.p.k + OK;
    return IFZ(kk >= 1 && kk < KBM1,
               0.25 * (AT(W, 0, 0, OK) + AT(W, 0, -1, OK)) * (AT(V, 0, 0, OK) + AT(V, 0, 0, OK - 1)));
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int k = A.p.k;
    const bool in = interior(A) && k < KBM1;
    return IFZ(in, ([&]() {   // :185-246
      auto arv = G2(GARV, 0, 0);
      auto vdiv = (vfl<0>(A) - vfl<1>(A)) * arv / A.dz(k);
      auto c1 = (G2(GCOR, 0, 0) * AT2(DT, 0, 0)) * (AT(U, 1, 0, 0) + AT(U, 0, 0, 0));
      auto c2 = (G2(GCOR, 0, -1) * AT2(DT, 0, -1)) * (AT(U, 1, -1, 0) + AT(U, 0, -1, 0));
      auto corio = arv * 0.25 * (c1 + c2);
      auto geg = (AT2(EGF, 0, 0) - AT2(EGF, 0, -1)) + (AT2(EGB, 0, 0) - AT2(EGB, 0, -1));
      auto gpa = 2.0 * (G2(GEATM, 0, 0) - G2(GEATM, 0, -1));
      auto pg = (prm.grav * 0.125) * ((AT2(DT, 0, 0) + AT2(DT, 0, -1)) * (geg + gpa) * (G2(GDX, 0, 0) + G2(GDX, 0, -1)));
      auto rhs = AT(ADVY, 0, 0, 0) + vdiv + corio + pg + AT(DRHOY, 0, 0, 0);
      auto hs = G2(GH, 0, 0) + G2(GH, 0, -1);
      auto dnew = (hs + AT2(ETF, 0, 0) + AT2(ETF, 0, -1)) * arv;
      auto dold = (hs + AT2(ETB, 0, 0) + AT2(ETB, 0, -1)) * arv;
      return (dold * AT(VB, 0, 0, 0) - (2.0 * prm.dti2) * rhs) / (dnew + 1e-20);
    }()));
  }
};

// =====================================================================
// dens: Mellor (1991) equation of state
// reference: src/neuralpom/cores/explicit/momentum.py:467-541
// =====================================================================
template <class X> __device__ __forceinline__ auto abs_pow15(X x) {
  // safe_abs_pow(x, 1.5) = clamp(|x|, 1e-8) ** 1.5  (utils/safety.py:29-30)
  using T = typename scalar_of<X>::type;
  auto a = clamp_min(pabs(x), T(1e-8));
  T r = pow(val(a), T(1.5));
  if constexpr (mask_of<decltype(a)>::value != 0u) {
    auto o = lanes_scaled(a, T(1.5) * sqrt(val(a)));
    o.v = r;
    return o;
  } else if constexpr (active_v<decltype(a)>) {
    return D<T>{r, T(1.5) * sqrt(val(a)) * a.d};
  } else {
    return P<T>{r};
  }
}

struct OpDens {
  enum { TT, SS, RHO, N3, NIN = N3 };
  enum { GH, GFSM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 0, RJ = 0, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int k = A.p.k;
    auto f = [&]() {
      auto tr = AT(TT, 0, 0, 0) + prm.tbias;
      auto sr = AT(SS, 0, 0, 0) + prm.sbias;
      auto tr2 = tr * tr;
      auto tr3 = tr2 * tr;
      auto tr4 = tr3 * tr;
      auto p = (prm.grav * prm.rhoref) * (-A.zz(k) * G2(GH, 0, 0)) * 1.0e-5;
      auto r0 = -0.157406 + 6.793952e-2 * tr - 9.095290e-3 * tr2 + 1.001685e-4 * tr3 - 1.120083e-6 * tr4 +
                6.536332e-9 * tr4 * tr;
      auto r1 = r0 + (0.824493 - 4.0899e-3 * tr + 7.6438e-5 * tr2 - 8.2467e-7 * tr3 + 5.3875e-9 * tr4) * sr;
      auto r2 = r1 + (-5.72466e-3 + 1.0227e-4 * tr - 1.6546e-6 * tr2) * abs_pow15(sr);
      auto r3 = r2 + 4.8314e-4 * sr * sr;
      auto cr = 1449.1 + 0.0821 * p + 4.55 * tr - 0.045 * tr2 + 1.34 * (sr - 35.0);
      auto cr2 = cr * cr;
      auto tp = 1.0e5 * p / cr2;
      auto r4 = r3 + tp * (1.0 - 2.0 * p / cr2);
      return (r4 / prm.rhoref) * G2(GFSM, 0, 0);
    };
    return BR(k < KBM1, f(), AT(RHO, 0, 0, 0));
  }
};

// =====================================================================
// bcond 3: internal velocity lateral boundaries + masks
// reference: src/neuralpom/cores/explicit/baroclinic.py:576-661.
// DEPTH: the case-adapter kernel copy, whose edge weight ga takes the total
// depth d (a differentiated 2-D input) instead of h
// (case_adapters/kernels/coastal_explicit_baroclinic.py:586, 602, 627, 641)
// =====================================================================
template <bool DEPTH> struct OpBc3G {
  enum { UF, VF, U, V, N3, D_ = N3, NIN = DEPTH ? N3 + 1 : N3 };
  enum { GH, GDUM, GDVM };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 2, RJ = 2, RK = 0;
  static constexpr int EDGE_BAND = 2;   // only points within 2 of the edge read neighbours
  template <int F, int OI, class AC> __device__ static auto smj(const AC& A) {
    return 0.25 * AT(F, OI, -1, 0) + 0.50 * AT(F, OI, 0, 0) + 0.25 * AT(F, OI, 1, 0);
  }
  template <int F, int OJ, class AC> __device__ static auto smi(const AC& A) {
    return 0.25 * AT(F, -1, OJ, 0) + 0.50 * AT(F, 0, OJ, 0) + 0.25 * AT(F, 1, OJ, 0);
  }
  // ga = safe_sqrt(depth / hmax) at (OI, OJ) (utils/safety.py:25-26)
  template <int OI, int OJ, class AC> __device__ static auto ga(const AC& A, double hmax) {
    if constexpr (DEPTH) {
      auto x = AT2(D_, OI, OJ) / hmax;
      using T = typename scalar_of<decltype(x)>::type;
      return BR(val(x) > T(1e-8), psqrt(x), P<T>{T(sqrt(T(1e-8)))});
    } else {
      auto h = G2(GH, OI, OJ);
      using T = typename scalar_of<decltype(h)>::type;
      return P<T>{sqrt(fmax(val(h) / T(hmax), T(1e-8)))};
    }
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    if constexpr (deep_v<AC>) {   // away from the edges: the masked interior values
      if constexpr (O == 0) return AT(UF, 0, 0, 0) * G2(GDUM, 0, 0);
      else return AT(VF, 0, 0, 0) * G2(GDVM, 0, 0);
    } else {
      return eval_edge<O>(A, prm);
    }
  }
  template <int O, class AC> __device__ static auto eval_edge(const AC& A, const Params& prm) {
    const int i = A.p.i, j = A.p.j, k = A.p.k, im = A.p.im, jm = A.p.jm;
    const bool kin = k < KBM1;
    const bool jin = j >= 1 && j <= jm - 2, iin = i >= 1 && i <= im - 2;
    if constexpr (O == 0) {
      auto east = [&]() {   // :585-599
        auto g = ga<0, 0>(A, prm.hmax);
        return g * smj<U, -1>(A) + (1.0 - g) * smj<U, 0>(A);
      };
      auto west1 = [&]() {  // :601-617 at i = 1
        auto g = ga<-1, 0>(A, prm.hmax);
        return g * smj<U, 1>(A) + (1.0 - g) * smj<U, 0>(A);
      };
      auto west0 = [&]() {  // :618 uf(0) = uf(1)
        auto g = ga<0, 0>(A, prm.hmax);
        return g * smj<U, 2>(A) + (1.0 - g) * smj<U, 1>(A);
      };
      using R = common_t<common_t<common_t<decltype(east()), decltype(west1())>, decltype(west0())>,
                         decltype(AT(UF, 0, 0, 0))>;
      R r = cvt<R>(AT(UF, 0, 0, 0));
      if (kin) {
        if (jin && i == im - 1) r = cvt<R>(east());
        else if (jin && i == 1) r = cvt<R>(west1());
        else if (jin && i == 0) r = cvt<R>(west0());
        else if (iin && (j == jm - 1 || j == 0)) r = R{};
      }
      return r * G2(GDUM, 0, 0);
    } else {
      auto north = [&]() {  // :626-638
        auto g = ga<0, 0>(A, prm.hmax);
        return g * smi<V, -1>(A) + (1.0 - g) * smi<V, 0>(A);
      };
      auto south1 = [&]() { // :640-652 at j = 1
        auto g = ga<0, -1>(A, prm.hmax);
        return g * smi<V, 1>(A) + (1.0 - g) * smi<V, 0>(A);
      };
      auto south0 = [&]() { // :653 vf(:,0) = vf(:,1)
        auto g = ga<0, 0>(A, prm.hmax);
        return g * smi<V, 2>(A) + (1.0 - g) * smi<V, 1>(A);
      };
      using R = common_t<common_t<common_t<decltype(north()), decltype(south1())>, decltype(south0())>,
                         decltype(AT(VF, 0, 0, 0))>;
      R r = cvt<R>(AT(VF, 0, 0, 0));
      if (kin) {
        if (jin && (i == im - 1 || i == 0)) r = R{};
        else if (iin && j == jm - 1) r = cvt<R>(north());
        else if (iin && j == 1) r = cvt<R>(south1());
        else if (iin && j == 0) r = cvt<R>(south0());
      }
      return r * G2(GDVM, 0, 0);
    }
  }
};
using OpBc3 = OpBc3G<false>;
using OpBc3D = OpBc3G<true>;

// =====================================================================
// bcond 4 (upstream T/S at the edges) fused with the T/S Asselin filter
// reference: baroclinic.py:666-844; time_loop.py:587-601; advt1's bottom
// copy f(kb) = f(kbm1) (tracers.py:462-465)
// outputs: 0 t_new, 1 tb_new, 2 s_new, 3 sb_new
// =====================================================================
// EXT: exterior tracer values tbe/tbw/tbn/tbs (sbe...) of open boundaries
// (explicit/base.py:181-189), as 3-D fields TX / SX holding each edge cell's
// exterior value; without EXT they are zero (closed basins).
template <bool EXT> struct OpBc4FilterG {
  enum { TF, SF, TT, SS, TB, SB, U, V, W, TX, SX, N3 = EXT ? SX + 1 : TX, DT = N3, NIN };
  enum { GDX, GDY, GFSM };
  static constexpr int NOUT = 4;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 1;
  static constexpr int EDGE_BAND = 1;   // only points within 1 of the edge read neighbours

  // edge value of tracer FC (TT or SS) for k < kbm1; side: 0 E, 1 W, 2 N, 3 S
  template <int FC, int SIDE, class AC> __device__ static auto edge(const AC& A, const Params& prm) {
    const int k = A.p.k;
    constexpr int II = (SIDE == 0) ? -1 : (SIDE == 1 ? 1 : 0);   // inner neighbour
    constexpr int JJ = (SIDE == 2) ? -1 : (SIDE == 3 ? 1 : 0);
    // u1 = 2*vel*dti/(dsum): E uses u(im-1), W uses u(1), N v(jm-1), S v(1)
    auto vel = [&]() {
      if constexpr (SIDE == 0) return AT(U, 0, 0, 0);
      else if constexpr (SIDE == 1) return AT(U, 1, 0, 0);
      else if constexpr (SIDE == 2) return AT(V, 0, 0, 0);
      else return AT(V, 0, 1, 0);
    }();
    auto dsum = [&]() {
      if constexpr (SIDE <= 1) return G2(GDX, 0, 0) + G2(GDX, II, 0);
      else return G2(GDY, 0, 0) + G2(GDY, 0, JJ);
    }();
    auto u1 = 2.0 * vel * prm.dti / dsum;
    auto fc = AT(FC, 0, 0, 0);
    auto fi = AT(FC, II, JJ, 0);
    auto fx = [&]() {
      if constexpr (EXT) return AT(FC == TT ? TX : SX, 0, 0, 0);
      else return 0.0;
    }();
    auto f_in = BR(SIDE == 0 || SIDE == 2, fc - u1 * (fx - fc), fc - u1 * (fc - fx));
    auto f_out = BR(SIDE == 0 || SIDE == 2, fc - u1 * (fc - fi), fc - u1 * (fi - fc));
    const bool inflow = (SIDE == 0 || SIDE == 2) ? (val(u1) <= 0) : (val(u1) >= 0);
    auto res = BR(inflow, f_in, f_out);
    // vertical correction for outflow on k in [1, kbm1-2]
    const bool outflow = (SIDE == 0 || SIDE == 2) ? (val(u1) > 0) : (val(u1) < 0);
    auto corr = [&]() {
      auto wm = 0.5 * (AT(W, II, JJ, 0) + AT(W, II, JJ, 1)) * prm.dti /
                ((A.zz(k - 1) - A.zz(k + 1)) * AT2(DT, II, JJ));
      return wm * (AT(FC, II, JJ, -1) - AT(FC, II, JJ, 1));
    };
    return BR(k >= 1 && k <= KBM1 - 2 && outflow, res - corr(), res);
  }

  template <int FNEW, int FC, class AC> __device__ static auto tnew(const AC& A, const Params& prm) {
    if constexpr (deep_v<AC>) {   // away from the edges: the masked new value
      return AT(FNEW, 0, 0, 0) * G2(GFSM, 0, 0);
    } else {
      const int i = A.p.i, j = A.p.j, k = A.p.k, im = A.p.im, jm = A.p.jm;
      using R = common_t<common_t<common_t<decltype(edge<FC, 0>(A, prm)), decltype(edge<FC, 1>(A, prm))>,
                                  common_t<decltype(edge<FC, 2>(A, prm)), decltype(edge<FC, 3>(A, prm))>>,
                         decltype(AT(FNEW, 0, 0, 0))>;
      R r = cvt<R>(AT(FNEW, 0, 0, 0));
      if (k < KBM1) {  // later writes win: S > N > W > E
        if (j == 0) r = cvt<R>(edge<FC, 3>(A, prm));
        else if (j == jm - 1) r = cvt<R>(edge<FC, 2>(A, prm));
        else if (i == 0) r = cvt<R>(edge<FC, 1>(A, prm));
        else if (i == im - 1) r = cvt<R>(edge<FC, 0>(A, prm));
      }
      return r * G2(GFSM, 0, 0);
    }
  }

  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    constexpr int FNEW = (O < 2) ? TF : SF;
    constexpr int FC = (O < 2) ? TT : SS;
    constexpr int FBK = (O < 2) ? TB : SB;
    if constexpr (O == 0 || O == 2) {
      return tnew<FNEW, FC>(A, prm);
    } else {
      // filtered old level; t', tb' carry the bottom copy f(kb) = f(kbm1)
      const bool bot = A.p.k == KB - 1;
      auto tc = BR(bot, AT(FC, 0, 0, -1), AT(FC, 0, 0, 0));
      auto tbk = BR(bot, AT(FBK, 0, 0, -1), AT(FBK, 0, 0, 0));
      return tc + (0.5 * prm.smoth) * (tnew<FNEW, FC>(A, prm) + tbk - 2.0 * tc);
    }
  }
};
using OpBc4Filter = OpBc4FilterG<false>;
using OpBc4FilterExt = OpBc4FilterG<true>;

// =====================================================================
// bcond 6 (upstream q2/q2l at the edges, +1e-10) fused with the q2/q2l
// Asselin filter.  reference: baroclinic.py:856-925; time_loop.py:525-540
// outputs: 0 q2_new, 1 q2b_new, 2 q2l_new, 3 q2lb_new
// =====================================================================
struct OpBc6Filter {
  enum { Q2F, Q2LF, Q2, Q2L, Q2B, Q2LB, U, V, N3, NIN = N3 };
  enum { GDX, GDY, GFSM };
  static constexpr int NOUT = 4;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  static constexpr int EDGE_BAND = 1;   // only points within 1 of the edge read neighbours
  template <int FC, int SIDE, class AC> __device__ static auto edge(const AC& A, const Params& prm) {
    constexpr int II = (SIDE == 0) ? -1 : (SIDE == 1 ? 1 : 0);
    constexpr int JJ = (SIDE == 2) ? -1 : (SIDE == 3 ? 1 : 0);
    auto vel = [&]() {
      if constexpr (SIDE == 0) return AT(U, 0, 0, 0);
      else if constexpr (SIDE == 1) return AT(U, 1, 0, 0);
      else if constexpr (SIDE == 2) return AT(V, 0, 0, 0);
      else return AT(V, 0, 1, 0);
    }();
    auto dsum = [&]() {
      if constexpr (SIDE <= 1) return G2(GDX, 0, 0) + G2(GDX, II, 0);
      else return G2(GDY, 0, 0) + G2(GDY, 0, JJ);
    }();
    auto u1 = 2.0 * vel * prm.dti / dsum;
    auto qc = AT(FC, 0, 0, 0);
    auto qi = AT(FC, II, JJ, 0);
    const double sm = prm.small;
    auto q_in = BR(SIDE == 0 || SIDE == 2, qc - u1 * (sm - qc), qc - u1 * (qc - sm));
    auto q_out = BR(SIDE == 0 || SIDE == 2, qc - u1 * (qc - qi), qc - u1 * (qi - qc));
    const bool inflow = (SIDE == 0 || SIDE == 2) ? (val(u1) <= 0) : (val(u1) >= 0);
    return BR(inflow, q_in, q_out);
  }
  template <int FNEW, int FC, class AC> __device__ static auto qnew(const AC& A, const Params& prm) {
    if constexpr (deep_v<AC>) {   // away from the edges: the masked new value
      return AT(FNEW, 0, 0, 0) * G2(GFSM, 0, 0) + 1.e-10;
    } else {
      const int i = A.p.i, j = A.p.j, im = A.p.im, jm = A.p.jm;
      using R = common_t<common_t<common_t<decltype(edge<FC, 0>(A, prm)), decltype(edge<FC, 1>(A, prm))>,
                                  common_t<decltype(edge<FC, 2>(A, prm)), decltype(edge<FC, 3>(A, prm))>>,
                         decltype(AT(FNEW, 0, 0, 0))>;
      R r = cvt<R>(AT(FNEW, 0, 0, 0));
      if (j == 0) r = cvt<R>(edge<FC, 3>(A, prm));
      else if (j == jm - 1) r = cvt<R>(edge<FC, 2>(A, prm));
      else if (i == 0) r = cvt<R>(edge<FC, 1>(A, prm));
      else if (i == im - 1) r = cvt<R>(edge<FC, 0>(A, prm));
      return r * G2(GFSM, 0, 0) + 1.e-10;
    }
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    constexpr int FNEW = (O < 2) ? Q2F : Q2LF;
    constexpr int FC = (O < 2) ? Q2 : Q2L;
    constexpr int FBK = (O < 2) ? Q2B : Q2LB;
    if constexpr (O == 0 || O == 2) {
      return qnew<FNEW, FC>(A, prm);
    } else {
      auto q = AT(FC, 0, 0, 0);
      return q + (0.5 * prm.smoth) * (qnew<FNEW, FC>(A, prm) + AT(FBK, 0, 0, 0) - 2.0 * q);
    }
  }
};

// =====================================================================
// Outputs O0 and O1 of an operator as a two-output operator (same fields,
// stencil and edge band): bcond 4 and bcond 6 alone are outputs 0 and 2 of
// the fused boundary + filter operators.
// =====================================================================
template <class Base, int O0, int O1> struct OpPick : Base {
  static constexpr int NOUT = 2;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    return Base::template eval<O == 0 ? O0 : O1>(A, prm);
  }
};
using OpBcond4 = OpPick<OpBc4Filter, 0, 2>;   // uf, vf = new t, s with bcond 4 (baroclinic.py:576-846)
using OpBcond6 = OpPick<OpBc6Filter, 0, 2>;   // uf, vf = new q2, q2l with bcond 6 (baroclinic.py:856-925)

// bcond 5 (baroclinic.py:849-851): w(:, :, 0:kbm1) * fsm
struct OpBcond5 {
  enum { W, N3, NIN = N3 };
  enum { GFSM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 0, RJ = 0, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    return BR(A.p.k < KBM1, AT(W, 0, 0, 0) * G2(GFSM, 0, 0), AT(W, 0, 0, 0));
  }
};

// =====================================================================
// km/kh lateral edge copy (tracers.py:425-429): arr(edge) = arr(inner)*fsm,
// applied E, W, N, S in that order.  outputs: 0 km, 1 kh
// =====================================================================
struct OpKEdge {
  enum { KM, KH, N3, NIN = N3 };
  enum { GFSM };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 2, RJ = 2, RK = 0;
  static constexpr int EDGE_BAND = 1;   // only points within 1 of the edge read neighbours
  // value after the E/W copies at column (i, j+OJ)
  template <int F, int OJ, class AC> __device__ static auto ew(const AC& A) {
    const int i = A.p.i, im = A.p.im;
    using R = common_t<common_t<decltype(AT(F, 0, OJ, 0)), decltype(AT(F, -1, OJ, 0))>, decltype(AT(F, 1, OJ, 0))>;
    R r = cvt<R>(AT(F, 0, OJ, 0));
    if (i == im - 1) r = cvt<R>(AT(F, -1, OJ, 0) * G2(GFSM, 0, OJ));
    else if (i == 0) r = cvt<R>(AT(F, 1, OJ, 0) * G2(GFSM, 0, OJ));
    return r;
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    constexpr int F = (O == 0) ? KM : KH;
    if constexpr (deep_v<AC>) {   // away from the edges: the point's own value
      return AT(F, 0, 0, 0);
    } else {
      const int j = A.p.j, jm = A.p.jm;
      using R = common_t<common_t<decltype(ew<F, 0>(A)), decltype(ew<F, -1>(A))>, decltype(ew<F, 1>(A))>;
      R r = cvt<R>(ew<F, 0>(A));
      if (j == jm - 1) r = cvt<R>(ew<F, -1>(A) * G2(GFSM, 0, 0));
      else if (j == 0) r = cvt<R>(ew<F, 1>(A) * G2(GFSM, 0, 0));
      return r;
    }
  }
};

// =====================================================================
// External (barotropic) mode, 2-D operators
// reference: src/neuralpom/cores/explicit/time_loop.py:275-431
// =====================================================================

// continuity update of the surface elevation on interior cells, zero
// elsewhere (time_loop.py:299-312)
struct OpExtElfRaw {
  enum { UA, VA, D_, ELB, N3 = 0, NIN = 4 };
  enum { GDX, GDY, GART, GVFLUXF };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int OI, int OJ, class AC> __device__ static auto fu(const AC& A) {   // :299-301
    return 0.25 * (AT2(D_, OI, OJ) + AT2(D_, OI - 1, OJ)) * (G2(GDY, OI, OJ) + G2(GDY, OI - 1, OJ)) * AT2(UA, OI, OJ);
  }
  template <int OI, int OJ, class AC> __device__ static auto fv(const AC& A) {   // :303-305
    return 0.25 * (AT2(D_, OI, OJ) + AT2(D_, OI, OJ - 1)) * (G2(GDX, OI, OJ) + G2(GDX, OI, OJ - 1)) * AT2(VA, OI, OJ);
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {   // :308-312
    return IFZ(interior(A), ([&]() {
      auto div = fu<1, 0>(A) - fu<0, 0>(A) + fv<0, 1>(A) - fv<0, 0>(A);
      auto term = -(div / G2(GART, 0, 0)) - G2(GVFLUXF, 0, 0);
      return AT2(ELB, 0, 0) + prm.dte2 * term;
    }()));
  }
};

// bcond 1 (baroclinic.py:510-524): the E, W, N, S edge copies make every
// edge cell take the value of its nearest interior cell; then * fsm
struct OpBc1 {
  enum { RAW, N3 = 0, NIN = 1 };
  enum { GFSM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  static constexpr int EDGE_BAND = 1;   // only points within 1 of the edge read neighbours
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    if constexpr (deep_v<AC>) {   // away from the edges: the point's own value
      return AT2(RAW, 0, 0) * G2(GFSM, 0, 0);
    } else {
      return eval_edge<O>(A);
    }
  }
  template <int O, class AC> __device__ static auto eval_edge(const AC& A) {
    const int i = A.p.i, j = A.p.j, im = A.p.im, jm = A.p.jm;
    const int di = (i == 0) ? 1 : (i == im - 1 ? -1 : 0);
    const int dj = (j == 0) ? 1 : (j == jm - 1 ? -1 : 0);
    using R = decltype(AT2(RAW, 0, 0) + AT2(RAW, 1, 1) + AT2(RAW, -1, -1) + AT2(RAW, 1, -1) + AT2(RAW, -1, 1) +
                       AT2(RAW, 1, 0) + AT2(RAW, -1, 0) + AT2(RAW, 0, 1) + AT2(RAW, 0, -1));
    // the nine candidates are plain reads: select instead of branching
    R r{};
#define POM_BC1_CASE(a, b) r = (di == a && dj == b) ? cvt<R>(AT2(RAW, a, b)) : r;
    POM_BC1_CASE(-1, -1) POM_BC1_CASE(-1, 1) POM_BC1_CASE(1, -1) POM_BC1_CASE(1, 1) POM_BC1_CASE(0, -1)
    POM_BC1_CASE(0, 1) POM_BC1_CASE(-1, 0) POM_BC1_CASE(1, 0) POM_BC1_CASE(0, 0)
#undef POM_BC1_CASE
    return r * G2(GFSM, 0, 0);
  }
};

// advave (mode 3): baroclinic.py:262-432.  outputs 0 advua, 1 advva,
// 2 fluxua, 3 fluxva, 4 tps (the last three are the reference's work arrays).
struct OpAdvave {
  enum { UA, VA, UAB, VAB, D_, AAM2D, N3 = 0, NIN = 6 };
  enum { GDX, GDY, GFLUXUA, GFLUXVA, GTPS };
  static constexpr int NOUT = 5;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 2, RJ = 2, RK = 0;
  // tps at (i+OI, j+OJ) (i, j >= 1): :334-349
  template <int OI, int OJ, class AC> __device__ static auto tps(const AC& A) {
    auto dsum = AT2(D_, OI, OJ) + AT2(D_, OI - 1, OJ) + AT2(D_, OI, OJ - 1) + AT2(D_, OI - 1, OJ - 1);
    auto asum = AT2(AAM2D, OI, OJ) + AT2(AAM2D, OI, OJ - 1) + AT2(AAM2D, OI - 1, OJ) + AT2(AAM2D, OI - 1, OJ - 1);
    auto dys = G2(GDY, OI, OJ) + G2(GDY, OI - 1, OJ) + G2(GDY, OI, OJ - 1) + G2(GDY, OI - 1, OJ - 1);
    auto dxs = G2(GDX, OI, OJ) + G2(GDX, OI - 1, OJ) + G2(GDX, OI, OJ - 1) + G2(GDX, OI - 1, OJ - 1);
    auto du_dy = (AT2(UAB, OI, OJ) - AT2(UAB, OI, OJ - 1)) / dys;
    auto dv_dx = (AT2(VAB, OI, OJ) - AT2(VAB, OI - 1, OJ)) / dxs;
    return 0.25 * dsum * asum * (du_dy + dv_dx);
  }
  // section-1 fluxua at (i+OI, j), i+OI in [1, im-2]: :294-306, :323-328, :352
  template <int OI, class AC> __device__ static auto fu1(const AC& A) {
    auto t1 = (AT2(D_, OI + 1, 0) + AT2(D_, OI, 0)) * AT2(UA, OI + 1, 0);
    auto t2 = (AT2(D_, OI, 0) + AT2(D_, OI - 1, 0)) * AT2(UA, OI, 0);
    auto f = 0.125 * (t1 + t2) * (AT2(UA, OI + 1, 0) + AT2(UA, OI, 0));
    auto diff = (AT2(UAB, OI + 1, 0) - AT2(UAB, OI, 0)) / G2(GDX, OI, 0);
    auto f2 = f - AT2(D_, OI, 0) * 2.0 * AT2(AAM2D, OI, 0) * diff;
    return f2 * G2(GDY, OI, 0);
  }
  // section-1 fluxva at (i, j+OJ), j+OJ in [1, jm-1]: :308-321, :355
  template <int OJ, class AC> __device__ static auto fv1(const AC& A) {
    auto t1 = (AT2(D_, 0, OJ) + AT2(D_, 0, OJ - 1)) * AT2(VA, 0, OJ);
    auto t2 = (AT2(D_, -1, OJ) + AT2(D_, -1, OJ - 1)) * AT2(VA, -1, OJ);
    auto f = 0.125 * (t1 + t2) * (AT2(UA, 0, OJ) + AT2(UA, 0, OJ - 1));
    auto dxs = G2(GDX, 0, OJ) + G2(GDX, -1, OJ) + G2(GDX, 0, OJ - 1) + G2(GDX, -1, OJ - 1);
    return (f - tps<0, OJ>(A)) * 0.25 * dxs;
  }
  // section-2 fluxua at (i+OI, j), i+OI in [1, im-1]: :383-388, :419-421
  template <int OI, class AC> __device__ static auto fu2(const AC& A) {
    auto t1 = (AT2(D_, OI, 0) + AT2(D_, OI - 1, 0)) * AT2(UA, OI, 0);
    auto t2 = (AT2(D_, OI, -1) + AT2(D_, OI - 1, -1)) * AT2(UA, OI, -1);
    auto f = 0.125 * (t1 + t2) * (AT2(VA, OI - 1, 0) + AT2(VA, OI, 0));
    auto dys = G2(GDY, OI, 0) + G2(GDY, OI - 1, 0) + G2(GDY, OI, -1) + G2(GDY, OI - 1, -1);
    return (f - tps<OI, 0>(A)) * 0.25 * dys;
  }
  // section-2 fluxva at (i, j+OJ), j+OJ in [1, jm-2]: :398-417
  template <int OJ, class AC> __device__ static auto fv2(const AC& A) {
    auto t1 = (AT2(D_, 0, OJ + 1) + AT2(D_, 0, OJ)) * AT2(VA, 0, OJ + 1);
    auto t2 = (AT2(D_, 0, OJ) + AT2(D_, 0, OJ - 1)) * AT2(VA, 0, OJ);
    auto f = 0.125 * (t1 + t2) * (AT2(VA, 0, OJ + 1) + AT2(VA, 0, OJ));
    auto diff = (AT2(VAB, 0, OJ + 1) - AT2(VAB, 0, OJ)) / G2(GDY, 0, OJ);
    auto f2 = f - AT2(D_, 0, OJ) * 2.0 * AT2(AAM2D, 0, OJ) * diff;
    return f2 * G2(GDX, 0, OJ);
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    const int i = A.p.i, j = A.p.j, im = A.p.im, jm = A.p.jm;
    const bool in = interior(A);
    if constexpr (O == 0) {   // advua :364-365 (fluxua(0, j) is the input work array)
      return IFZ(in, fu1<0>(A) - BR(i >= 2, fu1<-1>(A), G2(GFLUXUA, -1, 0)) + fv1<1>(A) - fv1<0>(A));
    } else if constexpr (O == 1) {   // advva :431-432 (fluxva(i, 0) is the input work array)
      return IFZ(in, fu2<1>(A) - fu2<0>(A) + fv2<0>(A) - BR(j >= 2, fv2<-1>(A), G2(GFLUXVA, 0, -1)));
    } else if constexpr (O == 2) {   // fluxua after section 2
      return BR(i >= 1 && j >= 1, fu2<0>(A), G2(GFLUXUA, 0, 0));
    } else if constexpr (O == 3) {   // fluxva after section 2 (row jm-1 keeps section 1 then *dx)
      auto s1 = fv1<0>(A) * G2(GDX, 0, 0);
      return BR(i >= 1 && j >= 1, BR(j <= jm - 2, fv2<0>(A), s1), G2(GFLUXVA, 0, 0));
    } else {
      return BR(i >= 1 && j >= 1, (tps<0, 0>(A)), G2(GTPS, 0, 0));
    }
    (void)im;
  }
};

// depth-mean momentum update + bcond 2 (time_loop.py:326-383;
// baroclinic.py:529-571).  outputs 0 uaf, 1 vaf.  BC_ONLY: bcond 2 alone on
// given uaf / vaf (UAFIN / VAFIN): interior points keep their input.
template <bool BC_ONLY> struct OpExtUVG {
  enum { EL, ELB, ELF, D_, UA, VA, UAB, VAB, ADVUA, ADVVA, ADX2D, ADY2D, DRX2D, DRY2D, WUBOT, WVBOT,
         UAFIN, VAFIN, N3 = 0, NIN = 18 };
  enum { GCOR, GDX, GDY, GH, GARU, GARV, GDUM, GDVM, GEATM, GWUSURF, GWVSURF };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  // a radiation form of one edge: zero, without evaluating it, when its
  // coefficient is zero (a closed edge; the coefficient is uniform over the
  // launch, so this is a scalar branch)
  template <class F> __device__ static auto radiation(double rf, F f) {
    using R = decltype(f());
    if (rf == 0.0) return R{};
    return f();
  }
  template <class X> __device__ static auto cwave(X h, double grav) {   // safe_sqrt(g / clamp(h))
    using T = typename scalar_of<X>::type;
    T hv = fmax(val(h), T(1e-8));
    return P<T>{sqrt(fmax(T(grav) / hv, T(1e-8)))};
  }
  template <class AC> __device__ static auto uaf_in(const AC& A, const Params& prm) {   // :327-353
    auto ct = 0.25 * (G2(GCOR, 0, 0) * AT2(D_, 0, 0) * (AT2(VA, 0, 1) + AT2(VA, 0, 0)) +
                      G2(GCOR, -1, 0) * AT2(D_, -1, 0) * (AT2(VA, -1, 1) + AT2(VA, -1, 0)));
    auto gt = 0.25 * prm.grav * (G2(GDY, 0, 0) + G2(GDY, -1, 0)) * (AT2(D_, 0, 0) + AT2(D_, -1, 0)) *
              ((1.0 - 2.0 * prm.alpha) * (AT2(EL, 0, 0) - AT2(EL, -1, 0)) +
               prm.alpha * (AT2(ELB, 0, 0) - AT2(ELB, -1, 0) + AT2(ELF, 0, 0) - AT2(ELF, -1, 0)) +
               G2(GEATM, 0, 0) - G2(GEATM, -1, 0));
    auto aru = G2(GARU, 0, 0);
    auto prov = AT2(ADX2D, 0, 0) + AT2(ADVUA, 0, 0) - aru * ct + gt + AT2(DRX2D, 0, 0) -
                aru * (-G2(GWUSURF, 0, 0) + AT2(WUBOT, 0, 0));
    auto den = (G2(GH, 0, 0) + AT2(ELF, 0, 0) + G2(GH, -1, 0) + AT2(ELF, -1, 0)) * aru;
    auto num = (G2(GH, 0, 0) + AT2(ELB, 0, 0) + G2(GH, -1, 0) + AT2(ELB, -1, 0)) * aru * AT2(UAB, 0, 0) -
               4.0 * prm.dte * prov;
    return num / (den + 1e-20);
  }
  template <class AC> __device__ static auto vaf_in(const AC& A, const Params& prm) {   // :356-379
    auto ct = 0.25 * (G2(GCOR, 0, 0) * AT2(D_, 0, 0) * (AT2(UA, 1, 0) + AT2(UA, 0, 0)) +
                      G2(GCOR, 0, -1) * AT2(D_, 0, -1) * (AT2(UA, 1, -1) + AT2(UA, 0, -1)));
    auto gt = 0.25 * prm.grav * (G2(GDX, 0, 0) + G2(GDX, 0, -1)) * (AT2(D_, 0, 0) + AT2(D_, 0, -1)) *
              ((1.0 - 2.0 * prm.alpha) * (AT2(EL, 0, 0) - AT2(EL, 0, -1)) +
               prm.alpha * (AT2(ELB, 0, 0) - AT2(ELB, 0, -1) + AT2(ELF, 0, 0) - AT2(ELF, 0, -1)) +
               G2(GEATM, 0, 0) - G2(GEATM, 0, -1));
    auto arv = G2(GARV, 0, 0);
    auto prov = AT2(ADY2D, 0, 0) + AT2(ADVVA, 0, 0) + arv * ct + gt + AT2(DRY2D, 0, 0) -
                arv * (-G2(GWVSURF, 0, 0) + AT2(WVBOT, 0, 0));
    auto den = (G2(GH, 0, 0) + AT2(ELF, 0, 0) + G2(GH, 0, -1) + AT2(ELF, 0, -1)) * arv;
    auto num = (G2(GH, 0, 0) + AT2(ELB, 0, 0) + G2(GH, 0, -1) + AT2(ELB, 0, -1)) * arv * AT2(VAB, 0, 0) -
               4.0 * prm.dte * prov;
    return num / (den + 1e-20);
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int i = A.p.i, j = A.p.j, im = A.p.im, jm = A.p.jm;
    const bool jin = j >= 1 && j <= jm - 2, iin = i >= 1 && i <= im - 2;
    const double ramp = A.rampv();
    if constexpr (O == 0) {
      // east (rfe) / west (rfw) open-boundary forms; exterior values are zero
      auto east = [&]() {
        return radiation(prm.rfe, [&]() { return (prm.rfe * cwave(G2(GH, -1, 0), prm.grav) * (AT2(EL, -1, 0) - 0.0)) * ramp; });
      };
      auto west1 = [&]() {
        return radiation(prm.rfw, [&]() { return (0.0 - prm.rfw * cwave(G2(GH, 0, 0), prm.grav) * (AT2(EL, 0, 0) - 0.0)) * ramp; });
      };
      auto west0 = [&]() {
        return radiation(prm.rfw, [&]() { return (0.0 - prm.rfw * cwave(G2(GH, 1, 0), prm.grav) * (AT2(EL, 1, 0) - 0.0)) * ramp; });
      };
      using R = common_t<common_t<common_t<decltype(uaf_in(A, prm)), decltype(east())>,
                                  common_t<decltype(west1()), decltype(west0())>>,
                         decltype(AT2(UAFIN, 0, 0))>;
      R r{};
      if constexpr (active_v<R>) {
        r = cvt<R>(AT2(UAFIN, 0, 0));
        if (jin && i == im - 1) r = cvt<R>(east());
        else if (jin && i == 1) r = cvt<R>(west1());
        else if (jin && i == 0) r = cvt<R>(west0());
        else if (iin && (j == jm - 1 || j == 0)) r = R{};
        else if (iin && jin && !BC_ONLY) r = cvt<R>(uaf_in(A, prm));
      } else {   // passive: every candidate evaluated (indices are clamped), then selected
        const R c_in = cvt<R>(AT2(UAFIN, 0, 0)), c_e = cvt<R>(east()), c_w1 = cvt<R>(west1()),
                c_w0 = cvt<R>(west0());
        R c_int = c_in;
        if constexpr (!BC_ONLY) c_int = cvt<R>(uaf_in(A, prm));
        r = (jin && i == im - 1) ? c_e
          : (jin && i == 1) ? c_w1
          : (jin && i == 0) ? c_w0
          : (iin && (j == jm - 1 || j == 0)) ? R{}
          : (iin && jin) ? c_int : c_in;
      }
      return r * G2(GDUM, 0, 0);
    } else {
      auto north = [&]() {
        return radiation(prm.rfn, [&]() { return (prm.rfn * cwave(G2(GH, 0, -1), prm.grav) * (AT2(EL, 0, -1) - 0.0)) * ramp; });
      };
      auto south1 = [&]() {
        return radiation(prm.rfs, [&]() { return (0.0 - prm.rfs * cwave(G2(GH, 0, 0), prm.grav) * (AT2(EL, 0, 0) - 0.0)) * ramp; });
      };
      auto south0 = [&]() {
        return radiation(prm.rfs, [&]() { return (0.0 - prm.rfs * cwave(G2(GH, 0, 1), prm.grav) * (AT2(EL, 0, 1) - 0.0)) * ramp; });
      };
      using R = common_t<common_t<common_t<decltype(vaf_in(A, prm)), decltype(north())>,
                                  common_t<decltype(south1()), decltype(south0())>>,
                         decltype(AT2(VAFIN, 0, 0))>;
      R r{};
      if constexpr (active_v<R>) {
        r = cvt<R>(AT2(VAFIN, 0, 0));
        if (jin && (i == im - 1 || i == 0)) r = R{};
        else if (iin && j == jm - 1) r = cvt<R>(north());
        else if (iin && j == 1) r = cvt<R>(south1());
        else if (iin && j == 0) r = cvt<R>(south0());
        else if (iin && jin && !BC_ONLY) r = cvt<R>(vaf_in(A, prm));
      } else {
        const R c_in = cvt<R>(AT2(VAFIN, 0, 0)), c_n = cvt<R>(north()), c_s1 = cvt<R>(south1()),
                c_s0 = cvt<R>(south0());
        R c_int = c_in;
        if constexpr (!BC_ONLY) c_int = cvt<R>(vaf_in(A, prm));
        r = (jin && (i == im - 1 || i == 0)) ? R{}
          : (iin && j == jm - 1) ? c_n
          : (iin && j == 1) ? c_s1
          : (iin && j == 0) ? c_s0
          : (iin && jin) ? c_int : c_in;
      }
      return r * G2(GDVM, 0, 0);
    }
  }
};

using OpExtUV = OpExtUVG<false>;
using OpBcond2 = OpExtUVG<true>;

// The depth-mean momentum update with the tide policy's bcond 2
// (case_adapters/coastal_boundary.py:23-29) in place of the closed one:
// radiation of the elevation difference across each open edge, added to the
// transport just inside it, with the wave speed sqrt(g / d) of the substep's
// d and no ramp.  Statement by statement, at interior j (i for v):
//   uaf(im-1) = uaf(im-2) + rfe c(im-2) (el(im-2) - el(im-1))
//   uaf(1)   -= rfw c(1) (el(1) - el(0)),   uaf(0) = uaf(1)
//   vaf(jm-1) = vaf(jm-2) + rfn c(jm-2) (el(jm-2) - el(jm-1))
//   vaf(1)   -= rfs c(1) (el(1) - el(0)),   vaf(0) = vaf(1)
// the other velocity zero on each edge, corners keep the given uaf / vaf;
// then * dum / dvm.  The transports inside are the interior formula at the
// neighbour (AccShift).  Forward only: its adjoint is not built.
struct OpExtUVTide : OpExtUVG<false> {
  using Base = OpExtUVG<false>;
  template <class X> __device__ static auto wave(X d, double grav) {   // sqrt(g / d)
    using T = typename scalar_of<X>::type;
    return P<T>{sqrt(T(grav) / val(d))};
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int i = A.p.i, j = A.p.j, im = A.p.im, jm = A.p.jm;
    const bool jin = j >= 1 && j <= jm - 2, iin = i >= 1 && i <= im - 2;
    if constexpr (O == 0) {
      using R = decltype(Base::uaf_in(A, prm));
      R r = cvt<R>(AT2(UAFIN, 0, 0));
      if (jin && i == im - 1) {
        const AccShift<AC, -1, 0> S{A};
        r = Base::uaf_in(S, prm) + prm.rfe * wave(AT2(D_, -1, 0), prm.grav) * (AT2(EL, -1, 0) - AT2(EL, 0, 0));
      } else if (jin && i == 0) {
        const AccShift<AC, 1, 0> S{A};
        r = Base::uaf_in(S, prm) - prm.rfw * wave(AT2(D_, 1, 0), prm.grav) * (AT2(EL, 1, 0) - AT2(EL, 0, 0));
      } else if (iin && (j == jm - 1 || j == 0)) {
        r = R{};
      } else if (iin && jin) {
        r = Base::uaf_in(A, prm);
        if (i == 1) r = r - prm.rfw * wave(AT2(D_, 0, 0), prm.grav) * (AT2(EL, 0, 0) - AT2(EL, -1, 0));
      }
      return r * G2(GDUM, 0, 0);
    } else {
      using R = decltype(Base::vaf_in(A, prm));
      R r = cvt<R>(AT2(VAFIN, 0, 0));
      if (jin && (i == im - 1 || i == 0)) {
        r = R{};
      } else if (iin && j == jm - 1) {
        const AccShift<AC, 0, -1> S{A};
        r = Base::vaf_in(S, prm) + prm.rfn * wave(AT2(D_, 0, -1), prm.grav) * (AT2(EL, 0, -1) - AT2(EL, 0, 0));
      } else if (iin && j == 0) {
        const AccShift<AC, 0, 1> S{A};
        r = Base::vaf_in(S, prm) - prm.rfs * wave(AT2(D_, 0, 1), prm.grav) * (AT2(EL, 0, 1) - AT2(EL, 0, 0));
      } else if (iin && jin) {
        r = Base::vaf_in(A, prm);
        if (j == 1) r = r - prm.rfs * wave(AT2(D_, 0, 0), prm.grav) * (AT2(EL, 0, 0) - AT2(EL, 0, -1));
      }
      return r * G2(GDVM, 0, 0);
    }
  }
};

// Robert-Asselin filter, level shift and time-mean accumulation of one
// external substep (time_loop.py:386-431).  The substep index selects the
// etf filter phase and whether egf/utf/vtf accumulate.
// outputs: 0 ua, 1 uab, 2 va, 3 vab, 4 el, 5 elb, 6 d, 7 egf, 8 utf, 9 vtf, 10 etf
struct OpExtFilter {
  enum { UA, UAB, VA, VAB, EL, ELB, UAF, VAF, ELF, EGF, UTF, VTF, ETF, N3 = 0, NIN = 13 };
  enum { GH, GFSM };
  static constexpr int NOUT = 11;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const int iext = prm.iext;   // substep 1..isplit
    const double f = 0.5 * prm.smoth;
    const bool acc = iext != prm.isplit;
    if constexpr (O == 0) return AT2(UAF, 0, 0);
    else if constexpr (O == 1) return AT2(UA, 0, 0) + f * (AT2(UAB, 0, 0) - 2.0 * AT2(UA, 0, 0) + AT2(UAF, 0, 0));
    else if constexpr (O == 2) return AT2(VAF, 0, 0);
    else if constexpr (O == 3) return AT2(VA, 0, 0) + f * (AT2(VAB, 0, 0) - 2.0 * AT2(VA, 0, 0) + AT2(VAF, 0, 0));
    else if constexpr (O == 4) return AT2(ELF, 0, 0);
    else if constexpr (O == 5) return AT2(EL, 0, 0) + f * (AT2(ELB, 0, 0) - 2.0 * AT2(EL, 0, 0) + AT2(ELF, 0, 0));
    else if constexpr (O == 6) return G2(GH, 0, 0) + AT2(ELF, 0, 0);
    else if constexpr (O == 7) return BR(acc, AT2(EGF, 0, 0) + AT2(ELF, 0, 0) * prm.ispi, AT2(EGF, 0, 0));
    else if constexpr (O == 8) {
      auto ds = (G2(GH, 0, 0) + AT2(ELF, 0, 0)) + (G2(GH, -1, 0) + AT2(ELF, -1, 0));
      return BR(acc && A.p.i >= 1, AT2(UTF, 0, 0) + AT2(UAF, 0, 0) * ds * prm.isp2i, AT2(UTF, 0, 0));
    } else if constexpr (O == 9) {
      auto ds = (G2(GH, 0, 0) + AT2(ELF, 0, 0)) + (G2(GH, 0, -1) + AT2(ELF, 0, -1));
      return BR(acc && A.p.j >= 1, AT2(VTF, 0, 0) + AT2(VAF, 0, 0) * ds * prm.isp2i, AT2(VTF, 0, 0));
    } else {
      using R = common_t<decltype(AT2(ETF, 0, 0)), decltype(AT2(ELF, 0, 0))>;
      R r = cvt<R>(AT2(ETF, 0, 0));
      if (iext == prm.isplit - 2) r = cvt<R>(0.25 * prm.smoth * AT2(ELF, 0, 0));
      else if (iext == prm.isplit - 1) r = cvt<R>(AT2(ETF, 0, 0) + 0.5 * (1.0 - 0.5 * prm.smoth) * AT2(ELF, 0, 0));
      else if (iext == prm.isplit) r = cvt<R>((AT2(ETF, 0, 0) + 0.5 * AT2(ELF, 0, 0)) * G2(GFSM, 0, 0));
      return r;
    }
  }
};

// Before the substeps (time_loop.py:269-285): adx2d -= advua, ady2d -= advva,
// egf = el/isplit, utf/vtf start.  outputs 0 adx2d, 1 ady2d, 2 egf, 3 utf, 4 vtf
struct OpExtPrep {
  enum { ADX2D, ADY2D, ADVUA, ADVVA, EL, UA, VA, D_, UTF, VTF, N3 = 0, NIN = 10 };
  static constexpr int NOUT = 5;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    if constexpr (O == 0) return AT2(ADX2D, 0, 0) - AT2(ADVUA, 0, 0);
    else if constexpr (O == 1) return AT2(ADY2D, 0, 0) - AT2(ADVVA, 0, 0);
    else if constexpr (O == 2) return AT2(EL, 0, 0) * prm.ispi;
    else if constexpr (O == 3)
      return BR(A.p.i >= 1, AT2(UA, 0, 0) * (AT2(D_, 0, 0) + AT2(D_, -1, 0)) * prm.isp2i, AT2(UTF, 0, 0));
    else
      return BR(A.p.j >= 1, AT2(VA, 0, 0) * (AT2(D_, 0, 0) + AT2(D_, 0, -1)) * prm.isp2i, AT2(VTF, 0, 0));
  }
};

}  // namespace pom
