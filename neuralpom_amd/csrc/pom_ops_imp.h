// Point functions of the semi-implicit (P0-fixed) POM step: the barotropic
// predictor / transport / post-solve / corrector / filter stages, the
// closed-basin boundary variants of the 3-D stages, the zero-subgradient
// Smagorinsky form and the elliptic operator A(D) (whose adjoint gather with
// respect to D is the eta-solve depth term).
//
// reference: src/neuralpom/cores/implicit_p0_fixed/core.py:121-199, 591-675, 795-943,
//            time_loop.py:152-201, 247-398, boundary/closed_basin.py:27-101,
//            numerics/double_gyre.py:78-184, autograd.py:6-29
#pragma once
#include "pom_ops.h"

namespace pom {

template <class AC> __device__ __forceinline__ bool edge2(const AC& A) { return !interior(A); }

// sqrt with the zero subgradient at 0 (autograd.py:6-29)
template <class T> __device__ __forceinline__ P<T> strain(P<T> a) { return {sqrt(a.v)}; }
template <class T> __device__ __forceinline__ D<T> strain(D<T> a) {
  T s = sqrt(a.v);
  return {s, s > T(0) ? a.d * (T(0.5) / s) : T(0)};
}
template <class T, uint32_t M> __device__ __forceinline__ DS<T, M> strain(const DS<T, M>& a) {
  const T s = sqrt(a.v);
  DS<T, M> r = lanes_scaled(a, s > T(0) ? T(0.5) / s : T(0));
  r.v = s;
  return r;
}

// =====================================================================
// Smagorinsky: aam = horcon dx dy sqrt(t1 + t2 + t3)  (time_loop.py:152-187)
// =====================================================================
struct OpSmagImp {
  enum { U, V, AAM, N3, NIN = N3 };
  enum { GDX, GDY };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const bool in = interior(A) && A.p.k < KBM1;
    auto f = [&]() {
      auto dx = G2(GDX, 0, 0), dy = G2(GDY, 0, 0);
      auto t1 = sq((AT(U, 1, 0, 0) - AT(U, 0, 0, 0)) / dx);
      auto t2 = sq((AT(V, 0, 1, 0) - AT(V, 0, 0, 0)) / dy);
      auto ut = 0.25 * (AT(U, 0, 1, 0) + AT(U, 1, 1, 0) - AT(U, 0, -1, 0) - AT(U, 1, -1, 0)) / dy;
      auto vt = 0.25 * (AT(V, 1, 0, 0) + AT(V, 1, 1, 0) - AT(V, -1, 0, 0) - AT(V, -1, 1, 0)) / dx;
      auto t3 = 0.5 * sq(ut + vt);
      auto visc = prm.horcon * dx * dy * strain(t1 + t2 + t3);
      return BR(prm.aam_min > 0.0, clamp_min(visc, prm.aam_min), visc);   // time_loop.py:186-188
    };
    return BR(in, f(), AT(AAM, 0, 0, 0));
  }
};

// =====================================================================
// closed-basin bcond 3: uf * dum, vf * dvm, zero on the edges
// =====================================================================
struct OpImpBc3 {
  enum { UF, VF, N3, NIN = N3 };
  enum { GDUM, GDVM };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 0, RJ = 0, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    if constexpr (O == 0) return IFZ(interior(A), AT(UF, 0, 0, 0) * G2(GDUM, 0, 0));
    else return IFZ(interior(A), AT(VF, 0, 0, 0) * G2(GDVM, 0, 0));
  }
};

// value of a field after the closed-basin bcond-4 edge copies: edge cells take
// their nearest interior cell (i edges first, then j edges)
template <int F, class AC> __device__ __forceinline__ auto edge_copy3(const AC& A) {
  if constexpr (deep_v<AC>) return AT(F, 0, 0, 0);   // away from the edges: the point's own value
  else {
  const int i = A.p.i, j = A.p.j, im = A.p.im, jm = A.p.jm;
  const int di = (i == 0) ? 1 : (i == im - 1 ? -1 : 0);
  const int dj = (j == 0) ? 1 : (j == jm - 1 ? -1 : 0);
  using R = decltype(AT(F, 0, 0, 0) + AT(F, 1, 1, 0) + AT(F, -1, -1, 0) + AT(F, 1, -1, 0) + AT(F, -1, 1, 0) +
                     AT(F, 1, 0, 0) + AT(F, -1, 0, 0) + AT(F, 0, 1, 0) + AT(F, 0, -1, 0));
  R r{};
#define POM_EC_CASE(a, b) if (di == a && dj == b) r = cvt<R>(AT(F, a, b, 0));
  POM_EC_CASE(0, 0) else POM_EC_CASE(1, 0) else POM_EC_CASE(-1, 0) else POM_EC_CASE(0, 1)
  else POM_EC_CASE(0, -1) else POM_EC_CASE(1, 1) else POM_EC_CASE(1, -1) else POM_EC_CASE(-1, 1)
  else POM_EC_CASE(-1, -1)
#undef POM_EC_CASE
  return r;
  }
}

// =====================================================================
// The closed-basin policy's bcond on its own, idx 1, 2, 4, 5, 6
// (implicit_p0_fixed/boundary/closed_basin.py:27-87; idx 3 is OpImpBc3):
//   1: elf * fsm, the edge ring zero          2: uaf * dum, vaf * dvm, edges zero
//   4: uf, vf (new T, S) edge cells copied from the nearest interior cell
//   5: w * fsm                                6: uf * fsm, vf * fsm (new q2, q2l)
// =====================================================================
template <int IDX> struct OpImpBcond {
  static constexpr bool OUT3 = IDX >= 3;
  static constexpr int NOUT = (IDX == 1 || IDX == 5) ? 1 : 2;
  static constexpr int NIN = NOUT, N3 = OUT3 ? NIN : 0;
  enum { A0, A1 };
  enum { GFSM, GDUM, GDVM };
  static constexpr int RI = IDX == 4 ? 1 : 0, RJ = IDX == 4 ? 1 : 0, RK = 0;
  static constexpr int EDGE_BAND = 1;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params&) {
    if constexpr (IDX == 1) {
      return IFZ(interior(A), AT2(A0, 0, 0) * G2(GFSM, 0, 0));
    } else if constexpr (IDX == 2) {
      if constexpr (O == 0) return IFZ(interior(A), AT2(A0, 0, 0) * G2(GDUM, 0, 0));
      else return IFZ(interior(A), AT2(A1, 0, 0) * G2(GDVM, 0, 0));
    } else if constexpr (IDX == 4) {
      return edge_copy3<O == 0 ? A0 : A1>(A);
    } else {   // 5 and 6
      return AT(O == 0 ? A0 : A1, 0, 0, 0) * G2(GFSM, 0, 0);
    }
  }
};


// =====================================================================
// closed-basin bcond 4 + T/S Asselin filter + tracer mask (time_loop.py:548-575,
// numerics/double_gyre.py:78-130 with gamma 0).  outputs t, tb, s, sb.
// =====================================================================
// MASKED = false: the new T, S are left unmasked, for the spatial Shapiro
// filter that follows them (OpShapiro multiplies by fsm after filtering)
template <bool MASKED> struct OpImpBc4FilterG {
  enum { TF, SF, TT, SS, TB, SB, N3, NIN = N3 };
  enum { GFSM };
  static constexpr int NOUT = 4;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 1;
  static constexpr int EDGE_BAND = 1;   // only edge points read neighbours (edge_copy3)
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    constexpr int FNEW = (O < 2) ? TF : SF;
    constexpr int FC = (O < 2) ? TT : SS;
    constexpr int FBK = (O < 2) ? TB : SB;
    if constexpr (O == 0 || O == 2) {
      if constexpr (MASKED) return edge_copy3<FNEW>(A) * G2(GFSM, 0, 0);
      else return edge_copy3<FNEW>(A);
    } else {
      // t, tb carry advt1's bottom copy f(kb) = f(kbm1)
      const bool bot = A.p.k == KB - 1;
      auto tc = BR(bot, AT(FC, 0, 0, -1), AT(FC, 0, 0, 0));
      auto tbk = BR(bot, AT(FBK, 0, 0, -1), AT(FBK, 0, 0, 0));
      return tc + (0.5 * prm.smoth) * (edge_copy3<FNEW>(A) + tbk - 2.0 * tc);
    }
  }
};
using OpImpBc4Filter = OpImpBc4FilterG<true>;

// =====================================================================
// spatial Shapiro filter of a 3-D field on the interior columns, levels
// 0-14, then the mask (numerics/double_gyre.py:78-130 for T, S with fsm,
// :132-184 for the filtered u, v with dum / dvm).  VEL selects the weight.
// =====================================================================
template <bool VEL> struct OpShapiro {
  enum { X, N3, NIN = N3 };
  enum { GMASK };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const double gamma = VEL ? prm.gamma_uv : prm.gamma_ts;
    const bool in = interior(A) && A.p.k < KBM1;
    auto f = BR(in, AT(X, 0, 0, 0) + (gamma / 4.0) * (AT(X, 1, 0, 0) + AT(X, -1, 0, 0) + AT(X, 0, 1, 0) +
                                                      AT(X, 0, -1, 0) - 4.0 * AT(X, 0, 0, 0)),
                AT(X, 0, 0, 0));
    return f * G2(GMASK, 0, 0);
  }
};

// =====================================================================
// closed-basin bcond 6 (q2f, q2lf * fsm) + turbulence Asselin filter
// (time_loop.py:471-487).  outputs q2, q2b, q2l, q2lb.
// =====================================================================
struct OpImpBc6Filter {
  enum { Q2F, Q2LF, Q2, Q2L, Q2B, Q2LB, N3, NIN = N3 };
  enum { GFSM };
  static constexpr int NOUT = 4;
  static constexpr bool OUT3 = true;
  static constexpr int RI = 0, RJ = 0, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    constexpr int FNEW = (O < 2) ? Q2F : Q2LF;
    constexpr int FC = (O < 2) ? Q2 : Q2L;
    constexpr int FBK = (O < 2) ? Q2B : Q2LB;
    auto qn = AT(FNEW, 0, 0, 0) * G2(GFSM, 0, 0);
    if constexpr (O == 0 || O == 2) {
      return qn;
    } else {
      auto q = AT(FC, 0, 0, 0);
      return q + (0.5 * prm.smoth) * (qn + AT(FBK, 0, 0, 0) - 2.0 * q);
    }
  }
};

// =====================================================================
// barotropic predictor u*, v* (+ bcond 2) and the denominators du*, dv*
// (core.py:795-856).  outputs 0 u*, 1 v*, 2 du*, 3 dv* (du*, dv* zero on edges)
// =====================================================================
struct OpImpPred {
  enum { UA, VA, EL, ELB, D_, UAB, VAB, ADX2D, ADY2D, ADVUA, ADVVA, DRX2D, DRY2D, WUBOT, WVBOT, N3 = 0,
         NIN = 15 };
  enum { GCOR, GDX, GDY, GH, GARU, GARV, GDUM, GDVM, GEATM, GWUSURF, GWVSURF };
  static constexpr int NOUT = 4;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <class AC> __device__ static auto du(const AC& A) {
    return (G2(GH, 0, 0) + AT2(EL, 0, 0) + G2(GH, -1, 0) + AT2(EL, -1, 0)) * G2(GARU, 0, 0);
  }
  template <class AC> __device__ static auto dv(const AC& A) {
    return (G2(GH, 0, 0) + AT2(EL, 0, 0) + G2(GH, 0, -1) + AT2(EL, 0, -1)) * G2(GARV, 0, 0);
  }
  template <class AC> __device__ static auto ustar(const AC& A, const Params& prm) {
    auto ct = 0.25 * (G2(GCOR, 0, 0) * AT2(D_, 0, 0) * (AT2(VA, 0, 1) + AT2(VA, 0, 0)) +
                      G2(GCOR, -1, 0) * AT2(D_, -1, 0) * (AT2(VA, -1, 1) + AT2(VA, -1, 0)));
    auto gt = 0.25 * prm.grav * (G2(GDY, 0, 0) + G2(GDY, -1, 0)) * (AT2(D_, 0, 0) + AT2(D_, -1, 0)) *
              ((1.0 - 2.0 * prm.alpha) * (AT2(EL, 0, 0) - AT2(EL, -1, 0)) +
               prm.alpha * (AT2(ELB, 0, 0) - AT2(ELB, -1, 0)) + G2(GEATM, 0, 0) - G2(GEATM, -1, 0));
    auto aru = G2(GARU, 0, 0);
    auto prov = AT2(ADX2D, 0, 0) + AT2(ADVUA, 0, 0) - aru * ct + gt + AT2(DRX2D, 0, 0) -
                aru * (-G2(GWUSURF, 0, 0) + AT2(WUBOT, 0, 0));
    auto num = (G2(GH, 0, 0) + AT2(ELB, 0, 0) + G2(GH, -1, 0) + AT2(ELB, -1, 0)) * aru * AT2(UAB, 0, 0) -
               4.0 * prm.dti * prov;
    return (num / (du(A) + 1e-20)) * G2(GDUM, 0, 0);
  }
  template <class AC> __device__ static auto vstar(const AC& A, const Params& prm) {
    auto ct = 0.25 * (G2(GCOR, 0, 0) * AT2(D_, 0, 0) * (AT2(UA, 1, 0) + AT2(UA, 0, 0)) +
                      G2(GCOR, 0, -1) * AT2(D_, 0, -1) * (AT2(UA, 1, -1) + AT2(UA, 0, -1)));
    auto gt = 0.25 * prm.grav * (G2(GDX, 0, 0) + G2(GDX, 0, -1)) * (AT2(D_, 0, 0) + AT2(D_, 0, -1)) *
              ((1.0 - 2.0 * prm.alpha) * (AT2(EL, 0, 0) - AT2(EL, 0, -1)) +
               prm.alpha * (AT2(ELB, 0, 0) - AT2(ELB, 0, -1)) + G2(GEATM, 0, 0) - G2(GEATM, 0, -1));
    auto arv = G2(GARV, 0, 0);
    auto prov = AT2(ADY2D, 0, 0) + AT2(ADVVA, 0, 0) + arv * ct + gt + AT2(DRY2D, 0, 0) -
                arv * (-G2(GWVSURF, 0, 0) + AT2(WVBOT, 0, 0));
    auto num = (G2(GH, 0, 0) + AT2(ELB, 0, 0) + G2(GH, 0, -1) + AT2(ELB, 0, -1)) * arv * AT2(VAB, 0, 0) -
               4.0 * prm.dti * prov;
    return (num / (dv(A) + 1e-20)) * G2(GDVM, 0, 0);
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const bool in = interior(A);
    if constexpr (O == 0) return IFZ(in, ustar(A, prm) * G2(GDUM, 0, 0));   // bcond 2: * dum again
    else if constexpr (O == 1) return IFZ(in, vstar(A, prm) * G2(GDVM, 0, 0));
    else if constexpr (O == 2) return IFZ(in, du(A));
    else return IFZ(in, dv(A));
  }
};

// =====================================================================
// face transports and the eta right-hand side (core.py:858-891).
// outputs 0 fluxua, 1 fluxva (zero at i = 0 / j = 0), 2 rhs (elb on the edges)
// =====================================================================
struct OpImpRhs {
  enum { US, VS, D_, ELB, N3 = 0, NIN = 4 };
  enum { GDX, GDY, GART, GVFLUXF };
  static constexpr int NOUT = 3;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int OI, class AC> __device__ static auto fu(const AC& A) {
    return 0.25 * (AT2(D_, OI, 0) + AT2(D_, OI - 1, 0)) * (G2(GDY, OI, 0) + G2(GDY, OI - 1, 0)) * AT2(US, OI, 0);
  }
  template <int OJ, class AC> __device__ static auto fv(const AC& A) {
    return 0.25 * (AT2(D_, 0, OJ) + AT2(D_, 0, OJ - 1)) * (G2(GDX, 0, OJ) + G2(GDX, 0, OJ - 1)) * AT2(VS, 0, OJ);
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    const bool f_ok = A.p.i >= 1 && A.p.j >= 1;
    if constexpr (O == 0) return IFZ(f_ok, fu<0>(A));
    else if constexpr (O == 1) return IFZ(f_ok, fv<0>(A));
    else {
      auto f = [&]() {
        auto div = fu<1>(A) - fu<0>(A) + fv<1>(A) - fv<0>(A);
        return AT2(ELB, 0, 0) - 2.0 * prm.dti * (div / (G2(GART, 0, 0) + 1e-20) + G2(GVFLUXF, 0, 0));
      };
      return BR(interior(A), f(), AT2(ELB, 0, 0));
    }
  }
};

// =====================================================================
// post-solve surface: Shapiro (gamma) * fsm on the interior, then bcond 1
// (elf * fsm, zero edges) (core.py:893-916, closed_basin.py:29-36)
// =====================================================================
struct OpImpSurf {
  enum { RAW, N3 = 0, NIN = 1 };
  enum { GFSM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    auto f = [&]() {
      auto sm = AT2(RAW, 0, 0) + (prm.gamma_baro / 4.0) * (AT2(RAW, 1, 0) + AT2(RAW, -1, 0) + AT2(RAW, 0, 1) +
                                                           AT2(RAW, 0, -1) - 4.0 * AT2(RAW, 0, 0));
      return sm * G2(GFSM, 0, 0) * G2(GFSM, 0, 0);
    };
    return IFZ(interior(A), f());
  }
};

// =====================================================================
// barotropic corrector (+ bcond 2) with D = h + elf (core.py:918-943).
// outputs 0 uaf, 1 vaf
// =====================================================================
struct OpImpCorr {
  enum { US, VS, DUS, DVS, ELF, N3 = 0, NIN = 5 };
  enum { GDX, GDY, GH, GARU, GARV, GDUM, GDVM };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    if constexpr (O == 0) {
      auto f = [&]() {
        auto dn = (G2(GH, 0, 0) + AT2(ELF, 0, 0)) + (G2(GH, -1, 0) + AT2(ELF, -1, 0));
        auto gt = 0.25 * prm.grav * (G2(GDY, 0, 0) + G2(GDY, -1, 0)) * dn * prm.alpha *
                  (AT2(ELF, 0, 0) - AT2(ELF, -1, 0));
        auto den = (G2(GH, 0, 0) + AT2(ELF, 0, 0) + G2(GH, -1, 0) + AT2(ELF, -1, 0)) * G2(GARU, 0, 0);
        return (AT2(US, 0, 0) * AT2(DUS, 0, 0) - 4.0 * prm.dti * gt) / (den + 1e-20) * G2(GDUM, 0, 0) *
               G2(GDUM, 0, 0);
      };
      return IFZ(interior(A), f());
    } else {
      auto f = [&]() {
        auto dn = (G2(GH, 0, 0) + AT2(ELF, 0, 0)) + (G2(GH, 0, -1) + AT2(ELF, 0, -1));
        auto gt = 0.25 * prm.grav * (G2(GDX, 0, 0) + G2(GDX, 0, -1)) * dn * prm.alpha *
                  (AT2(ELF, 0, 0) - AT2(ELF, 0, -1));
        auto den = (G2(GH, 0, 0) + AT2(ELF, 0, 0) + G2(GH, 0, -1) + AT2(ELF, 0, -1)) * G2(GARV, 0, 0);
        return (AT2(VS, 0, 0) * AT2(DVS, 0, 0) - 4.0 * prm.dti * gt) / (den + 1e-20) * G2(GDVM, 0, 0) *
               G2(GDVM, 0, 0);
      };
      return IFZ(interior(A), f());
    }
  }
};

// =====================================================================
// open-boundary tide policy (case_adapters/coastal_boundary.py:9-67).
//
// bcond 2 as radiation conditions on the depth-mean transports (:24-30),
// applied to the predictor's or the corrector's interior values:
//   east  uaf(E, j) = uaf(E-1, j) + rfe sqrt(g / d(E-1, j)) (el(E-1, j) - el(E, j)), vaf(E, j) = 0
//   west  uaf(1, j) -= rfw sqrt(g / d(1, j)) (el(1, j) - el(0, j)), uaf(0, j) = uaf(1, j), vaf(0, j) = 0
//   north vaf(i, N) = vaf(i, N-1) + rfn sqrt(g / d(i, N-1)) (el(i, N-1) - el(i, N)), uaf(i, N) = 0
//   south vaf(i, 1) -= rfs sqrt(g / d(i, 1)) (el(i, 1) - el(i, 0)), vaf(i, 0) = vaf(i, 1), uaf(i, 0) = 0
// for interior j (east / west) and interior i (north / south), then * dum / dvm.
// UI, VI: interior values (their edges are not read); UE, VE: the array the
// reference starts from, whose four corners no statement writes; EL: the
// current surface elevation; D_: the depth d (NEWD false: the predictor,
// core.py:853-856), or the new surface elf with d = h + elf (NEWD true: the
// corrector, which runs after the time loop sets d = h + elf).
// outputs 0 uaf, 1 vaf.
// =====================================================================
template <bool NEWD> struct OpImpRadG {
  enum { UI, VI, UE, VE, EL, D_, N3 = 0, NIN = 6 };
  enum { GDUM, GDVM, GH };
  static constexpr int NOUT = 2;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  static constexpr int EDGE_BAND = 1;   // only points within 1 of the edge read neighbours
  // rf sqrt(g / d(a)) (el(a) - el(b)) at offsets a = (AI, AJ), b = (BI, BJ)
  template <int AI, int AJ, class AC> __device__ static auto depth(const AC& A) {
    if constexpr (NEWD) return G2(GH, AI, AJ) + AT2(D_, AI, AJ);
    else return AT2(D_, AI, AJ);
  }
  template <int AI, int AJ, int BI, int BJ, class AC> __device__ static auto rad(const AC& A, double rf, double grav) {
    return rf * psqrt(grav / depth<AI, AJ>(A)) * (AT2(EL, AI, AJ) - AT2(EL, BI, BJ));
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    if constexpr (deep_v<AC>) {   // away from the edges: the interior values
      if constexpr (O == 0) return AT2(UI, 0, 0) * G2(GDUM, 0, 0);
      else return AT2(VI, 0, 0) * G2(GDVM, 0, 0);
    } else {
      return eval_edge<O>(A, prm);
    }
  }
  template <int O, class AC> __device__ static auto eval_edge(const AC& A, const Params& prm) {
    const int i = A.p.i, j = A.p.j, E = A.p.im - 1, N = A.p.jm - 1;
    const bool jin = j >= 1 && j <= N - 1, iin = i >= 1 && i <= E - 1;
    if constexpr (O == 0) {
      auto east = [&]() { return AT2(UI, -1, 0) + rad<-1, 0, 0, 0>(A, prm.rfe, prm.grav); };
      auto west1 = [&]() { return AT2(UI, 0, 0) - rad<0, 0, -1, 0>(A, prm.rfw, prm.grav); };
      auto west0 = [&]() { return AT2(UI, 1, 0) - rad<1, 0, 0, 0>(A, prm.rfw, prm.grav); };
      using R = common_t<common_t<common_t<decltype(east()), decltype(west1())>, decltype(west0())>,
                         decltype(AT2(UE, 0, 0))>;
      R r{};
      if (jin && i == E) r = cvt<R>(east());
      else if (jin && i == 1) r = cvt<R>(west1());
      else if (jin && i == 0) r = cvt<R>(west0());
      else if (iin && (j == N || j == 0)) r = R{};
      else if (iin && jin) r = cvt<R>(AT2(UI, 0, 0));
      else r = cvt<R>(AT2(UE, 0, 0));
      return r * G2(GDUM, 0, 0);
    } else {
      auto north = [&]() { return AT2(VI, 0, -1) + rad<0, -1, 0, 0>(A, prm.rfn, prm.grav); };
      auto south1 = [&]() { return AT2(VI, 0, 0) - rad<0, 0, 0, -1>(A, prm.rfs, prm.grav); };
      auto south0 = [&]() { return AT2(VI, 0, 1) - rad<0, 1, 0, 0>(A, prm.rfs, prm.grav); };
      using R = common_t<common_t<common_t<decltype(north()), decltype(south1())>, decltype(south0())>,
                         decltype(AT2(VE, 0, 0))>;
      R r{};
      if (iin && j == N) r = cvt<R>(north());
      else if (iin && j == 1) r = cvt<R>(south1());
      else if (iin && j == 0) r = cvt<R>(south0());
      else if (jin && (i == E || i == 0)) r = R{};
      else if (iin && jin) r = cvt<R>(AT2(VI, 0, 0));
      else r = cvt<R>(AT2(VE, 0, 0));
      return r * G2(GDVM, 0, 0);
    }
  }
};
using OpImpRad = OpImpRadG<false>;
using OpImpRadNew = OpImpRadG<true>;

// =====================================================================
// post-solve surface of the tide policy: Shapiro (gamma) * fsm on the
// interior (core.py:893-911), then bcond 1 of the adapter (coastal_boundary.py:
// 12-23): the tide at the nodes, then * fsm; edges keep the solution.
// GLIFT: the tide at the nodes at this step (zero elsewhere).
// =====================================================================
struct OpImpSurfTide {
  enum { RAW, N3 = 0, NIN = 1 };
  enum { GFSM, GNODE, GLIFT };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    auto sm = [&]() {
      return (AT2(RAW, 0, 0) + (prm.gamma_baro / 4.0) * (AT2(RAW, 1, 0) + AT2(RAW, -1, 0) + AT2(RAW, 0, 1) +
                                                         AT2(RAW, 0, -1) - 4.0 * AT2(RAW, 0, 0))) *
             G2(GFSM, 0, 0);
    };
    using R = common_t<decltype(sm()), decltype(G2(GLIFT, 0, 0))>;
    R r{};
    if (val(G2(GNODE, 0, 0)) != 0) r = cvt<R>(G2(GLIFT, 0, 0));
    else if (interior(A)) r = cvt<R>(sm());
    else r = cvt<R>(AT2(RAW, 0, 0));
    return r * G2(GFSM, 0, 0);
  }
};

// =====================================================================
// Asselin filter and level shift of the depth-mean state, egf, and the
// endpoint-mean face transports (time_loop.py:323-398).  outputs
// 0 ua, 1 uab, 2 va, 3 vab, 4 el, 5 elb, 6 d, 7 egf, 8 utf, 9 vtf, 10 etf
// =====================================================================
struct OpImpFilter {
  enum { UA, UAB, VA, VAB, EL, ELB, UAF, VAF, ELF, UTF0, VTF0, N3 = 0, NIN = 11 };
  enum { GH, GFSM };
  static constexpr int NOUT = 11;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int OI, int OJ, class AC> __device__ static auto uabf(const AC& A, const Params& prm) {
    return AT2(UA, OI, OJ) + (0.5 * prm.smoth) * (AT2(UAB, OI, OJ) - 2.0 * AT2(UA, OI, OJ) + AT2(UAF, OI, OJ));
  }
  template <int OI, int OJ, class AC> __device__ static auto vabf(const AC& A, const Params& prm) {
    return AT2(VA, OI, OJ) + (0.5 * prm.smoth) * (AT2(VAB, OI, OJ) - 2.0 * AT2(VA, OI, OJ) + AT2(VAF, OI, OJ));
  }
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    if constexpr (O == 0) return AT2(UAF, 0, 0);
    else if constexpr (O == 1) return uabf<0, 0>(A, prm);
    else if constexpr (O == 2) return AT2(VAF, 0, 0);
    else if constexpr (O == 3) return vabf<0, 0>(A, prm);
    else if constexpr (O == 4) return AT2(ELF, 0, 0);
    else if constexpr (O == 5)
      return AT2(EL, 0, 0) + (0.5 * prm.smoth) * (AT2(ELB, 0, 0) - 2.0 * AT2(EL, 0, 0) + AT2(ELF, 0, 0));
    else if constexpr (O == 6) return G2(GH, 0, 0) + AT2(ELF, 0, 0);
    else if constexpr (O == 7)
      return (AT2(EL, 0, 0) + (0.5 * prm.smoth) * (AT2(ELB, 0, 0) - 2.0 * AT2(EL, 0, 0) + AT2(ELF, 0, 0)) +
              AT2(ELF, 0, 0)) * 0.5;
    else if constexpr (O == 8) {
      auto f = [&]() {
        auto so = (G2(GH, 0, 0) + AT2(ELB, 0, 0)) + (G2(GH, -1, 0) + AT2(ELB, -1, 0));
        auto sn = (G2(GH, 0, 0) + AT2(ELF, 0, 0)) + (G2(GH, -1, 0) + AT2(ELF, -1, 0));
        return 0.25 * (uabf<0, 0>(A, prm) * so + AT2(UAF, 0, 0) * sn);
      };
      return BR(A.p.i >= 1, f(), AT2(UTF0, 0, 0));
    } else if constexpr (O == 9) {
      auto f = [&]() {
        auto so = (G2(GH, 0, 0) + AT2(ELB, 0, 0)) + (G2(GH, 0, -1) + AT2(ELB, 0, -1));
        auto sn = (G2(GH, 0, 0) + AT2(ELF, 0, 0)) + (G2(GH, 0, -1) + AT2(ELF, 0, -1));
        return 0.25 * (vabf<0, 0>(A, prm) * so + AT2(VAF, 0, 0) * sn);
      };
      return BR(A.p.j >= 1, f(), AT2(VTF0, 0, 0));
    } else {
      return AT2(ELF, 0, 0) * G2(GFSM, 0, 0);
    }
  }
};

// =====================================================================
// elliptic operator A(D) eta = eta - 4 dti^2 g alpha div(D grad eta)/art on the
// interior, eta on the edges (core.py:591-639 with the face metric ratios of
// core.py:121-178)
// =====================================================================
struct OpEll {
  enum { ETA, D_, N3 = 0, NIN = 2 };
  enum { GDX, GDY, GART, GDUM, GDVM };
  static constexpr int NOUT = 1;
  static constexpr bool OUT3 = false;
  static constexpr int RI = 1, RJ = 1, RK = 0;
  template <int O, class AC> __device__ static auto eval(const AC& A, const Params& prm) {
    auto f = [&]() {
      auto ex = (0.5 * (G2(GDY, 1, 0) + G2(GDY, 0, 0))) / (0.5 * (G2(GDX, 1, 0) + G2(GDX, 0, 0)) + 1e-20);
      auto wx = (0.5 * (G2(GDY, 0, 0) + G2(GDY, -1, 0))) / (0.5 * (G2(GDX, 0, 0) + G2(GDX, -1, 0)) + 1e-20);
      auto ny = (0.5 * (G2(GDX, 0, 1) + G2(GDX, 0, 0))) / (0.5 * (G2(GDY, 0, 1) + G2(GDY, 0, 0)) + 1e-20);
      auto sy = (0.5 * (G2(GDX, 0, 0) + G2(GDX, 0, -1))) / (0.5 * (G2(GDY, 0, 0) + G2(GDY, 0, -1)) + 1e-20);
      auto fe = 0.5 * (AT2(D_, 1, 0) + AT2(D_, 0, 0)) * (AT2(ETA, 1, 0) - AT2(ETA, 0, 0)) * ex * G2(GDUM, 1, 0);
      auto fw = 0.5 * (AT2(D_, 0, 0) + AT2(D_, -1, 0)) * (AT2(ETA, 0, 0) - AT2(ETA, -1, 0)) * wx * G2(GDUM, 0, 0);
      auto fn = 0.5 * (AT2(D_, 0, 1) + AT2(D_, 0, 0)) * (AT2(ETA, 0, 1) - AT2(ETA, 0, 0)) * ny * G2(GDVM, 0, 1);
      auto fs = 0.5 * (AT2(D_, 0, 0) + AT2(D_, 0, -1)) * (AT2(ETA, 0, 0) - AT2(ETA, 0, -1)) * sy * G2(GDVM, 0, 0);
      auto div = (fe - fw + fn - fs) * (1.0 / (G2(GART, 0, 0) + 1e-20));
      const double coeff = 4.0 * (prm.dti * prm.dti) * prm.grav * prm.alpha;
      return AT2(ETA, 0, 0) - coeff * div;
    };
    return BR(interior(A), f(), AT2(ETA, 0, 0));
  }
};

}  // namespace pom
