// Adjoint of the depth-mean advection advave (OpAdvave in pom_ops.h, mode 3;
// reference src/neuralpom/cores/explicit/baroclinic.py:262-432) with respect
// to its outputs advua and advva, as a transposed stencil over its face
// fluxes, in place of the generic gather (which evaluates the point function
// in dual numbers once per offset of its radius-2 stencil, 25 of them).  The
// outputs fluxua, fluxva and tps are the reference's work arrays and carry no
// cotangent.
//
// At an interior point p (zero elsewhere):
//   advua(p) = FU1(p) - [i>=2] FU1(p-i) + FV1(p+j) - FV1(p)
//   advva(p) = FU2(p+i) - FU2(p) + FV2(p) - [j>=2] FV2(p-j)
// with (every sum over the four cells of a corner b: b, b-i, b-j, b-i-j)
//   FU1(a) = (1/8 ((d(a+i) + d(a)) ua(a+i) + (d(a) + d(a-i)) ua(a)) (ua(a+i) + ua(a))
//             - 2 d(a) aam(a) (uab(a+i) - uab(a)) / dx(a)) dy(a)
//   FV2(b) = (1/8 ((d(b+j) + d(b)) va(b+j) + (d(b) + d(b-j)) va(b)) (va(b+j) + va(b))
//             - 2 d(b) aam(b) (vab(b+j) - vab(b)) / dy(b)) dx(b)
//   FV1(b) = (1/8 ((d(b) + d(b-j)) va(b) + (d(b-i) + d(b-i-j)) va(b-i)) (ua(b) + ua(b-j)) - TPS(b)) dx4(b) / 4
//   FU2(a) = (1/8 ((d(a) + d(a-i)) ua(a) + (d(a-j) + d(a-i-j)) ua(a-j)) (va(a-i) + va(a)) - TPS(a)) dy4(a) / 4
//   TPS(b) = 1/4 d4(b) aam4(b) ((uab(b) - uab(b-j)) / dy4(b) + (vab(b) - vab(b-i)) / dx4(b))
// The thread of point q sums the partials of every face that reads q, each
// weighted by the face's cotangent (a difference of the masked output
// cotangents across it; TPS takes those of FV1 and FU2 at its corner).  Every
// element is written by one thread, accumulated into the given arrays.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct AdvaveAdjArgs {
  const T *ua, *va, *uab, *vab, *d, *aam2d;          // primal (2-D)
  const T *dx, *dy;                                   // geometry
  const T *ob_u, *ob_v;                               // cotangents of advua, advva
  T *c_ua, *c_va, *c_uab, *c_vab, *c_d, *c_aam2d;     // accumulated; six distinct arrays
};

template <class T> void launch_adj_advave(const Ctx<T>& c, const AdvaveAdjArgs<T>& a);

}  // namespace pom
