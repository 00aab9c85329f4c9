// Adjoint of the horizontal momentum advection advct (OpAdvct in pom_ops.h;
// reference src/neuralpom/cores/explicit/baroclinic.py:14-260) as a
// transposed stencil over its face fluxes, in place of the generic gather
// (which evaluates the point function once per stencil offset, 25 of them).
//
// At an interior point p, level k < kb-1 (zero elsewhere):
//   advx(p) = Xu(p) - [i>=2] Xu(p-i) + Yu(p+j) - Yu(p) - [i>=2] aru(p)/4 (P1(p) + P1(p-i))
//   advy(p) = Xv(p+i) - Xv(p) + Yv(p) - [j>=2] Yv(p-j) + [j>=2] arv(p)/4 (P2(p) + P2(p-j))
// with the faces (every value at level k)
//   Xu(a) = (1/8 ((dt(a+i) + dt(a)) u(a+i) + (dt(a) + dt(a-i)) u(a)) (u(a+i) + u(a))
//            - 2 dt(a) aam(a) (ub(a+i) - ub(a)) / (dx(a) + 1e-20)) dy(a)
//   Yu(b) = (1/8 ((dt(b) + dt(b-j)) v(b) + (dt(b-i) + dt(b-i-j)) v(b-i)) (u(b) + u(b-j)) - CV(b)) sx4(b) / 4
//   Xv(a) = (1/8 ((dt(a) + dt(a-i)) u(a) + (dt(a-j) + dt(a-i-j)) u(a-j)) (v(a) + v(a-i)) - CV(a)) sy4(a) / 4
//   Yv(b) = (1/8 ((dt(b+j) + dt(b)) v(b+j) + (dt(b) + dt(b-j)) v(b)) (v(b+j) + v(b))
//            - 2 dt(b) aam(b) (vb(b+j) - vb(b)) / (dy(b) + 1e-20)) dx(b)
//   CV(b) = dt4(b)/4 aam4(b) ((ub(b) - ub(b-j)) / (sy4(b) + 1e-20) + (vb(b) - vb(b-i)) / (sx4(b) + 1e-20))
//   P1(a) = curv(a) dt(a) (v(a+j) + v(a)),  P2(a) = curv(a) dt(a) (u(a+i) + u(a))
//   curv(a) = ((v(a+j) + v(a)) (dy(a+i) - dy(a-i)) - (u(a+i) + u(a)) (dx(a+j) - dx(a-j))) / 4 / (dx(a) dy(a) + 1e-20)
// where f4(b) sums f over the corner's four cells b, b-i, b-j, b-i-j.  The
// thread of point q takes the partials of every face and product that reads
// q, each weighted by that face's cotangent (a difference of the masked
// output cotangents across it).  The 2-D input dt takes its partials per
// level, summed over the water column's 16 lanes.  Every element is written
// by one thread, accumulated into the given arrays.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct AdvctAdjArgs {
  const T *u, *v, *ub, *vb, *aam, *dt;   // primal (dt 2-D)
  const T *dx, *dy, *aru, *arv;          // geometry
  const T *ob_x, *ob_y;                  // cotangents of advx, advy
  T *c_u, *c_v, *c_ub, *c_vb, *c_aam, *c_dt;   // accumulated (c_dt 2-D)
};

template <class T> void launch_adj_advct(const Ctx<T>& c, const AdvctAdjArgs<T>& a);

}  // namespace pom
