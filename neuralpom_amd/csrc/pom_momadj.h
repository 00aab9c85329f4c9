// Adjoint of the internal-mode momentum updates advu and advv (OpAdvu,
// OpAdvv in pom_ops.h; reference src/neuralpom/cores/explicit/momentum.py:14-248)
// as one transposed stencil for both equations.
//
// At an interior point p, level k < kb-1, the u equation is
//   uf = (dold ub - 2 dti2 rhs) / (dnew + 1e-20)
//   rhs = advx + (vfl(k) - vfl(k+1)) aru / dz(k) + corio + pg + drhox
//   vfl(kk) = 0.25 (w(p, kk) + w(p-i, kk)) (u(p, kk) + u(p, kk-1))   for 1 <= kk < kb-1, else 0
//   corio = -aru/4 (cor(p) dt(p) (v(p+j) + v(p)) + cor(p-i) dt(p-i) (v(p-i+j) + v(p-i)))
//   pg = g/8 (dt(p) + dt(p-i)) (del egf + del egb + 2 del e_atmos) (dy(p) + dy(p-i))
//   dnew = (h(p) + h(p-i) + etf(p) + etf(p-i)) aru,  dold the same with etb
// and the v equation the same with i and j exchanged, +arv/4 for the
// Coriolis sign, dx for dy.  The thread of point q (one per 3-D point, k the
// fastest index) takes the partials of the u outputs at q, q+i, q-j, q+i-j and
// of the v outputs at q, q+j, q-i, q-i+j, all at its own level; the
// vertical-flux neighbours at k-1 and k+1 come from the lanes of the same
// water column by shuffles, and the 2-D inputs' partials are summed over the
// column's 16 lanes.  Every cotangent element is written by one thread (no
// atomics); the cotangents are accumulated into the given arrays.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct MomAdjArgs {
  const T *u, *v, *w, *ub, *vb, *uf, *vf;       // 3-D primal; uf, vf: the forward's outputs
  const T *dt, *egf, *egb, *etf, *etb;          // 2-D primal
  const T *aru, *arv, *cor, *dx, *dy, *h, *eatm;
  const T *ob_u, *ob_v;                         // cotangents of uf and vf
  T *c_advx, *c_drhox, *c_advy, *c_drhoy, *c_u, *c_v, *c_w, *c_ub, *c_vb;   // 3-D, accumulated
  T *c_dt, *c_egf, *c_egb, *c_etf, *c_etb;      // 2-D, accumulated
};

template <class T> void launch_adj_advuv(const Ctx<T>& c, const MomAdjArgs<T>& a);

}  // namespace pom
