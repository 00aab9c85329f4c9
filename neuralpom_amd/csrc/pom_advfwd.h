// Forward advection of a pair of 3-D fields that share their velocities,
// viscosity and depths: advt1 of T and S (OpAdvt1; reference
// src/neuralpom/cores/explicit/tracers.py:433-569) and advq of q2 and q2l
// (OpAdvq; tracers.py:14-156), one thread per cell with a water column's 16
// levels in 16 adjacent lanes.  The shared fields are read once for both,
// and every value at level k - 1 or k + 1 (of the cell's own column or of a
// neighbouring one) comes from the adjacent lane by a shuffle instead of a
// load (at the surface and the bottom the lane's own value, as the clamped
// level of the point-function form).  Each output is formed in the
// operator's own order of operations, so the values equal k_map's.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct AdvPairArgs {
  const T *fb[2], *f[2], *fclim[2];          // per field (fclim: advt1 only)
  const T *u, *v, *w, *aam;                  // shared 3-D
  const T *dt, *etb, *etf;                   // shared 2-D
  const T *dx, *dy, *h, *art, *dum, *dvm;    // geometry
  T* out[2];
};

// the momentum tendencies and leapfrog of u and v together (OpAdvu, OpAdvv;
// reference src/neuralpom/cores/explicit/momentum.py:14-248)
template <class T> struct AdvUVArgs {
  const T *advx, *drhox, *ub, *advy, *drhoy, *vb;   // per equation (3-D)
  const T *u, *v, *w;                               // shared 3-D
  const T *dt, *egf, *egb, *etf, *etb;              // shared 2-D
  const T *aru, *arv, *cor, *dx, *dy, *h, *eatm;    // geometry (e_atmos)
  T *uf, *vf;
};

template <class T> void launch_advt1_pair(const Ctx<T>& c, const AdvPairArgs<T>& a);
template <class T> void launch_advuv(const Ctx<T>& c, const AdvUVArgs<T>& a);
template <class T> void launch_advq_pair(const Ctx<T>& c, const AdvPairArgs<T>& a);

}  // namespace pom
