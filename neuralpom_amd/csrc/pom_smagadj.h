// Adjoint of the explicit core's Smagorinsky viscosity (OpSmag in pom_ops.h;
// reference src/neuralpom/cores/explicit/time_loop.py:10-16, 192-233) as a
// transposed stencil in two launches, in place of the generic gather (which
// evaluates the point function in dual numbers at each of its nine offsets).
//
// At an interior cell p (level k < kb-1):
//   aam(p) = horcon dx dy |(ux, vy, sh)|,  ux = (u(p+i) - u(p)) / dx,  vy = (v(p+j) - v(p)) / dy,
//   sh = (ut + vt) / sqrt(2),  ut = (u(p+j) + u(p+i+j) - u(p-j) - u(p+i-j)) / (4 dy),
//   vt = (v(p+i) + v(p+i+j) - v(p-i) - v(p-i+j)) / (4 dx);
// elsewhere aam(p) is the given aam(p).  With G = c(p) horcon dx dy / |.|
// (zero at a zero strain, torch.linalg.vector_norm's subgradient) the
// partials of c(p) aam(p) are G ux / dx at u(p+i) (minus at u(p)), G vy / dy
// at v(p+j) (minus at v(p)), H / dy at u(p+j), u(p+i+j) (minus at u(p-j),
// u(p+i-j)) and H / dx at v(p+i), v(p+i+j) (minus at v(p-i), v(p-i+j)),
// H = G sh / (4 sqrt(2)).
//   k_smag_fac  per cell p: the four factor planes FU = G ux / dx,
//               FV = G vy / dy, HY = H / dy, HX = H / dx (zero where p is not
//               an interior update), and aam's pass-through cotangent;
//   k_smag_gat  per cell q: the u and v cotangents from the planes at q and
//               its neighbours.
// Every element is written by one thread, accumulated into the given arrays.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct SmagAdjArgs {
  const T *u, *v, *dx, *dy;   // primal (u, v 3-D; dx, dy 2-D)
  const T* ob;                // cotangent of the new aam
  T *fu, *fv, *hy, *hx;       // factor planes (3-D scratch)
  T *c_u, *c_v, *c_aam;       // accumulated
};

template <class T> void launch_adj_smag(const Ctx<T>& c, const SmagAdjArgs<T>& a);

}  // namespace pom
