// Reverse of one explicit barotropic substep by transposed stencils.
//
// Forward substep n (explicit/time_loop.py:292-431): el_raw = continuity
// (OpExtElfRaw), elf = bcond 1 of el_raw (OpBc1), uaf / vaf = depth-mean
// momentum with bcond 2 (OpExtUV), then the filter, level shift and time
// means (OpExtFilter).  Its reverse runs in two launches instead of one
// generic gather per operator (each of which re-evaluates the operator's
// point function in dual numbers once per stencil offset):
//
//   k_ext_rev_uv   filter and depth-mean momentum.  The thread of point q
//                  takes the partials of the outputs that read q: uaf at q,
//                  q+i, q-j, q+i-j (and the west edge form at q-i), vaf at q,
//                  q+j, q-i, q-i+j (and the south edge form at q-j), the
//                  filter's outputs at q and the time means at q+i, q+j.  The
//                  filter's cotangents of uaf / vaf are recomputed at each of
//                  those points from the output cotangents.  Writes every
//                  input cotangent of the substep once (assign), completes
//                  the cotangent of elf, and accumulates the forcing terms.
//   k_ext_rev_elf  bcond 1 and continuity: el_raw's cotangent at q and its
//                  four neighbours (the edge copies folded onto the interior
//                  point they copy), then the face-flux partials.
//
// The cotangents of egf, utf and vtf pass through a substep unchanged (each is
// its own previous value plus a term), so both substep buffers name one array
// and nothing is copied; etf's is scaled in place on the three substeps that
// change it.  Every element is written by one thread: deterministic.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct ExtRevArgs {
  // primal: the substep's inputs (state n-1), its elf (after bcond 1), uaf, vaf
  const T *el, *elb, *d, *ua, *va, *uab, *vab;
  const T *elf, *uaf, *vaf;
  // geometry and the substep forcing that enters uaf / vaf as given fields
  const T *cor, *dx, *dy, *h, *aru, *arv, *dum, *dvm, *eatm, *fsm;
  // cotangents of the substep's outputs
  const T *c_ua, *c_uab, *c_va, *c_vab, *c_el, *c_elb, *c_d, *c_egf, *c_utf, *c_vtf, *c_uaf, *c_vaf;
  T* c_etf;   // in: of the output etf; out: of the input etf (in place)
  T* c_elf;   // of this substep's elf: completed here (assigned when elf_first)
  // cotangents of the substep's inputs (assigned)
  T *p_ua, *p_uab, *p_va, *p_vab, *p_el, *p_elb, *p_d, *p_uaf, *p_vaf;
  // accumulated over the substeps: advua / advva (the carried advave fields)
  // and the step's depth-mean forcing adx2d, ady2d, drx2d, dry2d, wubot, wvbot
  T *a_advua, *a_advva, *a_adx2d, *a_ady2d, *a_drx2d, *a_dry2d, *a_wubot, *a_wvbot;
  int32_t elf_first;
};

template <class T> struct ExtRevElfArgs {
  const T* c_elf;                                // complete cotangent of the substep's elf
  const T *d, *ua, *va, *dx, *dy, *art, *fsm;    // primal (state n-1) and geometry
  T *p_ua, *p_va, *p_d, *p_elb;                  // accumulated (k_ext_rev_uv assigned them)
};

// c.prm.iext is the substep
template <class T> void launch_ext_rev_uv(const Ctx<T>& c, const ExtRevArgs<T>& a);
template <class T> void launch_ext_rev_elf(const Ctx<T>& c, const ExtRevElfArgs<T>& a);

}  // namespace pom
