// Reverse of one explicit barotropic substep by transposed stencils.
//
// Forward substep n (explicit/time_loop.py:292-431): el_raw = continuity
// (OpExtElfRaw), elf = bcond 1 of el_raw (OpBc1), uaf / vaf = depth-mean
// momentum with bcond 2 (OpExtUV), then the filter, level shift and time
// means (OpExtFilter).  Its reverse runs in three launches instead of one
// generic gather per operator (each of which re-evaluates the operator's
// point function in dual numbers once per stencil offset):
//
//   k_ext_rev_fac  per point p, the partials of the outputs uaf(p) and
//                  vaf(p), each taken once, as 12 factor planes (below); the
//                  cotangents that only p's own outputs reach (uab, vab, the
//                  given uaf / vaf) are written directly, and the forcing
//                  partial into the advave-period accumulator.
//   k_ext_rev_gat  per point q, the sums of the factors of the outputs that
//                  read q (uaf at q, q+i, q-j, q+i-j, q-i; vaf at q, q+j, q-i,
//                  q-i+j, q-j) and the filter's own terms: the input
//                  cotangents ua va el elb d (assigned) and the complete
//                  cotangent of elf.
//   k_ext_rev_elf  bcond 1 and continuity: el_raw's cotangent at q and its
//                  four neighbours (the edge copies folded onto the interior
//                  point they copy), then the face-flux partials.
//
// The cotangents of egf, utf and vtf pass through a substep unchanged (each is
// its own previous value plus a term), so both substep buffers name one array
// and nothing is copied; etf's is scaled in place on the three substeps that
// change it.  The forcing terms adx2d, drx2d, wubot (ady2d, dry2d, wvbot)
// enter uaf (vaf) as advua (advva) does, so their cotangents are the sums of
// the advave-period accumulators, folded in (k_ext_fold) when a period's
// advave call consumes its accumulator.  Every element is written by one
// thread: deterministic.
#pragma once
#include "pom_kernels.h"

namespace pom {

// factor planes of the momentum outputs at a point p (zero where the form
// does not apply), as offsets into one array of 12 planes of n2 elements:
// per direction (u: e = i, v: e = j; ar = aru / arv, mask dum / dvm)
//   X = g ar uab(p)             elb(p) and elb(p-e)
//   Y = cgt G D                 el (+-(1-2 alpha)), elb and elf (+-alpha) at p, p-e
//   Z = cden + time-mean term   elf(p) and elf(p-e)
//   W = cgt G E                 d(p) and d(p-e)
//   C = -+ ar cgt               the Coriolis partials, scaled by cor d / 4 or cor (vel sum) / 4
//   R = radiation partial       el at p (forms at index 1), p - e (the far edge) or p + e (index 0)
// with g = ct dm / (den + 1e-20), cgt = -4 dte g, den, G, D, E, cden as in OpExtUVG.
enum { XF_XU, XF_YU, XF_ZU, XF_WU, XF_CU, XF_RU, XF_XV, XF_YV, XF_ZV, XF_WV, XF_CV, XF_RV, XF_N };

template <class T> struct ExtRevFacArgs {
  // cotangents of the substep's outputs read at the point
  const T *c_uaf, *c_ua, *c_uab, *c_utf, *c_vaf, *c_va, *c_vab, *c_vtf;
  // primal: the substep's inputs (state n-1), its elf (after bcond 1), uaf, vaf; geometry
  const T *h, *elf, *aru, *arv, *dum, *dvm, *dx, *dy, *d, *el, *elb, *eatm, *uab, *vab, *uaf, *vaf;
  T* fac;                   // XF_N planes of n2
  int64_t n2;
  T *p_uab, *p_vab, *p_uaf, *p_vaf;   // input cotangents reached from the point's own outputs only (assigned)
  T *a_advua, *a_advva;     // advave-period accumulators
};

template <class T> struct ExtRevGatArgs {
  const T* fac;
  int64_t n2;
  const T *cor, *d, *ua, *va, *fsm;   // primal (state n-1) and geometry
  const T *c_uab, *c_vab, *c_elb, *c_el, *c_d, *c_egf;
  T* c_etf;   // in: of the output etf; out: of the input etf (in place)
  T* c_elf;   // of this substep's elf: completed here (assigned when elf_first)
  T *p_ua, *p_va, *p_el, *p_elb, *p_d;   // input cotangents (assigned)
  int32_t elf_first;
};

template <class T> struct ExtRevElfArgs {
  const T* c_elf;                                // complete cotangent of the substep's elf
  const T *d, *ua, *va, *dx, *dy, *art, *fsm;    // primal (state n-1) and geometry
  T *p_ua, *p_va, *p_d, *p_elb;                  // accumulated (k_ext_rev_gat assigned them)
};

// the forcing cotangents from an advave-period accumulator: c_adx += acc_u,
// c_drx += acc_u, c_wubot -= aru acc_u (and the v terms); the accumulators
// are cleared when clear is set
template <class T> struct ExtFoldArgs {
  T *acc_u, *acc_v;
  const T *aru, *arv;
  T *c_adx, *c_drx, *c_wubot, *c_ady, *c_dry, *c_wvbot;
  int32_t clear;
};

// Forward: the substep's elf in one pass, bcond 1 of the continuity update:
// elf(p) = el_raw(m) fsm(p), m the interior point that p copies (p itself
// inside), el_raw(m) = elb(m) + dte2 (-(div(m) / art(m)) - vfluxf(m)) as
// OpExtElfRaw, and zero where m is not interior.  With node set (the tide
// policy's bcond 1, case_adapters/coastal_boundary.py:12-22) there is no
// copy: elf(p) = fsm(p) (amp(p) cos(wt - phase(p)) at a node, el_raw(p)
// inside, elf_prev(p), the previous substep's elf, on the other edge points).
template <class T> struct ExtElfArgs {
  const T *ua, *va, *d, *elb, *dx, *dy, *art, *vfluxf, *fsm;
  T* elf;
  const T *node = nullptr, *amp = nullptr, *phase = nullptr, *elf_prev = nullptr;
  double wt = 0.0;   // the substep's tide argument (speed pi / 180) t / 3600
};

// c.prm.iext is the substep
template <class T> void launch_ext_rev_fac(const Ctx<T>& c, const ExtRevFacArgs<T>& a);
template <class T> void launch_ext_rev_gat(const Ctx<T>& c, const ExtRevGatArgs<T>& a);
template <class T> void launch_ext_rev_elf(const Ctx<T>& c, const ExtRevElfArgs<T>& a);
template <class T> void launch_ext_fold(const Ctx<T>& c, const ExtFoldArgs<T>& a);
template <class T> void launch_ext_elf(const Ctx<T>& c, const ExtElfArgs<T>& a);

}  // namespace pom
