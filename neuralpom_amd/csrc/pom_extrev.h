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

// elf at the point (i, j) of q's member (in the member; ExtElfArgs): the
// continuity update at the interior point (i, j) copies, then bcond 1, or
// with the tide policy (TIDE) its node / interior / previous-value form
template <class T, bool TIDE>
__device__ __forceinline__ T ext_elf_at(const Pos& q, int i, int j, const ExtElfArgs<T>& a, double dte2) {
  const int im = q.im, jm = q.jm;
  auto L = [&](const T* f, int ii, int jj) { return ld_at(f, q.o2, lix2(q, clampi(ii, im), clampi(jj, jm))); };
  const int mi = i == 0 ? 1 : (i == im - 1 ? im - 2 : i);
  const int mj = j == 0 ? 1 : (j == jm - 1 ? jm - 2 : j);
  const T qa = T(0.25);
  auto fu = [&](int ii, int jj) {
    return qa * (L(a.d, ii, jj) + L(a.d, ii - 1, jj)) * (L(a.dy, ii, jj) + L(a.dy, ii - 1, jj)) * L(a.ua, ii, jj);
  };
  auto fv = [&](int ii, int jj) {
    return qa * (L(a.d, ii, jj) + L(a.d, ii, jj - 1)) * (L(a.dx, ii, jj) + L(a.dx, ii, jj - 1)) * L(a.va, ii, jj);
  };
  const T div = fu(mi + 1, mj) - fu(mi, mj) + fv(mi, mj + 1) - fv(mi, mj);
  const T term = -(div / L(a.art, mi, mj)) - L(a.vfluxf, mi, mj);
  const T raw = L(a.elb, mi, mj) + T(dte2) * term;
  const uint32_t ix = lix2(q, i, j);
  const T fsm = ld_at(a.fsm, q.o2, ix);
  if constexpr (TIDE) {
    const bool inside = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2;
    T v = inside ? raw : ld_at(a.elf_prev, q.o2, ix);
    if (ld_at(a.node, q.o2, ix) != T(0))
      v = T(double(ld_at(a.amp, q.o2, ix)) * cos(a.wt - double(ld_at(a.phase, q.o2, ix))));
    return v * fsm;
  } else {
    const bool in = mi >= 1 && mi <= im - 2 && mj >= 1 && mj <= jm - 2;
    return (in ? raw : T(0)) * fsm;
  }
}

// Forward: one substep of the closed-basin explicit core in one pass: elf at
// the point and at its (i-1) and (j-1) neighbours (ext_elf_at, recomputed from
// the substep's inputs instead of read back), the depth-mean momentum update
// with bcond 2 (OpExtUV, table tuv, outputs ouv = uaf, vaf) and the filter,
// level shift and time means (OpExtFilter, table tf, outputs of), elf stored
// to e.elf; with amax non-null the running max of |vaf| is accumulated there.
// Same values as k_ext_elf followed by k_map2 over the same tables.
template <class T>
void launch_ext_substep(const Ctx<T>& c, const ExtElfArgs<T>& e, const FieldTab<T>& tuv, const OutTab<T>& ouv,
                        const FieldTab<T>& tf, const OutTab<T>& of, T* amax);

// c.prm.iext is the substep
template <class T> void launch_ext_rev_fac(const Ctx<T>& c, const ExtRevFacArgs<T>& a);
template <class T> void launch_ext_rev_gat(const Ctx<T>& c, const ExtRevGatArgs<T>& a);
template <class T> void launch_ext_rev_elf(const Ctx<T>& c, const ExtRevElfArgs<T>& a);
template <class T> void launch_ext_fold(const Ctx<T>& c, const ExtFoldArgs<T>& a);
template <class T> void launch_ext_elf(const Ctx<T>& c, const ExtElfArgs<T>& a);

}  // namespace pom
