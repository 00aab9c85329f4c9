// Kernels and launchers of the reverse of the explicit barotropic substep by
// transposed stencils (pom_extrev.h).  The forward forms are OpExtElfRaw,
// OpBc1, OpExtUVG<false> and OpExtFilter (pom_ops.h); reference:
// src/neuralpom/cores/explicit/time_loop.py:292-431 and the bcond 1 / 2 forms
// of src/neuralpom/cores/explicit/baroclinic.py:510-571.
#include "pom_extrev.h"

namespace pom {
namespace {

// 2-D load at (i, j) of the thread's member, indices clamped into the member
// (a clamped read only feeds a term that is discarded or multiplied by zero)
template <class T> struct L2 {
  const Pos& q;
  __device__ __forceinline__ T operator()(const T* f, int i, int j) const {
    return ld_at(f, q.o2, lix2(q, clampi(i, q.im), clampi(j, q.jm)));
  }
};

template <class T> __device__ __forceinline__ void st2(const Pos& q, T* f, uint32_t ix, T v) { ref_at(f, q.o2, ix) = v; }

// wave speed of the radiation forms, safe_sqrt(g / clamp(h)) as OpExtUVG::cwave
template <class T> __device__ __forceinline__ T cwave(T h, double grav) {
  const T hv = fmax(h, T(1e-8));
  return sqrt(fmax(T(grav) / hv, T(1e-8)));
}

enum { K_NONE = 0, K_INT = 1, K_EDGE_HI = 2, K_EDGE_1 = 3, K_EDGE_0 = 4, K_PASS = 5 };

// the form OpExtUVG<false> output 0 (uaf) takes at (i, j), in its order of precedence
__device__ __forceinline__ int ukind(int i, int j, int im, int jm) {
  if (i < 0 || i >= im || j < 0 || j >= jm) return K_NONE;
  const bool jin = j >= 1 && j <= jm - 2, iin = i >= 1 && i <= im - 2;
  if (jin && i == im - 1) return K_EDGE_HI;   // east radiation, reads el(i-1, j)
  if (jin && i == 1) return K_EDGE_1;         // west radiation, reads el(1, j)
  if (jin && i == 0) return K_EDGE_0;         // west radiation, reads el(1, j)
  if (iin && (j == jm - 1 || j == 0)) return K_NONE;
  if (iin && jin) return K_INT;
  return K_PASS;                              // corners keep the given uaf
}
// output 1 (vaf)
__device__ __forceinline__ int vkind(int i, int j, int im, int jm) {
  if (i < 0 || i >= im || j < 0 || j >= jm) return K_NONE;
  const bool jin = j >= 1 && j <= jm - 2, iin = i >= 1 && i <= im - 2;
  if (jin && (i == im - 1 || i == 0)) return K_NONE;
  if (iin && j == jm - 1) return K_EDGE_HI;   // north radiation, reads el(i, j-1)
  if (iin && j == 1) return K_EDGE_1;         // south radiation, reads el(i, 1)
  if (iin && j == 0) return K_EDGE_0;         // south radiation, reads el(i, 1)
  if (iin && jin) return K_INT;
  return K_PASS;
}

// Partials of the interior momentum form at output point p with neighbour
// p - e (e = i for uaf, j for vaf), scaled by the output's cotangent ct:
//   out = dm num / (den + 1e-20),  num = Hb ar uab - 4 dte prov,  den = Hf ar,
//   Hb = h(p) + elb(p) + h(p-e) + elb(p-e),  Hf the same with elf,
//   prov = adx + adv -/+ ar ct + gt + drx - ar (-wsurf + wbot),
//   gt = G D E,  G = g/4 (ds(p) + ds(p-e)),  D = d(p) + d(p-e),
//   E = (1 - 2 alpha) del el + alpha (del elb + del elf) + del e_atmos.
template <class T> struct Mom {
  T g;      // cotangent of num
  T cgt;    // of prov (and so of gt, adx, adv, drx)
  T G, D, E;
  T cden;   // of elf(p) and of elf(p-e), through den
};

}  // namespace

// the factor planes fv[XF_N] of the outputs uaf(q), vaf(q) (pom_extrev.h);
// the input cotangents reached from q's own outputs only and the advave-period
// accumulators are written when own is set
template <class T>
__device__ __forceinline__ void ext_fac_point(const Pos& q, T rmp, const Params& prm, const ExtRevFacArgs<T>& a,
                                              bool own, T* fv) {
  const T f = T(0.5 * prm.smoth);
  const bool acc = prm.iext != prm.isplit;
  const T isp2i = T(prm.isp2i);
  const T al = T(prm.alpha), al12 = T(1.0 - 2.0 * prm.alpha);
  const T dte4 = T(4.0 * prm.dte), g4 = T(0.25 * prm.grav);
  const T rfe = T(prm.rfe), rfw = T(prm.rfw), rfn = T(prm.rfn), rfs = T(prm.rfs);
  const L2<T> L{q};
  const int im = q.im, jm = q.jm, i = q.i, j = q.j;
  auto sel = [](bool c, T x) { return c ? x : T(0); };
  const uint32_t ix = lix2(q, i, j);
  auto ld = [&](const T* p) { return ld_at(p, q.o2, ix); };
  // interior partials of the output at (i, j), neighbour (i - di, j - dj)
  auto mom = [&](int di, int dj, T ct, T ar, T dm, const T* ds_f, T out) {
    Mom<T> r;
    const T den = (ld(a.h) + ld(a.elf) + L(a.h, i - di, j - dj) + L(a.elf, i - di, j - dj)) * ar;
    const T rden = rcp_lane(den + T(1e-20));
    r.g = ct * dm * rden;
    r.cgt = -dte4 * r.g;
    r.G = g4 * (ld(ds_f) + L(ds_f, i - di, j - dj));
    r.D = ld(a.d) + L(a.d, i - di, j - dj);
    r.E = al12 * (ld(a.el) - L(a.el, i - di, j - dj)) +
          al * (ld(a.elb) - L(a.elb, i - di, j - dj) + ld(a.elf) - L(a.elf, i - di, j - dj)) +
          (ld(a.eatm) - L(a.eatm, i - di, j - dj));
    // d out / d den = -dm num / den^2 = -(dm num / den) / den, the stored output over den
    r.cden = -(ct * rden) * out * ar;
    return r;
  };
  const T h_q = ld(a.h), elf_q = ld(a.elf), elb_q = ld(a.elb);
  // ---- uaf(p): total cotangent = the next substep's (its given uaf) + the filter's (ua, uab, utf)
  {
    const int k = ukind(i, j, im, jm);
    const T ar = ld(a.aru), dm = ld(a.dum), uaf = ld(a.uaf);
    const T tu = ld(a.c_utf) * uaf * isp2i;   // d utf / d elf(p) = d utf / d elf(p-i), times c_utf
    const T ct = ld(a.c_uaf) + ld(a.c_ua) + f * ld(a.c_uab) +
                 sel(acc && i >= 1, ld(a.c_utf) * ((h_q + elf_q) + (L(a.h, i - 1, j) + L(a.elf, i - 1, j))) * isp2i);
    const Mom<T> P = mom(1, 0, ct, ar, dm, a.dy, uaf);
    const bool in = k == K_INT;
    const T s = ct * dm;
    fv[XF_XU] = sel(in, P.g * ar * ld(a.uab));
    fv[XF_YU] = sel(in, P.cgt * P.G * P.D);
    fv[XF_ZU] = sel(in, P.cden) + sel(acc && i >= 1, tu);
    fv[XF_WU] = sel(in, P.cgt * P.G * P.E);
    fv[XF_CU] = sel(in, -ar * P.cgt);
    // radiation: east form reads el(i-1, j), west forms el(1, j)
    fv[XF_RU] = sel(k == K_EDGE_HI && prm.rfe != 0.0, rfe * cwave(L(a.h, i - 1, j), prm.grav) * rmp * s) +
                sel(k == K_EDGE_1 && prm.rfw != 0.0, -(rfw * cwave(h_q, prm.grav)) * rmp * s) +
                sel(k == K_EDGE_0 && prm.rfw != 0.0, -(rfw * cwave(L(a.h, i + 1, j), prm.grav)) * rmp * s);
    if (own) st2(q, a.p_uab, ix, f * ld(a.c_uab) + sel(in, P.g * (h_q + elb_q + L(a.h, i - 1, j) + L(a.elb, i - 1, j)) * ar));
    if (own) st2(q, a.p_uaf, ix, sel(k == K_PASS, s));
    if (own) st2(q, a.a_advua, ix, ld(a.a_advua) + sel(in, P.cgt));
  }
  // ---- vaf(p)
  {
    const int k = vkind(i, j, im, jm);
    const T ar = ld(a.arv), dm = ld(a.dvm), vaf = ld(a.vaf);
    const T tv = ld(a.c_vtf) * vaf * isp2i;
    const T ct = ld(a.c_vaf) + ld(a.c_va) + f * ld(a.c_vab) +
                 sel(acc && j >= 1, ld(a.c_vtf) * ((h_q + elf_q) + (L(a.h, i, j - 1) + L(a.elf, i, j - 1))) * isp2i);
    const Mom<T> P = mom(0, 1, ct, ar, dm, a.dx, vaf);
    const bool in = k == K_INT;
    const T s = ct * dm;
    fv[XF_XV] = sel(in, P.g * ar * ld(a.vab));
    fv[XF_YV] = sel(in, P.cgt * P.G * P.D);
    fv[XF_ZV] = sel(in, P.cden) + sel(acc && j >= 1, tv);
    fv[XF_WV] = sel(in, P.cgt * P.G * P.E);
    fv[XF_CV] = sel(in, ar * P.cgt);
    // radiation: north form reads el(i, j-1), south forms el(i, 1)
    fv[XF_RV] = sel(k == K_EDGE_HI && prm.rfn != 0.0, rfn * cwave(L(a.h, i, j - 1), prm.grav) * rmp * s) +
                sel(k == K_EDGE_1 && prm.rfs != 0.0, -(rfs * cwave(h_q, prm.grav)) * rmp * s) +
                sel(k == K_EDGE_0 && prm.rfs != 0.0, -(rfs * cwave(L(a.h, i, j + 1), prm.grav)) * rmp * s);
    if (own) st2(q, a.p_vab, ix, f * ld(a.c_vab) + sel(in, P.g * (h_q + elb_q + L(a.h, i, j - 1) + L(a.elb, i, j - 1)) * ar));
    if (own) st2(q, a.p_vaf, ix, sel(k == K_PASS, s));
    if (own) st2(q, a.a_advva, ix, ld(a.a_advva) + sel(in, P.cgt));
  }
}

// factor planes of the outputs uaf(p), vaf(p) (pom_extrev.h)
template <class T>
__global__ void __launch_bounds__(TPB) k_ext_rev_fac(Batch B, const T* ramp, Params prm, ExtRevFacArgs<T> a) {
  const int nblk = B.blk2[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    T fv[XF_N];
    ext_fac_point(q, ramp[q.m], prm, a, true, fv);
    const uint32_t ix = lix2(q, q.i, q.j);
#pragma unroll
    for (int n = 0; n < XF_N; ++n) ref_at(a.fac, q.o2 + n * a.n2, ix) = fv[n];
  }
}

// the sums of the factors of the outputs that read q, and the filter's own
// terms; F(n, i, j) is plane n at (i, j) (any value where (i, j) is off the
// member: the existence tests below discard it)
template <class T, class FA>
__device__ __forceinline__ void ext_gat_point(const Pos& q, const Params& prm, const ExtRevGatArgs<T>& a, const FA& F) {
  const T f = T(0.5 * prm.smoth);
  const T f12 = T(1) + f * T(-2);   // d uab_new / d ua
  const bool acc = prm.iext != prm.isplit;
  const T ispi = T(prm.ispi);
  const T al = T(prm.alpha), al12 = T(1.0 - 2.0 * prm.alpha);
  const L2<T> L{q};
  const int im = q.im, jm = q.jm, i = q.i, j = q.j;
  auto sel = [](bool c, T x) { return c ? x : T(0); };
  const uint32_t ix = lix2(q, i, j);
  auto ld = [&](const T* p) { return ld_at(p, q.o2, ix); };
  // factors of the outputs that read q; a neighbour outside the member has
  // no output, and its clamped read is discarded by the existence test
  const bool ie = i + 1 <= im - 1, iw = i >= 1, jn = j + 1 <= jm - 1, js = j >= 1;
  const T YU0 = F(XF_YU, i, j), YU1 = sel(ie, F(XF_YU, i + 1, j));
  const T YV0 = F(XF_YV, i, j), YV1 = sel(jn, F(XF_YV, i, j + 1));
  const T XU = F(XF_XU, i, j) + sel(ie, F(XF_XU, i + 1, j));
  const T XV = F(XF_XV, i, j) + sel(jn, F(XF_XV, i, j + 1));
  const T ZU = F(XF_ZU, i, j) + sel(ie, F(XF_ZU, i + 1, j));
  const T ZV = F(XF_ZV, i, j) + sel(jn, F(XF_ZV, i, j + 1));
  const T WU = F(XF_WU, i, j) + sel(ie, F(XF_WU, i + 1, j));
  const T WV = F(XF_WV, i, j) + sel(jn, F(XF_WV, i, j + 1));
  const T CU0 = F(XF_CU, i, j) + sel(ie, F(XF_CU, i + 1, j));                       // va(q), d(q)
  const T CUs = sel(js, F(XF_CU, i, j - 1)) + sel(js && ie, F(XF_CU, i + 1, j - 1));   // va(q) via cor d at q-j
  const T CV0 = F(XF_CV, i, j) + sel(jn, F(XF_CV, i, j + 1));                       // ua(q), d(q)
  const T CVw = sel(iw, F(XF_CV, i - 1, j)) + sel(iw && jn, F(XF_CV, i - 1, j + 1));   // ua(q) via cor d at q-i
  const T rad = sel(ukind(i, j, im, jm) == K_EDGE_1, F(XF_RU, i, j)) +
                sel(ukind(i + 1, j, im, jm) == K_EDGE_HI, F(XF_RU, i + 1, j)) +
                sel(ukind(i - 1, j, im, jm) == K_EDGE_0, F(XF_RU, i - 1, j)) +
                sel(vkind(i, j, im, jm) == K_EDGE_1, F(XF_RV, i, j)) +
                sel(vkind(i, j + 1, im, jm) == K_EDGE_HI, F(XF_RV, i, j + 1)) +
                sel(vkind(i, j - 1, im, jm) == K_EDGE_0, F(XF_RV, i, j - 1));
  const T qa = T(0.25);
  const T cor_q = ld(a.cor), d_q = ld(a.d);
  const T cd_q = qa * cor_q * d_q, cd_s = qa * L(a.cor, i, j - 1) * L(a.d, i, j - 1),
          cd_w = qa * L(a.cor, i - 1, j) * L(a.d, i - 1, j);
  const T vsum = L(a.va, i, j + 1) + ld(a.va), usum = L(a.ua, i + 1, j) + ld(a.ua);
  const T c_uab = ld(a.c_uab), c_vab = ld(a.c_vab), c_elb = ld(a.c_elb);
  const T dY = (YU0 - YU1) + (YV0 - YV1);
  // filter's own terms of elf: el' = elf, elb' (f), d' = h + elf, egf' (ispi), etf'
  T celf = (a.elf_first ? T(0) : ld(a.c_elf)) + ld(a.c_el) + f * c_elb + ld(a.c_d) + sel(acc, ld(a.c_egf) * ispi) +
           ZU + ZV + al * dY;
  const int n = prm.iext, ns = prm.isplit;
  const T ce = ld(a.c_etf);
  if (n == ns - 2) {
    celf += T(0.25 * prm.smoth) * ce;
    st2(q, a.c_etf, ix, T(0));
  } else if (n == ns - 1) {
    celf += T(0.5 * (1.0 - 0.5 * prm.smoth)) * ce;
  } else if (n == ns) {
    const T fs = ld(a.fsm);
    celf += T(0.5) * ce * fs;
    st2(q, a.c_etf, ix, ce * fs);
  }
  st2(q, a.c_elf, ix, celf);
  st2(q, a.p_ua, ix, f12 * c_uab + CV0 * cd_q + CVw * cd_w);
  st2(q, a.p_va, ix, f12 * c_vab + CU0 * cd_q + CUs * cd_s);
  st2(q, a.p_el, ix, f12 * c_elb + al12 * dY + rad);
  st2(q, a.p_elb, ix, f * c_elb + XU + XV + al * dY);
  st2(q, a.p_d, ix, WU + WV + CU0 * (qa * cor_q * vsum) + CV0 * (qa * cor_q * usum));
}

// sums of the factor planes over the outputs that read q, and the filter's own terms
template <class T>
__global__ void __launch_bounds__(TPB) k_ext_rev_gat(Batch B, Params prm, ExtRevGatArgs<T> a) {
  const int nblk = B.blk2[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    const int im = q.im, jm = q.jm;
    auto F = [&](int n, int ii, int jj) {
      return ld_at(a.fac, q.o2 + n * a.n2, lix2(q, clampi(ii, im), clampi(jj, jm)));
    };
    ext_gat_point(q, prm, a, F);
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_ext_rev_elf(Batch B, Params prm, ExtRevElfArgs<T> a) {
  const int nblk = B.blk2[B.M];
  const T dte2 = T(prm.dte2);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    const L2<T> L{q};
    const int im = q.im, jm = q.jm, qi = q.i, qj = q.j;
    auto inner = [&](int i, int j) { return i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2; };
    auto sel = [](bool c, T x) { return c ? x : T(0); };
    auto ce = [&](int i, int j) { return L(a.fsm, i, j) * L(a.c_elf, i, j); };
    // bcond 1: every edge cell copies its nearest interior cell (corners the
    // diagonal one), so el_raw at an interior point takes its own elf's
    // cotangent and those of the edge cells that copy it.  All candidates are
    // read (clamped) and selected, with no branch around the reads.
    const T cnr00 = ce(0, 0), cnr0n = ce(0, jm - 1), cnrm0 = ce(im - 1, 0), cnrmn = ce(im - 1, jm - 1);
    auto craw = [&](int i, int j) {
      const bool w = i == 1, e = i == im - 2, s = j == 1, n = j == jm - 2;
      return ce(i, j) + sel(w, ce(0, j)) + sel(e, ce(im - 1, j)) + sel(s, ce(i, 0)) + sel(n, ce(i, jm - 1)) +
             sel(w && s, cnr00) + sel(w && n, cnr0n) + sel(e && s, cnrm0) + sel(e && n, cnrmn);
    };
    // cotangent of the flux divergence at (i, j): el_raw = elb - dte2 (div / art + vfluxf) inside
    auto wdiv = [&](int i, int j) { return sel(inner(i, j), -(dte2 * craw(i, j)) * rcp_lane(L(a.art, i, j))); };
    const T wq = wdiv(qi, qj), wim = wdiv(qi - 1, qj), wip = wdiv(qi + 1, qj), wjm = wdiv(qi, qj - 1),
            wjp = wdiv(qi, qj + 1);
    const T d_q = L(a.d, qi, qj), dy_q = L(a.dy, qi, qj), dx_q = L(a.dx, qi, qj);
    // fu(a) = 0.25 (d(a) + d(a-i)) (dy(a) + dy(a-i)) ua(a) enters div at a-i (+) and at a (-)
    const T sdy = dy_q + L(a.dy, qi - 1, qj), sdx = dx_q + L(a.dx, qi, qj - 1);
    const T cfu0 = wim - wq, cfu1 = wq - wip, cfv0 = wjm - wq, cfv1 = wq - wjp;
    const T cua = sel(qi >= 1, cfu0 * T(0.25) * (d_q + L(a.d, qi - 1, qj)) * sdy);
    const T cva = sel(qj >= 1, cfv0 * T(0.25) * (d_q + L(a.d, qi, qj - 1)) * sdx);
    const T cd = sel(qi >= 1, cfu0 * T(0.25) * sdy * L(a.ua, qi, qj)) +
                 sel(qi + 1 <= im - 1, cfu1 * T(0.25) * (L(a.dy, qi + 1, qj) + dy_q) * L(a.ua, qi + 1, qj)) +
                 sel(qj >= 1, cfv0 * T(0.25) * sdx * L(a.va, qi, qj)) +
                 sel(qj + 1 <= jm - 1, cfv1 * T(0.25) * (L(a.dx, qi, qj + 1) + dx_q) * L(a.va, qi, qj + 1));
    const uint32_t ix = lix2(q, qi, qj);
    auto ld = [&](const T* p) { return ld_at(p, q.o2, ix); };
    const T o_ua = ld(a.p_ua), o_va = ld(a.p_va), o_d = ld(a.p_d), o_elb = ld(a.p_elb);   // reads before stores
    st2(q, a.p_ua, ix, o_ua + cua);
    st2(q, a.p_va, ix, o_va + cva);
    st2(q, a.p_d, ix, o_d + cd);
    if (inner(qi, qj)) st2(q, a.p_elb, ix, o_elb + craw(qi, qj));   // the store alone is conditional
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_ext_fold(Batch B, ExtFoldArgs<T> a) {
  const int nblk = B.blk2[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    const uint32_t ix = lix2(q, q.i, q.j);
    T* const dst[6] = {a.c_adx, a.c_drx, a.c_wubot, a.c_ady, a.c_dry, a.c_wvbot};
    const T su = ld_at(a.acc_u, q.o2, ix), sv = ld_at(a.acc_v, q.o2, ix);
    const T add[6] = {su, su, -ld_at(a.aru, q.o2, ix) * su, sv, sv, -ld_at(a.arv, q.o2, ix) * sv};
    T old[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) old[n] = ld_at(dst[n], q.o2, ix);
#pragma unroll
    for (int n = 0; n < 6; ++n) ref_at(dst[n], q.o2, ix) = old[n] + add[n];
    if (a.clear) {
      ref_at(a.acc_u, q.o2, ix) = T(0);
      ref_at(a.acc_v, q.o2, ix) = T(0);
    }
  }
}

template <class T, bool TIDE>
__global__ void __launch_bounds__(TPB) k_ext_elf(Batch B, Params prm, ExtElfArgs<T> a) {
  const int nblk = B.blk2[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    st2(q, a.elf, lix2(q, q.i, q.j), ext_elf_at<T, TIDE>(q, q.i, q.j, a, prm.dte2));
  }
}

// the fused forward substep (pom_extrev.h launch_ext_substep)
template <class T>
__global__ void __launch_bounds__(TPB) k_ext_substep(Batch B, Vert<T> vert, const T* ramp, Params prm,
                                                     ExtElfArgs<T> e, FieldTab<T> ta, OutTab<T> oa, FieldTab<T> tb,
                                                     OutTab<T> ob, T* amax) {
  T vmax = T(0);
  for (int blk = logical_block(); blk < B.blk2[B.M]; blk += (int)gridDim.x) {
    Pos p;
    if (pos2_at(B, blk, p)) {
      const uint32_t li = lix2(p, p.i, p.j);
      // elf where OpExtUV and OpExtFilter read it; a neighbour off the member
      // is clamped onto it, as the memory reads of the two-launch form are
      const T e0 = ext_elf_at<T, false>(p, p.i, p.j, e, prm.dte2);
      const T ew = ext_elf_at<T, false>(p, clampi(p.i - 1, p.im), p.j, e, prm.dte2);
      const T es = ext_elf_at<T, false>(p, p.i, clampi(p.j - 1, p.jm), e, prm.dte2);
      T ra[2], rb[OpExtFilter::NOUT];
      {
        const AccElf<T, OpExtUV::ELF> A{{&ta, p, vert, ramp}, e0, ew, es, T(0), T(0)};
        ra[0] = val(OpExtUV::template eval<0>(A, prm));
        ra[1] = val(OpExtUV::template eval<1>(A, prm));
      }
      {
        const AccElf<T, OpExtFilter::ELF, OpExtFilter::UAF, OpExtFilter::VAF> A{{&tb, p, vert, ramp}, e0, ew, es,
                                                                               ra[0], ra[1]};
        static_for<0, OpExtFilter::NOUT - 1>([&](auto O) {
          constexpr int o = decltype(O)::value;
          rb[o] = val(OpExtFilter::template eval<o>(A, prm));
        });
      }
      vmax = nanmax(vmax, fabs(ra[1]));
      st2(p, e.elf, li, e0);
      if (oa.o[0]) st2(p, oa.o[0], li, ra[0]);
      if (oa.o[1]) st2(p, oa.o[1], li, ra[1]);
      static_for<0, OpExtFilter::NOUT - 1>([&](auto O) {
        constexpr int o = decltype(O)::value;
        if (ob.o[o]) st2(p, ob.o[o], li, rb[o]);
      });
    }
  }
  if (amax) block_amax(vmax, amax);   // uniform over the launch
}

template <class T>
void launch_ext_substep(const Ctx<T>& c, const ExtElfArgs<T>& e, const FieldTab<T>& tuv, const OutTab<T>& ouv,
                        const FieldTab<T>& tf, const OutTab<T>& of, T* amax) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_substep<T>), dim3(groups(c.nblk2, c.tune.map2_bpg, c.tune.map2_min)), dim3(TPB), 0,
                     c.stream, c.B, c.vert, c.ramp, c.prm, e, tuv, ouv, tf, of, amax);
}
template void launch_ext_substep<float>(const Ctx<float>&, const ExtElfArgs<float>&, const FieldTab<float>&,
                                        const OutTab<float>&, const FieldTab<float>&, const OutTab<float>&, float*);
template void launch_ext_substep<double>(const Ctx<double>&, const ExtElfArgs<double>&, const FieldTab<double>&,
                                         const OutTab<double>&, const FieldTab<double>&, const OutTab<double>&,
                                         double*);

template <class T> void launch_ext_elf(const Ctx<T>& c, const ExtElfArgs<T>& a) {
  if (c.nblk2 == 0) return;
  const dim3 grid(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min));
  if (a.node)
    hipLaunchKernelGGL((k_ext_elf<T, true>), grid, dim3(TPB), 0, c.stream, c.B, c.prm, a);
  else
    hipLaunchKernelGGL((k_ext_elf<T, false>), grid, dim3(TPB), 0, c.stream, c.B, c.prm, a);
}
template void launch_ext_elf<float>(const Ctx<float>&, const ExtElfArgs<float>&);
template void launch_ext_elf<double>(const Ctx<double>&, const ExtElfArgs<double>&);

template <class T> void launch_ext_rev_fac(const Ctx<T>& c, const ExtRevFacArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_rev_fac<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, c.ramp, c.prm, a);
}
template <class T> void launch_ext_rev_gat(const Ctx<T>& c, const ExtRevGatArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_rev_gat<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, c.prm, a);
}
template <class T> void launch_ext_rev_elf(const Ctx<T>& c, const ExtRevElfArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_rev_elf<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, c.prm, a);
}
template <class T> void launch_ext_fold(const Ctx<T>& c, const ExtFoldArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_fold<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, a);
}

template void launch_ext_rev_fac<float>(const Ctx<float>&, const ExtRevFacArgs<float>&);
template void launch_ext_rev_fac<double>(const Ctx<double>&, const ExtRevFacArgs<double>&);
template void launch_ext_rev_gat<float>(const Ctx<float>&, const ExtRevGatArgs<float>&);
template void launch_ext_rev_gat<double>(const Ctx<double>&, const ExtRevGatArgs<double>&);
template void launch_ext_fold<float>(const Ctx<float>&, const ExtFoldArgs<float>&);
template void launch_ext_fold<double>(const Ctx<double>&, const ExtFoldArgs<double>&);
template void launch_ext_rev_elf<float>(const Ctx<float>&, const ExtRevElfArgs<float>&);
template void launch_ext_rev_elf<double>(const Ctx<double>&, const ExtRevElfArgs<double>&);

}  // namespace pom
