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

template <class T>
__global__ void __launch_bounds__(TPB) k_ext_rev_uv(Batch B, const T* ramp, Params prm, ExtRevArgs<T> a) {
  const int nblk = B.blk2[B.M];
  const T f = T(0.5 * prm.smoth);
  const T f12 = T(1) + f * T(-2);   // d uab_new / d ua
  const bool acc = prm.iext != prm.isplit;
  const T isp2i = T(prm.isp2i), ispi = T(prm.ispi);
  const T al = T(prm.alpha), al12 = T(1.0 - 2.0 * prm.alpha);
  const T dte4 = T(4.0 * prm.dte), g4 = T(0.25 * prm.grav);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    const L2<T> L{q};
    const int im = q.im, jm = q.jm, qi = q.i, qj = q.j;
    const T rmp = ramp[q.m];
    // total cotangent of uaf (vaf) at (i, j): the next substep's, which read
    // it as its given uaf, plus the filter's (ua = uaf, uab, utf)
    auto ctu = [&](int i, int j) {
      const T c = L(a.c_uaf, i, j) + L(a.c_ua, i, j) + f * L(a.c_uab, i, j);
      const T t = L(a.c_utf, i, j) * ((L(a.h, i, j) + L(a.elf, i, j)) + (L(a.h, i - 1, j) + L(a.elf, i - 1, j))) * isp2i;
      return (acc && i >= 1) ? c + t : c;
    };
    auto ctv = [&](int i, int j) {
      const T c = L(a.c_vaf, i, j) + L(a.c_va, i, j) + f * L(a.c_vab, i, j);
      const T t = L(a.c_vtf, i, j) * ((L(a.h, i, j) + L(a.elf, i, j)) + (L(a.h, i, j - 1) + L(a.elf, i, j - 1))) * isp2i;
      return (acc && j >= 1) ? c + t : c;
    };
    auto mom = [&](int i, int j, int di, int dj, T ct, const T* ar_f, const T* dm_f, const T* ds_f, const T* out) {
      Mom<T> r;
      const T ar = L(ar_f, i, j);
      const T s = ct * L(dm_f, i, j);
      const T den = (L(a.h, i, j) + L(a.elf, i, j) + L(a.h, i - di, j - dj) + L(a.elf, i - di, j - dj)) * ar;
      const T rden = rcp_lane(den + T(1e-20));
      r.g = s * rden;
      r.cgt = -dte4 * r.g;
      r.G = g4 * (L(ds_f, i, j) + L(ds_f, i - di, j - dj));
      r.D = L(a.d, i, j) + L(a.d, i - di, j - dj);
      r.E = al12 * (L(a.el, i, j) - L(a.el, i - di, j - dj)) +
            al * (L(a.elb, i, j) - L(a.elb, i - di, j - dj) + L(a.elf, i, j) - L(a.elf, i - di, j - dj)) +
            (L(a.eatm, i, j) - L(a.eatm, i - di, j - dj));
      // d out / d den = -dm num / den^2 = -(dm num / den) / den, the stored output over den
      r.cden = -(ct * rden) * L(out, i, j) * ar;
      return r;
    };

    // Every candidate output is evaluated (clamped reads) and its terms are
    // selected by the form the output takes: with no branch around them, the
    // loads of all candidates are in flight together.
    auto sel = [](bool c, T x) { return c ? x : T(0); };
    const T aru_q = L(a.aru, qi, qj), arv_q = L(a.arv, qi, qj);
    const T cor_q = L(a.cor, qi, qj), d_q = L(a.d, qi, qj), h_q = L(a.h, qi, qj);
    const T cw_q = cwave(h_q, prm.grav);
    const T rfe = T(prm.rfe), rfw = T(prm.rfw), rfn = T(prm.rfn), rfs = T(prm.rfs);
    T r_ua, r_uab, r_va, r_vab, r_el, r_elb, r_elf, r_d, r_uafin, r_vafin, cg_u, cg_v;

    // ---- uaf outputs that read q: at q, q+i (q is its west neighbour), q-j
    // and q+i-j (va(q) in their Coriolis terms), q-i (west form at i = 0)
    {
      const int k0 = ukind(qi, qj, im, jm), k1 = ukind(qi + 1, qj, im, jm);
      const int k2 = ukind(qi, qj - 1, im, jm), k3 = ukind(qi + 1, qj - 1, im, jm);
      const int k4 = ukind(qi - 1, qj, im, jm);
      const T ct0 = ctu(qi, qj), ct1 = ctu(qi + 1, qj);
      const Mom<T> P0 = mom(qi, qj, 1, 0, ct0, a.aru, a.dum, a.dy, a.uaf);
      const Mom<T> P1 = mom(qi + 1, qj, 1, 0, ct1, a.aru, a.dum, a.dy, a.uaf);
      const Mom<T> P2 = mom(qi, qj - 1, 1, 0, ctu(qi, qj - 1), a.aru, a.dum, a.dy, a.uaf);
      const Mom<T> P3 = mom(qi + 1, qj - 1, 1, 0, ctu(qi + 1, qj - 1), a.aru, a.dum, a.dy, a.uaf);
      const T s0 = ct0 * L(a.dum, qi, qj), s1 = ct1 * L(a.dum, qi + 1, qj);
      const T s4 = ctu(qi - 1, qj) * L(a.dum, qi - 1, qj);
      const T ar1 = L(a.aru, qi + 1, qj);
      const T cv0 = T(0.25) * cor_q * (L(a.va, qi, qj + 1) + L(a.va, qi, qj));
      const T cd_s = T(0.25) * L(a.cor, qi, qj - 1) * L(a.d, qi, qj - 1);
      const bool i0 = k0 == K_INT, i1 = k1 == K_INT;
      const T gdd0 = P0.cgt * P0.G * P0.D, gdd1 = P1.cgt * P1.G * P1.D;
      const T cct0 = -aru_q * P0.cgt, cct1 = -ar1 * P1.cgt;
      r_uab = sel(i0, P0.g * (h_q + L(a.elb, qi, qj) + L(a.h, qi - 1, qj) + L(a.elb, qi - 1, qj)) * aru_q);
      r_elb = sel(i0, P0.g * aru_q * L(a.uab, qi, qj) + gdd0 * al) + sel(i1, P1.g * ar1 * L(a.uab, qi + 1, qj) - gdd1 * al);
      r_el = sel(i0, gdd0 * al12) - sel(i1, gdd1 * al12) +
             sel(k0 == K_EDGE_1 && prm.rfw != 0.0, -(rfw * cw_q) * rmp * s0) + sel(k1 == K_EDGE_HI && prm.rfe != 0.0, rfe * cw_q * rmp * s1) +
             sel(k4 == K_EDGE_0 && prm.rfw != 0.0, -(rfw * cw_q) * rmp * s4);
      r_elf = sel(i0, gdd0 * al + P0.cden) + sel(i1, P1.cden - gdd1 * al);
      r_d = sel(i0, P0.cgt * P0.G * P0.E + cct0 * cv0) + sel(i1, P1.cgt * P1.G * P1.E + cct1 * cv0);
      r_va = sel(i0, cct0 * T(0.25) * cor_q * d_q) + sel(i1, cct1 * T(0.25) * cor_q * d_q) +
             sel(k2 == K_INT, -L(a.aru, qi, qj - 1) * P2.cgt * cd_s) +
             sel(k3 == K_INT, -L(a.aru, qi + 1, qj - 1) * P3.cgt * cd_s);
      r_uafin = sel(k0 == K_PASS, s0);
      cg_u = sel(i0, P0.cgt);
    }
    // ---- vaf outputs that read q: at q, q+j (q is its south neighbour), q-i
    // and q-i+j (ua(q) in their Coriolis terms), q-j (south form at j = 0)
    {
      const int k0 = vkind(qi, qj, im, jm), k1 = vkind(qi, qj + 1, im, jm);
      const int k2 = vkind(qi - 1, qj, im, jm), k3 = vkind(qi - 1, qj + 1, im, jm);
      const int k4 = vkind(qi, qj - 1, im, jm);
      const T ct0 = ctv(qi, qj), ct1 = ctv(qi, qj + 1);
      const Mom<T> P0 = mom(qi, qj, 0, 1, ct0, a.arv, a.dvm, a.dx, a.vaf);
      const Mom<T> P1 = mom(qi, qj + 1, 0, 1, ct1, a.arv, a.dvm, a.dx, a.vaf);
      const Mom<T> P2 = mom(qi - 1, qj, 0, 1, ctv(qi - 1, qj), a.arv, a.dvm, a.dx, a.vaf);
      const Mom<T> P3 = mom(qi - 1, qj + 1, 0, 1, ctv(qi - 1, qj + 1), a.arv, a.dvm, a.dx, a.vaf);
      const T s0 = ct0 * L(a.dvm, qi, qj), s1 = ct1 * L(a.dvm, qi, qj + 1);
      const T s4 = ctv(qi, qj - 1) * L(a.dvm, qi, qj - 1);
      const T ar1 = L(a.arv, qi, qj + 1);
      const T cu0 = T(0.25) * cor_q * (L(a.ua, qi + 1, qj) + L(a.ua, qi, qj));
      const T cd_w = T(0.25) * L(a.cor, qi - 1, qj) * L(a.d, qi - 1, qj);
      const bool i0 = k0 == K_INT, i1 = k1 == K_INT;
      const T gdd0 = P0.cgt * P0.G * P0.D, gdd1 = P1.cgt * P1.G * P1.D;
      const T cct0 = arv_q * P0.cgt, cct1 = ar1 * P1.cgt;
      r_vab = sel(i0, P0.g * (h_q + L(a.elb, qi, qj) + L(a.h, qi, qj - 1) + L(a.elb, qi, qj - 1)) * arv_q);
      r_elb += sel(i0, P0.g * arv_q * L(a.vab, qi, qj) + gdd0 * al) + sel(i1, P1.g * ar1 * L(a.vab, qi, qj + 1) - gdd1 * al);
      r_el += sel(i0, gdd0 * al12) - sel(i1, gdd1 * al12) +
              sel(k0 == K_EDGE_1 && prm.rfs != 0.0, -(rfs * cw_q) * rmp * s0) + sel(k1 == K_EDGE_HI && prm.rfn != 0.0, rfn * cw_q * rmp * s1) +
              sel(k4 == K_EDGE_0 && prm.rfs != 0.0, -(rfs * cw_q) * rmp * s4);
      r_elf += sel(i0, gdd0 * al + P0.cden) + sel(i1, P1.cden - gdd1 * al);
      r_d += sel(i0, P0.cgt * P0.G * P0.E + cct0 * cu0) + sel(i1, P1.cgt * P1.G * P1.E + cct1 * cu0);
      r_ua = sel(i0, cct0 * T(0.25) * cor_q * d_q) + sel(i1, cct1 * T(0.25) * cor_q * d_q) +
             sel(k2 == K_INT, L(a.arv, qi - 1, qj) * P2.cgt * cd_w) +
             sel(k3 == K_INT, L(a.arv, qi - 1, qj + 1) * P3.cgt * cd_w);
      r_vafin = sel(k0 == K_PASS, s0);
      cg_v = sel(i0, P0.cgt);
    }

    // ---- filter, level shift and time means at q (OpExtFilter)
    const uint32_t ix = lix2(q, qi, qj);
    auto ld = [&](const T* p) { return ld_at(p, q.o2, ix); };
    const T c_uab = ld(a.c_uab), c_vab = ld(a.c_vab), c_elb = ld(a.c_elb);
    T celf = (a.elf_first ? T(0) : ld(a.c_elf)) + ld(a.c_el) + f * c_elb + ld(a.c_d) + r_elf;
    {
      const T t0 = ld(a.c_egf) * ispi, t1 = ld(a.c_utf) * ld(a.uaf) * isp2i;
      const T t2 = L(a.c_utf, qi + 1, qj) * L(a.uaf, qi + 1, qj) * isp2i, t3 = ld(a.c_vtf) * ld(a.vaf) * isp2i;
      const T t4 = L(a.c_vtf, qi, qj + 1) * L(a.vaf, qi, qj + 1) * isp2i;
      celf += sel(acc, t0) + sel(acc && qi >= 1, t1) + sel(acc && qi + 1 <= im - 1, t2) + sel(acc && qj >= 1, t3) +
              sel(acc && qj + 1 <= jm - 1, t4);
    }
    const int n = prm.iext, ns = prm.isplit;
    if (n == ns - 2) {
      celf += T(0.25 * prm.smoth) * ld(a.c_etf);
      st2(q, a.c_etf, ix, T(0));
    } else if (n == ns - 1) {
      celf += T(0.5 * (1.0 - 0.5 * prm.smoth)) * ld(a.c_etf);
    } else if (n == ns) {
      const T ce = ld(a.c_etf), fs = ld(a.fsm);
      celf += T(0.5) * ce * fs;
      st2(q, a.c_etf, ix, ce * fs);
    }
    st2(q, a.c_elf, ix, celf);
    st2(q, a.p_ua, ix, f12 * c_uab + r_ua);
    st2(q, a.p_uab, ix, f * c_uab + r_uab);
    st2(q, a.p_va, ix, f12 * c_vab + r_va);
    st2(q, a.p_vab, ix, f * c_vab + r_vab);
    st2(q, a.p_el, ix, f12 * c_elb + r_el);
    st2(q, a.p_elb, ix, f * c_elb + r_elb);
    st2(q, a.p_d, ix, r_d);
    st2(q, a.p_uaf, ix, r_uafin);
    st2(q, a.p_vaf, ix, r_vafin);
    // forcing cotangents: prov's partials 1 (adx, adv, drx) and -ar (wbot)
    st2(q, a.a_advua, ix, ld(a.a_advua) + cg_u);
    st2(q, a.a_adx2d, ix, ld(a.a_adx2d) + cg_u);
    st2(q, a.a_drx2d, ix, ld(a.a_drx2d) + cg_u);
    st2(q, a.a_wubot, ix, ld(a.a_wubot) + (-aru_q) * cg_u);
    st2(q, a.a_advva, ix, ld(a.a_advva) + cg_v);
    st2(q, a.a_ady2d, ix, ld(a.a_ady2d) + cg_v);
    st2(q, a.a_dry2d, ix, ld(a.a_dry2d) + cg_v);
    st2(q, a.a_wvbot, ix, ld(a.a_wvbot) + (-arv_q) * cg_v);
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
    st2(q, a.p_ua, ix, ld(a.p_ua) + cua);
    st2(q, a.p_va, ix, ld(a.p_va) + cva);
    st2(q, a.p_d, ix, ld(a.p_d) + cd);
    if (inner(qi, qj)) st2(q, a.p_elb, ix, ld(a.p_elb) + craw(qi, qj));   // the store alone is conditional
  }
}

template <class T> void launch_ext_rev_uv(const Ctx<T>& c, const ExtRevArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_rev_uv<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, c.ramp, c.prm, a);
}
template <class T> void launch_ext_rev_elf(const Ctx<T>& c, const ExtRevElfArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_ext_rev_elf<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, c.prm, a);
}

template void launch_ext_rev_uv<float>(const Ctx<float>&, const ExtRevArgs<float>&);
template void launch_ext_rev_uv<double>(const Ctx<double>&, const ExtRevArgs<double>&);
template void launch_ext_rev_elf<float>(const Ctx<float>&, const ExtRevElfArgs<float>&);
template void launch_ext_rev_elf<double>(const Ctx<double>&, const ExtRevElfArgs<double>&);

}  // namespace pom
