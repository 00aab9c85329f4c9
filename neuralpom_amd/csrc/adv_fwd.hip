// Kernels and launchers of the paired forward advection (pom_advfwd.h).
#include "pom_advfwd.h"

namespace pom {

template <class T>
__global__ void __launch_bounds__(TPB) k_advt1_pair(Batch B, Vert<T> vg, Params prm, AdvPairArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T q4 = T(0.25), h2 = T(0.5), tprni = T(prm.tprni), dti2 = T(prm.dti2), eps = T(1e-20);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q{};
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    const uint32_t ix = live ? lix3(q, q.i, q.j, k) : 0u;
    T f0[2] = {T(0), T(0)}, wk = T(0);
    if (live) {
      f0[0] = ld_at(a.f[0], q.o3, ix);
      f0[1] = ld_at(a.f[1], q.o3, ix);
      wk = ld_at(a.w, q.o3, ix);
    }
    const T fup[2] = {__shfl_up(f0[0], 1, KB), __shfl_up(f0[1], 1, KB)};       // level k - 1
    const T fdn[2] = {__shfl_down(f0[0], 1, KB), __shfl_down(f0[1], 1, KB)};   // level k + 1
    const T wdn = __shfl_down(wk, 1, KB);
    if (!live) continue;
    const int im = q.im, jm = q.jm, i = q.i, j = q.j;
    auto L3 = [&](const T* f, int di, int dj) { return ld_at(f, q.o3, lix3(q, clampi(i + di, im), clampi(j + dj, jm), k)); };
    auto L2 = [&](const T* f, int di, int dj) { return ld_at(f, q.o2, lix2(q, clampi(i + di, im), clampi(j + dj, jm))); };
    const bool in = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2 && k < KBM1;
    // shared values of the faces i, i+1, j, j+1
    const T dtc = L2(a.dt, 0, 0), dtw = L2(a.dt, -1, 0), dte = L2(a.dt, 1, 0), dts = L2(a.dt, 0, -1), dtn = L2(a.dt, 0, 1);
    const T amc = L3(a.aam, 0, 0), amw = L3(a.aam, -1, 0), ame = L3(a.aam, 1, 0), ams = L3(a.aam, 0, -1),
            amn = L3(a.aam, 0, 1);
    const T hc = L2(a.h, 0, 0), hw = L2(a.h, -1, 0), he = L2(a.h, 1, 0), hs = L2(a.h, 0, -1), hn = L2(a.h, 0, 1);
    const T dxc = L2(a.dx, 0, 0), dxw = L2(a.dx, -1, 0), dxe = L2(a.dx, 1, 0), dxs = L2(a.dx, 0, -1), dxn = L2(a.dx, 0, 1);
    const T dyc = L2(a.dy, 0, 0), dyw = L2(a.dy, -1, 0), dye = L2(a.dy, 1, 0), dys = L2(a.dy, 0, -1), dyn = L2(a.dy, 0, 1);
    const T u0 = L3(a.u, 0, 0), u1 = L3(a.u, 1, 0), v0 = L3(a.v, 0, 0), v1 = L3(a.v, 0, 1);
    const T dum0 = L2(a.dum, 0, 0), dum1 = L2(a.dum, 1, 0), dvm0 = L2(a.dvm, 0, 0), dvm1 = L2(a.dvm, 0, 1);
    const T art = L2(a.art, 0, 0);
    const T tb = (hc + L2(a.etb, 0, 0)) * art;
    const T tf = (hc + L2(a.etf, 0, 0)) * art;
    const T dz = vg.dz(q.m, k);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const T fc = f0[t], fw = L3(a.f[t], -1, 0), fe = L3(a.f[t], 1, 0), fs = L3(a.f[t], 0, -1), fn = L3(a.f[t], 0, 1);
      const T bc = L3(a.fb[t], 0, 0);
      const T ac = bc - L3(a.fclim[t], 0, 0);
      const T aw = L3(a.fb[t], -1, 0) - L3(a.fclim[t], -1, 0), ae = L3(a.fb[t], 1, 0) - L3(a.fclim[t], 1, 0);
      const T as = L3(a.fb[t], 0, -1) - L3(a.fclim[t], 0, -1), an = L3(a.fb[t], 0, 1) - L3(a.fclim[t], 0, 1);
      // x face between (hi) and (lo): OpAdvt1::xf
      auto xf = [&](T dthi, T dtlo, T fhi, T flo, T u, T amhi, T amlo, T hhi, T hlo, T ahi, T alo, T dm, T dxhi,
                    T dxlo, T dyhi, T dylo) {
        const T adv = q4 * (dthi + dtlo) * (fhi + flo) * u;
        const T dif = h2 * (amhi + amlo) * (hhi + hlo) * tprni * (ahi - alo) * dm / (dxhi + dxlo);
        return (adv - dif) * (h2 * (dyhi + dylo));
      };
      const T x1 = xf(dte, dtc, fe, fc, u1, ame, amc, he, hc, ae, ac, dum1, dxe, dxc, dye, dyc);
      const T x0 = xf(dtc, dtw, fc, fw, u0, amc, amw, hc, hw, ac, aw, dum0, dxc, dxw, dyc, dyw);
      // y faces: OpAdvt1::yf (the mask dvm, the spacing dy, the width dx)
      const T y1 = xf(dtn, dtc, fn, fc, v1, amn, amc, hn, hc, an, ac, dvm1, dyn, dyc, dxn, dxc);
      const T y0 = xf(dtc, dts, fc, fs, v0, amc, ams, hc, hs, ac, as, dvm0, dyc, dys, dxc, dxs);
      // vertical fluxes at k and k + 1: OpAdvt1::zf
      const T z0 = k < KBM1 ? (k == 0 ? fc * wk * art : h2 * (fup[t] + fc) * wk * art) : T(0);
      const T z1 = k + 1 < KBM1 ? h2 * (fc + fdn[t]) * wdn * art : T(0);
      const T hdiv = x1 - x0 + y1 - y0;
      const T vdiv = (z0 - z1) / dz;
      const T r = (tb * bc - dti2 * (hdiv + vdiv)) / (tf + eps);
      ref_at(a.out[t], q.o3, ix) = in ? r : T(0);
    }
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_advq_pair(Batch B, Vert<T> vg, Params prm, AdvPairArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T q4 = T(0.25), q8 = T(0.125), h2 = T(0.5), dti2 = T(prm.dti2), eps = T(1e-20);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q{};
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    const int im = q.im, jm = q.jm, i = q.i, j = q.j;
    const uint32_t ix = live ? lix3(q, i, j, k) : 0u;
    auto L3 = [&](const T* f, int di, int dj) {
      return live ? ld_at(f, q.o3, lix3(q, clampi(i + di, im), clampi(j + dj, jm), k)) : T(0);
    };
    auto L2 = [&](const T* f, int di, int dj) { return ld_at(f, q.o2, lix2(q, clampi(i + di, im), clampi(j + dj, jm))); };
    // values at level k, then those at k - 1 (and k + 1) from the adjacent lanes
    const T u0 = L3(a.u, 0, 0), u1 = L3(a.u, 1, 0), v0 = L3(a.v, 0, 0), v1 = L3(a.v, 0, 1), wk = L3(a.w, 0, 0);
    const T amc = L3(a.aam, 0, 0), amw = L3(a.aam, -1, 0), ame = L3(a.aam, 1, 0), ams = L3(a.aam, 0, -1),
            amn = L3(a.aam, 0, 1);
    T qc[2];
    qc[0] = L3(a.f[0], 0, 0);
    qc[1] = L3(a.f[1], 0, 0);
    const T u0m = __shfl_up(u0, 1, KB), u1m = __shfl_up(u1, 1, KB), v0m = __shfl_up(v0, 1, KB),
            v1m = __shfl_up(v1, 1, KB);
    const T amcm = __shfl_up(amc, 1, KB), amwm = __shfl_up(amw, 1, KB), amem = __shfl_up(ame, 1, KB),
            amsm = __shfl_up(ams, 1, KB), amnm = __shfl_up(amn, 1, KB);
    const T wup = __shfl_up(wk, 1, KB), wdn = __shfl_down(wk, 1, KB);
    const T qup[2] = {__shfl_up(qc[0], 1, KB), __shfl_up(qc[1], 1, KB)};
    const T qdn[2] = {__shfl_down(qc[0], 1, KB), __shfl_down(qc[1], 1, KB)};
    if (!live) continue;
    const bool in = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2 && k >= 1 && k < KBM1;
    const T dtc = L2(a.dt, 0, 0), dtw = L2(a.dt, -1, 0), dte = L2(a.dt, 1, 0), dts = L2(a.dt, 0, -1), dtn = L2(a.dt, 0, 1);
    const T hc = L2(a.h, 0, 0), hw = L2(a.h, -1, 0), he = L2(a.h, 1, 0), hs = L2(a.h, 0, -1), hn = L2(a.h, 0, 1);
    const T dxc = L2(a.dx, 0, 0), dxw = L2(a.dx, -1, 0), dxe = L2(a.dx, 1, 0), dxs = L2(a.dx, 0, -1), dxn = L2(a.dx, 0, 1);
    const T dyc = L2(a.dy, 0, 0), dyw = L2(a.dy, -1, 0), dye = L2(a.dy, 1, 0), dys = L2(a.dy, 0, -1), dyn = L2(a.dy, 0, 1);
    const T dum0 = L2(a.dum, 0, 0), dum1 = L2(a.dum, 1, 0), dvm0 = L2(a.dvm, 0, 0), dvm1 = L2(a.dvm, 0, 1);
    const T art = L2(a.art, 0, 0);
    const T tb = (hc + L2(a.etb, 0, 0)) * art;
    const T tf = (hc + L2(a.etf, 0, 0)) * art;
    const T dzs = vg.dz(q.m, k) + vg.dz(q.m, k - 1);
    // the four corner viscosities of each face, in the operator's order
    const T ax1 = ame + amc + amem + amcm, ax0 = amc + amw + amcm + amwm;
    const T ay1 = amn + amc + amnm + amcm, ay0 = amc + ams + amcm + amsm;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const T c = qc[t], w_ = L3(a.f[t], -1, 0), e = L3(a.f[t], 1, 0), s = L3(a.f[t], 0, -1), n = L3(a.f[t], 0, 1);
      const T bc = L3(a.fb[t], 0, 0), bw = L3(a.fb[t], -1, 0), be = L3(a.fb[t], 1, 0), bs = L3(a.fb[t], 0, -1),
              bn = L3(a.fb[t], 0, 1);
      // face between (hi) and (lo): OpAdvq::xf / yf
      auto fl = [&](T qhi, T qlo, T dthi, T dtlo, T vel, T velm, T asum, T hhi, T hlo, T bhi, T blo, T dm, T dshi,
                    T dslo, T dwhi, T dwlo) {
        const T adv = q8 * (qhi + qlo) * (dthi + dtlo) * (vel + velm);
        const T dif = q4 * asum * (hhi + hlo) * (bhi - blo) * dm / (dshi + dslo);
        return (adv - dif) * (h2 * (dwhi + dwlo));
      };
      const T x1 = fl(e, c, dte, dtc, u1, u1m, ax1, he, hc, be, bc, dum1, dxe, dxc, dye, dyc);
      const T x0 = fl(c, w_, dtc, dtw, u0, u0m, ax0, hc, hw, bc, bw, dum0, dxc, dxw, dyc, dyw);
      const T y1 = fl(n, c, dtn, dtc, v1, v1m, ay1, hn, hc, bn, bc, dvm1, dyn, dyc, dxn, dxc);
      const T y0 = fl(c, s, dtc, dts, v0, v0m, ay0, hc, hs, bc, bs, dvm0, dyc, dys, dxc, dxs);
      const T vert = (wup * qup[t] - wdn * qdn[t]) * art / dzs;
      const T hdiv = x1 - x0 + y1 - y0;
      const T r = (tb * bc - dti2 * (vert + hdiv)) / (tf + eps);
      ref_at(a.out[t], q.o3, ix) = in ? r : T(0);
    }
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_advuv(Batch B, Vert<T> vg, Params prm, AdvUVArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T q4 = T(0.25), eps = T(1e-20), g8 = T(prm.grav * 0.125), dti4 = T(2.0 * prm.dti2), two = T(2);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q{};
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    const int im = q.im, jm = q.jm, i = q.i, j = q.j;
    const uint32_t ix = live ? lix3(q, i, j, k) : 0u;
    auto L3 = [&](const T* f, int di, int dj) {
      return live ? ld_at(f, q.o3, lix3(q, clampi(i + di, im), clampi(j + dj, jm), k)) : T(0);
    };
    auto L2 = [&](const T* f, int di, int dj) { return ld_at(f, q.o2, lix2(q, clampi(i + di, im), clampi(j + dj, jm))); };
    const T uc = L3(a.u, 0, 0), vc = L3(a.v, 0, 0), wc = L3(a.w, 0, 0), ww = L3(a.w, -1, 0), ws = L3(a.w, 0, -1);
    const T uup = __shfl_up(uc, 1, KB), udn = __shfl_down(uc, 1, KB);   // u at k - 1, k + 1
    const T vup = __shfl_up(vc, 1, KB), vdn = __shfl_down(vc, 1, KB);
    const T wcd = __shfl_down(wc, 1, KB), wwd = __shfl_down(ww, 1, KB), wsd = __shfl_down(ws, 1, KB);   // w at k + 1
    if (!live) continue;
    const bool in = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2 && k < KBM1;
    const T dz = vg.dz(q.m, k);
    const T dtc = L2(a.dt, 0, 0), dtw = L2(a.dt, -1, 0), dts = L2(a.dt, 0, -1);
    const T hc = L2(a.h, 0, 0), hw = L2(a.h, -1, 0), hs_ = L2(a.h, 0, -1);
    const T ec = L2(a.eatm, 0, 0), ew = L2(a.eatm, -1, 0), es = L2(a.eatm, 0, -1);
    const T efc = L2(a.egf, 0, 0), efw = L2(a.egf, -1, 0), efs = L2(a.egf, 0, -1);
    const T ebc = L2(a.egb, 0, 0), ebw = L2(a.egb, -1, 0), ebs = L2(a.egb, 0, -1);
    const T tfc = L2(a.etf, 0, 0), tfw = L2(a.etf, -1, 0), tfs = L2(a.etf, 0, -1);
    const T tbc = L2(a.etb, 0, 0), tbw = L2(a.etb, -1, 0), tbs = L2(a.etb, 0, -1);
    const T corc = L2(a.cor, 0, 0), corw = L2(a.cor, -1, 0), cors = L2(a.cor, 0, -1);
    // ---- u: OpAdvu
    {
      const T aru = L2(a.aru, 0, 0);
      const T f0 = (k >= 1 && k < KBM1) ? q4 * (wc + ww) * (uc + uup) : T(0);       // vfl<0>
      const T f1 = (k + 1 < KBM1) ? q4 * (wcd + wwd) * (udn + uc) : T(0);          // vfl<1> (k + 1 >= 1)
      const T vdiv = (f0 - f1) * aru / dz;
      const T c1 = (corc * dtc) * (L3(a.v, 0, 1) + vc);
      const T c2 = (corw * dtw) * (L3(a.v, -1, 1) + L3(a.v, -1, 0));
      const T corio = -aru * q4 * (c1 + c2);
      const T geg = (efc - efw) + (ebc - ebw);
      const T gpa = two * (ec - ew);
      const T pg = g8 * ((dtc + dtw) * (geg + gpa) * (L2(a.dy, 0, 0) + L2(a.dy, -1, 0)));
      const T rhs = L3(a.advx, 0, 0) + vdiv + corio + pg + L3(a.drhox, 0, 0);
      const T hsum = hc + hw;
      const T dnew = (hsum + tfc + tfw) * aru;
      const T dold = (hsum + tbc + tbw) * aru;
      const T r = (dold * L3(a.ub, 0, 0) - dti4 * rhs) / (dnew + eps);
      ref_at(a.uf, q.o3, ix) = in ? r : T(0);
    }
    // ---- v: OpAdvv
    {
      const T arv = L2(a.arv, 0, 0);
      const T f0 = (k >= 1 && k < KBM1) ? q4 * (wc + ws) * (vc + vup) : T(0);
      const T f1 = (k + 1 < KBM1) ? q4 * (wcd + wsd) * (vdn + vc) : T(0);
      const T vdiv = (f0 - f1) * arv / dz;
      const T c1 = (corc * dtc) * (L3(a.u, 1, 0) + uc);
      const T c2 = (cors * dts) * (L3(a.u, 1, -1) + L3(a.u, 0, -1));
      const T corio = arv * q4 * (c1 + c2);
      const T geg = (efc - efs) + (ebc - ebs);
      const T gpa = two * (ec - es);
      const T pg = g8 * ((dtc + dts) * (geg + gpa) * (L2(a.dx, 0, 0) + L2(a.dx, 0, -1)));
      const T rhs = L3(a.advy, 0, 0) + vdiv + corio + pg + L3(a.drhoy, 0, 0);
      const T hsum = hc + hs_;
      const T dnew = (hsum + tfc + tfs) * arv;
      const T dold = (hsum + tbc + tbs) * arv;
      const T r = (dold * L3(a.vb, 0, 0) - dti4 * rhs) / (dnew + eps);
      ref_at(a.vf, q.o3, ix) = in ? r : T(0);
    }
  }
}

template <class T> void launch_advuv(const Ctx<T>& c, const AdvUVArgs<T>& a) {
  if (c.nblk3 == 0) return;
  hipLaunchKernelGGL((k_advuv<T>), dim3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)), dim3(TPB), 0,
                     c.stream, c.B, c.vert, c.prm, a);
}
template void launch_advuv<float>(const Ctx<float>&, const AdvUVArgs<float>&);
template void launch_advuv<double>(const Ctx<double>&, const AdvUVArgs<double>&);

template <class T> void launch_advt1_pair(const Ctx<T>& c, const AdvPairArgs<T>& a) {
  if (c.nblk3 == 0) return;
  hipLaunchKernelGGL((k_advt1_pair<T>), dim3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)), dim3(TPB), 0,
                     c.stream, c.B, c.vert, c.prm, a);
}
template <class T> void launch_advq_pair(const Ctx<T>& c, const AdvPairArgs<T>& a) {
  if (c.nblk3 == 0) return;
  hipLaunchKernelGGL((k_advq_pair<T>), dim3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)), dim3(TPB), 0,
                     c.stream, c.B, c.vert, c.prm, a);
}

template void launch_advt1_pair<float>(const Ctx<float>&, const AdvPairArgs<float>&);
template void launch_advt1_pair<double>(const Ctx<double>&, const AdvPairArgs<double>&);
template void launch_advq_pair<float>(const Ctx<float>&, const AdvPairArgs<float>&);
template void launch_advq_pair<double>(const Ctx<double>&, const AdvPairArgs<double>&);

}  // namespace pom
