// Kernel and launcher of the joint advu / advv adjoint (pom_momadj.h).
#include "pom_momadj.h"

namespace pom {
namespace {

template <class T> struct Eq {
  T s;      // output cotangent (zero where the output is not an interior update)
  T g;      // of the numerator: s / (dnew + 1e-20)
  T crhs;   // of rhs: -2 dti2 g
};

}  // namespace

template <class T>
__global__ void __launch_bounds__(TPB) k_adj_advuv(Batch B, Vert<T> vert, Params prm, MomAdjArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T dti4 = T(2.0 * prm.dti2), g8 = T(prm.grav * 0.125);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q;
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    auto sel = [](bool c, T x) { return c ? x : T(0); };
    // phase 1: the equations' cotangent factors at this level
    Eq<T> U0{}, U1{}, U2{}, U3{}, V0{}, V1{}, V2{}, V3{};
    T Au0 = T(0), Au1 = T(0), Av0 = T(0), Av1 = T(0);
    if (live) {
      const int im = q.im, jm = q.jm, qi = q.i, qj = q.j;
      auto L2 = [&](const T* f, int i, int j) { return ld_at(f, q.o2, lix2(q, clampi(i, im), clampi(j, jm))); };
      auto L3 = [&](const T* f, int i, int j, int kk) {
        return ld_at(f, q.o3, lix3(q, clampi(i, im), clampi(j, jm), clampi(kk, KB)));
      };
      auto eq = [&](int pi, int pj, int di, int dj, const T* ob, const T* ar_f) {
        Eq<T> e;
        const bool valid = pi >= 1 && pi <= im - 2 && pj >= 1 && pj <= jm - 2 && k < KBM1;
        e.s = sel(valid, L3(ob, pi, pj, k));
        const T dnew = ((L2(a.h, pi, pj) + L2(a.h, pi - di, pj - dj)) + L2(a.etf, pi, pj) + L2(a.etf, pi - di, pj - dj)) *
                       L2(ar_f, pi, pj);
        e.g = e.s * rcp_lane(dnew + T(1e-20));
        e.crhs = -dti4 * e.g;
        return e;
      };
      U0 = eq(qi, qj, 1, 0, a.ob_u, a.aru);
      U1 = eq(qi + 1, qj, 1, 0, a.ob_u, a.aru);
      U2 = eq(qi, qj - 1, 1, 0, a.ob_u, a.aru);
      U3 = eq(qi + 1, qj - 1, 1, 0, a.ob_u, a.aru);
      V0 = eq(qi, qj, 0, 1, a.ob_v, a.arv);
      V1 = eq(qi, qj + 1, 0, 1, a.ob_v, a.arv);
      V2 = eq(qi - 1, qj, 0, 1, a.ob_v, a.arv);
      V3 = eq(qi - 1, qj + 1, 0, 1, a.ob_v, a.arv);
      const T rdz = rcp_lane(vert.dz(q.m, k));
      // vdiv = (vfl(k) - vfl(k+1)) ar / dz(k): a source of vfl's cotangent at k (+) and k+1 (-)
      Au0 = U0.crhs * L2(a.aru, qi, qj) * rdz;
      Au1 = U1.crhs * L2(a.aru, qi + 1, qj) * rdz;
      Av0 = V0.crhs * L2(a.arv, qi, qj) * rdz;
      Av1 = V1.crhs * L2(a.arv, qi, qj + 1) * rdz;
    }
    // phase 2 (every lane): cotangent of vfl at (column, level k) and of the
    // own column's vfl at k + 1
    const bool vk = k >= 1 && k < KBM1;
    const T cvu0 = sel(vk, Au0 - __shfl_up(Au0, 1, KB)), cvu1 = sel(vk, Au1 - __shfl_up(Au1, 1, KB));
    const T cvv0 = sel(vk, Av0 - __shfl_up(Av0, 1, KB)), cvv1 = sel(vk, Av1 - __shfl_up(Av1, 1, KB));
    const T cvu0p = sel(k < KB - 1, __shfl_down(cvu0, 1, KB)), cvv0p = sel(k < KB - 1, __shfl_down(cvv0, 1, KB));
    // phase 3: 3-D cotangents and this level's share of the 2-D ones
    T p_dt = T(0), p_eg = T(0), p_etf = T(0), p_etb = T(0);
    if (live) {
      const int im = q.im, jm = q.jm, qi = q.i, qj = q.j;
      auto L2 = [&](const T* f, int i, int j) { return ld_at(f, q.o2, lix2(q, clampi(i, im), clampi(j, jm))); };
      auto L3 = [&](const T* f, int i, int j, int kk) {
        return ld_at(f, q.o3, lix3(q, clampi(i, im), clampi(j, jm), clampi(kk, KB)));
      };
      const T aru_q = L2(a.aru, qi, qj), arv_q = L2(a.arv, qi, qj), aru_e = L2(a.aru, qi + 1, qj),
              arv_n = L2(a.arv, qi, qj + 1);
      const T cor_q = L2(a.cor, qi, qj), dt_q = L2(a.dt, qi, qj);
      const T cdt_q = cor_q * dt_q, cdt_w = L2(a.cor, qi - 1, qj) * L2(a.dt, qi - 1, qj),
              cdt_s = L2(a.cor, qi, qj - 1) * L2(a.dt, qi, qj - 1);
      const T qa = T(0.25);
      // Coriolis partials: ccor = crhs (-aru / 4) for u, crhs (arv / 4) for v
      const T cu0 = U0.crhs * (-aru_q * qa), cu1 = U1.crhs * (-aru_e * qa);
      const T cu2 = U2.crhs * (-L2(a.aru, qi, qj - 1) * qa), cu3 = U3.crhs * (-L2(a.aru, qi + 1, qj - 1) * qa);
      const T cv0 = V0.crhs * (arv_q * qa), cv1 = V1.crhs * (arv_n * qa);
      const T cv2 = V2.crhs * (L2(a.arv, qi - 1, qj) * qa), cv3 = V3.crhs * (L2(a.arv, qi - 1, qj + 1) * qa);
      // v(q) in the u outputs at q (v(p)), q-j (v(p+j)), q+i (v(p-i)), q+i-j (v(p-i+j));
      // u(q) in the v outputs at q (u(p)), q-i (u(p+i)), q+j (u(p-j)), q-i+j (u(p+i-j))
      const T corV = (cu0 + cu1) * cdt_q + (cu2 + cu3) * cdt_s;
      const T corU = (cv0 + cv1) * cdt_q + (cv2 + cv3) * cdt_w;
      const T w_k = L3(a.w, qi, qj, k), w_k1 = L3(a.w, qi, qj, k + 1);
      const T wsu = w_k + L3(a.w, qi - 1, qj, k), wsu1 = w_k1 + L3(a.w, qi - 1, qj, k + 1);
      const T wsv = w_k + L3(a.w, qi, qj - 1, k), wsv1 = w_k1 + L3(a.w, qi, qj - 1, k + 1);
      const T u_q = L3(a.u, qi, qj, k), u_e = L3(a.u, qi + 1, qj, k);
      const T v_q = L3(a.v, qi, qj, k), v_n = L3(a.v, qi, qj + 1, k);
      const T su = u_q + L3(a.u, qi, qj, k - 1), su1 = u_e + L3(a.u, qi + 1, qj, k - 1);
      const T sv = v_q + L3(a.v, qi, qj, k - 1), sv1 = v_n + L3(a.v, qi, qj + 1, k - 1);
      const T cu_q = cvu0 * qa * wsu + cvu0p * qa * wsu1 + corU;
      const T cv_q = cvv0 * qa * wsv + cvv0p * qa * wsv1 + corV;
      const T cw_q = cvu0 * qa * su + cvu1 * qa * su1 + cvv0 * qa * sv + cvv1 * qa * sv1;
      const T h_q = L2(a.h, qi, qj), etb_q = L2(a.etb, qi, qj);
      const T dold_u = ((h_q + L2(a.h, qi - 1, qj)) + etb_q + L2(a.etb, qi - 1, qj)) * aru_q;
      const T dold_v = ((h_q + L2(a.h, qi, qj - 1)) + etb_q + L2(a.etb, qi, qj - 1)) * arv_q;
      const T ub_q = L3(a.ub, qi, qj, k), vb_q = L3(a.vb, qi, qj, k);
      const uint32_t ix = lix3(q, qi, qj, k);
      // accumulate: every old value read before the first store (the arrays may
      // not be assumed distinct by the compiler, so it would not move a read
      // above an earlier store)
      T* const dst[9] = {a.c_advx, a.c_drhox, a.c_advy, a.c_drhoy, a.c_ub, a.c_vb, a.c_u, a.c_v, a.c_w};
      const T add[9] = {U0.crhs, U0.crhs, V0.crhs, V0.crhs, U0.g * dold_u, V0.g * dold_v, cu_q, cv_q, cw_q};
      T old[9];
#pragma unroll
      for (int n = 0; n < 9; ++n) old[n] = ld_at(dst[n], q.o3, ix);
#pragma unroll
      for (int n = 0; n < 9; ++n) ref_at(dst[n], q.o3, ix) = old[n] + add[n];
      // 2-D inputs: dt (Coriolis and pressure gradient), egf = egb, etf (dnew), etb (dold)
      const T egf_q = L2(a.egf, qi, qj), egb_q = L2(a.egb, qi, qj), ea_q = L2(a.eatm, qi, qj);
      const T dy_q = L2(a.dy, qi, qj), dx_q = L2(a.dx, qi, qj);
      auto gg = [&](int pi, int pj, int di, int dj) {   // del egf + del egb + 2 del e_atmos across the face p-e | p
        return (L2(a.egf, pi, pj) - L2(a.egf, pi - di, pj - dj)) + (L2(a.egb, pi, pj) - L2(a.egb, pi - di, pj - dj)) +
               T(2) * (L2(a.eatm, pi, pj) - L2(a.eatm, pi - di, pj - dj));
      };
      const T gg_u0 = (egf_q - L2(a.egf, qi - 1, qj)) + (egb_q - L2(a.egb, qi - 1, qj)) +
                      T(2) * (ea_q - L2(a.eatm, qi - 1, qj));
      const T gg_v0 = (egf_q - L2(a.egf, qi, qj - 1)) + (egb_q - L2(a.egb, qi, qj - 1)) +
                      T(2) * (ea_q - L2(a.eatm, qi, qj - 1));
      const T gg_u1 = gg(qi + 1, qj, 1, 0), gg_v1 = gg(qi, qj + 1, 0, 1);
      const T dss_u0 = dy_q + L2(a.dy, qi - 1, qj), dss_u1 = L2(a.dy, qi + 1, qj) + dy_q;
      const T dss_v0 = dx_q + L2(a.dx, qi, qj - 1), dss_v1 = L2(a.dx, qi, qj + 1) + dx_q;
      const T dts_u0 = dt_q + L2(a.dt, qi - 1, qj), dts_u1 = L2(a.dt, qi + 1, qj) + dt_q;
      const T dts_v0 = dt_q + L2(a.dt, qi, qj - 1), dts_v1 = L2(a.dt, qi, qj + 1) + dt_q;
      const T pu0 = U0.crhs * g8, pu1 = U1.crhs * g8, pv0 = V0.crhs * g8, pv1 = V1.crhs * g8;
      p_dt = (cu0 + cu1) * cor_q * (v_n + v_q) + pu0 * gg_u0 * dss_u0 + pu1 * gg_u1 * dss_u1 +
             (cv0 + cv1) * cor_q * (u_e + u_q) + pv0 * gg_v0 * dss_v0 + pv1 * gg_v1 * dss_v1;
      p_eg = pu0 * dts_u0 * dss_u0 - pu1 * dts_u1 * dss_u1 + pv0 * dts_v0 * dss_v0 - pv1 * dts_v1 * dss_v1;
      p_etf = -(U0.g * L3(a.uf, qi, qj, k) * aru_q + U1.g * L3(a.uf, qi + 1, qj, k) * aru_e +
                V0.g * L3(a.vf, qi, qj, k) * arv_q + V1.g * L3(a.vf, qi, qj + 1, k) * arv_n);
      p_etb = U0.g * ub_q * aru_q + U1.g * L3(a.ub, qi + 1, qj, k) * aru_e + V0.g * vb_q * arv_q +
              V1.g * L3(a.vb, qi, qj + 1, k) * arv_n;
    }
    // phase 4 (every lane): sums over the water column
#pragma unroll
    for (int o = KB / 2; o > 0; o >>= 1) {
      p_dt += __shfl_xor(p_dt, o, KB);
      p_eg += __shfl_xor(p_eg, o, KB);
      p_etf += __shfl_xor(p_etf, o, KB);
      p_etb += __shfl_xor(p_etb, o, KB);
    }
    if (live && k == 0) {
      const uint32_t ix = lix2(q, q.i, q.j);
      T* const dst[5] = {a.c_dt, a.c_egf, a.c_egb, a.c_etf, a.c_etb};
      const T add[5] = {p_dt, p_eg, p_eg, p_etf, p_etb};
      T old[5];
#pragma unroll
      for (int n = 0; n < 5; ++n) old[n] = ld_at(dst[n], q.o2, ix);
#pragma unroll
      for (int n = 0; n < 5; ++n) ref_at(dst[n], q.o2, ix) = old[n] + add[n];
    }
  }
}

template <class T> void launch_adj_advuv(const Ctx<T>& c, const MomAdjArgs<T>& a) {
  if (c.nblk3 == 0) return;
  hipLaunchKernelGGL((k_adj_advuv<T>), dim3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)), dim3(TPB), 0,
                     c.stream, c.B, c.vert, c.prm, a);
}

template void launch_adj_advuv<float>(const Ctx<float>&, const MomAdjArgs<float>&);
template void launch_adj_advuv<double>(const Ctx<double>&, const MomAdjArgs<double>&);

}  // namespace pom
