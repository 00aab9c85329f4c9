// Kernel and launcher of the advct adjoint (pom_advctadj.h).
#include "pom_advctadj.h"

#include <type_traits>

namespace pom {

// PART 0: the cotangents of u and v and the advective and curvature share of
// dt's; PART 1: those of ub, vb, aam and the viscous share of dt's.  The two
// parts read the same output cotangents and hold half the live values each
// (one kernel with all five outputs held 175 VGPRs, two waves per SIMD).
template <class T, int PART>
__global__ void __launch_bounds__(TPB) k_adj_advct(Batch B, AdvctAdjArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T c8 = T(0.125), c4 = T(0.25), two = T(2), eps = T(1e-20);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q;
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    auto sel = [](bool c, T x) { return c ? x : T(0); };
    T pdt = T(0);   // this level's share of dt's cotangent
    // the body, once for columns at least 3 from every edge (FAST: every
    // neighbour read is in the member and every edge condition holds, so the
    // reads are a byte offset from the column's own element) and once for the
    // columns nearer the edge (clamped indices, edge conditions evaluated)
    auto body = [&](auto fast) {
      constexpr bool FAST = decltype(fast)::value;
      const int im = q.im, jm = q.jm, qi = q.i, qj = q.j;
      const uint32_t b3 = boff<T>(q.o3, lix3(q, qi, qj, k)), b2 = boff<T>(q.o2, lix2(q, qi, qj));
      const int rs3 = jm * KB * (int)sizeof(T), rs2 = jm * (int)sizeof(T);
      // values at q + (di, dj), level k; indices clamped into the member
      auto L3 = [&](const T* f, int di, int dj) {
        if constexpr (FAST) {
          return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(f) +
                                             (b3 + (uint32_t)(di * rs3 + dj * KB * (int)sizeof(T))));
        } else {
          return ld_at(f, q.o3, lix3(q, clampi(qi + di, im), clampi(qj + dj, jm), k));
        }
      };
      auto L2 = [&](const T* f, int di, int dj) {
        if constexpr (FAST) {
          return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(f) +
                                             (b2 + (uint32_t)(di * rs2 + dj * (int)sizeof(T))));
        } else {
          return ld_at(f, q.o2, lix2(q, clampi(qi + di, im), clampi(qj + dj, jm)));
        }
      };
      auto U = [&](int di, int dj) { return L3(a.u, di, dj); };
      auto V = [&](int di, int dj) { return L3(a.v, di, dj); };
      auto AM = [&](int di, int dj) { return L3(a.aam, di, dj); };
      auto DT = [&](int di, int dj) { return L2(a.dt, di, dj); };
      auto DX = [&](int di, int dj) { return L2(a.dx, di, dj); };
      auto DY = [&](int di, int dj) { return L2(a.dy, di, dj); };
      auto inner = [&](int di, int dj) {
        const int i = qi + di, j = qj + dj;
        return (FAST || (i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2)) && k < KBM1;
      };
      // output cotangents where the output is an interior update, else zero
      auto sx = [&](int di, int dj) { return sel(inner(di, dj), L3(a.ob_x, di, dj)); };
      auto sy = [&](int di, int dj) { return sel(inner(di, dj), L3(a.ob_y, di, dj)); };
      // face cotangents at a (or b) = q + (di, dj)
      auto cXu = [&](int di) { return sel(FAST || qi + di >= 1, sx(di, 0) - sx(di + 1, 0)); };
      auto cYu = [&](int di, int dj) { return sx(di, dj - 1) - sx(di, dj); };
      auto cXv = [&](int di, int dj) { return sy(di - 1, dj) - sy(di, dj); };
      auto cYv = [&](int dj) { return sel(FAST || qj + dj >= 1, sy(0, dj) - sy(0, dj + 1)); };
      // corner sums over b, b-i, b-j, b-i-j
      auto s4x = [&](int di, int dj) { return DX(di, dj) + DX(di - 1, dj) + DX(di, dj - 1) + DX(di - 1, dj - 1); };
      auto s4y = [&](int di, int dj) { return DY(di, dj) + DY(di - 1, dj) + DY(di, dj - 1) + DY(di - 1, dj - 1); };
      auto dt4 = [&](int di, int dj) { return c4 * (DT(di, dj) + DT(di - 1, dj) + DT(di, dj - 1) + DT(di - 1, dj - 1)); };
      auto am4 = [&](int di, int dj) { return AM(di, dj) + AM(di - 1, dj) + AM(di, dj - 1) + AM(di - 1, dj - 1); };
      auto rdy4 = [&](int di, int dj) { return rcp_lane(s4y(di, dj) + eps); };
      auto rdx4 = [&](int di, int dj) { return rcp_lane(s4x(di, dj) + eps); };
      auto duy = [&](int di, int dj) { return (L3(a.ub, di, dj) - L3(a.ub, di, dj - 1)) * rdy4(di, dj); };
      auto dvx = [&](int di, int dj) { return (L3(a.vb, di, dj) - L3(a.vb, di - 1, dj)) * rdx4(di, dj); };
      // corner viscosity CV(b): cotangent (it enters Yu and Xv) and its own partials
      auto cCV = [&](int di, int dj) { return -c4 * (cYu(di, dj) * s4x(di, dj) + cXv(di, dj) * s4y(di, dj)); };
      // curvature products P1 = curv dt (v(a+j) + v(a)) in advx, P2 = curv dt (u(a+i) + u(a)) in advy
      auto rdxy = [&](int di, int dj) { return rcp_lane(DX(di, dj) * DY(di, dj) + eps); };
      auto curv = [&](int di, int dj) {
        return c4 * ((V(di, dj + 1) + V(di, dj)) * (DY(di + 1, dj) - DY(di - 1, dj)) -
                     (U(di + 1, dj) + U(di, dj)) * (DX(di, dj + 1) - DX(di, dj - 1))) * rdxy(di, dj);
      };
      auto Ku = [&](int di, int dj) {   // cotangent of P1(a): advx(a) (i >= 2) and advx(a+i) read it
        return -c4 * (sel(FAST || qi + di >= 2, sx(di, dj) * L2(a.aru, di, dj)) +
                      sel(FAST || qi + di >= 1, sx(di + 1, dj) * L2(a.aru, di + 1, dj)));
      };
      auto Kv = [&](int di, int dj) {   // cotangent of P2(a): advy(a) (j >= 2) and advy(a+j) read it
        return c4 * (sel(FAST || qj + dj >= 2, sy(di, dj) * L2(a.arv, di, dj)) +
                     sel(FAST || qj + dj >= 1, sy(di, dj + 1) * L2(a.arv, di, dj + 1)));
      };
      auto ccurv = [&](int di, int dj) {   // cotangent of curv(a)
        const T d = DT(di, dj);
        return Ku(di, dj) * d * (V(di, dj + 1) + V(di, dj)) + Kv(di, dj) * d * (U(di + 1, dj) + U(di, dj));
      };
      auto dcu = [&](int di, int dj) { return -c4 * (DX(di, dj + 1) - DX(di, dj - 1)) * rdxy(di, dj); };   // d curv / d u
      auto dcv = [&](int di, int dj) { return c4 * (DY(di + 1, dj) - DY(di - 1, dj)) * rdxy(di, dj); };    // d curv / d v
      // face terms
      auto xuS = [&](int di) { return U(di + 1, 0) + U(di, 0); };
      auto xuT = [&](int di) {
        return (DT(di + 1, 0) + DT(di, 0)) * U(di + 1, 0) + (DT(di, 0) + DT(di - 1, 0)) * U(di, 0);
      };
      auto yvS = [&](int dj) { return V(0, dj + 1) + V(0, dj); };
      auto yvW = [&](int dj) {
        return (DT(0, dj + 1) + DT(0, dj)) * V(0, dj + 1) + (DT(0, dj) + DT(0, dj - 1)) * V(0, dj);
      };
      auto yuV = [&](int di, int dj) {   // the v flux of Yu(b)
        return (DT(di, dj) + DT(di, dj - 1)) * V(di, dj) + (DT(di - 1, dj) + DT(di - 1, dj - 1)) * V(di - 1, dj);
      };
      auto xvU = [&](int di, int dj) {   // the u flux of Xv(a)
        return (DT(di, dj) + DT(di - 1, dj)) * U(di, dj) + (DT(di, dj - 1) + DT(di - 1, dj - 1)) * U(di, dj - 1);
      };
      const T dtq = DT(0, 0), dtw = DT(-1, 0), dts = DT(0, -1);

      const uint32_t ix = lix3(q, qi, qj, k);
      if constexpr (PART == 0) {
        const T cXu0 = cXu(0), cXuw = cXu(-1), cYv0 = cYv(0), cYvs = cYv(-1);
        const T curv0 = curv(0, 0);
        // ---- u(q): Xu(q-i) (as u(a+i)), Xu(q) (as u(a)), Yu(q), Yu(q+j), Xv(q), Xv(q+j), curv and P2 at q-i, q
        const T cu = (cXuw * DY(-1, 0) * c8 * ((dtq + dtw) * xuS(-1) + xuT(-1)) +
                      cXu0 * DY(0, 0) * c8 * ((dtq + dtw) * xuS(0) + xuT(0))) +
                     c4 * c8 * (cYu(0, 0) * s4x(0, 0) * yuV(0, 0) + cYu(0, 1) * s4x(0, 1) * yuV(0, 1)) +
                     c4 * c8 * (dtq + dtw) *
                         (cXv(0, 0) * s4y(0, 0) * (V(0, 0) + V(-1, 0)) + cXv(0, 1) * s4y(0, 1) * (V(0, 1) + V(-1, 1))) +
                     ccurv(-1, 0) * dcu(-1, 0) + ccurv(0, 0) * dcu(0, 0) + Kv(-1, 0) * curv(-1, 0) * dtw +
                     Kv(0, 0) * curv0 * dtq;
        // ---- v(q): Yv(q-j) (as v(b+j)), Yv(q) (as v(b)), Xv(q), Xv(q+i), Yu(q), Yu(q+i), curv and P1 at q-j, q
        const T cv = (cYvs * DX(0, -1) * c8 * ((dtq + dts) * yvS(-1) + yvW(-1)) +
                      cYv0 * DX(0, 0) * c8 * ((dtq + dts) * yvS(0) + yvW(0))) +
                     c4 * c8 * (cXv(0, 0) * s4y(0, 0) * xvU(0, 0) + cXv(1, 0) * s4y(1, 0) * xvU(1, 0)) +
                     c4 * c8 * (dtq + dts) *
                         (cYu(0, 0) * s4x(0, 0) * (U(0, 0) + U(0, -1)) + cYu(1, 0) * s4x(1, 0) * (U(1, 0) + U(1, -1))) +
                     ccurv(0, -1) * dcv(0, -1) + ccurv(0, 0) * dcv(0, 0) + Ku(0, -1) * curv(0, -1) * dts +
                     Ku(0, 0) * curv0 * dtq;
        // ---- dt(q), this level's advective and curvature partials
        pdt = cXu0 * DY(0, 0) * c8 * xuS(0) * xuS(0) +
              cXuw * DY(-1, 0) * c8 * xuS(-1) * U(0, 0) + cXu(1) * DY(1, 0) * c8 * xuS(1) * U(1, 0) +
              c4 * c8 *
                  (cYu(0, 0) * s4x(0, 0) * (U(0, 0) + U(0, -1)) * V(0, 0) +
                   cYu(0, 1) * s4x(0, 1) * (U(0, 1) + U(0, 0)) * V(0, 1) +
                   cYu(1, 0) * s4x(1, 0) * (U(1, 0) + U(1, -1)) * V(0, 0) +
                   cYu(1, 1) * s4x(1, 1) * (U(1, 1) + U(1, 0)) * V(0, 1)) +
              c4 * c8 *
                  (cXv(0, 0) * s4y(0, 0) * (V(0, 0) + V(-1, 0)) * U(0, 0) +
                   cXv(1, 0) * s4y(1, 0) * (V(1, 0) + V(0, 0)) * U(1, 0) +
                   cXv(0, 1) * s4y(0, 1) * (V(0, 1) + V(-1, 1)) * U(0, 0) +
                   cXv(1, 1) * s4y(1, 1) * (V(1, 1) + V(0, 1)) * U(1, 0)) +
              cYv0 * DX(0, 0) * c8 * yvS(0) * yvS(0) +
              cYvs * DX(0, -1) * c8 * yvS(-1) * V(0, 0) + cYv(1) * DX(0, 1) * c8 * yvS(1) * V(0, 1) +
              Ku(0, 0) * curv0 * (V(0, 1) + V(0, 0)) + Kv(0, 0) * curv0 * (U(1, 0) + U(0, 0));
        // accumulate: every old value read before the first store
        const T ou = ld_at(a.c_u, q.o3, ix), ov = ld_at(a.c_v, q.o3, ix);
        ref_at(a.c_u, q.o3, ix) = ou + cu;
        ref_at(a.c_v, q.o3, ix) = ov + cv;
      } else {
        const T gu0 = (L3(a.ub, 1, 0) - L3(a.ub, 0, 0)) * rcp_lane(DX(0, 0) + eps);
        const T gv0 = (L3(a.vb, 0, 1) - L3(a.vb, 0, 0)) * rcp_lane(DY(0, 0) + eps);
        const T cXu0 = cXu(0), cXuw = cXu(-1), cYv0 = cYv(0), cYvs = cYv(-1);
        const T cCV00 = cCV(0, 0), cCV10 = cCV(1, 0), cCV01 = cCV(0, 1), cCV11 = cCV(1, 1);
        // ---- ub(q): the viscous part of Xu(q-i), Xu(q); CV(q), CV(q+j)
        const T cub = cXuw * (-DY(-1, 0) * dtw * AM(-1, 0) * two * rcp_lane(DX(-1, 0) + eps)) +
                      cXu0 * (DY(0, 0) * dtq * AM(0, 0) * two * rcp_lane(DX(0, 0) + eps)) +
                      cCV00 * dt4(0, 0) * am4(0, 0) * rdy4(0, 0) - cCV01 * dt4(0, 1) * am4(0, 1) * rdy4(0, 1);
        // ---- vb(q): the viscous part of Yv(q-j), Yv(q); CV(q), CV(q+i)
        const T cvb = cYvs * (-DX(0, -1) * dts * AM(0, -1) * two * rcp_lane(DY(0, -1) + eps)) +
                      cYv0 * (DX(0, 0) * dtq * AM(0, 0) * two * rcp_lane(DY(0, 0) + eps)) +
                      cCV00 * dt4(0, 0) * am4(0, 0) * rdx4(0, 0) - cCV10 * dt4(1, 0) * am4(1, 0) * rdx4(1, 0);
        // ---- aam(q): Xu(q), Yv(q), the four corners CV(q), CV(q+i), CV(q+j), CV(q+i+j)
        const T e00 = duy(0, 0) + dvx(0, 0), e10 = duy(1, 0) + dvx(1, 0), e01 = duy(0, 1) + dvx(0, 1),
                e11 = duy(1, 1) + dvx(1, 1);
        const T caam = cXu0 * (-DY(0, 0) * dtq * two * gu0) + cYv0 * (-DX(0, 0) * dtq * two * gv0) +
                       cCV00 * dt4(0, 0) * e00 + cCV10 * dt4(1, 0) * e10 + cCV01 * dt4(0, 1) * e01 +
                       cCV11 * dt4(1, 1) * e11;
        // ---- dt(q), this level's viscous partials
        pdt = -cXu0 * DY(0, 0) * two * AM(0, 0) * gu0 - cYv0 * DX(0, 0) * two * AM(0, 0) * gv0 +
              c4 * (cCV00 * am4(0, 0) * e00 + cCV10 * am4(1, 0) * e10 + cCV01 * am4(0, 1) * e01 +
                    cCV11 * am4(1, 1) * e11);
        // accumulate: every old value read before the first store
        const T oub = ld_at(a.c_ub, q.o3, ix), ovb = ld_at(a.c_vb, q.o3, ix), oam = ld_at(a.c_aam, q.o3, ix);
        ref_at(a.c_ub, q.o3, ix) = oub + cub;
        ref_at(a.c_vb, q.o3, ix) = ovb + cvb;
        ref_at(a.c_aam, q.o3, ix) = oam + caam;
      }
    };
    if (live) {
      if (q.i >= 3 && q.i <= q.im - 4 && q.j >= 3 && q.j <= q.jm - 4) body(std::true_type{});
      else body(std::false_type{});
    }
#pragma unroll
    for (int o = KB / 2; o > 0; o >>= 1) pdt += __shfl_xor(pdt, o, KB);
    if (live && k == 0) {
      T& r = ref_at(a.c_dt, q.o2, lix2(q, q.i, q.j));
      r = r + pdt;
    }
  }
}

template <class T> void launch_adj_advct(const Ctx<T>& c, const AdvctAdjArgs<T>& a) {
  if (c.nblk3 == 0) return;
  const dim3 grid(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min));
  hipLaunchKernelGGL((k_adj_advct<T, 0>), grid, dim3(TPB), 0, c.stream, c.B, a);
  hipLaunchKernelGGL((k_adj_advct<T, 1>), grid, dim3(TPB), 0, c.stream, c.B, a);
}

template void launch_adj_advct<float>(const Ctx<float>&, const AdvctAdjArgs<float>&);
template void launch_adj_advct<double>(const Ctx<double>&, const AdvctAdjArgs<double>&);

}  // namespace pom
