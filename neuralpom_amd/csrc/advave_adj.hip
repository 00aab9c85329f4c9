// Kernel and launcher of the advave adjoint (pom_advaveadj.h).
#include "pom_advaveadj.h"

namespace pom {

template <class T>
__global__ void __launch_bounds__(TPB) k_adj_advave(Batch B, AdvaveAdjArgs<T> a) {
  const int nblk = B.blk2[B.M];
  const T c8 = T(0.125), c4 = T(0.25), two = T(2);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos2_at(B, blk, q)) continue;
    const int im = q.im, jm = q.jm, qi = q.i, qj = q.j;
    auto sel = [](bool c, T x) { return c ? x : T(0); };
    // values at q + (di, dj); indices clamped into the member (a clamped value
    // only feeds a term whose weight is zero).  Quotients by the grid spacings
    // are products with their lane reciprocals (rcp_lane).
    auto L = [&](const T* f, int di, int dj) {
      return ld_at(f, q.o2, lix2(q, clampi(qi + di, im), clampi(qj + dj, jm)));
    };
    auto U = [&](int di, int dj) { return L(a.ua, di, dj); };
    auto V = [&](int di, int dj) { return L(a.va, di, dj); };
    auto UB = [&](int di, int dj) { return L(a.uab, di, dj); };
    auto VB = [&](int di, int dj) { return L(a.vab, di, dj); };
    auto D = [&](int di, int dj) { return L(a.d, di, dj); };
    auto AM = [&](int di, int dj) { return L(a.aam2d, di, dj); };
    auto DX = [&](int di, int dj) { return L(a.dx, di, dj); };
    auto DY = [&](int di, int dj) { return L(a.dy, di, dj); };
    auto inner = [&](int di, int dj) {
      const int i = qi + di, j = qj + dj;
      return i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2;
    };
    // output cotangents where the output is an interior update, else zero
    auto cu = [&](int di, int dj) { return sel(inner(di, dj), L(a.ob_u, di, dj)); };
    auto cv = [&](int di, int dj) { return sel(inner(di, dj), L(a.ob_v, di, dj)); };
    // face cotangents at q + (di, dj)
    auto wU1 = [&](int di) { return cu(di, 0) - sel(qi + di >= 1, cu(di + 1, 0)); };
    auto wV2 = [&](int dj) { return cv(0, dj) - sel(qj + dj >= 1, cv(0, dj + 1)); };
    auto wV1 = [&](int di, int dj) { return cu(di, dj - 1) - cu(di, dj); };
    auto wU2 = [&](int di, int dj) { return cv(di - 1, dj) - cv(di, dj); };
    // corner sums
    auto dx4 = [&](int di, int dj) { return DX(di, dj) + DX(di - 1, dj) + DX(di, dj - 1) + DX(di - 1, dj - 1); };
    auto dy4 = [&](int di, int dj) { return DY(di, dj) + DY(di - 1, dj) + DY(di, dj - 1) + DY(di - 1, dj - 1); };
    auto d4 = [&](int di, int dj) { return D(di, dj) + D(di - 1, dj) + D(di, dj - 1) + D(di - 1, dj - 1); };
    auto am4 = [&](int di, int dj) { return AM(di, dj) + AM(di - 1, dj) + AM(di, dj - 1) + AM(di - 1, dj - 1); };

    T gu = T(0), gv = T(0), gub = T(0), gvb = T(0), gd = T(0), gam = T(0);

    // ---- FU1 at a = q - i (q is a + i), q (q is a), q + i (q is a - i)
    {
      const T w = wU1(-1);   // a = q - i
      const T u1 = U(0, 0), u0 = U(-1, 0), dp = D(0, 0), d0 = D(-1, 0), dm = D(-2, 0);
      const T uu = u1 + u0, s = (dp + d0) * u1 + (d0 + dm) * u0, wy = w * DY(-1, 0);
      gu += wy * c8 * ((dp + d0) * uu + s);
      gd += wy * c8 * u1 * uu;
      gub -= wy * d0 * two * AM(-1, 0) * rcp_lane(DX(-1, 0));
    }
    {
      const T w = wU1(0);    // a = q
      const T u1 = U(1, 0), u0 = U(0, 0), dp = D(1, 0), d0 = D(0, 0), dm = D(-1, 0);
      const T uu = u1 + u0, s = (dp + d0) * u1 + (d0 + dm) * u0, wy = w * DY(0, 0);
      const T rdx = rcp_lane(DX(0, 0)), am = AM(0, 0), diff = (UB(1, 0) - UB(0, 0)) * rdx;
      gu += wy * c8 * ((d0 + dm) * uu + s);
      gd += wy * (c8 * uu * uu - two * am * diff);
      gam -= wy * d0 * two * diff;
      gub += wy * d0 * two * am * rdx;
    }
    {
      const T w = wU1(1);    // a = q + i
      const T u1 = U(2, 0), u0 = U(1, 0);
      gd += w * DY(1, 0) * c8 * u0 * (u1 + u0);
    }
    // ---- FV2 at b = q - j, q, q + j
    {
      const T w = wV2(-1);
      const T v1 = V(0, 0), v0 = V(0, -1), dp = D(0, 0), d0 = D(0, -1), dm = D(0, -2);
      const T vv = v1 + v0, s = (dp + d0) * v1 + (d0 + dm) * v0, wx = w * DX(0, -1);
      gv += wx * c8 * ((dp + d0) * vv + s);
      gd += wx * c8 * v1 * vv;
      gvb -= wx * d0 * two * AM(0, -1) * rcp_lane(DY(0, -1));
    }
    {
      const T w = wV2(0);
      const T v1 = V(0, 1), v0 = V(0, 0), dp = D(0, 1), d0 = D(0, 0), dm = D(0, -1);
      const T vv = v1 + v0, s = (dp + d0) * v1 + (d0 + dm) * v0, wx = w * DX(0, 0);
      const T rdy = rcp_lane(DY(0, 0)), am = AM(0, 0), diff = (VB(0, 1) - VB(0, 0)) * rdy;
      gv += wx * c8 * ((d0 + dm) * vv + s);
      gd += wx * (c8 * vv * vv - two * am * diff);
      gam -= wx * d0 * two * diff;
      gvb += wx * d0 * two * am * rdy;
    }
    {
      const T w = wV2(1);
      const T v1 = V(0, 2), v0 = V(0, 1);
      gd += w * DX(0, 1) * c8 * v0 * (v1 + v0);
    }
    // ---- FV1, FU2 and TPS at the corners b = q + (di, dj), di, dj in {0, 1}
#pragma unroll
    for (int di = 0; di <= 1; ++di) {
#pragma unroll
      for (int dj = 0; dj <= 1; ++dj) {
        const T sx = dx4(di, dj), sy = dy4(di, dj);
        const T w1 = wV1(di, dj) * c4 * sx;   // of FV1's bracket
        const T w2 = wU2(di, dj) * c4 * sy;   // of FU2's bracket
        // FV1(b): ua(b), ua(b-j) (q is b or b-j when di = 0); va(b) (q = b), va(b-i) (q = b-i)
        {
          const T vb0 = V(di, dj), vbi = V(di - 1, dj), wsum = U(di, dj) + U(di, dj - 1);
          const T db = D(di, dj) + D(di, dj - 1), dbi = D(di - 1, dj) + D(di - 1, dj - 1);
          if (di == 0) gu += w1 * c8 * (db * vb0 + dbi * vbi);
          if (dj == 0) gv += w1 * c8 * (di == 0 ? db : dbi) * wsum;
          gd += w1 * c8 * (di == 0 ? vb0 : vbi) * wsum;
        }
        // FU2(a): ua(a) (q = a), ua(a-j) (q = a-j); va(a), va(a-i) (q is a or a-i when dj = 0)
        {
          const T ua0 = U(di, dj), uaj = U(di, dj - 1), vsum = V(di - 1, dj) + V(di, dj);
          const T da = D(di, dj) + D(di - 1, dj), daj = D(di, dj - 1) + D(di - 1, dj - 1);
          if (di == 0) gu += w2 * c8 * (dj == 0 ? da : daj) * vsum;
          if (dj == 0) gv += w2 * c8 * (da * ua0 + daj * uaj);
          gd += w2 * c8 * (dj == 0 ? ua0 : uaj) * vsum;
        }
        // TPS(b) enters FV1(b) and FU2(b) with the factor -1
        {
          const T wt = -(w1 + w2);
          const T ds = d4(di, dj), as = am4(di, dj), rsy = rcp_lane(sy), rsx = rcp_lane(sx);
          const T qq = (UB(di, dj) - UB(di, dj - 1)) * rsy + (VB(di, dj) - VB(di - 1, dj)) * rsx;
          gd += wt * c4 * as * qq;
          gam += wt * c4 * ds * qq;
          const T k = wt * c4 * ds * as;
          if (di == 0) gub += (dj == 0 ? k : -k) * rsy;
          if (dj == 0) gvb += (di == 0 ? k : -k) * rsx;
        }
      }
    }
    // accumulate: every old value read before the first store
    const uint32_t ix = lix2(q, qi, qj);
    T* dst[6] = {a.c_ua, a.c_va, a.c_uab, a.c_vab, a.c_d, a.c_aam2d};
    const T add[6] = {gu, gv, gub, gvb, gd, gam};
    T old[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) old[n] = ld_at(dst[n], q.o2, ix);
#pragma unroll
    for (int n = 0; n < 6; ++n) ref_at(dst[n], q.o2, ix) = old[n] + add[n];
  }
}

template <class T> void launch_adj_advave(const Ctx<T>& c, const AdvaveAdjArgs<T>& a) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_adj_advave<T>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)), dim3(TPB), 0,
                     c.stream, c.B, a);
}

template void launch_adj_advave<float>(const Ctx<float>&, const AdvaveAdjArgs<float>&);
template void launch_adj_advave<double>(const Ctx<double>&, const AdvaveAdjArgs<double>&);

}  // namespace pom
