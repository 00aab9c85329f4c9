// Kernels and launcher of the Smagorinsky adjoint (pom_smagadj.h).
#include "pom_smagadj.h"

namespace pom {

template <class T>
__global__ void __launch_bounds__(TPB) k_smag_fac(Batch B, Params prm, SmagAdjArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T q4 = T(0.25), s2 = T(0.70710678118654757);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos p;
    if (!pos3_at(B, blk, p)) continue;
    const int im = p.im, jm = p.jm, i = p.i, j = p.j, k = p.k;
    auto U = [&](int di, int dj) { return ld_at(a.u, p.o3, lix3(p, clampi(i + di, im), clampi(j + dj, jm), k)); };
    auto V = [&](int di, int dj) { return ld_at(a.v, p.o3, lix3(p, clampi(i + di, im), clampi(j + dj, jm), k)); };
    const uint32_t ix = lix3(p, i, j, k), i2 = lix2(p, i, j);
    const bool in = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2 && k < KBM1;
    const T c = ld_at(a.ob, p.o3, ix);
    const T dx = ld_at(a.dx, p.o2, i2), dy = ld_at(a.dy, p.o2, i2);
    const T ux = (U(1, 0) - U(0, 0)) / dx;
    const T vy = (V(0, 1) - V(0, 0)) / dy;
    const T ut = q4 * (U(0, 1) + U(1, 1) - U(0, -1) - U(1, -1)) / dy;
    const T vt = q4 * (V(1, 0) + V(1, 1) - V(-1, 0) - V(-1, 1)) / dx;
    const T sh = (ut + vt) * s2;
    const T n = sqrt(ux * ux + vy * vy + sh * sh);
    const T g = (in && n > T(0)) ? c * T(prm.horcon) * dx * dy / n : T(0);
    const T h = g * sh * s2 * q4;
    ref_at(a.fu, p.o3, ix) = g * ux / dx;
    ref_at(a.fv, p.o3, ix) = g * vy / dy;
    ref_at(a.hy, p.o3, ix) = h / dy;
    ref_at(a.hx, p.o3, ix) = h / dx;
    if (!in) {   // aam passes through
      T& r = ref_at(a.c_aam, p.o3, ix);
      r = r + c;
    }
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_smag_gat(Batch B, SmagAdjArgs<T> a) {
  const int nblk = B.blk3[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    if (!pos3_at(B, blk, q)) continue;
    const int im = q.im, jm = q.jm, i = q.i, j = q.j, k = q.k;
    // planes at q + (di, dj); a neighbour off the member is clamped onto an
    // edge cell, whose factors are zero
    auto F = [&](const T* f, int di, int dj) {
      return ld_at(f, q.o3, lix3(q, clampi(i + di, im), clampi(j + dj, jm), k));
    };
    const uint32_t ix = lix3(q, i, j, k);
    const T gu = F(a.fu, -1, 0) - F(a.fu, 0, 0) + F(a.hy, 0, -1) + F(a.hy, -1, -1) - F(a.hy, 0, 1) - F(a.hy, -1, 1);
    const T gv = F(a.fv, 0, -1) - F(a.fv, 0, 0) + F(a.hx, -1, 0) + F(a.hx, -1, -1) - F(a.hx, 1, 0) - F(a.hx, 1, -1);
    const T ou = ld_at(a.c_u, q.o3, ix), ov = ld_at(a.c_v, q.o3, ix);
    ref_at(a.c_u, q.o3, ix) = ou + gu;
    ref_at(a.c_v, q.o3, ix) = ov + gv;
  }
}

template <class T> void launch_adj_smag(const Ctx<T>& c, const SmagAdjArgs<T>& a) {
  if (c.nblk3 == 0) return;
  const dim3 grid(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min));
  hipLaunchKernelGGL((k_smag_fac<T>), grid, dim3(TPB), 0, c.stream, c.B, c.prm, a);
  hipLaunchKernelGGL((k_smag_gat<T>), grid, dim3(TPB), 0, c.stream, c.B, a);
}

template void launch_adj_smag<float>(const Ctx<float>&, const SmagAdjArgs<float>&);
template void launch_adj_smag<double>(const Ctx<double>&, const SmagAdjArgs<double>&);

}  // namespace pom
