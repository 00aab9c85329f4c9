// Flux-form adjoints of the tracer-advection operators: transposed stencils
// that take each face flux's partials once, in place of the generic gather
// (pom_kernels.h), which re-evaluates the whole point function once per
// stencil offset in dual numbers.
//
// The operator's output at an interior point p is
//   out(p) = (tb(p) fb(p) - dti2 (X(i+1) - X(i) + Y(j+1) - Y(j) + vdiv(p))) / (tf(p) + 1e-20)
// with X, Y the horizontal face fluxes (X(a): the face between columns a-1
// and a) and vdiv the vertical flux divergence.  With r(p) = obar(p) /
// (tf(p) + 1e-20) on the interior (zero elsewhere), a face's cotangent is a
// difference of r across it, and the thread of input point q sums the
// partials of the faces that touch q.  Vertical neighbours are lanes of the
// same 16-lane column group (k is the fastest index of a 3-D field), read
// with shuffles; the 2-D inputs (dt, etb, etf) sum their per-level partials
// over the column group.  Every cotangent element is written by one thread:
// no atomics, deterministic.
#pragma once
#include "pom_kernels.h"

namespace pom {

template <class T> struct FluxLoad {
  const FieldTab<T>* tab;
  Pos q;
  __device__ __forceinline__ T f3(int fid, int i, int j, int k) const {
    return ld_local(tab->f[fid] + q.o3, lix3(q, i, j, k));
  }
  __device__ __forceinline__ T f2(int fid, int i, int j) const { return ld_local(tab->f[fid] + q.o2, lix2(q, i, j)); }
  __device__ __forceinline__ T g(int gid, int i, int j) const { return ld_local(tab->g[gid] + q.o2, lix2(q, i, j)); }
};

// partials of one horizontal face flux of the centred tracer scheme
// (tracers.py:479-522): flux = (0.25 (dt0 + dt1) (f0 + f1) vel0
//   - 0.5 (aam0 + aam1) (h0 + h1) tprni ((fb0 - fc0) - (fb1 - fc1)) mask0 / (ds0 + ds1)) * 0.5 (dn0 + dn1)
// where 0 is the column on the face's high side, 1 the low side, ds the
// spacing across the face and dn along it.  Scaled by the face cotangent c.
template <class T> struct FaceT1 {
  T flux;                      // the flux itself
  T d_f, d_vel, d_dt, d_fb, d_aam;   // d flux / d (f0 = f1, vel0, dt0 = dt1, fb0 = -fb1 = -fc0 = fc1, aam0 = aam1) * c
};

template <class T>
__device__ __forceinline__ FaceT1<T> face_t1(T dt0, T dt1, T f0, T f1, T vel0, T fa0, T fa1, T am0, T am1, T h0, T h1,
                                             T mask0, T ds0, T ds1, T dn0, T dn1, double tprni, T c) {
  FaceT1<T> r;
  const T sdt = dt0 + dt1, sf = f0 + f1;
  const T dnh = T(0.5) * (dn0 + dn1);
  const T ids = rcp_lane(ds0 + ds1);
  const T g = T(0.5) * (h0 + h1) * T(tprni) * mask0 * ids;
  const T kd = (am0 + am1) * g;   // dif = kd (fa0 - fa1)
  const T adv = T(0.25) * sdt * sf * vel0;
  r.flux = (adv - kd * (fa0 - fa1)) * dnh;
  const T cd = c * dnh;
  r.d_f = T(0.25) * sdt * vel0 * cd;
  r.d_vel = T(0.25) * sdt * sf * cd;
  r.d_dt = T(0.25) * sf * vel0 * cd;
  r.d_fb = -kd * cd;
  r.d_aam = -g * (fa0 - fa1) * cd;
  return r;
}

// the transposed advt1 (OpAdvt1 in pom_ops.h; tracers.py:433-569).
// Branch-free: neighbour indices are clamped into the member and every load
// is issued before the arithmetic, so a wave has all of its loads in flight
// at once; faces and points that do not exist are removed by selects.
template <class T>
__global__ void __launch_bounds__(TPB) k_flux_advt1(Batch B, Vert<T> vert, Params prm, FieldTab<T> tab,
                                                     COutTab<T> obar, InTab<T> ib) {
  using O = OpAdvt1;
  Pos q;
  if (!pos3(B, q)) return;   // whole 16-lane column groups are live or not
  const FluxLoad<T> L{&tab, q};
  const int i = q.i, j = q.j, k = q.k, im = q.im, jm = q.jm;
  const int iw = i > 0 ? i - 1 : 0, ie = i < im - 1 ? i + 1 : im - 1;
  const int js = j > 0 ? j - 1 : 0, jn = j < jm - 1 ? j + 1 : jm - 1;
  const double dti2 = prm.dti2, tprni = prm.tprni;
  const T* ob = obar.o[0] + q.o3;

  // ---- loads: the point, its four horizontal neighbours
  const T fq = L.f3(O::F, i, j, k), fbq = L.f3(O::FB, i, j, k), fcq = L.f3(O::FCLIM, i, j, k);
  const T aamq = L.f3(O::AAM, i, j, k), uq = L.f3(O::U, i, j, k), vq = L.f3(O::V, i, j, k);
  const T wq = L.f3(O::W, i, j, k), obq = ld_local(ob, lix3(q, i, j, k));
  const T dtq = L.f2(O::DT, i, j), etfq = L.f2(O::ETF, i, j), etbq = L.f2(O::ETB, i, j);
  const T hq = L.g(O::GH, i, j), dxq = L.g(O::GDX, i, j), dyq = L.g(O::GDY, i, j), artq = L.g(O::GART, i, j);
  const T dumq = L.g(O::GDUM, i, j), dvmq = L.g(O::GDVM, i, j);
  // west
  const T f_w = L.f3(O::F, iw, j, k), fa_w = L.f3(O::FB, iw, j, k) - L.f3(O::FCLIM, iw, j, k);
  const T am_w = L.f3(O::AAM, iw, j, k), ob_w = ld_local(ob, lix3(q, iw, j, k));
  const T dt_w = L.f2(O::DT, iw, j), etf_w = L.f2(O::ETF, iw, j);
  const T h_w = L.g(O::GH, iw, j), dx_w = L.g(O::GDX, iw, j), dy_w = L.g(O::GDY, iw, j), art_w = L.g(O::GART, iw, j);
  // east
  const T f_e = L.f3(O::F, ie, j, k), fa_e = L.f3(O::FB, ie, j, k) - L.f3(O::FCLIM, ie, j, k);
  const T am_e = L.f3(O::AAM, ie, j, k), u_e = L.f3(O::U, ie, j, k), ob_e = ld_local(ob, lix3(q, ie, j, k));
  const T dt_e = L.f2(O::DT, ie, j), etf_e = L.f2(O::ETF, ie, j);
  const T h_e = L.g(O::GH, ie, j), dx_e = L.g(O::GDX, ie, j), dy_e = L.g(O::GDY, ie, j), art_e = L.g(O::GART, ie, j);
  const T dum_e = L.g(O::GDUM, ie, j);
  // south
  const T f_s = L.f3(O::F, i, js, k), fa_s = L.f3(O::FB, i, js, k) - L.f3(O::FCLIM, i, js, k);
  const T am_s = L.f3(O::AAM, i, js, k), ob_s = ld_local(ob, lix3(q, i, js, k));
  const T dt_s = L.f2(O::DT, i, js), etf_s = L.f2(O::ETF, i, js);
  const T h_s = L.g(O::GH, i, js), dx_s = L.g(O::GDX, i, js), dy_s = L.g(O::GDY, i, js), art_s = L.g(O::GART, i, js);
  // north
  const T f_n = L.f3(O::F, i, jn, k), fa_n = L.f3(O::FB, i, jn, k) - L.f3(O::FCLIM, i, jn, k);
  const T am_n = L.f3(O::AAM, i, jn, k), v_n = L.f3(O::V, i, jn, k), ob_n = ld_local(ob, lix3(q, i, jn, k));
  const T dt_n = L.f2(O::DT, i, jn), etf_n = L.f2(O::ETF, i, jn);
  const T h_n = L.g(O::GH, i, jn), dx_n = L.g(O::GDX, i, jn), dy_n = L.g(O::GDY, i, jn), art_n = L.g(O::GART, i, jn);
  const T dvm_n = L.g(O::GDVM, i, jn);
  const T idzk = rcp_lane(vert.dz(q.m, k)), idzu = rcp_lane(vert.dz(q.m, k - 1));
  const T idzd = rcp_lane(vert.dz(q.m, k + 1));

  // ---- r = obar / (tf + 1e-20) on the interior, zero elsewhere
  const bool kin = k < KBM1;
  auto inter = [&](int ii, int jj) { return kin && ii >= 1 && ii <= im - 2 && jj >= 1 && jj <= jm - 2; };
  auto rsel = [&](bool in, T o, T h, T etf, T art) { return in ? o * rcp_lane((h + etf) * art + T(1e-20)) : T(0); };
  const bool in = inter(i, j);
  const T rc = rsel(in, obq, hq, etfq, artq);
  const T rw = rsel(i >= 1 && inter(i - 1, j), ob_w, h_w, etf_w, art_w);
  const T re = rsel(i + 1 <= im - 1 && inter(i + 1, j), ob_e, h_e, etf_e, art_e);
  const T rs = rsel(j >= 1 && inter(i, j - 1), ob_s, h_s, etf_s, art_s);
  const T rn = rsel(j + 1 <= jm - 1 && inter(i, j + 1), ob_n, h_n, etf_n, art_n);

  // ---- horizontal faces; a face outside the member has a zero cotangent
  const T dti2t = T(dti2);
  const FaceT1<T> Xw = face_t1(dtq, dt_w, fq, f_w, uq, fbq - fcq, fa_w, aamq, am_w, hq, h_w, dumq, dxq, dx_w, dyq,
                               dy_w, tprni, i >= 1 ? dti2t * (rc - rw) : T(0));
  const FaceT1<T> Xe = face_t1(dt_e, dtq, f_e, fq, u_e, fa_e, fbq - fcq, am_e, aamq, h_e, hq, dum_e, dx_e, dxq, dy_e,
                               dyq, tprni, i + 1 <= im - 1 ? dti2t * (re - rc) : T(0));
  const FaceT1<T> Ys = face_t1(dtq, dt_s, fq, f_s, vq, fbq - fcq, fa_s, aamq, am_s, hq, h_s, dvmq, dyq, dy_s, dxq,
                               dx_s, tprni, j >= 1 ? dti2t * (rc - rs) : T(0));
  const FaceT1<T> Yn = face_t1(dt_n, dtq, f_n, fq, v_n, fa_n, fbq - fcq, am_n, aamq, h_n, hq, dvm_n, dy_n, dyq, dx_n,
                               dxq, tprni, j + 1 <= jm - 1 ? dti2t * (rn - rc) : T(0));
  T b_f = Xw.d_f + Xe.d_f + Ys.d_f + Yn.d_f;
  // b_fd: the diffusion partials of fb, which fclim takes with the opposite sign
  const T b_fd = Xw.d_fb - Xe.d_fb + Ys.d_fb - Yn.d_fb;
  const T b_aam = Xw.d_aam + Xe.d_aam + Ys.d_aam + Yn.d_aam;
  T b_dt = Xw.d_dt + Xe.d_dt + Ys.d_dt + Yn.d_dt;

  // ---- vertical faces of this column: Z(kk) = f(0) w(0) art (kk = 0),
  // 0.5 (f(kk-1) + f(kk)) w(kk) art (0 < kk < KBM1), 0 from KBM1 on
  const T f_up = __shfl_up(fq, 1, 16), w_dn = __shfl_down(wq, 1, 16), f_dn = __shfl_down(fq, 1, 16);
  const T r_up = __shfl_up(rc, 1, 16), r_dn = __shfl_down(rc, 1, 16);
  // Zbar(kk) = -dti2 (r(kk) / dz(kk) - r(kk-1) / dz(kk-1))
  const T zb_k = -dti2t * (rc * idzk - (k >= 1 ? r_up * idzu : T(0)));
  const T zb_k1 = -dti2t * ((k + 1 < KB ? r_dn : T(0)) * idzd - rc * idzk);
  const T fz = k == 0 ? fq : T(0.5) * (f_up + fq);   // face value at level k
  const T wz = k == 0 ? T(1) : T(0.5);               // d fz / d f(k)
  const T zk = kin ? fz * wq * artq : T(0);
  const T zk1 = k + 1 < KBM1 ? T(0.5) * (fq + f_dn) * w_dn * artq : T(0);
  b_f += kin ? wz * wq * artq * zb_k : T(0);
  b_f += k + 1 < KBM1 ? T(0.5) * w_dn * artq * zb_k1 : T(0);
  const T b_w = kin ? fz * artq * zb_k : T(0);

  // ---- the point itself: tb fb / (tf + 1e-20) and the depth terms
  const T tb = (hq + etbq) * artq;
  const T tf = (hq + etfq) * artq;
  const T hdiv = Xe.flux - Xw.flux + Yn.flux - Ys.flux;
  const T num = tb * fbq - dti2t * (hdiv + (zk - zk1) * idzk);
  const T b_fb = b_fd + (in ? tb * rc : T(0));
  T b_etb = in ? artq * fbq * rc : T(0);
  T b_etf = in ? -artq * rc * num * rcp_lane(tf + T(1e-20)) : T(0);

  const int64_t c3 = ix3(q, i, j, k);
  if (ib.o[O::FB]) put_in(ib, O::FB, c3, b_fb);
  if (ib.o[O::F]) put_in(ib, O::F, c3, b_f);
  if (ib.o[O::FCLIM]) put_in(ib, O::FCLIM, c3, -b_fd);
  if (ib.o[O::U]) put_in(ib, O::U, c3, Xw.d_vel);
  if (ib.o[O::V]) put_in(ib, O::V, c3, Ys.d_vel);
  if (ib.o[O::W]) put_in(ib, O::W, c3, b_w);
  if (ib.o[O::AAM]) put_in(ib, O::AAM, c3, b_aam);
  // 2-D inputs: sums over the column group
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) {
    b_dt += __shfl_xor(b_dt, o, 16);
    b_etb += __shfl_xor(b_etb, o, 16);
    b_etf += __shfl_xor(b_etf, o, 16);
  }
  if (k == 0) {
    const int64_t c2 = ix2(q, i, j);
    if (ib.o[O::DT]) put_in(ib, O::DT, c2, b_dt);
    if (ib.o[O::ETB]) put_in(ib, O::ETB, c2, b_etb);
    if (ib.o[O::ETF]) put_in(ib, O::ETF, c2, b_etf);
  }
}

}  // namespace pom
