// Flux-form adjoints of the tracer and turbulence advection operators: transposed stencils
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
    return ld_at(tab->f[fid], q.o3, lix3(q, i, j, k));
  }
  __device__ __forceinline__ T f2(int fid, int i, int j) const { return ld_at(tab->f[fid], q.o2, lix2(q, i, j)); }
  __device__ __forceinline__ T g(int gid, int i, int j) const { return ld_at(tab->g[gid], q.o2, lix2(q, i, j)); }
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

// The tracers of one launch (T and S for advt1, q2 and q2l for advq): each
// has its own fields and output cotangent, and they share the velocities,
// aam, w, the depths and the geometry, read from tab[0].  The shared input
// cotangents are ib[0]'s (the launcher checks that every tracer names the
// same arrays); they take the tracers' partials in tracer order, each added
// on its own, so the sums round as with one launch per tracer.
template <class T, int NT> struct FluxSet {
  FieldTab<T> tab[NT];
  const T* ob[NT];
  InTab<T> ib[NT];
};

// o[f][i] (+)= a[0], then += a[1], ... (assign: the first partial replaces the old value)
template <class T, int NT>
__device__ __forceinline__ void put_in_seq(const InTab<T>& t, int f, int64_t o, uint32_t li, const T (&a)[NT]) {
  T& r = ref_at(t.o[f], o, li);
  T v = ((t.assign >> f) & 1u) ? a[0] : r + a[0];
#pragma unroll
  for (int n = 1; n < NT; ++n) v += a[n];
  r = v;
}

// Accumulation in two phases: old values of every field first (each under
// its own uniform null check, with no use in between, so that the reads are
// in flight together), then the stores.  The arrays of one launch are distinct.
template <class T>
__device__ __forceinline__ T old_in(const InTab<T>& t, int f, int64_t o, uint32_t li) {
  return (t.o[f] && !((t.assign >> f) & 1u)) ? ld_at(t.o[f], o, li) : T(0);
}
template <class T, int NT>
__device__ __forceinline__ void put_seq(const InTab<T>& t, int f, int64_t o, uint32_t li, T old, const T (&a)[NT]) {
  if (!t.o[f]) return;
  T v = old + a[0];
#pragma unroll
  for (int n = 1; n < NT; ++n) v += a[n];
  ref_at(t.o[f], o, li) = v;
}

// 2-D partials summed over the 16-lane column group
template <class T> __device__ __forceinline__ T column_sum(T v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
  return v;
}

// the transposed advt1 (OpAdvt1 in pom_ops.h; tracers.py:433-569) for NT tracers.
// Branch-free: neighbour indices are clamped into the member and every load
// is issued before the arithmetic, so a wave has all of its loads in flight
// at once; faces and points that do not exist are removed by selects.
template <class T, int NT>
__global__ void __launch_bounds__(TPB) k_flux_advt1(Batch B, Vert<T> vert, Params prm, FluxSet<T, NT> fs) {
  using O = OpAdvt1;
  Pos q;
  if (!pos3(B, q)) return;   // whole 16-lane column groups are live or not
  const FluxLoad<T> L{&fs.tab[0], q};
  const int i = q.i, j = q.j, k = q.k, im = q.im, jm = q.jm;
  const int iw = i > 0 ? i - 1 : 0, ie = i < im - 1 ? i + 1 : im - 1;
  const int js = j > 0 ? j - 1 : 0, jn = j < jm - 1 ? j + 1 : jm - 1;
  const double tprni = prm.tprni;

  // ---- shared loads: the point and its four horizontal neighbours
  const T aamq = L.f3(O::AAM, i, j, k), uq = L.f3(O::U, i, j, k), vq = L.f3(O::V, i, j, k);
  const T wq = L.f3(O::W, i, j, k);
  const T dtq = L.f2(O::DT, i, j), etfq = L.f2(O::ETF, i, j), etbq = L.f2(O::ETB, i, j);
  const T hq = L.g(O::GH, i, j), dxq = L.g(O::GDX, i, j), dyq = L.g(O::GDY, i, j), artq = L.g(O::GART, i, j);
  const T dumq = L.g(O::GDUM, i, j), dvmq = L.g(O::GDVM, i, j);
  const T am_w = L.f3(O::AAM, iw, j, k), dt_w = L.f2(O::DT, iw, j), etf_w = L.f2(O::ETF, iw, j);
  const T h_w = L.g(O::GH, iw, j), dx_w = L.g(O::GDX, iw, j), dy_w = L.g(O::GDY, iw, j), art_w = L.g(O::GART, iw, j);
  const T am_e = L.f3(O::AAM, ie, j, k), u_e = L.f3(O::U, ie, j, k);
  const T dt_e = L.f2(O::DT, ie, j), etf_e = L.f2(O::ETF, ie, j);
  const T h_e = L.g(O::GH, ie, j), dx_e = L.g(O::GDX, ie, j), dy_e = L.g(O::GDY, ie, j), art_e = L.g(O::GART, ie, j);
  const T dum_e = L.g(O::GDUM, ie, j);
  const T am_s = L.f3(O::AAM, i, js, k), dt_s = L.f2(O::DT, i, js), etf_s = L.f2(O::ETF, i, js);
  const T h_s = L.g(O::GH, i, js), dx_s = L.g(O::GDX, i, js), dy_s = L.g(O::GDY, i, js), art_s = L.g(O::GART, i, js);
  const T am_n = L.f3(O::AAM, i, jn, k), v_n = L.f3(O::V, i, jn, k);
  const T dt_n = L.f2(O::DT, i, jn), etf_n = L.f2(O::ETF, i, jn);
  const T h_n = L.g(O::GH, i, jn), dx_n = L.g(O::GDX, i, jn), dy_n = L.g(O::GDY, i, jn), art_n = L.g(O::GART, i, jn);
  const T dvm_n = L.g(O::GDVM, i, jn);
  const T idzk = rcp_lane(vert.dz(q.m, k)), idzu = rcp_lane(vert.dz(q.m, k - 1));
  const T idzd = rcp_lane(vert.dz(q.m, k + 1));
  // ---- per-tracer loads
  T fq[NT], fbq[NT], fcq[NT], obq[NT], f_w[NT], fa_w[NT], ob_w[NT], f_e[NT], fa_e[NT], ob_e[NT];
  T f_s[NT], fa_s[NT], ob_s[NT], f_n[NT], fa_n[NT], ob_n[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const FluxLoad<T> Lt{&fs.tab[t], q};
    const T* ob = fs.ob[t];
    fq[t] = Lt.f3(O::F, i, j, k); fbq[t] = Lt.f3(O::FB, i, j, k); fcq[t] = Lt.f3(O::FCLIM, i, j, k);
    obq[t] = ld_at(ob, q.o3, lix3(q, i, j, k));
    f_w[t] = Lt.f3(O::F, iw, j, k); fa_w[t] = Lt.f3(O::FB, iw, j, k) - Lt.f3(O::FCLIM, iw, j, k);
    ob_w[t] = ld_at(ob, q.o3, lix3(q, iw, j, k));
    f_e[t] = Lt.f3(O::F, ie, j, k); fa_e[t] = Lt.f3(O::FB, ie, j, k) - Lt.f3(O::FCLIM, ie, j, k);
    ob_e[t] = ld_at(ob, q.o3, lix3(q, ie, j, k));
    f_s[t] = Lt.f3(O::F, i, js, k); fa_s[t] = Lt.f3(O::FB, i, js, k) - Lt.f3(O::FCLIM, i, js, k);
    ob_s[t] = ld_at(ob, q.o3, lix3(q, i, js, k));
    f_n[t] = Lt.f3(O::F, i, jn, k); fa_n[t] = Lt.f3(O::FB, i, jn, k) - Lt.f3(O::FCLIM, i, jn, k);
    ob_n[t] = ld_at(ob, q.o3, lix3(q, i, jn, k));
  }

  // ---- r = obar / (tf + 1e-20) on the interior, zero elsewhere
  const bool kin = k < KBM1;
  auto inter = [&](int ii, int jj) { return kin && ii >= 1 && ii <= im - 2 && jj >= 1 && jj <= jm - 2; };
  const bool in = inter(i, j), in_w = i >= 1 && inter(i - 1, j), in_e = i + 1 <= im - 1 && inter(i + 1, j);
  const bool in_s = j >= 1 && inter(i, j - 1), in_n = j + 1 <= jm - 1 && inter(i, j + 1);
  const T den = rcp_lane((hq + etfq) * artq + T(1e-20)), den_w = rcp_lane((h_w + etf_w) * art_w + T(1e-20));
  const T den_e = rcp_lane((h_e + etf_e) * art_e + T(1e-20)), den_s = rcp_lane((h_s + etf_s) * art_s + T(1e-20));
  const T den_n = rcp_lane((h_n + etf_n) * art_n + T(1e-20));
  const T dti2t = T(prm.dti2);
  const T w_dn = __shfl_down(wq, 1, 16);
  const T tb = (hq + etbq) * artq;
  const T tf = (hq + etfq) * artq;
  const T fz_w = k == 0 ? T(1) : T(0.5);   // d (face value at level k) / d f(k)

  T p_u[NT], p_v[NT], p_w[NT], p_aam[NT], p_dt[NT], p_etb[NT], p_etf[NT];
  T s_fb[NT], s_f[NT], s_fc[NT];   // the per-field partials, stored after every old value is read
  const uint32_t c3 = lix3(q, i, j, k);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const T rc = in ? obq[t] * den : T(0), rw = in_w ? ob_w[t] * den_w : T(0), re = in_e ? ob_e[t] * den_e : T(0);
    const T rs = in_s ? ob_s[t] * den_s : T(0), rn = in_n ? ob_n[t] * den_n : T(0);
    const T faq = fbq[t] - fcq[t];
    // ---- horizontal faces; a face outside the member has a zero cotangent
    const FaceT1<T> Xw = face_t1(dtq, dt_w, fq[t], f_w[t], uq, faq, fa_w[t], aamq, am_w, hq, h_w, dumq, dxq, dx_w,
                                 dyq, dy_w, tprni, i >= 1 ? dti2t * (rc - rw) : T(0));
    const FaceT1<T> Xe = face_t1(dt_e, dtq, f_e[t], fq[t], u_e, fa_e[t], faq, am_e, aamq, h_e, hq, dum_e, dx_e, dxq,
                                 dy_e, dyq, tprni, i + 1 <= im - 1 ? dti2t * (re - rc) : T(0));
    const FaceT1<T> Ys = face_t1(dtq, dt_s, fq[t], f_s[t], vq, faq, fa_s[t], aamq, am_s, hq, h_s, dvmq, dyq, dy_s,
                                 dxq, dx_s, tprni, j >= 1 ? dti2t * (rc - rs) : T(0));
    const FaceT1<T> Yn = face_t1(dt_n, dtq, f_n[t], fq[t], v_n, fa_n[t], faq, am_n, aamq, h_n, hq, dvm_n, dy_n, dyq,
                                 dx_n, dxq, tprni, j + 1 <= jm - 1 ? dti2t * (rn - rc) : T(0));
    T b_f = Xw.d_f + Xe.d_f + Ys.d_f + Yn.d_f;
    // b_fd: the diffusion partials of fb, which fclim takes with the opposite sign
    const T b_fd = Xw.d_fb - Xe.d_fb + Ys.d_fb - Yn.d_fb;
    p_aam[t] = Xw.d_aam + Xe.d_aam + Ys.d_aam + Yn.d_aam;
    p_dt[t] = Xw.d_dt + Xe.d_dt + Ys.d_dt + Yn.d_dt;
    p_u[t] = Xw.d_vel;
    p_v[t] = Ys.d_vel;

    // ---- vertical faces of this column: Z(kk) = f(0) w(0) art (kk = 0),
    // 0.5 (f(kk-1) + f(kk)) w(kk) art (0 < kk < KBM1), 0 from KBM1 on
    const T f_up = __shfl_up(fq[t], 1, 16), f_dn = __shfl_down(fq[t], 1, 16);
    const T r_up = __shfl_up(rc, 1, 16), r_dn = __shfl_down(rc, 1, 16);
    // Zbar(kk) = -dti2 (r(kk) / dz(kk) - r(kk-1) / dz(kk-1))
    const T zb_k = -dti2t * (rc * idzk - (k >= 1 ? r_up * idzu : T(0)));
    const T zb_k1 = -dti2t * ((k + 1 < KB ? r_dn : T(0)) * idzd - rc * idzk);
    const T fz = k == 0 ? fq[t] : T(0.5) * (f_up + fq[t]);   // face value at level k
    const T zk = kin ? fz * wq * artq : T(0);
    const T zk1 = k + 1 < KBM1 ? T(0.5) * (fq[t] + f_dn) * w_dn * artq : T(0);
    b_f += kin ? fz_w * wq * artq * zb_k : T(0);
    b_f += k + 1 < KBM1 ? T(0.5) * w_dn * artq * zb_k1 : T(0);
    p_w[t] = kin ? fz * artq * zb_k : T(0);

    // ---- the point itself: tb fb / (tf + 1e-20) and the depth terms
    const T hdiv = Xe.flux - Xw.flux + Yn.flux - Ys.flux;
    const T num = tb * fbq[t] - dti2t * (hdiv + (zk - zk1) * idzk);
    const T b_fb = b_fd + (in ? tb * rc : T(0));
    p_etb[t] = in ? artq * fbq[t] * rc : T(0);
    p_etf[t] = in ? -artq * rc * num * rcp_lane(tf + T(1e-20)) : T(0);

    s_fb[t] = b_fb;
    s_f[t] = b_f;
    s_fc[t] = -b_fd;
  }
  const InTab<T>& ib = fs.ib[0];
  // 2-D inputs: sums over the column group
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    p_dt[t] = column_sum(p_dt[t]);
    p_etb[t] = column_sum(p_etb[t]);
    p_etf[t] = column_sum(p_etf[t]);
  }
  const uint32_t c2 = lix2(q, i, j);
  T o_fb[NT], o_f[NT], o_fc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    o_fb[t] = old_in(fs.ib[t], O::FB, q.o3, c3);
    o_f[t] = old_in(fs.ib[t], O::F, q.o3, c3);
    o_fc[t] = old_in(fs.ib[t], O::FCLIM, q.o3, c3);
  }
  const T o_u = old_in(ib, O::U, q.o3, c3), o_v = old_in(ib, O::V, q.o3, c3), o_w = old_in(ib, O::W, q.o3, c3);
  const T o_aam = old_in(ib, O::AAM, q.o3, c3);
  T o_dt = T(0), o_etb = T(0), o_etf = T(0);
  if (k == 0) {
    o_dt = old_in(ib, O::DT, q.o2, c2);
    o_etb = old_in(ib, O::ETB, q.o2, c2);
    o_etf = old_in(ib, O::ETF, q.o2, c2);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const InTab<T>& it = fs.ib[t];
    if (it.o[O::FB]) ref_at(it.o[O::FB], q.o3, c3) = o_fb[t] + s_fb[t];
    if (it.o[O::F]) ref_at(it.o[O::F], q.o3, c3) = o_f[t] + s_f[t];
    if (it.o[O::FCLIM]) ref_at(it.o[O::FCLIM], q.o3, c3) = o_fc[t] + s_fc[t];
  }
  put_seq(ib, O::U, q.o3, c3, o_u, p_u);
  put_seq(ib, O::V, q.o3, c3, o_v, p_v);
  put_seq(ib, O::W, q.o3, c3, o_w, p_w);
  put_seq(ib, O::AAM, q.o3, c3, o_aam, p_aam);
  if (k == 0) {
    put_seq(ib, O::DT, q.o2, c2, o_dt, p_dt);
    put_seq(ib, O::ETB, q.o2, c2, o_etb, p_etb);
    put_seq(ib, O::ETF, q.o2, c2, o_etf, p_etf);
  }
}

// partials of one horizontal face flux of the turbulence advection at a
// w level k (tracers.py:41-99): flux = (0.125 (q0 + q1) (dt0 + dt1) (vel0(k) + vel0(k-1))
//   - 0.25 (aam0(k) + aam1(k) + aam0(k-1) + aam1(k-1)) (h0 + h1) (qb0 - qb1) mask0 / (ds0 + ds1))
//   * 0.5 (dn0 + dn1); sides and spacings as for face_t1.  Scaled by the face cotangent c.
template <class T> struct FaceQ {
  T flux;
  T d_q, d_dt, d_vel, d_qb, d_aam;   // (q0 = q1, dt0 = dt1, both vel levels, qb0 = -qb1, each of the four aam)
};

template <class T>
__device__ __forceinline__ FaceQ<T> face_q(T q0, T q1, T dt0, T dt1, T vk, T vkm, T a0k, T a1k, T a0m, T a1m, T qb0,
                                           T qb1, T h0, T h1, T mask0, T ds0, T ds1, T dn0, T dn1, T c) {
  FaceQ<T> r;
  const T sq = q0 + q1, sdt = dt0 + dt1, sv = vk + vkm;
  const T dnh = T(0.5) * (dn0 + dn1);
  const T g = T(0.25) * (h0 + h1) * mask0 * rcp_lane(ds0 + ds1);
  const T asum = a0k + a1k + a0m + a1m;
  r.flux = (T(0.125) * sq * sdt * sv - asum * g * (qb0 - qb1)) * dnh;
  const T cd = c * dnh;
  r.d_q = T(0.125) * sdt * sv * cd;
  r.d_dt = T(0.125) * sq * sv * cd;
  r.d_vel = T(0.125) * sq * sdt * cd;
  r.d_qb = -asum * g * cd;
  r.d_aam = -g * (qb0 - qb1) * cd;
  return r;
}

// the transposed advq (OpAdvq in pom_ops.h; tracers.py:14-156) for NT
// fields, branch-free as k_flux_advt1.  A face at level k also reads vel and
// aam at level k - 1: those partials go to the lane of level k - 1 by a shuffle.
template <class T, int NT>
__global__ void __launch_bounds__(TPB) k_flux_advq(Batch B, Vert<T> vert, Params prm, FluxSet<T, NT> fs) {
  using O = OpAdvq;
  Pos q;
  if (!pos3(B, q)) return;   // whole 16-lane column groups are live or not
  const FluxLoad<T> L{&fs.tab[0], q};
  const int i = q.i, j = q.j, k = q.k, im = q.im, jm = q.jm;
  const int iw = i > 0 ? i - 1 : 0, ie = i < im - 1 ? i + 1 : im - 1;
  const int js = j > 0 ? j - 1 : 0, jn = j < jm - 1 ? j + 1 : jm - 1;
  const T dti2t = T(prm.dti2);

  // ---- shared loads
  const T uq = L.f3(O::U, i, j, k), vq = L.f3(O::V, i, j, k);
  const T wq = L.f3(O::W, i, j, k), amq = L.f3(O::AAM, i, j, k);
  const T dtq = L.f2(O::DT, i, j), etfq = L.f2(O::ETF, i, j), etbq = L.f2(O::ETB, i, j);
  const T hq = L.g(O::GH, i, j), dxq = L.g(O::GDX, i, j), dyq = L.g(O::GDY, i, j), artq = L.g(O::GART, i, j);
  const T dumq = L.g(O::GDUM, i, j), dvmq = L.g(O::GDVM, i, j);
  const T am_w = L.f3(O::AAM, iw, j, k), dt_w = L.f2(O::DT, iw, j), etf_w = L.f2(O::ETF, iw, j);
  const T h_w = L.g(O::GH, iw, j), dx_w = L.g(O::GDX, iw, j), dy_w = L.g(O::GDY, iw, j), art_w = L.g(O::GART, iw, j);
  const T am_e = L.f3(O::AAM, ie, j, k), u_e = L.f3(O::U, ie, j, k);
  const T dt_e = L.f2(O::DT, ie, j), etf_e = L.f2(O::ETF, ie, j);
  const T h_e = L.g(O::GH, ie, j), dx_e = L.g(O::GDX, ie, j), dy_e = L.g(O::GDY, ie, j), art_e = L.g(O::GART, ie, j);
  const T dum_e = L.g(O::GDUM, ie, j);
  const T am_s = L.f3(O::AAM, i, js, k), dt_s = L.f2(O::DT, i, js), etf_s = L.f2(O::ETF, i, js);
  const T h_s = L.g(O::GH, i, js), dx_s = L.g(O::GDX, i, js), dy_s = L.g(O::GDY, i, js), art_s = L.g(O::GART, i, js);
  const T am_n = L.f3(O::AAM, i, jn, k), v_n = L.f3(O::V, i, jn, k);
  const T dt_n = L.f2(O::DT, i, jn), etf_n = L.f2(O::ETF, i, jn);
  const T h_n = L.g(O::GH, i, jn), dx_n = L.g(O::GDX, i, jn), dy_n = L.g(O::GDY, i, jn), art_n = L.g(O::GART, i, jn);
  const T dvm_n = L.g(O::GDVM, i, jn);
  const T dzk = vert.dz(q.m, k), dzu = vert.dz(q.m, k - 1);
  // ---- per-field loads
  T qq[NT], qbq[NT], obq[NT], q_w[NT], qb_w[NT], ob_w[NT], q_e[NT], qb_e[NT], ob_e[NT];
  T q_s[NT], qb_s[NT], ob_s[NT], q_n[NT], qb_n[NT], ob_n[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const FluxLoad<T> Lt{&fs.tab[t], q};
    const T* ob = fs.ob[t];
    qq[t] = Lt.f3(O::Q, i, j, k); qbq[t] = Lt.f3(O::QB, i, j, k); obq[t] = ld_at(ob, q.o3, lix3(q, i, j, k));
    q_w[t] = Lt.f3(O::Q, iw, j, k); qb_w[t] = Lt.f3(O::QB, iw, j, k); ob_w[t] = ld_at(ob, q.o3, lix3(q, iw, j, k));
    q_e[t] = Lt.f3(O::Q, ie, j, k); qb_e[t] = Lt.f3(O::QB, ie, j, k); ob_e[t] = ld_at(ob, q.o3, lix3(q, ie, j, k));
    q_s[t] = Lt.f3(O::Q, i, js, k); qb_s[t] = Lt.f3(O::QB, i, js, k); ob_s[t] = ld_at(ob, q.o3, lix3(q, i, js, k));
    q_n[t] = Lt.f3(O::Q, i, jn, k); qb_n[t] = Lt.f3(O::QB, i, jn, k); ob_n[t] = ld_at(ob, q.o3, lix3(q, i, jn, k));
  }
  // values at level k - 1 of this lane's columns (level 0 takes its own: its
  // faces carry a zero cotangent, since no output lives on level 0)
  const T u_km = __shfl_up(uq, 1, 16), v_km = __shfl_up(vq, 1, 16), am_km = __shfl_up(amq, 1, 16);
  const T u_e_km = __shfl_up(u_e, 1, 16), v_n_km = __shfl_up(v_n, 1, 16);
  const T am_w_km = __shfl_up(am_w, 1, 16), am_e_km = __shfl_up(am_e, 1, 16);
  const T am_s_km = __shfl_up(am_s, 1, 16), am_n_km = __shfl_up(am_n, 1, 16);

  // ---- r = obar / (tf + 1e-20) on the interior w levels 1 .. KBM1 - 1, zero elsewhere
  const bool kin = k >= 1 && k < KBM1;
  auto inter = [&](int ii, int jj) { return kin && ii >= 1 && ii <= im - 2 && jj >= 1 && jj <= jm - 2; };
  const bool in = inter(i, j), in_w = i >= 1 && inter(i - 1, j), in_e = i + 1 <= im - 1 && inter(i + 1, j);
  const bool in_s = j >= 1 && inter(i, j - 1), in_n = j + 1 <= jm - 1 && inter(i, j + 1);
  const T den = rcp_lane((hq + etfq) * artq + T(1e-20)), den_w = rcp_lane((h_w + etf_w) * art_w + T(1e-20));
  const T den_e = rcp_lane((h_e + etf_e) * art_e + T(1e-20)), den_s = rcp_lane((h_s + etf_s) * art_s + T(1e-20));
  const T den_n = rcp_lane((h_n + etf_n) * art_n + T(1e-20));
  const bool below = k + 1 < KB;
  const T w_up = __shfl_up(wq, 1, 16), w_dn = __shfl_down(wq, 1, 16);
  const T idzs = rcp_lane(dzk + dzu);
  const T tb = (hq + etbq) * artq;
  const T tf = (hq + etfq) * artq;

  T p_u[NT], p_v[NT], p_w[NT], p_aam[NT], p_dt[NT], p_etb[NT], p_etf[NT];
  T s_qb[NT], s_q[NT];   // the per-field partials, stored after every old value is read
  const uint32_t c3 = lix3(q, i, j, k);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const T rc = in ? obq[t] * den : T(0), rw = in_w ? ob_w[t] * den_w : T(0), re = in_e ? ob_e[t] * den_e : T(0);
    const T rs = in_s ? ob_s[t] * den_s : T(0), rn = in_n ? ob_n[t] * den_n : T(0);
    // ---- horizontal faces at this level
    const FaceQ<T> Xw = face_q(qq[t], q_w[t], dtq, dt_w, uq, u_km, amq, am_w, am_km, am_w_km, qbq[t], qb_w[t], hq, h_w,
                               dumq, dxq, dx_w, dyq, dy_w, i >= 1 ? dti2t * (rc - rw) : T(0));
    const FaceQ<T> Xe = face_q(q_e[t], qq[t], dt_e, dtq, u_e, u_e_km, am_e, amq, am_e_km, am_km, qb_e[t], qbq[t], h_e,
                               hq, dum_e, dx_e, dxq, dy_e, dyq, i + 1 <= im - 1 ? dti2t * (re - rc) : T(0));
    const FaceQ<T> Ys = face_q(qq[t], q_s[t], dtq, dt_s, vq, v_km, amq, am_s, am_km, am_s_km, qbq[t], qb_s[t], hq, h_s,
                               dvmq, dyq, dy_s, dxq, dx_s, j >= 1 ? dti2t * (rc - rs) : T(0));
    const FaceQ<T> Yn = face_q(q_n[t], qq[t], dt_n, dtq, v_n, v_n_km, am_n, amq, am_n_km, am_km, qb_n[t], qbq[t], h_n,
                               hq, dvm_n, dy_n, dyq, dx_n, dxq, j + 1 <= jm - 1 ? dti2t * (rn - rc) : T(0));
    // this level's inputs
    T b_q = Xw.d_q + Xe.d_q + Ys.d_q + Yn.d_q;
    T b_qb = Xw.d_qb - Xe.d_qb + Ys.d_qb - Yn.d_qb;
    p_dt[t] = Xw.d_dt + Xe.d_dt + Ys.d_dt + Yn.d_dt;
    const T am_here = Xw.d_aam + Xe.d_aam + Ys.d_aam + Yn.d_aam;
    // the partials the faces at level k + 1 take of this level (vel, aam at their k - 1)
    const T up_u = __shfl_down(Xw.d_vel, 1, 16), up_v = __shfl_down(Ys.d_vel, 1, 16);
    const T up_am = __shfl_down(am_here, 1, 16);
    p_u[t] = Xw.d_vel + (below ? up_u : T(0));
    p_v[t] = Ys.d_vel + (below ? up_v : T(0));
    p_aam[t] = am_here + (below ? up_am : T(0));

    // ---- vertical term vert(k) = (w(k-1) q(k-1) - w(k+1) q(k+1)) art / (dz(k) + dz(k-1))
    const T q_up = __shfl_up(qq[t], 1, 16), q_dn = __shfl_down(qq[t], 1, 16);
    const T vb = -dti2t * rc * artq * idzs;   // cotangent of w(k-1) q(k-1) (and minus that of w(k+1) q(k+1))
    const T vb_dn = __shfl_down(vb, 1, 16), vb_up = __shfl_up(vb, 1, 16);
    const T from_dn = below ? vb_dn : T(0);          // vert(k + 1) reads this level as its k - 1
    const T from_up = k >= 1 ? vb_up : T(0);         // vert(k - 1) reads this level as its k + 1
    b_q += (from_dn - from_up) * wq;
    p_w[t] = (from_dn - from_up) * qq[t];

    // ---- the point itself: tb qb / (tf + 1e-20) and the depth terms
    const T vert_k = (w_up * q_up - w_dn * q_dn) * artq * idzs;
    const T hdiv = Xe.flux - Xw.flux + Yn.flux - Ys.flux;
    const T num = tb * qbq[t] - dti2t * (vert_k + hdiv);
    b_qb += in ? tb * rc : T(0);
    p_etb[t] = in ? artq * qbq[t] * rc : T(0);
    p_etf[t] = in ? -artq * rc * num * rcp_lane(tf + T(1e-20)) : T(0);

    s_qb[t] = b_qb;
    s_q[t] = b_q;
  }
  const InTab<T>& ib = fs.ib[0];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    p_dt[t] = column_sum(p_dt[t]);
    p_etb[t] = column_sum(p_etb[t]);
    p_etf[t] = column_sum(p_etf[t]);
  }
  const uint32_t c2 = lix2(q, i, j);
  T o_qb[NT], o_q[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    o_qb[t] = old_in(fs.ib[t], O::QB, q.o3, c3);
    o_q[t] = old_in(fs.ib[t], O::Q, q.o3, c3);
  }
  const T o_u = old_in(ib, O::U, q.o3, c3), o_v = old_in(ib, O::V, q.o3, c3), o_w = old_in(ib, O::W, q.o3, c3);
  const T o_aam = old_in(ib, O::AAM, q.o3, c3);
  T o_dt = T(0), o_etb = T(0), o_etf = T(0);
  if (k == 0) {
    o_dt = old_in(ib, O::DT, q.o2, c2);
    o_etb = old_in(ib, O::ETB, q.o2, c2);
    o_etf = old_in(ib, O::ETF, q.o2, c2);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const InTab<T>& it = fs.ib[t];
    if (it.o[O::QB]) ref_at(it.o[O::QB], q.o3, c3) = o_qb[t] + s_qb[t];
    if (it.o[O::Q]) ref_at(it.o[O::Q], q.o3, c3) = o_q[t] + s_q[t];
  }
  put_seq(ib, O::U, q.o3, c3, o_u, p_u);
  put_seq(ib, O::V, q.o3, c3, o_v, p_v);
  put_seq(ib, O::W, q.o3, c3, o_w, p_w);
  put_seq(ib, O::AAM, q.o3, c3, o_aam, p_aam);
  if (k == 0) {
    put_seq(ib, O::DT, q.o2, c2, o_dt, p_dt);
    put_seq(ib, O::ETB, q.o2, c2, o_etb, p_etb);
    put_seq(ib, O::ETF, q.o2, c2, o_etf, p_etf);
  }
}

}  // namespace pom
