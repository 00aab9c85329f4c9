// Mellor-Yamada 2.5 turbulence closure (profq), forward and reverse sweep.
// reference: src/neuralpom/cores/explicit/tracers.py:158-431
//
// One thread per water column.  Inputs read from neighbouring columns (shear
// uses u(i+1), v(j+1); the surface/bottom stress averages use i+1, j+1) make
// the reverse sweep write "neighbour" cotangents to scratch columns that a
// second pass gathers.
#pragma once
#include "pom_column.h"

namespace pom {

namespace myc {
constexpr double A1 = 0.92, B1 = 16.6, A2 = 0.74, B2 = 10.1, C1 = 0.08;
constexpr double E1 = 1.8, E2 = 1.33, SEF = 1.0, CBCNST = 100.0, SURFL = 2.0e5;
constexpr double COEF4 = 18.0 * A1 * A1 + 9.0 * A1 * A2;
constexpr double COEF5 = 9.0 * A1 * A2;
constexpr double COEF1 = A2 * (1.0 - 6.0 * A1 / B1);
constexpr double COEF2 = 3.0 * A2 * B2 + 18.0 * A1 * A2;
constexpr double COEF3 = A1 * (1.0 - 3.0 * C1 - 6.0 * A1 / B1);
}  // namespace myc

// What the reverse kernels need of the forward sweep and do not recompute:
// the recurrences of both tridiagonal solves (ee, gg), their unclipped
// solutions, the unclamped stability argument and the turbulence production.
// k_profq writes it when ee1 is non-null (the adjoint's recompute pass).
template <class T> struct ProfqTape {
  T *ee1, *gg1, *q2n, *ee2, *gg2, *q2ln, *ghr, *kn;
};

template <class T> struct ProfqArgs {
  const T *q2f, *q2lf, *q2, *q2b, *q2lb, *km, *kh, *kq, *u, *v, *t, *s, *rho;
  const T *wusurf, *wvsurf, *wubot, *wvbot, *etf, *dt, *h;
  T *q2o, *q2lo, *lo, *kqo, *kmo, *kho;
  ProfqTape<T> tape;
  // the reference operator's work arrays, for callers of the operator on its
  // own (tracers.py:423-431 returns them): stability functions sm, sh, sound
  // speed cc, diffusion coefficients a, c (3-D) and dh = h + etf (2-D); each
  // may be null
  T *smo, *sho, *cco, *ao, *co, *dho;
};

template <class T> __device__ __forceinline__ T sgn(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

// all per-column intermediates of the forward sweep
template <class T> struct ProfqCol {
  T a[KB], c[KB], cc[KB], boygr[KB], lraw[KB], l[KB], ghr[KB], gh[KB], kn[KB], dtef[KB];
  T ee1[KB], gg1[KB], q2n[KB], ee2[KB], gg2[KB], q2ln[KB];
  T dh, utau2, l0, ubt, vbt, qbm, q2bot;
  bool has_s, inner;
};

// Forward sweep of one column.  Own-column 3-D inputs are loaded through the
// block's LDS tile (every thread of the block executes this function).
template <class T>
__device__ __forceinline__ void profq_fwd_col(const Pos& p, const ColIO<T>& io, const Vert<T>& vg, const Params& prm,
                                              const ProfqArgs<T>& s, ProfqCol<T>& C) {
  using namespace myc;
  const int i = p.i, j = p.j, m = p.m;
  const int64_t c2 = ix2(p, i, j);
  const T dti2 = T(prm.dti2), grav = T(prm.grav), kappa = T(prm.kappa), small = T(prm.small);
  const T weps = T(prm.core ? 1e-8 : 1e-20);   // wall guard: explicit tracers.py:362, implicit :360
  C.dh = s.h[c2] + s.etf[c2];
  const T dh = C.dh;
  T x[KB], y[KB];
  col_ld(io, s.kq, x);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.a[k] = C.c[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    C.a[k] = -dti2 * (x[k + 1] + x[k] + T(2.0 * prm.umol)) * T(0.5) /
             (vg.dzz(m, k - 1) * vg.dz(m, k) * dh * dh);
    C.c[k] = -dti2 * (x[k - 1] + x[k] + T(2.0 * prm.umol)) * T(0.5) /
             (vg.dzz(m, k - 1) * vg.dz(m, k - 1) * dh * dh);
  }
  C.has_s = (i < p.im - 1) && (j < p.jm - 1);
  C.inner = i >= 1 && i <= p.im - 2 && j >= 1 && j <= p.jm - 2;
  C.utau2 = T(0);
  C.l0 = T(0);
  C.q2bot = T(0);
  C.ubt = C.vbt = C.qbm = T(0);
  if (C.has_s) {
    const T ust = T(0.5) * (s.wusurf[c2] + s.wusurf[ix2(p, i + 1, j)]);
    const T vst = T(0.5) * (s.wvsurf[c2] + s.wvsurf[ix2(p, i, j + 1)]);
    C.utau2 = sqrt(ust * ust + vst * vst + T(1e-8));
    C.l0 = T(SURFL) * C.utau2 / grav;
    C.ubt = T(0.5) * (s.wubot[c2] + s.wubot[ix2(p, i + 1, j)]);
    C.vbt = T(0.5) * (s.wvbot[c2] + s.wvbot[ix2(p, i, j + 1)]);
    C.qbm = sqrt(C.ubt * C.ubt + C.vbt * C.vbt + T(1e-8));
    C.q2bot = C.qbm * T(pow(16.6, 2.0 / 3.0) * SEF);
  }
  // sound speed (tracers.py:245-255)
  col_ld(io, s.t, x);
  col_ld(io, s.s, y);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.cc[k] = T(0);
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) {
    const T tp = x[k] + T(prm.tbias), sp = y[k] + T(prm.sbias);
    const T pp = T(prm.grav * prm.rhoref) * (-vg.zz(m, k) * dh) * T(1.e-4);
    T cc0 = T(1449.1) + T(0.00821) * pp + T(4.55) * tp - T(0.045) * (tp * tp) + T(1.34) * (sp - T(35.0));
    const T ccs = fabs(cc0) < T(1e-6) ? T(1e-6) : cc0;
    const T rad = (T(1) - T(0.01642) * pp / ccs) * (T(1) - T(0.40) * pp / (ccs * ccs));
    C.cc[k] = ccs / sqrt(fmax(rad, T(1e-8)));
  }
  col_ld(io, s.rho, x);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.boygr[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k)
    C.boygr[k] = grav * (x[k - 1] - x[k]) / (vg.dzz(m, k - 1) * dh) +
                 T(prm.grav * prm.grav * 2.0) / (C.cc[k - 1] * C.cc[k - 1] + C.cc[k] * C.cc[k]);
  // length scale (tracers.py:263-285); x = q2b, y = q2lb
  col_ld(io, s.q2b, x);
  col_ld(io, s.q2lb, y);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.lraw[k] = C.ghr[k] = C.gh[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) C.lraw[k] = fabs(y[k]) / (fabs(x[k]) + T(1e-20));
  const T kl0 = kappa * C.l0;
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) C.l[k] = (vg.z(m, k) > T(-0.5)) ? (C.lraw[k] >= kl0 ? C.lraw[k] : kl0) : C.lraw[k];
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    C.ghr[k] = (C.l[k] * C.l[k]) * C.boygr[k] / (fabs(x[k]) + T(1e-20));
    C.gh[k] = C.ghr[k] <= T(0.028) ? C.ghr[k] : T(0.028);
  }
  C.l[0] = kappa * C.l0;
  C.l[KB - 1] = T(0);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.dtef[k] = sqrt(fabs(x[k]) + T(1e-8)) * T(1) / (T(B1) * C.l[k] + small);
  // shear production (tracers.py:287-300): du, dv over the (i+1, j) / (i, j+1) neighbours
  T du[KB], dv[KB];
  col_ld(io, s.u, x);
  col_ld_shift(io, s.u, (int64_t)p.jm, y);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) du[k] = x[k] - x[k - 1] + y[k] - y[k - 1];
  col_ld(io, s.v, x);
  col_ld_shift(io, s.v, (int64_t)1, y);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) dv[k] = x[k] - x[k - 1] + y[k] - y[k - 1];
  col_ld(io, s.km, x);
  col_ld(io, s.kh, y);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.kn[k] = T(0);
  if (C.inner) {
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) {
      const T dz2 = vg.dzz(m, k - 1) * dh;
      const T val = x[k] * T(0.25) * T(SEF) * (du[k] * du[k] + dv[k] * dv[k]) / (dz2 * dz2);
      C.kn[k] = val + y[k] * C.boygr[k];
    }
  }
  // q2 tridiagonal solve (tracers.py:302-346)
  col_ld(io, s.q2f, x);
  C.ee1[0] = T(0);
  C.gg1[0] = C.has_s ? T(pow(15.8 * CBCNST, 2.0 / 3.0)) * C.utau2 : T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T den = C.a[k] + C.c[k] * (T(1) - C.ee1[k - 1]) - (T(2) * dti2 * C.dtef[k] + T(1));
    C.ee1[k] = C.a[k] / (den + T(1e-20));
    C.gg1[k] = (T(-2.0 * prm.dti2) * C.kn[k] + C.c[k] * C.gg1[k - 1] - x[k]) / (den + T(1e-20));
  }
  C.ee1[KB - 1] = C.gg1[KB - 1] = T(0);
  C.q2n[KB - 1] = C.q2bot;
  #pragma unroll
  for (int k = KBM1 - 1; k >= 0; --k) C.q2n[k] = C.ee1[k] * C.q2n[k + 1] + C.gg1[k];
  // q2l tridiagonal solve with wall proximity (tracers.py:348-384)
  col_ld(io, s.q2lf, x);
  C.ee2[0] = C.gg2[0] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T W = T(1) / (fabs(vg.z(m, k) - vg.z(m, 0)) + weps) + T(1) / (fabs(vg.z(m, k) - vg.z(m, KB - 1)) + weps);
    const T wall = W * C.l[k] / (dh * kappa);
    const T dtk = C.dtef[k] * (T(1) + T(E2) * (wall * wall));
    const T den = C.a[k] + C.c[k] * (T(1) - C.ee2[k - 1]) - (dti2 * dtk + T(1));
    C.ee2[k] = C.a[k] / (den + T(1e-20));
    C.gg2[k] = (dti2 * (-C.kn[k] * C.l[k] * T(E1)) + C.c[k] * C.gg2[k - 1] - x[k]) / (den + T(1e-20));
  }
  C.ee2[KB - 1] = C.gg2[KB - 1] = T(0);
  C.q2ln[KB - 1] = T(0);
  #pragma unroll
  for (int k = KBM1 - 1; k >= 0; --k) C.q2ln[k] = C.ee2[k] * C.q2ln[k + 1] + C.gg2[k];
}

template <class T>
__global__ void __launch_bounds__(TPB, 2) k_profq(Batch B, Vert<T> vg, Params prm, ProfqArgs<T> s) {
  using namespace myc;
  __shared__ T lds[TILE];
  Pos p;
  const ColIO<T> io = col_pos(B, p, lds);
  const int64_t c2 = ix2(p, p.i, p.j);
  ProfqCol<T> C;
  profq_fwd_col(p, io, vg, prm, s, C);
  const T small = T(prm.small);
  T o1[KB], o2[KB];
  // clipped q2, q2l and the length scale
  #pragma unroll
  for (int k = 0; k < KB; ++k) {
    T q2c = C.q2n[k], q2lc = C.q2ln[k];
    if (k >= 1 && k < KBM1 && (C.q2n[k] <= small || C.q2ln[k] <= small)) {
      q2c = small;
      q2lc = T(0.1) * s.dt[c2] * small;
    }
    o1[k] = q2c;
    o2[k] = q2lc;
  }
  col_st(io, s.q2o, o1);
  col_st(io, s.q2lo, o2);
  col_st(io, s.lo, C.l);
  if (s.smo || s.sho) {   // uniform over the launch
    #pragma unroll
    for (int k = 0; k < KB; ++k) {
      const T sh = T(COEF1) / (T(1) - T(COEF2) * C.gh[k] + T(1e-20));
      o1[k] = sh;
      o2[k] = (T(COEF3) + sh * T(COEF4) * C.gh[k]) / (T(1) - T(COEF5) * C.gh[k] + T(1e-20));
    }
    if (s.sho) col_st(io, s.sho, o1);
    if (s.smo) col_st(io, s.smo, o2);
  }
  if (s.cco) col_st(io, s.cco, C.cc);
  if (s.ao) col_st(io, s.ao, C.a);
  if (s.co) col_st(io, s.co, C.c);
  if (s.dho && io.valid) s.dho[c2] = C.dh;
  if (s.tape.ee1) {   // uniform over the launch
    col_st(io, s.tape.ee1, C.ee1);
    col_st(io, s.tape.gg1, C.gg1);
    col_st(io, s.tape.q2n, C.q2n);
    col_st(io, s.tape.ee2, C.ee2);
    col_st(io, s.tape.gg2, C.gg2);
    col_st(io, s.tape.q2ln, C.q2ln);
    col_st(io, s.tape.ghr, C.ghr);
    col_st(io, s.tape.kn, C.kn);
  }
  // stability functions and the time-averaged mixing coefficients (tracers.py:405-423)
  T knf[KB];
  col_ld(io, s.q2, knf);
  #pragma unroll
  for (int k = 0; k < KB; ++k) knf[k] = C.l[k] * sqrt(fabs(knf[k]) + T(1e-8));
  col_ld(io, s.kq, o1);
  #pragma unroll
  for (int k = 0; k < KB; ++k) {
    const T sh = T(COEF1) / (T(1) - T(COEF2) * C.gh[k] + T(1e-20));
    o1[k] = (knf[k] * T(0.41) * sh + o1[k]) * T(0.5);
  }
  col_st(io, s.kqo, o1);
  col_ld(io, s.km, o1);
  #pragma unroll
  for (int k = 0; k < KB; ++k) {
    const T sh = T(COEF1) / (T(1) - T(COEF2) * C.gh[k] + T(1e-20));
    const T sm = (T(COEF3) + sh * T(COEF4) * C.gh[k]) / (T(1) - T(COEF5) * C.gh[k] + T(1e-20));
    o1[k] = (knf[k] * sm + o1[k]) * T(0.5);
  }
  col_st(io, s.kmo, o1);
  col_ld(io, s.kh, o1);
  #pragma unroll
  for (int k = 0; k < KB; ++k) {
    const T sh = T(COEF1) / (T(1) - T(COEF2) * C.gh[k] + T(1e-20));
    o1[k] = (knf[k] * sh + o1[k]) * T(0.5);
  }
  col_st(io, s.kho, o1);
}

template <class T> struct ProfqAdjArgs {
  ProfqArgs<T> fw;
  const T *b_q2o, *b_q2lo, *b_lo, *b_kqo, *b_kmo, *b_kho;   // output cotangents (any may be null)
  // accumulated input cotangents (own column)
  T *b_q2f, *b_q2lf, *b_q2, *b_q2b, *b_q2lb, *b_km, *b_kh, *b_kq, *b_u, *b_v, *b_t, *b_s, *b_rho;
  T *b_wubot, *b_wvbot, *b_etf, *b_dt;
  // scratch for cotangents of the (i+1, j) / (i, j+1) neighbours
  T *s_unb, *s_vnb, *s_wubnb, *s_wvbnb;
  // scratch between the reverse kernels: cotangents of l, gh, q2l, a, c, kn,
  // dtef (3-D) and of the bottom value q2bot and of dh (2-D)
  T *r_bl, *r_bgh, *r_bq2ln, *r_ba, *r_bc, *r_bkn, *r_bdtef, *r_bboy, *r_bq2bot, *r_bdh;
  int assign_q2f;   // 1: b_q2f and b_q2lf are assigned (their first writer), 0: accumulated
};


// own-column load of an optional array (null: zeros); the null test is uniform
template <class T> __device__ __forceinline__ void ldz_col(const T* f, int64_t c, T* x) {
  if (f) {
    ld_col(f + c, x);
  } else {
    #pragma unroll
    for (int k = 0; k < KB; ++k) x[k] = T(0);
  }
}


// ---- reverse sweep, in three column kernels --------------------------------
// The forward (recomputed by the adjoint pass with the tape) is not rebuilt
// here beyond what is cheap (a, c from kq; dtef from q2b and l; the sound
// speed and buoyancy from t, s, rho).  Each kernel holds only its own phase's
// columns, so none needs more registers than the hardware has:
//   a: stability functions, clipping, q2 solve reversed
//   b: q2l solve reversed (its cotangents of a, c, kn, dtef, l go to their
//      own scratch columns; c adds them)
//   c1: dtef, shear production, length scale (its buoyancy cotangent and
//       its share of dh's go to scratch)
//   c2: buoyancy, sound speed, bottom value, diffusion coefficients
// Columns are read and written by their own thread (16 contiguous values;
// the loads of a wavefront cover whole lines between them), with no block
// barriers, so the kernels run at the occupancy their registers allow.

template <class T> __device__ __forceinline__ void ac_coef(const Vert<T>& vg, int m, const Params& prm, T dh,
                                                           const T* kq, T* a, T* c) {
  const T dti2 = T(prm.dti2);
  #pragma unroll
  for (int k = 0; k < KB; ++k) a[k] = c[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    a[k] = -dti2 * (kq[k + 1] + kq[k] + T(2.0 * prm.umol)) * T(0.5) / (vg.dzz(m, k - 1) * vg.dz(m, k) * dh * dh);
    c[k] = -dti2 * (kq[k - 1] + kq[k] + T(2.0 * prm.umol)) * T(0.5) / (vg.dzz(m, k - 1) * vg.dz(m, k - 1) * dh * dh);
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_profq_adj_a(Batch B, Vert<T> vg, Params prm, ProfqAdjArgs<T> S) {
  using namespace myc;
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c3 = ix3(p, p.i, p.j, 0);
  const ProfqArgs<T>& s = S.fw;
  const int m = p.m;
  const int64_t c2 = ix2(p, p.i, p.j);
  const T dti2 = T(prm.dti2), small = T(prm.small);
  const T dh = s.h[c2] + s.etf[c2];
  // ---- mixing coefficients (tracers.py:405-423), reversed
  T gh[KB], l[KB], q2[KB], bl[KB], bgh[KB], bq2[KB], gq[KB], gm[KB], gk[KB];
  ld_col(s.tape.ghr + c3, gh);
  #pragma unroll
  for (int k = 0; k < KB; ++k) gh[k] = gh[k] <= T(0.028) ? gh[k] : T(0.028);
  ld_col(s.lo + c3, l);
  ld_col(s.q2 + c3, q2);
  ldz_col(S.b_lo, c3, bl);
  ldz_col(S.b_kqo, c3, gq);
  ldz_col(S.b_kmo, c3, gm);
  ldz_col(S.b_kho, c3, gk);
  #pragma unroll
  for (int k = 0; k < KB; ++k) {
    const T g = gh[k];
    const T Dh = T(1) - T(COEF2) * g + T(1e-20);
    const T sh = T(COEF1) / Dh;
    const T N = T(COEF3) + sh * T(COEF4) * g;
    const T Dm = T(1) - T(COEF5) * g + T(1e-20);
    const T sm = N / Dm;
    const T sq2 = sqrt(fabs(q2[k]) + T(1e-8));
    const T knf = l[k] * sq2;
    gq[k] *= T(0.5);
    gm[k] *= T(0.5);
    gk[k] *= T(0.5);
    const T bknf = gq[k] * T(0.41) * sh + gm[k] * sm + gk[k] * sh;
    T bsh = gq[k] * knf * T(0.41) + gk[k] * knf;
    const T bsm = gm[k] * knf;
    bl[k] += bknf * sq2;
    bq2[k] = bknf * l[k] * sgn(q2[k]) / (T(2) * sq2);
    const T bN = bsm / Dm, bDm = -bsm * sm / Dm;
    bsh += bN * T(COEF4) * g;
    bgh[k] = bN * sh * T(COEF4) - bDm * T(COEF5);
    const T bDh = -bsh * sh / Dh;
    bgh[k] += -bDh * T(COEF2);
  }
  if (S.b_kq) add_col(S.b_kq + c3, gq);
  if (S.b_km) add_col(S.b_km + c3, gm);
  if (S.b_kh) add_col(S.b_kh + c3, gk);
  if (S.b_q2) add_col(S.b_q2 + c3, bq2);
  st_col(S.r_bl + c3, bl);
  st_col(S.r_bgh + c3, bgh);
  // ---- clipping of q2, q2l (tracers.py:386-403), reversed
  T q2n[KB], q2ln[KB], bq2n[KB], bq2ln[KB];
  ld_col(s.tape.q2n + c3, q2n);
  ld_col(s.tape.q2ln + c3, q2ln);
  ldz_col(S.b_q2o, c3, bq2n);
  ldz_col(S.b_q2lo, c3, bq2ln);
  T bdt2 = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    if (q2n[k] <= small || q2ln[k] <= small) {
      bdt2 += bq2ln[k] * T(0.1) * small;
      bq2n[k] = T(0);
      bq2ln[k] = T(0);
    }
  }
  st_col(S.r_bq2ln + c3, bq2ln);
  if (S.b_dt) S.b_dt[c2] += bdt2;
  // ---- q2 tridiagonal solve (tracers.py:302-346), reversed
  T ee[KB], gg[KB], cc_[KB], den[KB];
  {
    T a[KB], x[KB];
    ld_col(s.kq + c3, x);
    ac_coef(vg, m, prm, dh, x, a, cc_);
    ld_col(s.q2b + c3, x);
    ld_col(s.tape.ee1 + c3, ee);
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) {
      const T dtef = sqrt(fabs(x[k]) + T(1e-8)) * T(1) / (T(B1) * l[k] + small);
      den[k] = a[k] + cc_[k] * (T(1) - ee[k - 1]) - (T(2) * dti2 * dtef + T(1));
    }
  }
  ld_col(s.tape.gg1 + c3, gg);
  T bee[KB], bgg[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) bee[k] = bgg[k] = T(0);
  #pragma unroll
  for (int k = 0; k <= KBM1 - 1; ++k) {
    bee[k] += bq2n[k] * q2n[k + 1];
    bq2n[k + 1] += bq2n[k] * ee[k];
    bgg[k] += bq2n[k];
  }
  S.r_bq2bot[c2] = bq2n[KB - 1];
  T ba[KB], bc[KB], bkn[KB], bdtef[KB], bq2f[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) ba[k] = bc[k] = bkn[k] = bdtef[k] = bq2f[k] = T(0);
  #pragma unroll
  for (int k = KBM1 - 1; k >= 1; --k) {
    const T de = den[k] + T(1e-20);
    const T bR = bgg[k] / de;
    const T bde = -(bgg[k] * gg[k] + bee[k] * ee[k]) / de;
    ba[k] += bee[k] / de + bde;
    bkn[k] += bR * T(-2.0 * prm.dti2);
    bc[k] += bR * gg[k - 1] + bde * (T(1) - ee[k - 1]);
    bgg[k - 1] += bR * cc_[k];
    bq2f[k] -= bR;
    bee[k - 1] -= bde * cc_[k];
    bdtef[k] += -bde * T(2) * dti2;
  }
  // gg1(0) = const * utau2: utau2 depends only on the passive surface stress
  if (S.b_q2f) put_col(S.b_q2f + c3, bq2f, S.assign_q2f != 0);
  st_col(S.r_ba + c3, ba);
  st_col(S.r_bc + c3, bc);
  st_col(S.r_bkn + c3, bkn);
  st_col(S.r_bdtef + c3, bdtef);
}

template <class T>
__global__ void __launch_bounds__(TPB) k_profq_adj_b(Batch B, Vert<T> vg, Params prm, ProfqAdjArgs<T> S) {
  using namespace myc;
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c3 = ix3(p, p.i, p.j, 0);
  const ProfqArgs<T>& s = S.fw;
  const int m = p.m;
  const int64_t c2 = ix2(p, p.i, p.j);
  const T dti2 = T(prm.dti2), kappa = T(prm.kappa), small = T(prm.small);
  const T weps = T(prm.core ? 1e-8 : 1e-20);   // wall guard: explicit tracers.py:362, implicit :360
  const T dh = s.h[c2] + s.etf[c2];
  // ---- q2l tridiagonal solve with wall proximity (tracers.py:348-384), reversed
  T ee[KB], bee[KB], bgg[KB];
  {
    T bq2ln[KB], q2ln[KB];
    ld_col(S.r_bq2ln + c3, bq2ln);
    ld_col(s.tape.q2ln + c3, q2ln);
    ld_col(s.tape.ee2 + c3, ee);
    #pragma unroll
    for (int k = 0; k < KB; ++k) bee[k] = bgg[k] = T(0);
    #pragma unroll
    for (int k = 0; k <= KBM1 - 1; ++k) {
      bee[k] += bq2ln[k] * q2ln[k + 1];
      bq2ln[k + 1] += bq2ln[k] * ee[k];
      bgg[k] += bq2ln[k];
    }
  }
  T l[KB], dtef[KB], cc_[KB], den[KB];
  ld_col(s.lo + c3, l);
  {
    T a[KB], x[KB];
    ld_col(s.kq + c3, x);
    ac_coef(vg, m, prm, dh, x, a, cc_);
    ld_col(s.q2b + c3, x);
    #pragma unroll
    for (int k = 0; k < KB; ++k) dtef[k] = sqrt(fabs(x[k]) + T(1e-8)) * T(1) / (T(B1) * l[k] + small);
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) {
      const T W = T(1) / (fabs(vg.z(m, k) - vg.z(m, 0)) + weps) + T(1) / (fabs(vg.z(m, k) - vg.z(m, KB - 1)) + weps);
      const T wall = W * l[k] / (dh * kappa);
      const T dtk = dtef[k] * (T(1) + T(E2) * (wall * wall));
      den[k] = a[k] + cc_[k] * (T(1) - ee[k - 1]) - (dti2 * dtk + T(1));
    }
  }
  T gg[KB], kn[KB];
  ld_col(s.tape.gg2 + c3, gg);
  ld_col(s.tape.kn + c3, kn);
  T ba[KB], bc[KB], bkn[KB], bdtef[KB], bl[KB], bq2lf[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) ba[k] = bc[k] = bkn[k] = bdtef[k] = bl[k] = bq2lf[k] = T(0);
  T bdh = T(0);
  #pragma unroll
  for (int k = KBM1 - 1; k >= 1; --k) {
    const T W = T(1) / (fabs(vg.z(m, k) - vg.z(m, 0)) + weps) + T(1) / (fabs(vg.z(m, k) - vg.z(m, KB - 1)) + weps);
    const T wall = W * l[k] / (dh * kappa);
    const T de = den[k] + T(1e-20);
    const T bR = bgg[k] / de;
    const T bde = -(bgg[k] * gg[k] + bee[k] * ee[k]) / de;
    ba[k] += bee[k] / de + bde;
    bkn[k] += bR * dti2 * (-l[k] * T(E1));
    bl[k] += bR * dti2 * (-kn[k] * T(E1));
    bc[k] += bR * gg[k - 1] + bde * (T(1) - ee[k - 1]);
    bgg[k - 1] += bR * cc_[k];
    bq2lf[k] -= bR;
    bee[k - 1] -= bde * cc_[k];
    const T bdtk = -bde * dti2;
    bdtef[k] += bdtk * (T(1) + T(E2) * (wall * wall));
    const T bwall = bdtk * dtef[k] * T(E2) * T(2) * wall;
    bl[k] += bwall * W / (dh * kappa);
    bdh += -bwall * wall / dh;
  }
  if (S.b_q2lf) put_col(S.b_q2lf + c3, bq2lf, S.assign_q2f != 0);
  add_col(S.r_ba + c3, ba);
  add_col(S.r_bc + c3, bc);
  add_col(S.r_bkn + c3, bkn);
  add_col(S.r_bdtef + c3, bdtef);
  add_col(S.r_bl + c3, bl);
  S.r_bdh[c2] = bdh;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_profq_adj_c1(Batch B, Vert<T> vg, Params prm, ProfqAdjArgs<T> S) {
  using namespace myc;
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c3 = ix3(p, p.i, p.j, 0);
  const ProfqArgs<T>& s = S.fw;
  const int i = p.i, j = p.j, m = p.m;
  const int64_t c2 = ix2(p, i, j);
  const T dti2 = T(prm.dti2), grav = T(prm.grav), kappa = T(prm.kappa), small = T(prm.small);
  const T dh = s.h[c2] + s.etf[c2];
  const bool has_s = (i < p.im - 1) && (j < p.jm - 1);
  const bool inner = i >= 1 && i <= p.im - 2 && j >= 1 && j <= p.jm - 2;
  T bdh = T(0);
  T l[KB], q2b[KB], bl[KB], bq2b[KB];
  ld_col(s.lo + c3, l);
  ld_col(s.q2b + c3, q2b);
  ld_col(S.r_bl + c3, bl);
  // ---- dtef = sqrt(|q2b| + 1e-8) / (B1 l + small), reversed
  {
    T bdtef[KB];
    ld_col(S.r_bdtef + c3, bdtef);
    #pragma unroll
    for (int k = 0; k < KB; ++k) {
      const T qb = q2b[k];
      const T sqb = sqrt(fabs(qb) + T(1e-8));
      const T dd = T(B1) * l[k] + small;
      const T dtef = sqb * T(1) / dd;
      bq2b[k] = bdtef[k] / dd * sgn(qb) / (T(2) * sqb);
      bl[k] += -bdtef[k] * dtef * T(B1) / dd;
    }
  }
  // ---- sound speed (tracers.py:245-255) and buoyancy frequency
  T cc[KB], boygr[KB], rho[KB];
  {
    T t[KB], sl[KB];
    ld_col(s.t + c3, t);
    ld_col(s.s + c3, sl);
    #pragma unroll
    for (int k = 0; k < KB; ++k) cc[k] = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) {
      const T tp = t[k] + T(prm.tbias), sp = sl[k] + T(prm.sbias);
      const T pp = T(prm.grav * prm.rhoref) * (-vg.zz(m, k) * dh) * T(1.e-4);
      T cc0 = T(1449.1) + T(0.00821) * pp + T(4.55) * tp - T(0.045) * (tp * tp) + T(1.34) * (sp - T(35.0));
      const T ccs = fabs(cc0) < T(1e-6) ? T(1e-6) : cc0;
      const T rad = (T(1) - T(0.01642) * pp / ccs) * (T(1) - T(0.40) * pp / (ccs * ccs));
      cc[k] = ccs / sqrt(fmax(rad, T(1e-8)));
    }
  }
  ld_col(s.rho + c3, rho);
  #pragma unroll
  for (int k = 0; k < KB; ++k) boygr[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k)
    boygr[k] = grav * (rho[k - 1] - rho[k]) / (vg.dzz(m, k - 1) * dh) +
               T(prm.grav * prm.grav * 2.0) / (cc[k - 1] * cc[k - 1] + cc[k] * cc[k]);
  // ---- shear production (tracers.py:287-300), reversed
  T bboy[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) bboy[k] = T(0);
  {
    T bkn[KB], km[KB], bkm[KB], bkh[KB];
    ld_col(S.r_bkn + c3, bkn);
    ld_col(s.km + c3, km);
    #pragma unroll
    for (int k = 0; k < KB; ++k) bkm[k] = bkh[k] = T(0);
    T x[KB], y[KB];   // u (own, i+1) then v (own, j+1)
    T du[KB], dv[KB];
    ld_col(s.u + c3, x);
    ld_col(s.u + ix3(p, p.i + 1 < p.im ? p.i + 1 : p.i, p.j, 0), y);
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) du[k] = x[k] - x[k - 1] + y[k] - y[k - 1];
    ld_col(s.v + c3, x);
    ld_col(s.v + ix3(p, p.i, p.j + 1 < p.jm ? p.j + 1 : p.j, 0), y);
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) dv[k] = x[k] - x[k - 1] + y[k] - y[k - 1];
    ld_col(s.kh + c3, y);   // kh
    T bu[KB], bun[KB], bv[KB], bvn[KB];
    #pragma unroll
    for (int k = 0; k < KB; ++k) bu[k] = bun[k] = bv[k] = bvn[k] = T(0);
    if (inner) {
      #pragma unroll
      for (int k = 1; k < KBM1; ++k) {
        const T dz2 = vg.dzz(m, k - 1) * dh;
        const T D2 = dz2 * dz2;
        const T S2 = du[k] * du[k] + dv[k] * dv[k];
        const T val = km[k] * T(0.25) * T(SEF) * S2 / D2;
        const T g = bkn[k];
        bkh[k] += g * boygr[k];
        bboy[k] += g * y[k];
        bkm[k] += g * T(0.25) * T(SEF) * S2 / D2;
        const T bS = g * km[k] * T(0.25) * T(SEF) / D2;
        bdh += -g * val / D2 * T(2) * dz2 * vg.dzz(m, k - 1);
        const T bdu = bS * T(2) * du[k], bdv = bS * T(2) * dv[k];
        bu[k] += bdu;
        bu[k - 1] -= bdu;
        bun[k] += bdu;
        bun[k - 1] -= bdu;
        bv[k] += bdv;
        bv[k - 1] -= bdv;
        bvn[k] += bdv;
        bvn[k - 1] -= bdv;
      }
    }
    if (S.b_km) add_col(S.b_km + c3, bkm);
    if (S.b_kh) add_col(S.b_kh + c3, bkh);
    if (S.b_u) add_col(S.b_u + c3, bu);
    if (S.b_v) add_col(S.b_v + c3, bv);
    st_col(S.s_unb + c3, bun);
    st_col(S.s_vnb + c3, bvn);
  }
  // ---- stability-function argument gh and length scale, reversed
  T bq2lb[KB];
  {
    T bgh[KB], ghr[KB], q2lb[KB];
    ld_col(S.r_bgh + c3, bgh);
    ld_col(s.tape.ghr + c3, ghr);
    ld_col(s.q2lb + c3, q2lb);
    #pragma unroll
    for (int k = 0; k < KB; ++k) bq2lb[k] = T(0);
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) {
      if (ghr[k] <= T(0.028)) {
        const T qb = q2b[k];
        const T Dq = fabs(qb) + T(1e-20);
        const T g = bgh[k];
        bl[k] += g * T(2) * l[k] * boygr[k] / Dq;
        bboy[k] += g * l[k] * l[k] / Dq;
        bq2b[k] += -g * ghr[k] / Dq * sgn(qb);
      }
    }
    // l = max(lraw, kappa l0) near the surface; l(0) = kappa l0 (l0 depends only
    // on the passive surface stress)
    T l0 = T(0);
    if (has_s) {
      const T ust = T(0.5) * (s.wusurf[c2] + s.wusurf[ix2(p, i + 1, j)]);
      const T vst = T(0.5) * (s.wvsurf[c2] + s.wvsurf[ix2(p, i, j + 1)]);
      l0 = T(SURFL) * sqrt(ust * ust + vst * vst + T(1e-8)) / grav;
    }
    const T kl0 = kappa * l0;
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) {
      const T qb = q2b[k], qlb = q2lb[k];
      const T Dq = fabs(qb) + T(1e-20);
      const T lraw = fabs(qlb) / Dq;
      T blr = bl[k];
      if (vg.z(m, k) > T(-0.5) && !(lraw >= kl0)) blr = T(0);
      bq2lb[k] += blr * sgn(qlb) / Dq;
      bq2b[k] += -blr * lraw / Dq * sgn(qb);
    }
  }
  if (S.b_q2b) add_col(S.b_q2b + c3, bq2b);
  if (S.b_q2lb) add_col(S.b_q2lb + c3, bq2lb);
  st_col(S.r_bboy + c3, bboy);
  S.r_bdh[c2] += bdh;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_profq_adj_c2(Batch B, Vert<T> vg, Params prm, ProfqAdjArgs<T> S) {
  using namespace myc;
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c3 = ix3(p, p.i, p.j, 0);
  const ProfqArgs<T>& s = S.fw;
  const int i = p.i, j = p.j, m = p.m;
  const int64_t c2 = ix2(p, i, j);
  const T dti2 = T(prm.dti2), grav = T(prm.grav);
  const T dh = s.h[c2] + s.etf[c2];
  const bool has_s = (i < p.im - 1) && (j < p.jm - 1);
  T bdh = S.r_bdh[c2];
  // sound speed (tracers.py:245-255)
  T cc[KB], rho[KB], bboy[KB];
  {
    T t[KB], sl[KB];
    ld_col(s.t + c3, t);
    ld_col(s.s + c3, sl);
    #pragma unroll
    for (int k = 0; k < KB; ++k) cc[k] = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) {
      const T tp = t[k] + T(prm.tbias), sp = sl[k] + T(prm.sbias);
      const T pp = T(prm.grav * prm.rhoref) * (-vg.zz(m, k) * dh) * T(1.e-4);
      T cc0 = T(1449.1) + T(0.00821) * pp + T(4.55) * tp - T(0.045) * (tp * tp) + T(1.34) * (sp - T(35.0));
      const T ccs = fabs(cc0) < T(1e-6) ? T(1e-6) : cc0;
      const T rad = (T(1) - T(0.01642) * pp / ccs) * (T(1) - T(0.40) * pp / (ccs * ccs));
      cc[k] = ccs / sqrt(fmax(rad, T(1e-8)));
    }
  }
  ld_col(s.rho + c3, rho);
  ld_col(S.r_bboy + c3, bboy);
  // ---- buoyancy, reversed
  T bcc[KB], brho[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) bcc[k] = brho[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T g = bboy[k];
    const T dd = vg.dzz(m, k - 1) * dh;
    const T b1t = grav * (rho[k - 1] - rho[k]) / dd;
    brho[k - 1] += g * grav / dd;
    brho[k] -= g * grav / dd;
    bdh += -g * b1t / dh;
    const T Sc = cc[k - 1] * cc[k - 1] + cc[k] * cc[k];
    const T b2t = T(prm.grav * prm.grav * 2.0) / Sc;
    const T bSc = -g * b2t / Sc;
    bcc[k - 1] += bSc * T(2) * cc[k - 1];
    bcc[k] += bSc * T(2) * cc[k];
  }
  if (S.b_rho) add_col(S.b_rho + c3, brho);
  // ---- sound speed, reversed
  {
    T t[KB], sl[KB], bt[KB], bs[KB];
    ld_col(s.t + c3, t);
    ld_col(s.s + c3, sl);
    #pragma unroll
    for (int k = 0; k < KB; ++k) bt[k] = bs[k] = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) {
      const T tp = t[k] + T(prm.tbias), sp = sl[k] + T(prm.sbias);
      const T cz = T(prm.grav * prm.rhoref) * (-vg.zz(m, k));
      const T pp = cz * dh * T(1.e-4);
      const T cc0 = T(1449.1) + T(0.00821) * pp + T(4.55) * tp - T(0.045) * (tp * tp) + T(1.34) * (sp - T(35.0));
      const bool rep = fabs(cc0) < T(1e-6);
      const T ccs = rep ? T(1e-6) : cc0;
      const T X = T(1) - T(0.01642) * pp / ccs;
      const T Y = T(1) - T(0.40) * pp / (ccs * ccs);
      const T rad = X * Y;
      const T sq = sqrt(fmax(rad, T(1e-8)));
      const T g = bcc[k];
      T bccs = g / sq;
      const T bsq = -g * cc[k] / sq;
      const T brad = rad >= T(1e-8) ? bsq / (T(2) * sq) : T(0);
      const T bX = brad * Y, bY = brad * X;
      T bp = -bX * T(0.01642) / ccs - bY * T(0.40) / (ccs * ccs);
      bccs += bX * T(0.01642) * pp / (ccs * ccs) + bY * T(0.40) * pp * T(2) / (ccs * ccs * ccs);
      if (!rep) {
        bp += bccs * T(0.00821);
        bt[k] += bccs * (T(4.55) - T(0.045) * T(2) * tp);
        bs[k] += bccs * T(1.34);
      }
      bdh += bp * cz * T(1.e-4);
    }
    if (S.b_t) add_col(S.b_t + c3, bt);
    if (S.b_s) add_col(S.b_s + c3, bs);
  }
  // ---- bottom boundary value q2bot = const * sqrt(ubt^2 + vbt^2 + 1e-8), reversed
  T bwub = T(0), bwvb = T(0);
  if (has_s) {
    const T ubt = T(0.5) * (s.wubot[c2] + s.wubot[ix2(p, i + 1, j)]);
    const T vbt = T(0.5) * (s.wvbot[c2] + s.wvbot[ix2(p, i, j + 1)]);
    const T qbm = sqrt(ubt * ubt + vbt * vbt + T(1e-8));
    const T g = S.r_bq2bot[c2] * T(pow(16.6, 2.0 / 3.0) * SEF) / qbm;
    bwub = T(0.5) * g * ubt;
    bwvb = T(0.5) * g * vbt;
  }
  // ---- diffusion coefficients a, c from kq and dh, reversed
  {
    T ba[KB], bc[KB], kq[KB], a[KB], c[KB], bkq[KB];
    ld_col(S.r_ba + c3, ba);
    ld_col(S.r_bc + c3, bc);
    ld_col(s.kq + c3, kq);
    ac_coef(vg, m, prm, dh, kq, a, c);
    #pragma unroll
    for (int k = 0; k < KB; ++k) bkq[k] = T(0);
    #pragma unroll
    for (int k = 1; k < KBM1; ++k) {
      const T Da = vg.dzz(m, k - 1) * vg.dz(m, k) * dh * dh;
      const T Dc = vg.dzz(m, k - 1) * vg.dz(m, k - 1) * dh * dh;
      const T ga = ba[k] * (-dti2) * T(0.5) / Da;
      const T gc = bc[k] * (-dti2) * T(0.5) / Dc;
      bkq[k + 1] += ga;
      bkq[k] += ga + gc;
      bkq[k - 1] += gc;
      bdh += -ba[k] * a[k] * T(2) / dh - bc[k] * c[k] * T(2) / dh;
    }
    if (S.b_kq) add_col(S.b_kq + c3, bkq);
  }
  S.s_wubnb[c2] = bwub;
  S.s_wvbnb[c2] = bwvb;
  if (S.b_wubot) S.b_wubot[c2] += bwub;
  if (S.b_wvbot) S.b_wvbot[c2] += bwvb;
  if (S.b_etf) S.b_etf[c2] += bdh;
}

// gather the neighbour cotangents: point (i, j, k) receives what column
// (i-1, j) wrote for u / wubot and what column (i, j-1) wrote for v / wvbot
// (3-D launch, one thread per point: unit-stride on every array)
template <class T>
__global__ void __launch_bounds__(TPB) k_profq_adj2(Batch B, ProfqAdjArgs<T> S) {
  Pos p;
  if (!pos3(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, p.k);
  if (p.i >= 1) {
    if (S.b_u) S.b_u[c] += S.s_unb[c - (int64_t)p.jm * KB];
    if (S.b_wubot && p.k == 0) S.b_wubot[ix2(p, p.i, p.j)] += S.s_wubnb[ix2(p, p.i - 1, p.j)];
  }
  if (p.j >= 1) {
    if (S.b_v) S.b_v[c] += S.s_vnb[c - KB];
    if (S.b_wvbot && p.k == 0) S.b_wvbot[ix2(p, p.i, p.j)] += S.s_wvbnb[ix2(p, p.i, p.j - 1)];
  }
}

}  // namespace pom
