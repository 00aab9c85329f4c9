// The corrector hook between physics intervals, fused: the network's raw
// increment (60 planes: u, v, t, s on the 15 active levels) is scaled, the
// velocity increments lose their dz-weighted column mean and are masked, the
// tracer increments lose their volume-weighted layer mean (wet cells only),
// each pair is capped by one shared factor, and both time levels of the four
// fields receive the increment; the ghost level 15 is copied.
// reference: src/neuralpom/hybrid/sparse_coupling.py:18-67 (volume_weights,
// budget_project, paired_limit, apply_joint_correction).  The density refresh
// that follows (:61) is the EOS entry point, called by the host after this.
// The same file holds the density-75 projection that follows the correction
// in the DND policy (src/neuralpom/evaluation/density.py:14-64).
//
// Per member, five launches: column pass (velocity column means, layer sums
// of T w, S w, w, velocity peak) -> member reduction -> column pass (tracer
// projections, tracer peak, budget sums) -> member reduction -> column pass
// applying the increments.  The adjoint is the same shape in reverse.  Block
// partials are reduced in a fixed order, in double: results are
// deterministic.  All passes are one thread per water column; the raw planes
// are read coalesced (consecutive threads, consecutive columns).
#include "../../include/neuralpom_amd.h"
#include "pom_batch.h"
#include "pom_ops.h"

using namespace pom;

namespace {

constexpr int LEV = 15;           // corrected levels (the reference's LEVELS)
constexpr int NCH = 4 * LEV;      // raw channels
// per-member statistics (double)
enum {
  MS_MT = 0,            // layer means of T (15)
  MS_MS = LEV,          // layer means of S (15)
  MS_W = 2 * LEV,       // layer weights sum_ij w (15)
  MS_WTOT = 3 * LEV,    // sum_ijk w
  MS_PUV, MS_FUV,       // velocity peak and factor
  MS_PTS, MS_FTS,       // tracer peak and factor
  MS_BT, MS_BS,         // sum_ijk w T', sum_ijk w S'
  MS_GUV, MS_GTS,       // adjoint: cotangent per tied peak element
  MS_DB1, MS_DB2,       // adjoint: cotangents of the two budgets
  MS_ST = MS_DB2 + 1,   // adjoint: sum_ij T'bar wet per level (15)
  MS_SS = MS_ST + LEV,  // adjoint: sum_ij S'bar wet per level (15)
  MS_N = MS_SS + LEV
};
constexpr int NP = 48;  // block partials per logical block

template <class T> struct CorrArgs {
  const T* raw;          // 60 planes per member at 60 * off2[m]
  const T *dum, *dvm, *fsm, *art, *h, *et;
  const T* vg;           // [M][4][16]: dz is row 2
  double su, sv, st, ss, cap;
  double* part;          // NP per logical 2-D block
  double* ms;            // MS_N per member
};

template <class T> __device__ __forceinline__ T dzk(const CorrArgs<T>& a, int m, int k) {
  return a.vg[(m * 4 + 2) * KB + k];
}

// raw channel ch of column `cell` of the member at p
template <class T> __device__ __forceinline__ T rawc(const CorrArgs<T>& a, const Pos& p, int ch) {
  const int64_t n2 = (int64_t)p.im * p.jm;
  return a.raw[(int64_t)NCH * p.o2 + (int64_t)ch * n2 + ((int64_t)p.i * p.jm + p.j)];
}

// velocity increments of a column: scale, minus the dz-weighted column mean, times the mask
template <class T> __device__ __forceinline__ void vel_inc(const CorrArgs<T>& a, const Pos& p, int base, double sc,
                                                          T mask, T* x) {
  T num = T(0), den = T(0);
#pragma unroll
  for (int k = 0; k < LEV; ++k) {
    x[k] = rawc(a, p, base + k) * T(sc);
    num += x[k] * dzk(a, p.m, k);
    den += dzk(a, p.m, k);
  }
  const T mean = num / den;
#pragma unroll
  for (int k = 0; k < LEV; ++k) x[k] = (x[k] - mean) * mask;
}

// block reduction of NV per-thread doubles (sums, or maxima when MAX) into
// part[lb * NP + off ...]; fixed order, so deterministic
template <int NV, bool MAX, int STRIDE = NP>
__device__ __forceinline__ void block_reduce(double* v, double* part, int off) {
  __shared__ double red[TPB / 64][NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o);
      x = MAX ? fmax(x, y) : x + y;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    double x = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < TPB / 64; ++w) x = MAX ? fmax(x, red[w][threadIdx.x]) : x + red[w][threadIdx.x];
    part[(int64_t)logical_block() * STRIDE + off + threadIdx.x] = x;
  }
  __syncthreads();
}

// per-member reduction of block partials [off, off + nv) (one block per member)
__device__ __forceinline__ void member_reduce(const Batch& B, const double* part, int off, int nv, bool mx,
                                              double* out, int stride = NP) {
  const int m = blockIdx.x;
  const int b0 = B.blk2[m], b1 = B.blk2[m + 1];
  for (int q = threadIdx.x; q < nv; q += TPB) {
    double x = 0.0;   // sums from zero; maxima of absolute values from zero
    for (int b = b0; b < b1; ++b) {
      const double y = part[(int64_t)b * stride + off + q];
      x = mx ? fmax(x, y) : x + y;
    }
    out[q] = x;
  }
}

// tracer increments of a column after the layer projection (needs MS_MT / MS_MS)
template <class T> __device__ __forceinline__ void tracer_inc(const CorrArgs<T>& a, const Pos& p, T* tt, T* ss, T* w,
                                                             bool* wet) {
  const int64_t c2 = ix2(p, p.i, p.j);
  const T col = a.art[c2] * a.fsm[c2] * (a.h[c2] + a.et[c2]);
  const double* ms = a.ms + (int64_t)p.m * MS_N;
#pragma unroll
  for (int k = 0; k < LEV; ++k) {
    w[k] = col * dzk(a, p.m, k);
    wet[k] = w[k] > T(0);
    tt[k] = wet[k] ? rawc(a, p, 2 * LEV + k) * T(a.st) - T(ms[MS_MT + k]) : T(0);
    ss[k] = wet[k] ? rawc(a, p, 3 * LEV + k) * T(a.ss) - T(ms[MS_MS + k]) : T(0);
  }
}

// ---------------------------------------------------------------- forward
template <class T> __global__ void __launch_bounds__(TPB) k_corr_pass1(Batch B, CorrArgs<T> a) {
  Pos p;
  const bool live = pos2(B, p);
  double v[3 * LEV + 2];
#pragma unroll
  for (int q = 0; q < 3 * LEV + 2; ++q) v[q] = 0.0;
  if (live) {
    const int64_t c2 = ix2(p, p.i, p.j);
    T x[LEV];
    vel_inc(a, p, 0, a.su, a.dum[c2], x);
    double pk = 0.0;
#pragma unroll
    for (int k = 0; k < LEV; ++k) pk = fmax(pk, fabs(double(x[k] / T(a.su))));
    vel_inc(a, p, LEV, a.sv, a.dvm[c2], x);
#pragma unroll
    for (int k = 0; k < LEV; ++k) pk = fmax(pk, fabs(double(x[k] / T(a.sv))));
    v[3 * LEV] = pk;
    const T col = a.art[c2] * a.fsm[c2] * (a.h[c2] + a.et[c2]);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      const T w = col * dzk(a, p.m, k);
      v[k] = double(rawc(a, p, 2 * LEV + k) * T(a.st) * w);
      v[LEV + k] = double(rawc(a, p, 3 * LEV + k) * T(a.ss) * w);
      v[2 * LEV + k] = double(w);
    }
  }
  block_reduce<3 * LEV, false>(v, a.part, 0);
  block_reduce<1, true>(v + 3 * LEV, a.part, 3 * LEV);
}

__device__ __forceinline__ double pair_factor(double peak, double cap) {   // clamp(cap / clamp(peak, min=cap), max=1)
  const double f = cap / fmax(peak, cap);
  return f < 1.0 ? f : 1.0;
}

template <class T> __global__ void __launch_bounds__(TPB) k_corr_member1(Batch B, CorrArgs<T> a) {
  double* ms = a.ms + (int64_t)blockIdx.x * MS_N;
  __shared__ double s[3 * LEV + 1];
  member_reduce(B, a.part, 0, 3 * LEV, false, s);
  member_reduce(B, a.part, 3 * LEV, 1, true, s + 3 * LEV);
  __syncthreads();
  if (threadIdx.x == 0) {
    double wtot = 0.0;
    for (int k = 0; k < LEV; ++k) {
      const double wk = s[2 * LEV + k];
      ms[MS_MT + k] = s[k] / fmax(wk, 1e-20);
      ms[MS_MS + k] = s[LEV + k] / fmax(wk, 1e-20);
      ms[MS_W + k] = wk;
      wtot += wk;
    }
    ms[MS_WTOT] = wtot;
    ms[MS_PUV] = s[3 * LEV];
    ms[MS_FUV] = pair_factor(s[3 * LEV], a.cap);
  }
}

template <class T> __global__ void __launch_bounds__(TPB) k_corr_pass2(Batch B, CorrArgs<T> a) {
  Pos p;
  const bool live = pos2(B, p);
  double v[3] = {0.0, 0.0, 0.0};
  if (live) {
    T tt[LEV], ss[LEV], w[LEV];
    bool wet[LEV];
    tracer_inc(a, p, tt, ss, w, wet);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      v[0] = fmax(v[0], fmax(fabs(double(tt[k] / T(a.st))), fabs(double(ss[k] / T(a.ss)))));
      v[1] += double(w[k] * tt[k]);
      v[2] += double(w[k] * ss[k]);
    }
  }
  block_reduce<1, true>(v, a.part, 0);
  block_reduce<2, false>(v + 1, a.part, 1);
}

template <class T>
__global__ void __launch_bounds__(TPB) k_corr_member2(Batch B, CorrArgs<T> a, T* __restrict__ diag) {
  double* ms = a.ms + (int64_t)blockIdx.x * MS_N;
  __shared__ double s[3];
  member_reduce(B, a.part, 0, 1, true, s);
  member_reduce(B, a.part, 1, 2, false, s + 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double f = pair_factor(s[0], a.cap);
    ms[MS_PTS] = s[0];
    ms[MS_FTS] = f;
    ms[MS_BT] = s[1];
    ms[MS_BS] = s[2];
    if (diag) {
      diag[3 * blockIdx.x + 0] = T(f);
      diag[3 * blockIdx.x + 1] = T(f * s[1] / ms[MS_WTOT]);
      diag[3 * blockIdx.x + 2] = T(f * s[2] / ms[MS_WTOT]);
    }
  }
}

template <class T> struct ApplyArgs {
  const T* in[8];   // u ub v vb t tb s sb
  T* out[8];
};

template <class T> __global__ void __launch_bounds__(TPB) k_corr_apply(Batch B, CorrArgs<T> a, ApplyArgs<T> f) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c2 = ix2(p, p.i, p.j), c = ix3(p, p.i, p.j, 0);
  const double* ms = a.ms + (int64_t)p.m * MS_N;
  const T fuv = T(ms[MS_FUV]), fts = T(ms[MS_FTS]);
  T inc[4][LEV];
  vel_inc(a, p, 0, a.su, a.dum[c2], inc[0]);
  vel_inc(a, p, LEV, a.sv, a.dvm[c2], inc[1]);
  {
    T w[LEV];
    bool wet[LEV];
    tracer_inc(a, p, inc[2], inc[3], w, wet);
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const T fac = g < 2 ? fuv : fts;
#pragma unroll
    for (int lvl = 0; lvl < 2; ++lvl) {
      const T* src = f.in[2 * g + lvl] + c;
      T* dst = f.out[2 * g + lvl] + c;
      T y[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) y[k] = src[k];
#pragma unroll
      for (int k = 0; k < KB; ++k) dst[k] = k < LEV ? y[k] + fac * inc[g][k] : y[k];
    }
  }
}

// ---------------------------------------------------------------- adjoint
template <class T> struct AdjArgs {
  const T* cot[8];     // cotangents of the 8 outputs (may be null)
  const T* cdiag;      // [M][3] cotangents of (factor, t budget, s budget), or null
  T* graw;             // written: 60 planes per member
  T* get;              // accumulated cotangent of et, or null
};

// G = cotangent of the increment of field g: the sum over its two time levels
template <class T> __device__ __forceinline__ void inc_cot(const AdjArgs<T>& d, int g, int64_t c, T* G) {
#pragma unroll
  for (int k = 0; k < LEV; ++k) G[k] = T(0);
#pragma unroll
  for (int lvl = 0; lvl < 2; ++lvl) {
    const T* x = d.cot[2 * g + lvl];
    if (x) {
#pragma unroll
      for (int k = 0; k < LEV; ++k) G[k] += x[c + k];
    }
  }
}

// dot products of increments with their cotangents, and the elements at each pair's peak
template <class T> __global__ void __launch_bounds__(TPB) k_corr_adj1(Batch B, CorrArgs<T> a, AdjArgs<T> d) {
  Pos p;
  const bool live = pos2(B, p);
  double v[4] = {0.0, 0.0, 0.0, 0.0};   // sum U'G_U + V'G_V, sum T'G_T + S'G_S, ties uv, ties ts
  if (live) {
    const int64_t c2 = ix2(p, p.i, p.j), c = ix3(p, p.i, p.j, 0);
    const double* ms = a.ms + (int64_t)p.m * MS_N;
    T x[LEV], G[LEV];
    const double sc[4] = {a.su, a.sv, a.st, a.ss};
    for (int g = 0; g < 2; ++g) {
      vel_inc(a, p, g * LEV, sc[g], g ? a.dvm[c2] : a.dum[c2], x);
      inc_cot(d, g, c, G);
#pragma unroll
      for (int k = 0; k < LEV; ++k) {
        v[0] += double(x[k] * G[k]);
        v[2] += fabs(double(x[k] / T(sc[g]))) == ms[MS_PUV] ? 1.0 : 0.0;
      }
    }
    T tt[LEV], ss[LEV], w[LEV];
    bool wet[LEV];
    tracer_inc(a, p, tt, ss, w, wet);
    inc_cot(d, 2, c, G);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      v[1] += double(tt[k] * G[k]);
      v[3] += fabs(double(tt[k] / T(a.st))) == ms[MS_PTS] ? 1.0 : 0.0;
    }
    inc_cot(d, 3, c, G);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      v[1] += double(ss[k] * G[k]);
      v[3] += fabs(double(ss[k] / T(a.ss))) == ms[MS_PTS] ? 1.0 : 0.0;
    }
  }
  block_reduce<4, false>(v, a.part, 0);
}

// cotangent of each factor, through clamp(cap / clamp(peak, min=cap), max=1)
// to the peak (where peak >= cap), shared evenly by the elements at the peak
template <class T> __global__ void __launch_bounds__(TPB) k_corr_adj_member1(Batch B, CorrArgs<T> a, AdjArgs<T> d) {
  double* ms = a.ms + (int64_t)blockIdx.x * MS_N;
  __shared__ double s[4];
  member_reduce(B, a.part, 0, 4, false, s);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double wtot = ms[MS_WTOT];
    double db0 = 0.0, db1 = 0.0, db2 = 0.0;
    if (d.cdiag) {
      db0 = double(d.cdiag[3 * blockIdx.x + 0]);
      db1 = double(d.cdiag[3 * blockIdx.x + 1]);
      db2 = double(d.cdiag[3 * blockIdx.x + 2]);
    }
    const double fbar_uv = s[0];
    const double fbar_ts = s[1] + db0 + db1 * ms[MS_BT] / wtot + db2 * ms[MS_BS] / wtot;
    const double puv = ms[MS_PUV], pts = ms[MS_PTS];
    ms[MS_GUV] = (puv >= a.cap && s[2] > 0.0) ? -fbar_uv * a.cap / (puv * puv) / s[2] : 0.0;
    ms[MS_GTS] = (pts >= a.cap && s[3] > 0.0) ? -fbar_ts * a.cap / (pts * pts) / s[3] : 0.0;
    ms[MS_DB1] = db1;
    ms[MS_DB2] = db2;
  }
}

template <class T> __device__ __forceinline__ T sgn_t(T x) { return x > T(0) ? T(1) : (x < T(0) ? T(-1) : T(0)); }

// cotangents of the capped-increment inputs; the velocity channels are
// finished here (the column mean is column-local), the tracer layer sums are
// reduced for the projection's adjoint
template <class T> __global__ void __launch_bounds__(TPB) k_corr_adj2(Batch B, CorrArgs<T> a, AdjArgs<T> d) {
  Pos p;
  const bool live = pos2(B, p);
  double v[2 * LEV];
#pragma unroll
  for (int q = 0; q < 2 * LEV; ++q) v[q] = 0.0;
  if (live) {
    const int64_t c2 = ix2(p, p.i, p.j), c = ix3(p, p.i, p.j, 0);
    const double* ms = a.ms + (int64_t)p.m * MS_N;
    const int64_t n2 = (int64_t)p.im * p.jm;
    T* gout = d.graw + (int64_t)NCH * p.o2 + ((int64_t)p.i * p.jm + p.j);
    T x[LEV], G[LEV];
    const double sc[2] = {a.su, a.sv};
    for (int g = 0; g < 2; ++g) {
      const T mask = g ? a.dvm[c2] : a.dum[c2];
      vel_inc(a, p, g * LEV, sc[g], mask, x);
      inc_cot(d, g, c, G);
      T sum = T(0), den = T(0);
#pragma unroll
      for (int k = 0; k < LEV; ++k) {
        T b = T(ms[MS_FUV]) * G[k];
        if (fabs(double(x[k] / T(sc[g]))) == ms[MS_PUV]) b += T(ms[MS_GUV]) * sgn_t(x[k] / T(sc[g])) / T(sc[g]);
        G[k] = b * mask;   // cotangent of the centred, unmasked increment
        sum += G[k];
        den += dzk(a, p.m, k);
      }
#pragma unroll
      for (int k = 0; k < LEV; ++k) gout[(g * LEV + k) * n2] = (G[k] - dzk(a, p.m, k) / den * sum) * T(sc[g]);
    }
    T tt[LEV], ss[LEV], w[LEV];
    bool wet[LEV];
    tracer_inc(a, p, tt, ss, w, wet);
    const double wtot = ms[MS_WTOT];
    const T fts = T(ms[MS_FTS]);
    inc_cot(d, 2, c, G);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      T b = fts * G[k] + T(ms[MS_DB1] * double(fts) / wtot) * w[k];
      if (fabs(double(tt[k] / T(a.st))) == ms[MS_PTS]) b += T(ms[MS_GTS]) * sgn_t(tt[k] / T(a.st)) / T(a.st);
      v[k] = wet[k] ? double(b) : 0.0;
    }
    inc_cot(d, 3, c, G);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      T b = fts * G[k] + T(ms[MS_DB2] * double(fts) / wtot) * w[k];
      if (fabs(double(ss[k] / T(a.ss))) == ms[MS_PTS]) b += T(ms[MS_GTS]) * sgn_t(ss[k] / T(a.ss)) / T(a.ss);
      v[LEV + k] = wet[k] ? double(b) : 0.0;
    }
  }
  block_reduce<2 * LEV, false>(v, a.part, 0);
}

template <class T> __global__ void __launch_bounds__(TPB) k_corr_adj_member2(Batch B, CorrArgs<T> a) {
  double* ms = a.ms + (int64_t)blockIdx.x * MS_N;
  __shared__ double s[2 * LEV];
  member_reduce(B, a.part, 0, 2 * LEV, false, s);
  __syncthreads();
  if ((int)threadIdx.x < 2 * LEV) ms[MS_ST + threadIdx.x] = s[threadIdx.x];
}

// tracer channels: the adjoint of the layer projection (x - mean) * wet with
// mean = sum_ij x w / max(sum_ij w, 1e-20); and the cotangent of et through
// the weights w = art fsm (h + et) dz (layer means and budgets)
template <class T> __global__ void __launch_bounds__(TPB) k_corr_adj3(Batch B, CorrArgs<T> a, AdjArgs<T> d) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c2 = ix2(p, p.i, p.j), c = ix3(p, p.i, p.j, 0);
  const double* ms = a.ms + (int64_t)p.m * MS_N;
  const int64_t n2 = (int64_t)p.im * p.jm;
  T* gout = d.graw + (int64_t)NCH * p.o2 + ((int64_t)p.i * p.jm + p.j);
  T tt[LEV], ss[LEV], w[LEV];
  bool wet[LEV];
  tracer_inc(a, p, tt, ss, w, wet);
  T G[LEV];
  const T fts = T(ms[MS_FTS]);
  const double wtot = ms[MS_WTOT];
  T bet = T(0);
  const T colw = a.art[c2] * a.fsm[c2];
  for (int g = 0; g < 2; ++g) {
    const T* x = g ? ss : tt;
    const double sc = g ? a.ss : a.st;
    const double gp = ms[MS_GTS], db = g ? ms[MS_DB2] : ms[MS_DB1];
    inc_cot(d, 2 + g, c, G);
#pragma unroll
    for (int k = 0; k < LEV; ++k) {
      T b = fts * G[k] + T(db * double(fts) / wtot) * w[k];
      if (fabs(double(x[k] / T(sc))) == ms[MS_PTS]) b += T(gp) * sgn_t(x[k] / T(sc)) / T(sc);
      const double wk = ms[MS_W + k];
      const double sum = ms[(g ? MS_SS : MS_ST) + k];
      const T wc = T(fmax(wk, 1e-20));
      const T bx = (wet[k] ? b : T(0)) - w[k] / wc * T(sum);
      gout[((2 + g) * LEV + k) * n2] = bx * T(sc);
      // weights: the layer mean m = N / max(W, 1e-20) and the budget (fts sum w x') / sum w
      const T raw_x = rawc(a, p, (2 + g) * LEV + k) * T(sc);
      const T mk = T(ms[(g ? MS_MS : MS_MT) + k]);
      const T dm_dw = (raw_x - (wk >= 1e-20 ? mk : T(0))) / wc;
      T bw = -T(sum) * dm_dw;
      const double budget = double(fts) * (g ? ms[MS_BS] : ms[MS_BT]) / wtot;
      bw += T(db / wtot) * (fts * x[k] - T(budget));
      bet += bw * colw * dzk(a, p.m, k);
    }
  }
  if (d.get) d.get[c2] += bet;
}

template <class T>
int run_corr(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid, const double* scales,
             double cap, void* work, const void* const* in, void* const* out, void* diag, hipStream_t st) {
  CorrArgs<T> a{};
  a.raw = static_cast<const T*>(raw);
  const T* g[6];
  for (int q = 0; q < 6; ++q) g[q] = static_cast<const T*>(geom[q]);
  a.dum = g[0]; a.dvm = g[1]; a.fsm = g[2]; a.art = g[3]; a.h = g[4]; a.et = g[5];
  a.vg = static_cast<const T*>(vgrid);
  a.su = scales[0]; a.sv = scales[1]; a.st = scales[2]; a.ss = scales[3]; a.cap = cap;
  a.part = static_cast<double*>(work);
  a.ms = a.part + (int64_t)NP * b->nblk2;
  const Batch B{b->d_mem, b->d_blk3, b->d_blk2, b->M};
  const dim3 g2(b->nblk2), gm(b->M), tb(TPB);
  hipLaunchKernelGGL(k_corr_pass1<T>, g2, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_member1<T>, gm, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_pass2<T>, g2, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_member2<T>, gm, tb, 0, st, B, a, static_cast<T*>(diag));
  ApplyArgs<T> f{};
  for (int q = 0; q < 8; ++q) {
    f.in[q] = static_cast<const T*>(in[q]);
    f.out[q] = static_cast<T*>(out[q]);
  }
  hipLaunchKernelGGL(k_corr_apply<T>, g2, tb, 0, st, B, a, f);
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

template <class T>
int run_corr_adj(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid,
                 const double* scales, double cap, void* work, const void* const* cot_out, const void* cot_diag,
                 void* grad_raw, void* grad_et, hipStream_t st) {
  CorrArgs<T> a{};
  a.raw = static_cast<const T*>(raw);
  const T* g[6];
  for (int q = 0; q < 6; ++q) g[q] = static_cast<const T*>(geom[q]);
  a.dum = g[0]; a.dvm = g[1]; a.fsm = g[2]; a.art = g[3]; a.h = g[4]; a.et = g[5];
  a.vg = static_cast<const T*>(vgrid);
  a.su = scales[0]; a.sv = scales[1]; a.st = scales[2]; a.ss = scales[3]; a.cap = cap;
  a.part = static_cast<double*>(work);
  a.ms = a.part + (int64_t)NP * b->nblk2;
  AdjArgs<T> d{};
  for (int q = 0; q < 8; ++q) d.cot[q] = cot_out ? static_cast<const T*>(cot_out[q]) : nullptr;
  d.cdiag = static_cast<const T*>(cot_diag);
  d.graw = static_cast<T*>(grad_raw);
  d.get = static_cast<T*>(grad_et);
  const Batch B{b->d_mem, b->d_blk3, b->d_blk2, b->M};
  const dim3 g2(b->nblk2), gm(b->M), tb(TPB);
  // the forward statistics (layer means, peaks, factors, budgets)
  hipLaunchKernelGGL(k_corr_pass1<T>, g2, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_member1<T>, gm, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_pass2<T>, g2, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_member2<T>, gm, tb, 0, st, B, a, static_cast<T*>(nullptr));
  // reverse
  hipLaunchKernelGGL(k_corr_adj1<T>, g2, tb, 0, st, B, a, d);
  hipLaunchKernelGGL(k_corr_adj_member1<T>, gm, tb, 0, st, B, a, d);
  hipLaunchKernelGGL(k_corr_adj2<T>, g2, tb, 0, st, B, a, d);
  hipLaunchKernelGGL(k_corr_adj_member2<T>, gm, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_corr_adj3<T>, g2, tb, 0, st, B, a, d);
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// ---------------------------------------------------------------- density-75
// The DND projection that follows the joint correction
// (src/neuralpom/evaluation/density.py:14-64).  Per point of the 15 active
// levels: x = (after - before) / scale for T and S; the direction d = g / |g|
// of g = (drho/dT scale_T, drho/dS scale_S), the EOS partials at the state
// before the correction (the model EOS in float64, T and S seeded at the
// point); the volume weight w; the amplitude a = x . d.  Per member and
// level: the weighted 2 x 2 Gram matrix of d and the moments of d a (weights
// normalised by the level's sum), whose pseudo-inverse (eigenvalues at or
// below 1e-10 of the largest dropped) gives coefficients c.  Then
// adjusted = a - d . c, proposed = (x - 0.25 adjusted d) scale, one cap
// factor for the pair, and T, S and their history levels become before +
// factor proposed (ghost level from before).  A member whose T/S did not
// change keeps the corrected state (the reference's early return).
// Launches: column pass -> member solve -> column pass -> member factor ->
// apply.  Reductions in double, fixed order.
constexpr int NPD = 96;   // block partials per logical block (90 sums + 1 max)
enum {
  D_C0 = 0,             // coefficient c_0 per level (15)
  D_C1 = LEV,           // coefficient c_1 per level (15)
  D_WC = 2 * LEV,       // max(level weight sum, 1e-30) (15)
  D_ANY = 3 * LEV,      // max |x| over the member
  D_FAC,                // cap factor of the projected increments
  D_N
};

template <class T> struct D75Args {
  const T *bt, *bs;      // T and S before the joint correction
  const T *at, *as;      // after it
  const T *fsm, *art, *h, *et;
  const T* vg;           // [M][4][16]: zz row 1, dz row 2
  double st, ss, cap;
  Params prm;            // tbias, sbias, grav, rhoref of the EOS
  double* part;
  double* ms;            // D_N per member
};

// OpDens at one point with T and S both seeded, in double
struct DensAcc {
  Pos p;
  double t, s, h, fsm, zzk;   // the density input is read only below the active levels
  template <int F, int DI, int DJ, int DK> __device__ __forceinline__ auto at() const {
    static_assert(DI == 0 && DJ == 0 && DK == 0, "the EOS is pointwise");
    if constexpr (F == OpDens::TT) return DS<double, (1u << OpDens::TT)>{t, {1.0}};
    else if constexpr (F == OpDens::SS) return DS<double, (1u << OpDens::SS)>{s, {1.0}};
    else return P<double>{0.0};
  }
  template <int G, int DI, int DJ> __device__ __forceinline__ P<double> g() const {
    return P<double>{G == OpDens::GH ? h : fsm};
  }
  __device__ __forceinline__ double zz(int) const { return zzk; }
};

struct D75Pt {
  double x0, x1, d0, d1, w;
};

template <class T>
__device__ __forceinline__ D75Pt d75_point(const D75Args<T>& a, Pos p, int64_t c2, int64_t c, int k) {
  D75Pt r;
  r.x0 = double(a.at[c] - a.bt[c]) / a.st;
  r.x1 = double(a.as[c] - a.bs[c]) / a.ss;
  p.k = k;
  const DensAcc A{p, double(a.bt[c]), double(a.bs[c]), double(a.h[c2]), double(a.fsm[c2]),
                  double(a.vg[(p.m * 4 + 1) * KB + k])};
  const auto e = OpDens::template eval<0>(A, a.prm);
  constexpr uint32_t M = mask_of<std::decay_t<decltype(e)>>::value;
  static_assert(((M >> OpDens::TT) & 1u) && ((M >> OpDens::SS) & 1u), "both partials");
  const double g0 = e.d[slot(M, OpDens::TT)] * a.st, g1 = e.d[slot(M, OpDens::SS)] * a.ss;
  const double n = fmax(sqrt(g0 * g0 + g1 * g1), 1e-30);
  r.d0 = g0 / n;
  r.d1 = g1 / n;
  const T col = a.art[c2] * a.fsm[c2] * (a.h[c2] + a.et[c2]);
  r.w = double(col * a.vg[(p.m * 4 + 2) * KB + k]);
  return r;
}

template <class T> __global__ void __launch_bounds__(TPB) k_d75_pass1(Batch B, D75Args<T> a) {
  Pos p;
  const bool live = pos2(B, p);
  const int64_t c2 = live ? ix2(p, p.i, p.j) : 0, c = live ? ix3(p, p.i, p.j, 0) : 0;
  double any = 0.0;
  for (int k = 0; k < LEV; ++k) {
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (live) {
      const D75Pt q = d75_point(a, p, c2, c + k, k);
      const double amp = q.x0 * q.d0 + q.x1 * q.d1;
      v[0] = q.w;
      v[1] = q.w * q.d0 * q.d0;
      v[2] = q.w * q.d0 * q.d1;
      v[3] = q.w * q.d1 * q.d1;
      v[4] = q.w * q.d0 * amp;
      v[5] = q.w * q.d1 * amp;
      any = fmax(any, fmax(fabs(q.x0), fabs(q.x1)));
    }
    block_reduce<6, false, NPD>(v, a.part, 6 * k);
  }
  block_reduce<1, true, NPD>(&any, a.part, 6 * LEV);
}

// coefficients V diag(1/lambda, lambda > 1e-10 lambda_max) V^T r of the
// symmetric [[g00, g01], [g01, g11]]: closed-form eigenpairs
__device__ __forceinline__ void pinv2(double g00, double g01, double g11, double r0, double r1, double& c0,
                                      double& c1) {
  const double mean = 0.5 * (g00 + g11), diff = 0.5 * (g00 - g11);
  const double rad = hypot(diff, g01);
  const double l1 = mean + rad;                                   // largest
  const double det = g00 * g11 - g01 * g01;
  const double l0 = l1 > 0.0 ? det / l1 : mean - rad;             // smallest, without cancellation
  double v0, v1;                                                  // unit eigenvector of l1
  if (rad == 0.0) {
    v0 = 1.0;
    v1 = 0.0;
  } else if (diff >= 0.0) {
    v0 = diff + rad;
    v1 = g01;
  } else {
    v0 = g01;
    v1 = rad - diff;
  }
  const double nv = hypot(v0, v1);
  v0 /= nv;
  v1 /= nv;
  const double cut = l1 * 1e-10;
  const double i1 = l1 > cut ? 1.0 / fmax(l1, 1e-30) : 0.0;
  const double i0 = l0 > cut ? 1.0 / fmax(l0, 1e-30) : 0.0;
  const double p1 = (v0 * r0 + v1 * r1) * i1;                      // along (v0, v1)
  const double p0 = (-v1 * r0 + v0 * r1) * i0;                     // along (-v1, v0)
  c0 = v0 * p1 - v1 * p0;
  c1 = v1 * p1 + v0 * p0;
}

template <class T> __global__ void __launch_bounds__(TPB) k_d75_member1(Batch B, D75Args<T> a) {
  double* ms = a.ms + (int64_t)blockIdx.x * D_N;
  __shared__ double s[6 * LEV + 1];
  member_reduce(B, a.part, 0, 6 * LEV, false, s, NPD);
  member_reduce(B, a.part, 6 * LEV, 1, true, s + 6 * LEV, NPD);
  __syncthreads();
  const int k = threadIdx.x;
  if (k < LEV) {
    const double* q = s + 6 * k;
    const double wc = fmax(q[0], 1e-30);
    double c0, c1;
    pinv2(q[1] / wc, q[2] / wc, q[3] / wc, q[4] / wc, q[5] / wc, c0, c1);
    ms[D_C0 + k] = c0;
    ms[D_C1 + k] = c1;
    ms[D_WC + k] = wc;
  }
  if (threadIdx.x == 0) ms[D_ANY] = s[6 * LEV];
}

// proposed increments (already times the scales) of one point
template <class T>
__device__ __forceinline__ void d75_proposed(const D75Args<T>& a, const double* ms, const D75Pt& q, int k,
                                             double& adj, double& p0, double& p1) {
  adj = (q.x0 * q.d0 + q.x1 * q.d1) - (q.d0 * ms[D_C0 + k] + q.d1 * ms[D_C1 + k]);
  const double ch = (0.75 - 1.0) * adj;
  p0 = (q.x0 + ch * q.d0) * a.st;
  p1 = (q.x1 + ch * q.d1) * a.ss;
}

template <class T> __global__ void __launch_bounds__(TPB) k_d75_pass2(Batch B, D75Args<T> a) {
  Pos p;
  const bool live = pos2(B, p);
  const int64_t c2 = live ? ix2(p, p.i, p.j) : 0, c = live ? ix3(p, p.i, p.j, 0) : 0;
  const double* ms = a.ms + (int64_t)(live ? p.m : 0) * D_N;
  double peak = 0.0;
  for (int k = 0; k < LEV; ++k) {
    double v[2] = {0.0, 0.0};
    if (live) {
      const D75Pt q = d75_point(a, p, c2, c + k, k);
      double adj, p0, p1;
      d75_proposed(a, ms, q, k, adj, p0, p1);
      v[0] = q.w * q.d0 * adj;
      v[1] = q.w * q.d1 * adj;
      peak = fmax(peak, fmax(fabs(p0 / a.st), fabs(p1 / a.ss)));
    }
    block_reduce<2, false, NPD>(v, a.part, 2 * k);
  }
  block_reduce<1, true, NPD>(&peak, a.part, 2 * LEV);
}

__global__ void __launch_bounds__(TPB) k_d75_member2(Batch B, double* ms_all, const double* part, double cap,
                                                      double* diag) {
  double* ms = ms_all + (int64_t)blockIdx.x * D_N;
  __shared__ double s[2 * LEV + 1];
  member_reduce(B, part, 0, 2 * LEV, false, s, NPD);
  member_reduce(B, part, 2 * LEV, 1, true, s + 2 * LEV, NPD);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double f = pair_factor(s[2 * LEV], cap);
    double res = 0.0;
    for (int k = 0; k < LEV; ++k)
      res = fmax(res, fmax(fabs(s[2 * k] / ms[D_WC + k]), fabs(s[2 * k + 1] / ms[D_WC + k])));
    const bool changed = ms[D_ANY] != 0.0;
    ms[D_FAC] = f;
    if (diag) {
      diag[3 * blockIdx.x + 0] = changed ? f : 1.0;
      diag[3 * blockIdx.x + 1] = changed ? res : 0.0;
      diag[3 * blockIdx.x + 2] = changed ? 1.0 : 0.0;
    }
  }
}

template <class T> struct D75Out {
  const T* before[4];   // t tb s sb
  const T* after[4];
  T* out[4];
};

template <class T> __global__ void __launch_bounds__(TPB) k_d75_apply(Batch B, D75Args<T> a, D75Out<T> f) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c2 = ix2(p, p.i, p.j), c = ix3(p, p.i, p.j, 0);
  const double* ms = a.ms + (int64_t)p.m * D_N;
  if (ms[D_ANY] == 0.0) {   // no T/S change: the corrected state stands
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      T y[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) y[k] = f.after[q][c + k];
#pragma unroll
      for (int k = 0; k < KB; ++k) f.out[q][c + k] = y[k];
    }
    return;
  }
  const double fac = ms[D_FAC];
  T dt[LEV], ds[LEV];
  for (int k = 0; k < LEV; ++k) {
    const D75Pt q = d75_point(a, p, c2, c + k, k);
    double adj, p0, p1;
    d75_proposed(a, ms, q, k, adj, p0, p1);
    dt[k] = T(fac * p0);
    ds[k] = T(fac * p1);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const T* src = f.before[q] + c;
    T* dst = f.out[q] + c;
    const T* inc = q < 2 ? dt : ds;
    T y[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) y[k] = src[k];
#pragma unroll
    for (int k = 0; k < KB; ++k) dst[k] = k < LEV ? y[k] + inc[k] : y[k];
  }
}

template <class T>
int run_d75(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* geom,
            const double* scales, double cap, void* work, const void* const* before, const void* const* after,
            void* const* out, double* diag, hipStream_t st) {
  D75Args<T> a{};
  a.bt = static_cast<const T*>(before[0]);
  a.bs = static_cast<const T*>(before[2]);
  a.at = static_cast<const T*>(after[0]);
  a.as = static_cast<const T*>(after[2]);
  a.fsm = static_cast<const T*>(geom[0]);
  a.art = static_cast<const T*>(geom[1]);
  a.h = static_cast<const T*>(geom[2]);
  a.et = static_cast<const T*>(geom[3]);
  a.vg = static_cast<const T*>(vgrid);
  a.st = scales[0];
  a.ss = scales[1];
  a.cap = cap;
  a.prm.tbias = prm->tbias;
  a.prm.sbias = prm->sbias;
  a.prm.grav = prm->grav;
  a.prm.rhoref = prm->rhoref;
  a.part = static_cast<double*>(work);
  a.ms = a.part + (int64_t)NPD * b->nblk2;
  D75Out<T> f{};
  for (int q = 0; q < 4; ++q) {
    f.before[q] = static_cast<const T*>(before[q]);
    f.after[q] = static_cast<const T*>(after[q]);
    f.out[q] = static_cast<T*>(out[q]);
  }
  const Batch B{b->d_mem, b->d_blk3, b->d_blk2, b->M};
  const dim3 g2(b->nblk2), gm(b->M), tb(TPB);
  hipLaunchKernelGGL(k_d75_pass1<T>, g2, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_d75_member1<T>, gm, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_d75_pass2<T>, g2, tb, 0, st, B, a);
  hipLaunchKernelGGL(k_d75_member2, gm, tb, 0, st, B, a.ms, a.part, cap, diag);
  hipLaunchKernelGGL(k_d75_apply<T>, g2, tb, 0, st, B, a, f);
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

bool corr_args_ok(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid,
                  const double* scales, double cap, void* work) {
  if (!b || !raw || !geom || !vgrid || !scales || !work || !(cap > 0.0)) return false;
  for (int q = 0; q < 6; ++q)
    if (!geom[q]) return false;
  for (int q = 0; q < 4; ++q)
    if (!(scales[q] != 0.0)) return false;
  return true;
}

}  // namespace

extern "C" {

int pom_correction_workspace(const pom_batch* b, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  *bytes = sizeof(double) * ((size_t)NP * b->nblk2 + (size_t)MS_N * b->M);
  return POM_OK;
}

int pom_joint_correction(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid,
                         const double* scales, double cap, void* work, const void* const* state_in,
                         void* const* state_out, void* diag, void* stream) {
  if (!corr_args_ok(b, raw, geom, vgrid, scales, cap, work) || !state_in || !state_out) return POM_E_ARG;
  for (int q = 0; q < 8; ++q)
    if (!state_in[q] || !state_out[q]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_corr<float>(b, raw, geom, vgrid, scales, cap, work, state_in, state_out, diag, st);
  return run_corr<double>(b, raw, geom, vgrid, scales, cap, work, state_in, state_out, diag, st);
}

int pom_joint_correction_adj(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid,
                             const double* scales, double cap, void* work, const void* const* cot_out,
                             const void* cot_diag, void* grad_raw, void* grad_et, void* stream) {
  if (!corr_args_ok(b, raw, geom, vgrid, scales, cap, work) || !grad_raw) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_corr_adj<float>(b, raw, geom, vgrid, scales, cap, work, cot_out, cot_diag, grad_raw, grad_et, st);
  return run_corr_adj<double>(b, raw, geom, vgrid, scales, cap, work, cot_out, cot_diag, grad_raw, grad_et, st);
}

int pom_density75_workspace(const pom_batch* b, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  *bytes = sizeof(double) * ((size_t)NPD * b->nblk2 + (size_t)D_N * b->M);
  return POM_OK;
}

int pom_density75(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* geom,
                  const double* scales, double cap, void* work, const void* const* before,
                  const void* const* after, void* const* state_out, void* diag, void* stream) {
  if (!b || !prm || !vgrid || !geom || !scales || !work || !before || !after || !state_out || !(cap > 0.0))
    return POM_E_ARG;
  if (!(scales[0] != 0.0) || !(scales[1] != 0.0)) return POM_E_ARG;
  for (int q = 0; q < 4; ++q)
    if (!geom[q] || !before[q] || !after[q] || !state_out[q]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* dg = static_cast<double*>(diag);
  if (b->dtype == POM_F32)
    return run_d75<float>(b, prm, vgrid, geom, scales, cap, work, before, after, state_out, dg, st);
  return run_d75<double>(b, prm, vgrid, geom, scales, cap, work, before, after, state_out, dg, st);
}

}  // extern "C"
