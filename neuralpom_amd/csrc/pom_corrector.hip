// The corrector hook between physics intervals, fused: the network's raw
// increment (60 planes: u, v, t, s on the 15 active levels) is scaled, the
// velocity increments lose their dz-weighted column mean and are masked, the
// tracer increments lose their volume-weighted layer mean (wet cells only),
// each pair is capped by one shared factor, and both time levels of the four
// fields receive the increment; the ghost level 15 is copied.
// reference: src/neuralpom/hybrid/sparse_coupling.py:18-67 (volume_weights,
// budget_project, paired_limit, apply_joint_correction).  The density refresh
// that follows (:61) is the EOS entry point, called by the host after this.
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
template <int NV, bool MAX>
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
    part[(int64_t)logical_block() * NP + off + threadIdx.x] = x;
  }
  __syncthreads();
}

// per-member reduction of block partials [off, off + nv) (one block per member)
__device__ __forceinline__ void member_reduce(const Batch& B, const double* part, int off, int nv, bool mx,
                                              double* out) {
  const int m = blockIdx.x;
  const int b0 = B.blk2[m], b1 = B.blk2[m + 1];
  for (int q = threadIdx.x; q < nv; q += TPB) {
    double x = 0.0;   // sums from zero; maxima of absolute values from zero
    for (int b = b0; b < b1; ++b) {
      const double y = part[(int64_t)b * NP + off + q];
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

}  // extern "C"
