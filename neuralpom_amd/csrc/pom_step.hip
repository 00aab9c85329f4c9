// Host orchestration of the explicit and semi-implicit POM steps and their
// adjoints, and the C ABI.
//
// Forward: the stages of src/neuralpom/cores/explicit/time_loop.py:75-696
// (explicit) or src/neuralpom/cores/implicit_p0_fixed/time_loop.py:95-688
// (semi-implicit) in order, every intermediate in its own workspace buffer
// (static single assignment), so the same forward doubles as the recompute
// pass of the adjoint.  Adjoint: the stages reversed; stencil stages through the generated
// gathers of pom_kernels.h, column stages through the hand-written reverse
// sweeps of pom_column.h / pom_profq.h.
#include <vector>
#include <mutex>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <type_traits>

#include "../../include/neuralpom_amd.h"
#include "pom_kernels.h"
#include "pom_ops_imp.h"
#include "pom_pcg.h"
#include "pom_profq.h"
#include "pom_extrev.h"
#include "pom_momadj.h"
#include "pom_advctadj.h"
#include "pom_advaveadj.h"
#include "pom_smagadj.h"
#include "pom_advfwd.h"

#include "pom_batch.h"

using namespace pom;

namespace {

constexpr int MAXSPLIT = 256;

// ---------------------------------------------------------------- tracing
// HIP events around every launch of one selected stage (pom_profile_stage);
// pom_profile_read sums their durations.  Off by default: no events recorded.
struct StageProf {
  int stage = -1;
  std::vector<hipEvent_t> pool;   // pairs (begin, end)
  size_t used = 0;
};
StageProf g_prof;
constexpr size_t PROF_POOL = 512;   // events created up front (256 launches)

struct ProfScope {
  hipEvent_t b = nullptr, e = nullptr;
  hipStream_t s;
  ProfScope(int stage, hipStream_t st) : s(st) {
    if (g_prof.stage != stage) return;
    if (g_prof.used + 2 > g_prof.pool.size()) {
      hipEvent_t x, y;
      if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) return;
      g_prof.pool.push_back(x);
      g_prof.pool.push_back(y);
    }
    b = g_prof.pool[g_prof.used];
    e = g_prof.pool[g_prof.used + 1];
    g_prof.used += 2;
    (void)hipEventRecord(b, s);
  }
  ~ProfScope() {
    if (e) (void)hipEventRecord(e, s);
  }
};

// ---------------------------------------------------------------- workspace
struct Arena {
  char* base;
  size_t off, cap;
  bool measure;
  template <class T> T* get(int64_t n) {
    off = (off + 255) & ~size_t(255);
    T* p = measure ? nullptr : reinterpret_cast<T*>(base + off);
    off += size_t(n) * sizeof(T);
    return p;
  }
};

// flat elementwise kernels: grid-stride loops over a capped grid (block
// dispatch, not bandwidth, limits a one-element-per-thread launch of a 2-D field)
template <class T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
// zero up to 32 arrays in one launch (grid.y selects the array)
template <class T> struct FillTab {
  T* p[32];
  int64_t n[32];
};
template <class T>
__global__ void k_fill_multi(FillTab<T> t) {
  const int e = blockIdx.y;
  T* p = t.p[e];
  const int64_t n = t.n[e];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = T(0);
}
// up to 24 sums d = s0 + s1 + s2 of equal-length arrays in one launch
// (grid.y selects the entry; null sources are skipped, d may be one of them)
template <class T> struct SumTab {
  T* d[24];
  const T* s[24][3];
  int64_t n;
};
template <class T>
__global__ void k_sum_multi(SumTab<T> t) {
  const int e = blockIdx.y;
  T* d = t.d[e];
  const T *s0 = t.s[e][0], *s1 = t.s[e][1], *s2 = t.s[e][2];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < t.n; i += (int64_t)gridDim.x * blockDim.x) {
    T v = s0[i];
    if (s1) v += s1[i];
    if (s2) v += s2[i];
    d[i] = v;
  }
}
template <class T>
__global__ void k_copy(T* __restrict__ d, const T* __restrict__ s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}
// the copy in 16-byte words (both arrays 16-byte aligned, n16 words)
__global__ void k_copy16(uint4* __restrict__ d, const uint4* __restrict__ s, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}
template <class T>
__global__ void k_add(T* __restrict__ d, const T* __restrict__ s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] += s[i];
}
// dt = h + et (time_loop.py:693)
template <class T>
__global__ void k_hplus(T* __restrict__ d, const T* __restrict__ h, const T* __restrict__ e, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = h[i] + e[i];
}
// running max |vaf| (time_loop.py:393)
// grid-stride max over |vaf| (NaN propagates), one atomic per block
template <class T>
__global__ void __launch_bounds__(TPB) k_vamax(const T* __restrict__ vaf, int64_t n, T* out) {
  T v = T(0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v = nanmax(v, fabs(vaf[i]));
  block_amax(v, out);
}

inline int blocks(int64_t n) { return (int)((n + 255) / 256); }
// grid of a grid-stride flat kernel: 4 elements per thread, at most 2048 blocks
inline int flat_blocks(int64_t n) {
  const int64_t b = (n + 1023) / 1024;
  return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

template <class T> struct X {   // state of one barotropic substep
  T *ua, *uab, *va, *vab, *el, *elb, *d, *egf, *utf, *vtf, *etf, *advua, *advva, *uaf, *vaf, *elf, *fluxua, *fluxva,
      *tps;
};

// Pinned host words for the eta solver's convergence polls: one per solve in
// flight, taken from a pool under a lock, so solves running at the same time
// (other streams, other host threads) never share one.  A word goes back to
// the pool when its solve returns; by then its last copy has completed (the
// solve synchronised on it).  Words live until the process ends.
std::mutex g_pin_mu;
std::vector<int*> g_pin_free;
struct PinnedWord {
  int* w = nullptr;
  hipError_t acquire() {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (!g_pin_free.empty()) {
      w = g_pin_free.back();
      g_pin_free.pop_back();
      return hipSuccess;
    }
    return hipHostMalloc(reinterpret_cast<void**>(&w), sizeof(int));
  }
  ~PinnedWord() {
    if (!w) return;
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.push_back(w);
  }
};

// work arrays and driver of the batched eta solve
template <class T> struct Pcg {
  T *r, *p, *z, *ap, *t1, *t2;
  double *part, *sc, *tol;
  int *act, *its, *nact;
  const SpecMember* sp = nullptr;   // spectral cache (device)
  const T* cache = nullptr;
  int maxni = 0, maxnj = 0;
  bool spectral = false;
  bool capture = false;             // no host polls: a fixed launch count (capture-safe)
  int status = POM_OK;              // POM_E_HIP after a failed poll
  static constexpr int POLL = 4;    // iterations between convergence polls

  void plan(Arena& A, int64_t N2, int nblk2, int M) {
    r = A.get<T>(N2); p = A.get<T>(N2); z = A.get<T>(N2); ap = A.get<T>(N2);
    t1 = A.get<T>(N2); t2 = A.get<T>(N2);
    part = A.get<double>(2 * (int64_t)nblk2 + 2);
    sc = A.get<double>((int64_t)PC_N * M);
    tol = A.get<double>(M);
    act = A.get<int>(M); its = A.get<int>(M); nact = A.get<int>(1);
  }

  // z = M^-1 r and partial sums of r.z (slot 0).  With test set (an
  // iteration's preconditioner), the kernel that forms the partial sums first
  // applies the convergence test (pcg_test): members that stop write zeros.
  void precond(const Ctx<T>& c, const EllGeom<T>& g, const T* d, bool test) {
    const dim3 g2(c.nblk2), tb(TPB);
    if (spectral) {
      GemmArg<T> a1{nullptr, r, t1, 0, 2, 3, 0};    // T1 = S_ni R
      GemmArg<T> a2{t1, nullptr, t2, 3, 1, 3, 1};   // T2 = (T1 S_nj) * L
      GemmArg<T> a3{nullptr, t2, t1, 0, 3, 3, 0};   // T1 = S_ni T2
      GemmArg<T> a4{t1, nullptr, z, 3, 1, 2, 2};    // Z = T1 S_nj / scale (interior)
      if constexpr (std::is_same<T, float>::value) {
        const dim3 gg((maxnj + 31) / 32, (maxni + 31) / 32, c.B.M);
        hipLaunchKernelGGL(k_spec_gemm_mfma, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a1, 1);
        hipLaunchKernelGGL(k_spec_gemm_mfma, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a2, 0);
        hipLaunchKernelGGL(k_spec_gemm_mfma, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a3, 1);
        hipLaunchKernelGGL(k_spec_gemm_mfma, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a4, 0);
      } else {
        const dim3 gg((maxnj + 15) / 16, (maxni + 15) / 16, c.B.M);
        hipLaunchKernelGGL(k_spec_gemm<T>, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a1, 1);
        hipLaunchKernelGGL(k_spec_gemm<T>, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a2, 0);
        hipLaunchKernelGGL(k_spec_gemm<T>, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a3, 1);
        hipLaunchKernelGGL(k_spec_gemm<T>, gg, tb, 0, c.stream, c.B.mem, sp, cache, act, a4, 0);
      }
      if (test)
        hipLaunchKernelGGL(k_pcg_dot_test<T>, g2, tb, 0, c.stream, c.B, r, z, act, part, sc, its, nact);
      else
        hipLaunchKernelGGL(k_pcg_dot<T>, g2, tb, 0, c.stream, c.B, r, z, act, part, 0);
    } else if (test) {
      hipLaunchKernelGGL(k_pcg_jacobi_test<T>, g2, tb, 0, c.stream, c.B, g, d, r, z, act, part, sc, its, nact);
    } else {
      hipLaunchKernelGGL(k_pcg_jacobi<T>, g2, tb, 0, c.stream, c.B, g, d, r, z, act, part);
    }
  }

  // x = A(d)^-1 rhs from x0 (null: zero), per-member absolute tolerance tol[]
  // (device, set by the caller) and relative tolerance rel_tol.  The edges and
  // the points g.fixed marks keep the lift (null: zero).  Without capture the
  // host reads the count of unconverged members every POLL iterations and
  // stops at zero; a failed poll sets status to POM_E_HIP and abandons the
  // solve.  With capture all max_iter iterations are launched and converged
  // members' kernels do nothing, so x is the same either way.  An iteration
  // is four launches (Jacobi) or eight (spectral: four of them the sine
  // products): A p with the p.ap sums; alpha and the x, r update with the
  // r.r sums; the convergence test with the preconditioner's r.z sums; beta
  // and the p update (pom_pcg.h).
  void solve(const Ctx<T>& c, const EllGeom<T>& g, const T* rhs, const T* d, const T* x0, T* x, double rel_tol,
             int max_iter, const T* lift = nullptr) {
    const dim3 g2(c.nblk2), tb(TPB), gm(c.B.M);
    PinnedWord pin;
    if (!capture && pin.acquire() != hipSuccess) {
      status = POM_E_HIP;
      return;
    }
    int* hw = pin.w;
    hipLaunchKernelGGL(k_pcg_x0<T>, g2, tb, 0, c.stream, c.B, x0, g.fixed, lift, x, z);
    hipLaunchKernelGGL(k_pcg_resid<T>, g2, tb, 0, c.stream, c.B, g, d, x, rhs, r, p);
    hipLaunchKernelGGL(k_pcg_act1, dim3(1), tb, 0, c.stream, act, c.B.M);
    precond(c, g, d, false);
    hipLaunchKernelGGL(k_pcg_dot<T>, g2, tb, 0, c.stream, c.B, r, r, act, part, 1);
    hipLaunchKernelGGL(k_pcg_start, gm, tb, 0, c.stream, c.B, part, sc, act, its, tol, rel_tol);
    hipLaunchKernelGGL(k_pcg_p<T>, g2, tb, 0, c.stream, c.B, p, z, g.fixed, sc, act, part, 1, 0);
    for (int it = 1; it <= max_iter; ++it) {
      const int par = (it - 1) & 1;   // rho's slot in this iteration
      hipLaunchKernelGGL(k_pcg_ap<T>, g2, tb, 0, c.stream, c.B, g, d, p, ap, act, part, nact);
      hipLaunchKernelGGL(k_pcg_xr<T>, g2, tb, 0, c.stream, c.B, x, r, p, ap, sc, act, part, par);
      precond(c, g, d, true);
      if (!capture && (it % POLL == 0 || it == max_iter)) {
        if (hipMemcpyAsync(hw, nact, sizeof(int), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess) {
          status = POM_E_HIP;
          return;
        }
        if (*hw == 0) break;
      }
      if (it == max_iter) break;
      hipLaunchKernelGGL(k_pcg_p<T>, g2, tb, 0, c.stream, c.B, p, z, g.fixed, sc, act, part, 0, par);
    }
  }
};

template <class T> struct Step {
  const pom_batch* b;
  Ctx<T> c;
  const T* in[POM_NSTATE];
  T* out[POM_NSTATE];
  const T* g[POM_NGEOM];
  int isplit;
  bool internal;
  T* vamax;
  // intermediates
  T *rho1, *adx2d_raw, *ady2d_raw, *advua0, *advva0, *fluxua0, *fluxva0, *tps0, *egf0, *utf0, *vtf0;
  X<T> xs[MAXSPLIT + 1];
  // semi-implicit core
  bool imp = false;
  const pom_params_t* hp = nullptr;   // host parameters (eta solver settings)
  const void* icache = nullptr;       // spectral preconditioner cache (pom_implicit_prepare)
  int32_t* iters_out = nullptr;       // per-member eta-solve iterations (device), may be null
  X<T> xi;                            // depth-mean state after the semi-implicit barotropic step
  T *us, *vs, *dus, *dvs, *rhs, *elfraw;
  // open-boundary tide policy (case_adapters/coastal_boundary.py:9-67)
  bool open = false;
  double tide_wt = 0.0;               // (speed pi / 180) t / 3600 of the step's tide
  T *lift = nullptr, *us2, *vs2, *uac, *vac;   // tide at the nodes; radiated predictor; closed corrector
  Pcg<T> pw;
  T *u1, *v1, *q2f0, *q2lf0, *q2f1, *q2lf1, *kmraw, *khraw, *tf0, *sf0, *tf1, *sf1, *uf0, *vf0, *uf1, *vf1, *uf2,
      *vf2, *tpsu;
  bool tape = false;     // the forward keeps profq's tape for the reverse kernels (adjoint recompute)
  ProfqTape<T> qt{};
  // spatial Shapiro filters of the closed-basin numerics (weights > 0): the
  // unfiltered new T, S and filtered u, v go through these before the filter
  bool shap_ts = false, shap_uv = false;
  T *tsh = nullptr, *ssh = nullptr, *ush = nullptr, *vsh = nullptr;

  int64_t n3() const { return b->n3; }
  int64_t n2() const { return b->n2; }

  void plan(Arena& A, bool own_out) {
    const int64_t N3 = n3(), N2 = n2();
    if (own_out)
      for (int f = 0; f < POM_NSTATE; ++f) out[f] = A.get<T>(f < POM_N3D ? N3 : N2);
    rho1 = A.get<T>(N3);
    adx2d_raw = A.get<T>(N2);
    ady2d_raw = A.get<T>(N2);
    advua0 = A.get<T>(N2);
    advva0 = A.get<T>(N2);
    fluxua0 = A.get<T>(N2);
    fluxva0 = A.get<T>(N2);
    tps0 = A.get<T>(N2);
    egf0 = A.get<T>(N2);
    utf0 = A.get<T>(N2);
    vtf0 = A.get<T>(N2);
    if (imp) {
      us = A.get<T>(N2); vs = A.get<T>(N2); dus = A.get<T>(N2); dvs = A.get<T>(N2);
      rhs = A.get<T>(N2); elfraw = A.get<T>(N2);
      if (open) {
        lift = A.get<T>(N2); us2 = A.get<T>(N2); vs2 = A.get<T>(N2); uac = A.get<T>(N2); vac = A.get<T>(N2);
      }
      pw.plan(A, N2, b->nblk2, b->M);
      xi = X<T>{};
      xi.ua = out[POM_UA]; xi.uab = out[POM_UAB]; xi.va = out[POM_VA]; xi.vab = out[POM_VAB];
      xi.el = out[POM_EL]; xi.elb = out[POM_ELB]; xi.d = out[POM_D]; xi.egf = out[POM_EGF];
      xi.utf = out[POM_UTF]; xi.vtf = out[POM_VTF]; xi.etf = out[POM_ETF]; xi.uaf = out[POM_UAF];
      xi.vaf = out[POM_VAF]; xi.elf = out[POM_ELF]; xi.advua = out[POM_ADVUA]; xi.advva = out[POM_ADVVA];
      xi.fluxua = out[POM_FLUXUA]; xi.fluxva = out[POM_FLUXVA]; xi.tps = tps0;
    }
    for (int n = 1; n <= (imp ? 0 : isplit); ++n) {
      X<T>& x = xs[n];
      const bool last = n == isplit;
      // the level shift makes ua, va, el of substep n its uaf, vaf, elf: one
      // array each, except for the step's outputs
      x.uaf = last ? out[POM_UAF] : A.get<T>(N2);
      x.vaf = last ? out[POM_VAF] : A.get<T>(N2);
      x.elf = last ? out[POM_ELF] : A.get<T>(N2);
      x.ua = last ? out[POM_UA] : x.uaf;
      x.uab = last ? out[POM_UAB] : A.get<T>(N2);
      x.va = last ? out[POM_VA] : x.vaf;
      x.vab = last ? out[POM_VAB] : A.get<T>(N2);
      x.el = last ? out[POM_EL] : x.elf;
      x.elb = last ? out[POM_ELB] : A.get<T>(N2);
      x.d = last ? out[POM_D] : A.get<T>(N2);
      x.egf = last ? out[POM_EGF] : A.get<T>(N2);
      x.utf = last ? out[POM_UTF] : A.get<T>(N2);
      x.vtf = last ? out[POM_VTF] : A.get<T>(N2);
      x.etf = last ? out[POM_ETF] : A.get<T>(N2);
      if (n % c.prm.ispadv == 0) {
        const bool lastadv = n + c.prm.ispadv > isplit;
        x.advua = lastadv ? out[POM_ADVUA] : A.get<T>(N2);
        x.advva = lastadv ? out[POM_ADVVA] : A.get<T>(N2);
        x.fluxua = lastadv ? out[POM_FLUXUA] : A.get<T>(N2);
        x.fluxva = lastadv ? out[POM_FLUXVA] : A.get<T>(N2);
        x.tps = A.get<T>(N2);
      } else {
        x.advua = x.advva = x.fluxua = x.fluxva = x.tps = nullptr;   // carried
      }
    }
    u1 = A.get<T>(N3);
    v1 = A.get<T>(N3);
    q2f0 = A.get<T>(N3);
    q2lf0 = A.get<T>(N3);
    q2f1 = A.get<T>(N3);
    q2lf1 = A.get<T>(N3);
    kmraw = A.get<T>(N3);
    khraw = A.get<T>(N3);
    tf0 = A.get<T>(N3);
    sf0 = A.get<T>(N3);
    tf1 = A.get<T>(N3);
    sf1 = A.get<T>(N3);
    uf0 = A.get<T>(N3);
    vf0 = A.get<T>(N3);
    uf1 = A.get<T>(N3);
    vf1 = A.get<T>(N3);
    uf2 = A.get<T>(N3);
    vf2 = A.get<T>(N3);
    tpsu = A.get<T>(N2);
    if (shap_ts) { tsh = A.get<T>(N3); ssh = A.get<T>(N3); }
    if (shap_uv) { ush = A.get<T>(N3); vsh = A.get<T>(N3); }
    if (tape) {
      qt.ee1 = A.get<T>(N3); qt.gg1 = A.get<T>(N3); qt.q2n = A.get<T>(N3); qt.ee2 = A.get<T>(N3);
      qt.gg2 = A.get<T>(N3); qt.q2ln = A.get<T>(N3); qt.ghr = A.get<T>(N3); qt.kn = A.get<T>(N3);
    }
  }

  void copy(T* d, const T* s, int64_t n) {
    if (d == s) return;
    const int64_t bytes = n * (int64_t)sizeof(T);
    if ((((uintptr_t)d | (uintptr_t)s) & 15u) == 0 && bytes % 16 == 0) {
      // one 16-byte word per thread (a grid-stride loop beyond 65 536 workgroups)
      const int64_t n16 = bytes / 16, blocks = (n16 + 255) / 256;
      hipLaunchKernelGGL(k_copy16, dim3((unsigned)(blocks < 1 ? 1 : (blocks > 65536 ? 65536 : blocks))), dim3(256), 0,
                         c.stream, reinterpret_cast<uint4*>(d), reinterpret_cast<const uint4*>(s), n16);
      return;
    }
    hipLaunchKernelGGL(k_copy<T>, dim3(flat_blocks(n)), dim3(256), 0, c.stream, d, s, n);
  }

  // X_0: the state entering the substeps
  X<T> x0() const {
    X<T> x;
    x.ua = (T*)in[POM_UA];
    x.uab = (T*)in[POM_UAB];
    x.va = (T*)in[POM_VA];
    x.vab = (T*)in[POM_VAB];
    x.el = (T*)in[POM_EL];
    x.elb = (T*)in[POM_ELB];
    x.d = (T*)in[POM_D];
    x.egf = egf0;
    x.utf = utf0;
    x.vtf = vtf0;
    x.etf = (T*)in[POM_ETF];
    x.advua = advua0;
    x.advva = advva0;
    x.fluxua = fluxua0;
    x.fluxva = fluxva0;
    x.tps = tps0;
    x.uaf = (T*)in[POM_UAF];
    x.vaf = (T*)in[POM_VAF];
    x.elf = (T*)in[POM_ELF];
    return x;
  }
  // resolve carried advave fields of substep n
  void resolve(int n, X<T>& x) const {
    if (!x.advua) {
      const X<T>& p = n == 1 ? x0_ : xs[n - 1];
      x.advua = p.advua;
      x.advva = p.advva;
      x.fluxua = p.fluxua;
      x.fluxva = p.fluxva;
      x.tps = p.tps;
    }
  }
  X<T> x0_;
  // The host-side pointers the forward settles on its way (x0_ and the
  // carried advave fields of every substep), for a reverse that runs without
  // its forward in the same call; they depend on the plan only.
  void replay_host() {
    x0_ = x0();
    if (!imp)
      for (int n = 1; n <= isplit; ++n) resolve(n, xs[n]);
  }

  // ------------------------------------------------------------ field tables
  FieldTab<T> t_advct() const {
    FieldTab<T> t{};
    t.f[OpAdvct::U] = in[POM_U];
    t.f[OpAdvct::V] = in[POM_V];
    t.f[OpAdvct::UB] = in[POM_UB];
    t.f[OpAdvct::VB] = in[POM_VB];
    t.f[OpAdvct::AAM] = in[POM_AAM];
    t.f[OpAdvct::DT] = in[POM_DT];
    t.g[OpAdvct::GDX] = g[POM_G_DX];
    t.g[OpAdvct::GDY] = g[POM_G_DY];
    t.g[OpAdvct::GARU] = g[POM_G_ARU];
    t.g[OpAdvct::GARV] = g[POM_G_ARV];
    return t;
  }
  FieldTab<T> t_smag() const {
    FieldTab<T> t{};
    t.f[OpSmag::U] = in[POM_U];
    t.f[OpSmag::V] = in[POM_V];
    t.f[OpSmag::AAM] = in[POM_AAM];
    t.g[OpSmag::GDX] = g[POM_G_DX];
    t.g[OpSmag::GDY] = g[POM_G_DY];
    return t;
  }
  FieldTab<T> t_advave(const X<T>& x) const {
    FieldTab<T> t{};
    t.f[OpAdvave::UA] = x.ua;
    t.f[OpAdvave::VA] = x.va;
    t.f[OpAdvave::UAB] = x.uab;
    t.f[OpAdvave::VAB] = x.vab;
    t.f[OpAdvave::D_] = x.d;
    t.f[OpAdvave::AAM2D] = out[POM_AAM2D];
    t.g[OpAdvave::GDX] = g[POM_G_DX];
    t.g[OpAdvave::GDY] = g[POM_G_DY];
    t.g[OpAdvave::GFLUXUA] = in[POM_FLUXUA];
    t.g[OpAdvave::GFLUXVA] = in[POM_FLUXVA];
    t.g[OpAdvave::GTPS] = in[POM_TPS];
    return t;
  }
  FieldTab<T> t_prep() const {
    FieldTab<T> t{};
    t.f[OpExtPrep::ADX2D] = adx2d_raw;
    t.f[OpExtPrep::ADY2D] = ady2d_raw;
    t.f[OpExtPrep::ADVUA] = advua0;
    t.f[OpExtPrep::ADVVA] = advva0;
    t.f[OpExtPrep::EL] = in[POM_EL];
    t.f[OpExtPrep::UA] = in[POM_UA];
    t.f[OpExtPrep::VA] = in[POM_VA];
    t.f[OpExtPrep::D_] = in[POM_D];
    t.f[OpExtPrep::UTF] = in[POM_UTF];
    t.f[OpExtPrep::VTF] = in[POM_VTF];
    return t;
  }
  FieldTab<T> t_uv(const X<T>& xp, const X<T>& xn) const {
    FieldTab<T> t{};
    t.f[OpExtUV::EL] = xp.el;
    t.f[OpExtUV::ELB] = xp.elb;
    t.f[OpExtUV::ELF] = xn.elf;
    t.f[OpExtUV::D_] = xp.d;
    t.f[OpExtUV::UA] = xp.ua;
    t.f[OpExtUV::VA] = xp.va;
    t.f[OpExtUV::UAB] = xp.uab;
    t.f[OpExtUV::VAB] = xp.vab;
    t.f[OpExtUV::ADVUA] = xn.advua;
    t.f[OpExtUV::ADVVA] = xn.advva;
    t.f[OpExtUV::ADX2D] = out[POM_ADX2D];
    t.f[OpExtUV::ADY2D] = out[POM_ADY2D];
    t.f[OpExtUV::DRX2D] = out[POM_DRX2D];
    t.f[OpExtUV::DRY2D] = out[POM_DRY2D];
    t.f[OpExtUV::WUBOT] = in[POM_WUBOT];
    t.f[OpExtUV::WVBOT] = in[POM_WVBOT];
    t.f[OpExtUV::UAFIN] = xp.uaf;
    t.f[OpExtUV::VAFIN] = xp.vaf;
    t.g[OpExtUV::GCOR] = g[POM_G_COR];
    t.g[OpExtUV::GDX] = g[POM_G_DX];
    t.g[OpExtUV::GDY] = g[POM_G_DY];
    t.g[OpExtUV::GH] = g[POM_G_H];
    t.g[OpExtUV::GARU] = g[POM_G_ARU];
    t.g[OpExtUV::GARV] = g[POM_G_ARV];
    t.g[OpExtUV::GDUM] = g[POM_G_DUM];
    t.g[OpExtUV::GDVM] = g[POM_G_DVM];
    t.g[OpExtUV::GEATM] = out[POM_E_ATMOS];
    t.g[OpExtUV::GWUSURF] = g[POM_G_WUSURF];
    t.g[OpExtUV::GWVSURF] = g[POM_G_WVSURF];
    return t;
  }
  FieldTab<T> t_filter(const X<T>& xp, const X<T>& xn) const {
    FieldTab<T> t{};
    t.f[OpExtFilter::UA] = xp.ua;
    t.f[OpExtFilter::UAB] = xp.uab;
    t.f[OpExtFilter::VA] = xp.va;
    t.f[OpExtFilter::VAB] = xp.vab;
    t.f[OpExtFilter::EL] = xp.el;
    t.f[OpExtFilter::ELB] = xp.elb;
    t.f[OpExtFilter::UAF] = xn.uaf;
    t.f[OpExtFilter::VAF] = xn.vaf;
    t.f[OpExtFilter::ELF] = xn.elf;
    t.f[OpExtFilter::EGF] = xp.egf;
    t.f[OpExtFilter::UTF] = xp.utf;
    t.f[OpExtFilter::VTF] = xp.vtf;
    t.f[OpExtFilter::ETF] = xp.etf;
    t.g[OpExtFilter::GH] = g[POM_G_H];
    t.g[OpExtFilter::GFSM] = g[POM_G_FSM];
    return t;
  }
  const X<T>& xf() const { return imp ? xi : xs[isplit]; }
  EllGeom<T> ell() const {
    return EllGeom<T>{g[POM_G_DX], g[POM_G_DY], g[POM_G_ART], g[POM_G_DUM], g[POM_G_DVM],
                      4.0 * (c.prm.dti * c.prm.dti) * c.prm.grav * c.prm.alpha,
                      open ? g[POM_G_ETA_FIXED] : nullptr};
  }
  // predictor values the transport and the corrector read (after bcond 2)
  const T* usr() const { return open ? us2 : us; }
  const T* vsr() const { return open ? vs2 : vs; }
  // ---- semi-implicit tables
  FieldTab<T> t_pred() const {
    FieldTab<T> t{};
    t.f[OpImpPred::UA] = in[POM_UA]; t.f[OpImpPred::VA] = in[POM_VA]; t.f[OpImpPred::EL] = in[POM_EL];
    t.f[OpImpPred::ELB] = in[POM_ELB]; t.f[OpImpPred::D_] = in[POM_D]; t.f[OpImpPred::UAB] = in[POM_UAB];
    t.f[OpImpPred::VAB] = in[POM_VAB]; t.f[OpImpPred::ADX2D] = out[POM_ADX2D];
    t.f[OpImpPred::ADY2D] = out[POM_ADY2D]; t.f[OpImpPred::ADVUA] = advua0; t.f[OpImpPred::ADVVA] = advva0;
    t.f[OpImpPred::DRX2D] = out[POM_DRX2D]; t.f[OpImpPred::DRY2D] = out[POM_DRY2D];
    t.f[OpImpPred::WUBOT] = in[POM_WUBOT]; t.f[OpImpPred::WVBOT] = in[POM_WVBOT];
    t.g[OpImpPred::GCOR] = g[POM_G_COR]; t.g[OpImpPred::GDX] = g[POM_G_DX]; t.g[OpImpPred::GDY] = g[POM_G_DY];
    t.g[OpImpPred::GH] = g[POM_G_H]; t.g[OpImpPred::GARU] = g[POM_G_ARU]; t.g[OpImpPred::GARV] = g[POM_G_ARV];
    t.g[OpImpPred::GDUM] = g[POM_G_DUM]; t.g[OpImpPred::GDVM] = g[POM_G_DVM];
    t.g[OpImpPred::GEATM] = out[POM_E_ATMOS]; t.g[OpImpPred::GWUSURF] = g[POM_G_WUSURF];
    t.g[OpImpPred::GWVSURF] = g[POM_G_WVSURF];
    return t;
  }
  FieldTab<T> t_rhs() const {
    FieldTab<T> t{};
    t.f[OpImpRhs::US] = usr(); t.f[OpImpRhs::VS] = vsr(); t.f[OpImpRhs::D_] = in[POM_D];
    t.f[OpImpRhs::ELB] = in[POM_ELB];
    t.g[OpImpRhs::GDX] = g[POM_G_DX]; t.g[OpImpRhs::GDY] = g[POM_G_DY]; t.g[OpImpRhs::GART] = g[POM_G_ART];
    t.g[OpImpRhs::GVFLUXF] = out[POM_VFLUXF];
    return t;
  }
  FieldTab<T> t_surf() const {
    FieldTab<T> t{};
    t.f[OpImpSurf::RAW] = elfraw;
    t.g[OpImpSurf::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_corr() const {
    FieldTab<T> t{};
    t.f[OpImpCorr::US] = usr(); t.f[OpImpCorr::VS] = vsr(); t.f[OpImpCorr::DUS] = dus;
    t.f[OpImpCorr::DVS] = dvs;
    t.f[OpImpCorr::ELF] = out[POM_ELF];
    t.g[OpImpCorr::GDX] = g[POM_G_DX]; t.g[OpImpCorr::GDY] = g[POM_G_DY]; t.g[OpImpCorr::GH] = g[POM_G_H];
    t.g[OpImpCorr::GARU] = g[POM_G_ARU]; t.g[OpImpCorr::GARV] = g[POM_G_ARV];
    t.g[OpImpCorr::GDUM] = g[POM_G_DUM]; t.g[OpImpCorr::GDVM] = g[POM_G_DVM];
    return t;
  }
  // radiation bcond 2 on interior values ui, vi; corners from ue, ve
  // dd: the depth d (predictor) or the new surface elf, with d = h + elf (corrector)
  FieldTab<T> t_rad(const T* ui, const T* vi, const T* ue, const T* ve, const T* dd) const {
    FieldTab<T> t{};
    t.f[OpImpRad::UI] = ui; t.f[OpImpRad::VI] = vi; t.f[OpImpRad::UE] = ue; t.f[OpImpRad::VE] = ve;
    t.f[OpImpRad::EL] = in[POM_EL]; t.f[OpImpRad::D_] = dd;
    t.g[OpImpRad::GDUM] = g[POM_G_DUM]; t.g[OpImpRad::GDVM] = g[POM_G_DVM]; t.g[OpImpRad::GH] = g[POM_G_H];
    return t;
  }
  FieldTab<T> t_surf_tide() const {
    FieldTab<T> t{};
    t.f[OpImpSurfTide::RAW] = elfraw;
    t.g[OpImpSurfTide::GFSM] = g[POM_G_FSM];
    t.g[OpImpSurfTide::GNODE] = g[POM_G_TIDE_NODE];
    t.g[OpImpSurfTide::GLIFT] = lift;
    return t;
  }
  FieldTab<T> t_ifilter() const {
    FieldTab<T> t{};
    t.f[OpImpFilter::UA] = in[POM_UA]; t.f[OpImpFilter::UAB] = in[POM_UAB]; t.f[OpImpFilter::VA] = in[POM_VA];
    t.f[OpImpFilter::VAB] = in[POM_VAB]; t.f[OpImpFilter::EL] = in[POM_EL]; t.f[OpImpFilter::ELB] = in[POM_ELB];
    t.f[OpImpFilter::UAF] = out[POM_UAF]; t.f[OpImpFilter::VAF] = out[POM_VAF];
    t.f[OpImpFilter::ELF] = out[POM_ELF]; t.f[OpImpFilter::UTF0] = utf0; t.f[OpImpFilter::VTF0] = vtf0;
    t.g[OpImpFilter::GH] = g[POM_G_H]; t.g[OpImpFilter::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_ell() const {
    FieldTab<T> t{};
    t.f[OpEll::ETA] = elfraw; t.f[OpEll::D_] = in[POM_D];
    t.g[OpEll::GDX] = g[POM_G_DX]; t.g[OpEll::GDY] = g[POM_G_DY]; t.g[OpEll::GART] = g[POM_G_ART];
    t.g[OpEll::GDUM] = g[POM_G_DUM]; t.g[OpEll::GDVM] = g[POM_G_DVM];
    return t;
  }
  FieldTab<T> t_ibc6() const {
    FieldTab<T> t{};
    t.f[OpImpBc6Filter::Q2F] = q2f1; t.f[OpImpBc6Filter::Q2LF] = q2lf1; t.f[OpImpBc6Filter::Q2] = in[POM_Q2];
    t.f[OpImpBc6Filter::Q2L] = in[POM_Q2L]; t.f[OpImpBc6Filter::Q2B] = in[POM_Q2B];
    t.f[OpImpBc6Filter::Q2LB] = in[POM_Q2LB];
    t.g[OpImpBc6Filter::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_ibc4() const {
    FieldTab<T> t{};
    t.f[OpImpBc4Filter::TF] = tf1; t.f[OpImpBc4Filter::SF] = sf1; t.f[OpImpBc4Filter::TT] = in[POM_T];
    t.f[OpImpBc4Filter::SS] = in[POM_S]; t.f[OpImpBc4Filter::TB] = in[POM_TB]; t.f[OpImpBc4Filter::SB] = in[POM_SB];
    t.g[OpImpBc4Filter::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_shap(const T* x, const T* mask) const {
    FieldTab<T> t{};
    t.f[OpShapiro<false>::X] = x;
    t.g[OpShapiro<false>::GMASK] = mask;
    return t;
  }
  FieldTab<T> t_ibc3() const {
    FieldTab<T> t{};
    t.f[OpImpBc3::UF] = uf1; t.f[OpImpBc3::VF] = vf1;
    t.g[OpImpBc3::GDUM] = g[POM_G_DUM]; t.g[OpImpBc3::GDVM] = g[POM_G_DVM];
    return t;
  }
  FieldTab<T> t_smag_imp() const {
    FieldTab<T> t{};
    t.f[OpSmagImp::U] = in[POM_U]; t.f[OpSmagImp::V] = in[POM_V]; t.f[OpSmagImp::AAM] = in[POM_AAM];
    t.g[OpSmagImp::GDX] = g[POM_G_DX]; t.g[OpSmagImp::GDY] = g[POM_G_DY];
    return t;
  }
  // the fields advt1 and advq share between their two tracers (t_advt1, t_advq)
  void adv_shared(AdvPairArgs<T>& a) const {
    a.u = u1; a.v = v1; a.w = out[POM_W]; a.aam = out[POM_AAM];
    a.dt = in[POM_DT]; a.etb = in[POM_ETB]; a.etf = xf().etf;
    a.dx = g[POM_G_DX]; a.dy = g[POM_G_DY]; a.h = g[POM_G_H]; a.art = g[POM_G_ART];
    a.dum = g[POM_G_DUM]; a.dvm = g[POM_G_DVM];
  }
  FieldTab<T> t_advq(bool l2) const {
    FieldTab<T> t{};
    t.f[OpAdvq::QB] = in[l2 ? POM_Q2LB : POM_Q2B];
    t.f[OpAdvq::Q] = in[l2 ? POM_Q2L : POM_Q2];
    t.f[OpAdvq::U] = u1;
    t.f[OpAdvq::V] = v1;
    t.f[OpAdvq::W] = out[POM_W];
    t.f[OpAdvq::AAM] = out[POM_AAM];
    t.f[OpAdvq::DT] = in[POM_DT];
    t.f[OpAdvq::ETB] = in[POM_ETB];
    t.f[OpAdvq::ETF] = xf().etf;
    t.g[OpAdvq::GDX] = g[POM_G_DX];
    t.g[OpAdvq::GDY] = g[POM_G_DY];
    t.g[OpAdvq::GH] = g[POM_G_H];
    t.g[OpAdvq::GART] = g[POM_G_ART];
    t.g[OpAdvq::GDUM] = g[POM_G_DUM];
    t.g[OpAdvq::GDVM] = g[POM_G_DVM];
    return t;
  }
  FieldTab<T> t_kedge() const {
    FieldTab<T> t{};
    t.f[OpKEdge::KM] = kmraw;
    t.f[OpKEdge::KH] = khraw;
    t.g[OpKEdge::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_bc6() const {
    FieldTab<T> t{};
    t.f[OpBc6Filter::Q2F] = q2f1;
    t.f[OpBc6Filter::Q2LF] = q2lf1;
    t.f[OpBc6Filter::Q2] = in[POM_Q2];
    t.f[OpBc6Filter::Q2L] = in[POM_Q2L];
    t.f[OpBc6Filter::Q2B] = in[POM_Q2B];
    t.f[OpBc6Filter::Q2LB] = in[POM_Q2LB];
    t.f[OpBc6Filter::U] = u1;
    t.f[OpBc6Filter::V] = v1;
    t.g[OpBc6Filter::GDX] = g[POM_G_DX];
    t.g[OpBc6Filter::GDY] = g[POM_G_DY];
    t.g[OpBc6Filter::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_advt1(bool sal) const {
    FieldTab<T> t{};
    t.f[OpAdvt1::FB] = in[sal ? POM_SB : POM_TB];
    t.f[OpAdvt1::F] = in[sal ? POM_S : POM_T];
    t.f[OpAdvt1::FCLIM] = g[sal ? POM_G_SCLIM : POM_G_TCLIM];
    t.f[OpAdvt1::U] = u1;
    t.f[OpAdvt1::V] = v1;
    t.f[OpAdvt1::W] = out[POM_W];
    t.f[OpAdvt1::AAM] = out[POM_AAM];
    t.f[OpAdvt1::DT] = in[POM_DT];
    t.f[OpAdvt1::ETB] = in[POM_ETB];
    t.f[OpAdvt1::ETF] = xf().etf;
    t.g[OpAdvt1::GDX] = g[POM_G_DX];
    t.g[OpAdvt1::GDY] = g[POM_G_DY];
    t.g[OpAdvt1::GH] = g[POM_G_H];
    t.g[OpAdvt1::GART] = g[POM_G_ART];
    t.g[OpAdvt1::GDUM] = g[POM_G_DUM];
    t.g[OpAdvt1::GDVM] = g[POM_G_DVM];
    return t;
  }
  FieldTab<T> t_bc4() const {
    FieldTab<T> t{};
    t.f[OpBc4Filter::TF] = tf1;
    t.f[OpBc4Filter::SF] = sf1;
    t.f[OpBc4Filter::TT] = in[POM_T];
    t.f[OpBc4Filter::SS] = in[POM_S];
    t.f[OpBc4Filter::TB] = in[POM_TB];
    t.f[OpBc4Filter::SB] = in[POM_SB];
    t.f[OpBc4Filter::U] = u1;
    t.f[OpBc4Filter::V] = v1;
    t.f[OpBc4Filter::W] = out[POM_W];
    t.f[OpBc4Filter::DT] = in[POM_DT];
    t.g[OpBc4Filter::GDX] = g[POM_G_DX];
    t.g[OpBc4Filter::GDY] = g[POM_G_DY];
    t.g[OpBc4Filter::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_dens() const {
    FieldTab<T> t{};
    t.f[OpDens::TT] = out[POM_T];
    t.f[OpDens::SS] = out[POM_S];
    t.f[OpDens::RHO] = rho1;
    t.g[OpDens::GH] = g[POM_G_H];
    t.g[OpDens::GFSM] = g[POM_G_FSM];
    return t;
  }
  FieldTab<T> t_advu() const {
    FieldTab<T> t{};
    t.f[OpAdvu::ADVX] = out[POM_ADVX];
    t.f[OpAdvu::DRHOX] = out[POM_DRHOX];
    t.f[OpAdvu::U] = u1;
    t.f[OpAdvu::V] = v1;
    t.f[OpAdvu::W] = out[POM_W];
    t.f[OpAdvu::UB] = in[POM_UB];
    t.f[OpAdvu::DT] = in[POM_DT];
    t.f[OpAdvu::EGF] = xf().egf;
    t.f[OpAdvu::EGB] = in[POM_EGB];
    t.f[OpAdvu::ETF] = xf().etf;
    t.f[OpAdvu::ETB] = in[POM_ETB];
    t.g[OpAdvu::GARU] = g[POM_G_ARU];
    t.g[OpAdvu::GCOR] = g[POM_G_COR];
    t.g[OpAdvu::GDY] = g[POM_G_DY];
    t.g[OpAdvu::GH] = g[POM_G_H];
    t.g[OpAdvu::GEATM] = out[POM_E_ATMOS];
    return t;
  }
  FieldTab<T> t_advv() const {
    FieldTab<T> t{};
    t.f[OpAdvv::ADVY] = out[POM_ADVY];
    t.f[OpAdvv::DRHOY] = out[POM_DRHOY];
    t.f[OpAdvv::U] = u1;
    t.f[OpAdvv::V] = v1;
    t.f[OpAdvv::W] = out[POM_W];
    t.f[OpAdvv::VB] = in[POM_VB];
    t.f[OpAdvv::DT] = in[POM_DT];
    t.f[OpAdvv::EGF] = xf().egf;
    t.f[OpAdvv::EGB] = in[POM_EGB];
    t.f[OpAdvv::ETF] = xf().etf;
    t.f[OpAdvv::ETB] = in[POM_ETB];
    t.g[OpAdvv::GARV] = g[POM_G_ARV];
    t.g[OpAdvv::GCOR] = g[POM_G_COR];
    t.g[OpAdvv::GDX] = g[POM_G_DX];
    t.g[OpAdvv::GH] = g[POM_G_H];
    t.g[OpAdvv::GEATM] = out[POM_E_ATMOS];
    return t;
  }
  FieldTab<T> t_bc3() const {
    FieldTab<T> t{};
    t.f[OpBc3::UF] = uf1;
    t.f[OpBc3::VF] = vf1;
    t.f[OpBc3::U] = u1;
    t.f[OpBc3::V] = v1;
    t.g[OpBc3::GH] = g[POM_G_H];
    t.g[OpBc3::GDUM] = g[POM_G_DUM];
    t.g[OpBc3::GDVM] = g[POM_G_DVM];
    return t;
  }
  FieldTab<T> t_bc3d() const {   // edge weight from the depth after the barotropic step
    FieldTab<T> t = t_bc3();
    t.f[OpBc3D::D_] = xf().d;
    return t;
  }
  FieldTab<T> t_bc4ext() const {
    FieldTab<T> t = t_bc4();
    t.f[OpBc4FilterExt::TX] = g[POM_G_TBND];
    t.f[OpBc4FilterExt::SX] = g[POM_G_SBND];
    t.f[OpBc4FilterExt::DT] = in[POM_DT];
    return t;
  }

  ProfqArgs<T> a_profq() const {
    ProfqArgs<T> s{};
    s.q2f = q2f0;
    s.q2lf = q2lf0;
    s.q2 = in[POM_Q2];
    s.q2b = in[POM_Q2B];
    s.q2lb = in[POM_Q2LB];
    s.km = in[POM_KM];
    s.kh = in[POM_KH];
    s.kq = in[POM_KQ];
    s.u = u1;
    s.v = v1;
    s.t = in[POM_T];
    s.s = in[POM_S];
    s.rho = rho1;
    s.wusurf = g[POM_G_WUSURF];
    s.wvsurf = g[POM_G_WVSURF];
    s.wubot = in[POM_WUBOT];
    s.wvbot = in[POM_WVBOT];
    s.etf = xf().etf;
    s.dt = in[POM_DT];
    s.h = g[POM_G_H];
    s.q2o = q2f1;
    s.q2lo = q2lf1;
    s.lo = out[POM_L];
    s.kqo = out[POM_KQ];
    s.kmo = kmraw;
    s.kho = khraw;
    s.tape = qt;   // null unless the adjoint's recompute
    return s;
  }
  ProftArgs<T> a_proft(bool sal) const {
    ProftArgs<T> s{};
    s.f = sal ? sf0 : tf0;
    s.kh = out[POM_KH];
    s.etf = xf().etf;
    s.h = g[POM_G_H];
    s.wfsurf = out[sal ? POM_WSSURF : POM_WTSURF];
    s.fsurf = g[sal ? POM_G_SSURF : POM_G_TSURF];
    s.out = sal ? sf1 : tf1;
    s.nbc = sal ? c.prm.nbcs : c.prm.nbct;
    return s;
  }
  ProfuvArgs<T> a_profuv(bool vdir) const {
    ProfuvArgs<T> s{};
    s.f = vdir ? vf0 : uf0;
    s.km = out[POM_KM];
    s.etf = xf().etf;
    s.h = g[POM_G_H];
    s.wsurf = g[vdir ? POM_G_WVSURF : POM_G_WUSURF];
    s.wbot = in[vdir ? POM_WVBOT : POM_WUBOT];
    s.cbc = g[POM_G_CBC];
    s.ub = in[POM_UB];
    s.vb = in[POM_VB];
    s.mask = g[vdir ? POM_G_DVM : POM_G_DUM];
    s.out = vdir ? vf1 : uf1;
    s.wbot_out = out[vdir ? POM_WVBOT : POM_WUBOT];
    s.tps_out = vdir ? out[POM_TPS] : tpsu;
    return s;
  }
  VertvlArgs<T> a_vertvl() const {
    VertvlArgs<T> s{};
    s.u = u1;
    s.v = v1;
    s.w = in[POM_W];
    s.dt = in[POM_DT];
    s.etf = xf().etf;
    s.etb = in[POM_ETB];
    s.vfluxb = in[POM_VFLUXB];
    s.vfluxf = out[POM_VFLUXF];
    s.dx = g[POM_G_DX];
    s.dy = g[POM_G_DY];
    s.fsm = g[POM_G_FSM];
    s.wout = out[POM_W];
    s.all_levels = imp && !open;
    return s;
  }

  // ------------------------------------------------------------ forward
  void forward() {
    const dim3 g2(c.nblk2), tb(TPB);
    if (vamax) hipLaunchKernelGGL(k_fill<T>, dim3(1), dim3(64), 0, c.stream, vamax, (int64_t)1, T(0));
    {   // surface forcing reset (time_loop.py:118-164)
      ForcingArgs<T> a{in[POM_E_ATMOS], in[POM_VFLUXF], in[POM_SWRAD], in[POM_WTSURF], in[POM_WSSURF],
                       in[POM_T], in[POM_S], out[POM_E_ATMOS], out[POM_VFLUXF], out[POM_SWRAD],
                       out[POM_WTSURF], out[POM_WSSURF]};
      hipLaunchKernelGGL(k_forcing<T>, g2, tb, 0, c.stream, c.B, c.prm, a);
    }
    // 3-D preparation (time_loop.py:169-270)
    {
      ProfScope ps(POM_STAGE_ADVCT, c.stream);
      launch_map<T, OpAdvct>(c, t_advct(), OutTab<T>{{out[POM_ADVX], out[POM_ADVY]}});
    }
    {
      BaropgArgs<T> a{in[POM_RHO], g[POM_G_RMEAN], in[POM_DT], g[POM_G_DX], g[POM_G_DY], g[POM_G_DUM],
                      g[POM_G_DVM], rho1, out[POM_DRHOX], out[POM_DRHOY]};
      hipLaunchKernelGGL(k_baropg<T>, g2, tb, 0, c.stream, c.B, c.vert, c.ramp, c.prm, a);
    }
    if (imp) launch_map<T, OpSmagImp>(c, t_smag_imp(), OutTab<T>{{out[POM_AAM]}});
    else launch_map<T, OpSmag>(c, t_smag(), OutTab<T>{{out[POM_AAM]}});
    {
      ColsumArgs<T> a{};
      a.f[0] = out[POM_ADVX]; a.out[0] = adx2d_raw;
      a.f[1] = out[POM_ADVY]; a.out[1] = ady2d_raw;
      a.f[2] = out[POM_DRHOX]; a.out[2] = out[POM_DRX2D];
      a.f[3] = out[POM_DRHOY]; a.out[3] = out[POM_DRY2D];
      a.f[4] = out[POM_AAM]; a.out[4] = out[POM_AAM2D];
      a.n = 5;
      hipLaunchKernelGGL(k_colsum<T>, g2, tb, 0, c.stream, c.B, c.vert, a);
    }
    x0_ = x0();
    x0_.fluxua = (T*)in[POM_FLUXUA];
    x0_.fluxva = (T*)in[POM_FLUXVA];
    x0_.tps = (T*)in[POM_TPS];
    {
      X<T> xin = x0_;
      launch_map<T, OpAdvave>(c, t_advave(xin), OutTab<T>{{advua0, advva0, fluxua0, fluxva0, tps0}});
    }
    x0_ = x0();
    launch_map<T, OpExtPrep>(c, t_prep(), OutTab<T>{{out[POM_ADX2D], out[POM_ADY2D], egf0, utf0, vtf0}});
    if (imp) {
      external_implicit();
      if (pw.status != POM_OK) return;
    } else {
      external_explicit();
    }
    internal_and_shifts();
  }

  // semi-implicit barotropic step (implicit_p0_fixed/time_loop.py:247-398)
  void external_implicit() {
    const int64_t N2 = n2();
    launch_map<T, OpImpPred>(c, t_pred(), OutTab<T>{{us, vs, dus, dvs}});
    if (open) launch_map<T, OpImpRad>(c, t_rad(us, vs, in[POM_UA], in[POM_VA], in[POM_D]), OutTab<T>{{us2, vs2}});
    launch_map<T, OpImpRhs>(c, t_rhs(), OutTab<T>{{out[POM_FLUXUA], out[POM_FLUXVA], rhs}});
    if (open)
      hipLaunchKernelGGL(k_tide_lift<T>, dim3(c.nblk2), dim3(TPB), 0, c.stream, c.B, g[POM_G_TIDE_NODE],
                         g[POM_G_TIDE_AMP], g[POM_G_TIDE_PHASE], tide_wt, lift);
    hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(256), 0, c.stream, pw.tol, (int64_t)b->M, hp->eta_tol);
    pw.solve(c, ell(), rhs, in[POM_D], in[POM_ELF], elfraw, hp->eta_rel_tol, hp->eta_max_iter, lift);
    if (pw.status != POM_OK) return;
    if (iters_out)
      hipLaunchKernelGGL(k_copy<int32_t>, dim3(1), dim3(256), 0, c.stream, iters_out, pw.its, (int64_t)b->M);
    if (open) {
      launch_map<T, OpImpSurfTide>(c, t_surf_tide(), OutTab<T>{{out[POM_ELF]}});
      launch_map<T, OpImpCorr>(c, t_corr(), OutTab<T>{{uac, vac}});
      launch_map<T, OpImpRadNew>(c, t_rad(uac, vac, in[POM_UAF], in[POM_VAF], out[POM_ELF]),
                                 OutTab<T>{{out[POM_UAF], out[POM_VAF]}});
    } else {
      launch_map<T, OpImpSurf>(c, t_surf(), OutTab<T>{{out[POM_ELF]}});
      launch_map<T, OpImpCorr>(c, t_corr(), OutTab<T>{{out[POM_UAF], out[POM_VAF]}});
    }
    if (vamax)
      hipLaunchKernelGGL(k_vamax<T>, dim3(std::min(blocks(N2), 1024)), dim3(256), 0, c.stream, out[POM_VAF], N2,
                         vamax);
    launch_map<T, OpImpFilter>(c, t_ifilter(),
                               OutTab<T>{{out[POM_UA], out[POM_UAB], out[POM_VA], out[POM_VAB], out[POM_EL],
                                          out[POM_ELB], out[POM_D], out[POM_EGF], out[POM_UTF], out[POM_VTF],
                                          out[POM_ETF]}});
    copy(out[POM_ADVUA], advua0, N2);
    copy(out[POM_ADVVA], advva0, N2);
  }

  // explicit barotropic substeps (explicit/time_loop.py:292-431)
  void external_explicit() {
    const int64_t N2 = n2();
    for (int n = 1; n <= isplit; ++n) {
      const X<T>& xp = n == 1 ? x0_ : xs[n - 1];
      X<T>& xn = xs[n];
      Params prm = c.prm;
      prm.iext = n;
      Ctx<T> cn = c;
      cn.prm = prm;
      ExtElfArgs<T> ea{xp.ua, xp.va, xp.d, xp.elb, g[POM_G_DX], g[POM_G_DY], g[POM_G_ART], out[POM_VFLUXF],
                       g[POM_G_FSM], xn.elf};
      if (open) {   // the adapter's time of substep n (coastal_boundary.py:14)
        ea.node = g[POM_G_TIDE_NODE];
        ea.amp = g[POM_G_TIDE_AMP];
        ea.phase = g[POM_G_TIDE_PHASE];
        ea.elf_prev = xp.elf;
        const double sec = hp->tide_sec + hp->dte * double(hp->tide_substep0 + n - 1);
        ea.wt = hp->tide_coef * sec / 3600.0;
        launch_ext_elf(cn, ea);   // continuity with the tide's bcond 1
      }
      if (n % c.prm.ispadv == 0) {   // reads the substep's inputs only
        launch_map<T, OpAdvave>(cn, t_advave(xp), OutTab<T>{{xn.advua, xn.advva, xn.fluxua, xn.fluxva, xn.tps}});
      }
      resolve(n, xn);
      {   // depth-mean momentum, then its filter / shift / time means at the same point, and max |vaf|
        ProfScope ps(POM_STAGE_EXTUV, c.stream);
        // ua, va, el are not stored where they share uaf's, vaf's, elf's array
        const OutTab<T> of{{xn.ua == xn.uaf ? nullptr : xn.ua, xn.uab, xn.va == xn.vaf ? nullptr : xn.va, xn.vab,
                            xn.el == xn.elf ? nullptr : xn.el, xn.elb, xn.d, xn.egf, xn.utf, xn.vtf, xn.etf}};
        if (open)   // the tide's bcond 2 reads elf beyond the three points the fused pass forms
          launch_map2<T, OpExtUVTide, OpExtFilter, OpExtFilter::UAF, OpExtFilter::VAF>(
              cn, t_uv(xp, xn), OutTab<T>{{xn.uaf, xn.vaf}}, t_filter(xp, xn), of, vamax);
        else   // continuity and bcond 1 fused in: elf recomputed where it is read
          launch_ext_substep(cn, ea, t_uv(xp, xn), OutTab<T>{{xn.uaf, xn.vaf}}, t_filter(xp, xn), of, vamax);
      }
    }
    const X<T>& xl = xs[isplit];
    copy(out[POM_ADVUA], xl.advua, N2);
    copy(out[POM_ADVVA], xl.advva, N2);
    copy(out[POM_FLUXUA], xl.fluxua, N2);
    copy(out[POM_FLUXVA], xl.fluxva, N2);
  }

  void internal_and_shifts() {
    const int64_t N2 = n2(), N3 = n3();
    const dim3 g2(c.nblk2), g3(c.nblk3), tb(TPB);
    const X<T>& xl = xf();
    if (internal) {
      {
        AdjustArgs<T> a{in[POM_U], in[POM_V], in[POM_UTB], xl.utf, in[POM_VTB], xl.vtf, in[POM_DT], u1, v1};
        hipLaunchKernelGGL(k_adjust<T>, g2, tb, 0, c.stream, c.B, c.vert, a);
      }
      hipLaunchKernelGGL(k_vertvl<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a_vertvl());
      {   // advq of q2 and q2l in one pass
        AdvPairArgs<T> a{};
        a.fb[0] = in[POM_Q2B]; a.f[0] = in[POM_Q2]; a.fb[1] = in[POM_Q2LB]; a.f[1] = in[POM_Q2L];
        adv_shared(a);
        a.out[0] = q2f0; a.out[1] = q2lf0;
        launch_advq_pair(c, a);
      }
      {
        ProfScope ps(POM_STAGE_PROFQ, c.stream);
        hipLaunchKernelGGL(k_profq<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a_profq());
      }
      launch_map<T, OpKEdge>(c, t_kedge(), OutTab<T>{{out[POM_KM], out[POM_KH]}});
      if (imp && !open)
        launch_map<T, OpImpBc6Filter>(c, t_ibc6(),
                                      OutTab<T>{{out[POM_Q2], out[POM_Q2B], out[POM_Q2L], out[POM_Q2LB]}});
      else
        launch_map<T, OpBc6Filter>(c, t_bc6(), OutTab<T>{{out[POM_Q2], out[POM_Q2B], out[POM_Q2L], out[POM_Q2LB]}});
      {   // advt1 of T and S in one pass
        AdvPairArgs<T> a{};
        a.fb[0] = in[POM_TB]; a.f[0] = in[POM_T]; a.fclim[0] = g[POM_G_TCLIM];
        a.fb[1] = in[POM_SB]; a.f[1] = in[POM_S]; a.fclim[1] = g[POM_G_SCLIM];
        adv_shared(a);
        a.out[0] = tf0; a.out[1] = sf0;
        launch_advt1_pair(c, a);
      }
      hipLaunchKernelGGL(k_proft<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a_proft(false));
      hipLaunchKernelGGL(k_proft<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a_proft(true));
      if (open)
        launch_map<T, OpBc4FilterExt>(c, t_bc4ext(), OutTab<T>{{out[POM_T], out[POM_TB], out[POM_S], out[POM_SB]}});
      else if (imp && shap_ts) {   // numerics/double_gyre.py:78-130 with gamma > 0
        launch_map<T, OpImpBc4FilterG<false>>(c, t_ibc4(), OutTab<T>{{tsh, out[POM_TB], ssh, out[POM_SB]}});
        launch_map<T, OpShapiro<false>>(c, t_shap(tsh, g[POM_G_FSM]), OutTab<T>{{out[POM_T]}});
        launch_map<T, OpShapiro<false>>(c, t_shap(ssh, g[POM_G_FSM]), OutTab<T>{{out[POM_S]}});
      } else if (imp)
        launch_map<T, OpImpBc4Filter>(c, t_ibc4(), OutTab<T>{{out[POM_T], out[POM_TB], out[POM_S], out[POM_SB]}});
      else
        launch_map<T, OpBc4Filter>(c, t_bc4(), OutTab<T>{{out[POM_T], out[POM_TB], out[POM_S], out[POM_SB]}});
      launch_map<T, OpDens>(c, t_dens(), OutTab<T>{{out[POM_RHO]}});
      {   // advu and advv in one pass
        AdvUVArgs<T> a{};
        a.advx = out[POM_ADVX]; a.drhox = out[POM_DRHOX]; a.ub = in[POM_UB];
        a.advy = out[POM_ADVY]; a.drhoy = out[POM_DRHOY]; a.vb = in[POM_VB];
        a.u = u1; a.v = v1; a.w = out[POM_W];
        a.dt = in[POM_DT]; a.egf = xf().egf; a.egb = in[POM_EGB]; a.etf = xf().etf; a.etb = in[POM_ETB];
        a.aru = g[POM_G_ARU]; a.arv = g[POM_G_ARV]; a.cor = g[POM_G_COR]; a.dx = g[POM_G_DX]; a.dy = g[POM_G_DY];
        a.h = g[POM_G_H]; a.eatm = out[POM_E_ATMOS];
        a.uf = uf0; a.vf = vf0;
        launch_advuv(c, a);
      }
      hipLaunchKernelGGL((k_profuv<T, 0>), g2, tb, 0, c.stream, c.B, c.vert, c.prm, a_profuv(false));
      hipLaunchKernelGGL((k_profuv<T, 1>), g2, tb, 0, c.stream, c.B, c.vert, c.prm, a_profuv(true));
      if (open) launch_map<T, OpBc3D>(c, t_bc3d(), OutTab<T>{{uf2, vf2}});
      else if (imp) launch_map<T, OpImpBc3>(c, t_ibc3(), OutTab<T>{{uf2, vf2}});
      else launch_map<T, OpBc3>(c, t_bc3(), OutTab<T>{{uf2, vf2}});
      if (shap_uv) {   // numerics/double_gyre.py:132-184 with gamma > 0: filter, then mask
        UVFiltArgs<T> a{uf2, in[POM_UB], u1, ush, out[POM_U], nullptr};
        hipLaunchKernelGGL(k_uvfilt<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        UVFiltArgs<T> bv{vf2, in[POM_VB], v1, vsh, out[POM_V], nullptr};
        hipLaunchKernelGGL(k_uvfilt<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, bv);
        launch_map<T, OpShapiro<true>>(c, t_shap(ush, g[POM_G_DUM]), OutTab<T>{{out[POM_UB]}});
        launch_map<T, OpShapiro<true>>(c, t_shap(vsh, g[POM_G_DVM]), OutTab<T>{{out[POM_VB]}});
      } else {
        UVFiltArgs<T> a{uf2, in[POM_UB], u1, out[POM_UB], out[POM_U], imp ? g[POM_G_DUM] : nullptr};
        hipLaunchKernelGGL(k_uvfilt<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        UVFiltArgs<T> bv{vf2, in[POM_VB], v1, out[POM_VB], out[POM_V], imp ? g[POM_G_DVM] : nullptr};
        hipLaunchKernelGGL(k_uvfilt<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, bv);
      }
    } else {
      for (int f : {POM_U, POM_UB, POM_V, POM_VB, POM_T, POM_TB, POM_S, POM_SB, POM_Q2, POM_Q2B, POM_Q2L, POM_Q2LB,
                    POM_KM, POM_KH, POM_KQ, POM_L})
        copy(out[f], in[f], N3);
      copy(out[POM_RHO], rho1, N3);
      hipLaunchKernelGGL(k_wsurf<T>, g3, tb, 0, c.stream, c.B, in[POM_W], out[POM_VFLUXF], out[POM_W]);
      copy(out[POM_WUBOT], in[POM_WUBOT], N2);
      copy(out[POM_WVBOT], in[POM_WVBOT], N2);
      copy(out[POM_TPS], xl.tps, N2);
    }
    // shifts (time_loop.py:680-696)
    // (one launch: egb, etb, et, dt = h + et, utb, vtb, vfluxb)
    {
      SumTab<T> st{};
      const T* src[7][2] = {{xl.egf, nullptr}, {in[POM_ET], nullptr}, {xl.etf, nullptr}, {g[POM_G_H], xl.etf},
                            {xl.utf, nullptr}, {xl.vtf, nullptr}, {out[POM_VFLUXF], nullptr}};
      T* dst[7] = {out[POM_EGB], out[POM_ETB], out[POM_ET], out[POM_DT], out[POM_UTB], out[POM_VTB], out[POM_VFLUXB]};
      for (int e = 0; e < 7; ++e) {
        st.d[e] = dst[e];
        st.s[e][0] = src[e][0];
        st.s[e][1] = src[e][1];
        st.s[e][2] = nullptr;
      }
      st.n = N2;
      hipLaunchKernelGGL(k_sum_multi<T>, dim3(flat_blocks(N2), 7), dim3(256), 0, c.stream, st);
    }
  }
};

// ---------------------------------------------------------------- adjoint
template <class T> struct Adj {
  Step<T>& S;
  Ctx<T>& c;
  // Cotangents of the step outputs, read through bo.  An output that the step
  // itself reads later (T, S, km, kh, aam, w, the advection and density
  // gradient terms) collects those uses' cotangents as well: it has an
  // accumulator in bw, filled with the caller's cotangent first.  Every other
  // output's cotangent is read from the caller's array as it is, or from a
  // shared zero array when the caller passes none.
  const T* bo[POM_NSTATE];
  T* bw[POM_NSTATE];
  T *zero3, *zero2;
  T* bi[POM_NSTATE];    // cotangents of step inputs (accumulated)
  // intermediates
  T *b_rho1, *b_adx2d_raw, *b_ady2d_raw, *b_advua0, *b_advva0, *b_egf0, *b_utf0, *b_vtf0;
  X<T> bx[2];           // cotangents of substep states (current / previous)
  T *b_u1, *b_v1, *b_q2f0, *b_q2lf0, *b_q2f1, *b_q2lf1, *b_kmraw, *b_khraw, *b_tf0, *b_sf0, *b_tf1, *b_sf1;
  T *b_uf0, *b_vf0, *b_uf1, *b_vf1, *b_uf2, *b_vf2;
  // scratch
  T *s3a, *s3b, *s3c, *s3d, *s2a, *s2b, *s2c, *s2d, *s2e;
  T* xfac = nullptr;    // factor planes of the substep reverse (XF_N x n2)
  // semi-implicit barotropic step
  T *b_us = nullptr, *b_vs, *b_dus, *b_dvs, *b_raw, *b_rhs, *nlam, *arhs, *xadj;
  T *b_us0 = nullptr, *b_vs0, *b_uc, *b_vc;   // tide policy: closed predictor / corrector outputs
  T *q_bl, *q_bgh, *q_bq2ln, *q_ba, *q_bc, *q_bkn, *q_bdtef, *q_bboy, *q_bq2bot, *q_bdh;   // between profq's reverse kernels

  Adj(Step<T>& s, Ctx<T>& cc) : S(s), c(cc) {}

  static constexpr bool accumulated(int f) {
    switch (f) {
      case POM_T: case POM_S: case POM_KM: case POM_KH: case POM_AAM: case POM_W: case POM_ADVX: case POM_ADVY:
      case POM_DRHOX: case POM_DRHOY: case POM_ADX2D: case POM_ADY2D: case POM_DRX2D: case POM_DRY2D:
      case POM_AAM2D:
        return true;
      default:
        return false;
    }
  }
  // the accumulator of output F: the only writable view of an output cotangent
  template <int F> T* w() const {
    static_assert(accumulated(F), "this output's cotangent has no accumulator");
    return bw[F];
  }

  void plan(Arena& A, T* const* grad_in) {
    const int64_t N3 = S.n3(), N2 = S.n2();
    for (int f = 0; f < POM_NSTATE; ++f) {
      const int64_t n = f < POM_N3D ? N3 : N2;
      bw[f] = accumulated(f) ? A.get<T>(n) : nullptr;
      bo[f] = bw[f];
      bi[f] = grad_in && grad_in[f] ? grad_in[f] : A.get<T>(n);
    }
    zero3 = A.get<T>(N3);
    zero2 = A.get<T>(N2);
    b_rho1 = A.get<T>(N3);
    b_adx2d_raw = A.get<T>(N2);
    b_ady2d_raw = A.get<T>(N2);
    b_advua0 = A.get<T>(N2);
    b_advva0 = A.get<T>(N2);
    b_egf0 = A.get<T>(N2);
    b_utf0 = A.get<T>(N2);
    b_vtf0 = A.get<T>(N2);
    // advua / advva are carried unchanged between advave calls, so their
    // cotangent is one accumulator shared by both substep buffers: the
    // substeps that read a carried value add into it, and the advave call
    // that produced the value consumes it and clears it
    T* b_advua_c = A.get<T>(N2);
    T* b_advva_c = A.get<T>(N2);
    // egf, utf, vtf and etf: a substep passes their cotangents on unchanged
    // (etf's is scaled in place on the last three), so the two substep
    // buffers share one array each
    T* b_egf_c = A.get<T>(N2);
    T* b_utf_c = A.get<T>(N2);
    T* b_vtf_c = A.get<T>(N2);
    T* b_etf_c = A.get<T>(N2);
    for (int q = 0; q < 2; ++q) {
      X<T>& x = bx[q];
      x.ua = A.get<T>(N2); x.uab = A.get<T>(N2); x.va = A.get<T>(N2); x.vab = A.get<T>(N2);
      x.el = A.get<T>(N2); x.elb = A.get<T>(N2); x.d = A.get<T>(N2); x.egf = b_egf_c;
      x.utf = b_utf_c; x.vtf = b_vtf_c; x.etf = b_etf_c; x.advua = b_advua_c;
      x.advva = b_advva_c; x.uaf = A.get<T>(N2); x.vaf = A.get<T>(N2); x.elf = A.get<T>(N2);
      x.fluxua = x.fluxva = x.tps = nullptr;
    }
    b_u1 = A.get<T>(N3); b_v1 = A.get<T>(N3); b_q2f0 = A.get<T>(N3); b_q2lf0 = A.get<T>(N3);
    b_q2f1 = A.get<T>(N3); b_q2lf1 = A.get<T>(N3); b_kmraw = A.get<T>(N3); b_khraw = A.get<T>(N3);
    b_tf0 = A.get<T>(N3); b_sf0 = A.get<T>(N3); b_tf1 = A.get<T>(N3); b_sf1 = A.get<T>(N3);
    b_uf0 = A.get<T>(N3); b_vf0 = A.get<T>(N3); b_uf1 = A.get<T>(N3); b_vf1 = A.get<T>(N3);
    b_uf2 = A.get<T>(N3); b_vf2 = A.get<T>(N3);
    s3a = A.get<T>(N3); s3b = A.get<T>(N3); s3c = A.get<T>(N3); s3d = A.get<T>(N3);
    q_bl = A.get<T>(N3); q_bgh = A.get<T>(N3); q_bq2ln = A.get<T>(N3); q_ba = A.get<T>(N3);
    q_bc = A.get<T>(N3); q_bkn = A.get<T>(N3); q_bdtef = A.get<T>(N3); q_bboy = A.get<T>(N3);
    q_bq2bot = A.get<T>(N2); q_bdh = A.get<T>(N2);
    s2a = A.get<T>(N2); s2b = A.get<T>(N2); s2c = A.get<T>(N2); s2d = A.get<T>(N2); s2e = A.get<T>(N2);
    xfac = A.get<T>((int64_t)XF_N * N2);
    if (S.imp) {
      b_us = A.get<T>(N2); b_vs = A.get<T>(N2); b_dus = A.get<T>(N2); b_dvs = A.get<T>(N2);
      b_raw = A.get<T>(N2); b_rhs = A.get<T>(N2); nlam = A.get<T>(N2); arhs = A.get<T>(N2); xadj = A.get<T>(N2);
      if (S.open) { b_us0 = A.get<T>(N2); b_vs0 = A.get<T>(N2); b_uc = A.get<T>(N2); b_vc = A.get<T>(N2); }
    }
  }

  void zero(T* p, int64_t n) {
    hipLaunchKernelGGL(k_fill<T>, dim3(flat_blocks(n)), dim3(256), 0, c.stream, p, n, T(0));
  }
  // batched zeroing: queue arrays, launch once per 32
  FillTab<T> ft{};
  int nft = 0;
  int64_t ftmax = 0;
  void zq(T* p, int64_t n) {
    ft.p[nft] = p;
    ft.n[nft] = n;
    ftmax = n > ftmax ? n : ftmax;
    if (++nft == 32) zflush();
  }
  void zflush() {
    if (nft == 0) return;
    const int64_t per = ftmax / 8;   // the nft arrays share the block budget
    hipLaunchKernelGGL(k_fill_multi<T>, dim3(flat_blocks(per > 0 ? per : 1), nft), dim3(256), 0, c.stream, ft);
    nft = 0;
    ftmax = 0;
  }
  // the forcing cotangents adx2d, drx2d, wubot, ady2d, dry2d, wvbot from the
  // advave-period accumulators of x (cleared when clear is set)
  void fold_forcing(X<T>& x, bool clear) {
    ExtFoldArgs<T> a{x.advua, x.advva, S.g[POM_G_ARU], S.g[POM_G_ARV], w<POM_ADX2D>(), w<POM_DRX2D>(),
                     bi[POM_WUBOT], w<POM_ADY2D>(), w<POM_DRY2D>(), bi[POM_WVBOT], clear ? 1 : 0};
    launch_ext_fold(c, a);
  }
  void add(T* d, const T* s, int64_t n) {
    hipLaunchKernelGGL(k_add<T>, dim3(flat_blocks(n)), dim3(256), 0, c.stream, d, s, n);
  }
  // advave's input cotangents from those of its advua / advva (its flux and
  // tps outputs are work arrays without one); aam2d's accumulates across calls
  void advave_rev(const X<T>& x, const T* c_advua, const T* c_advva, T* p_ua, T* p_va, T* p_uab, T* p_vab,
                  T* p_d) {
    AdvaveAdjArgs<T> a{x.ua, x.va, x.uab, x.vab, x.d, S.out[POM_AAM2D], S.g[POM_G_DX], S.g[POM_G_DY],
                       c_advua, c_advva, p_ua, p_va, p_uab, p_vab, p_d, w<POM_AAM2D>()};
    launch_adj_advave(c, a);
  }

  void run(const T* const* cot_out, T* const* grad_in) {
    const int64_t N3 = S.n3(), N2 = S.n2();
    const dim3 g2(c.nblk2), tb(TPB);
    // output cotangents: copied where they are accumulated into below, read in place elsewhere
    zq(zero3, N3);
    zq(zero2, N2);
    for (int f = 0; f < POM_NSTATE; ++f) {
      const int64_t n = f < POM_N3D ? N3 : N2;
      const T* given = cot_out ? cot_out[f] : nullptr;
      if (bw[f]) {
        if (given) S.copy(bw[f], given, n);
        else zq(bw[f], n);
      } else {
        bo[f] = given ? given : (f < POM_N3D ? zero3 : zero2);
      }
      if (!(grad_in && grad_in[f])) zq(bi[f], n);
    }
    // with the internal mode the first writer of each of these assigns it (internal_rev)
    if (!S.internal)
      for (T* p : {b_rho1, b_u1, b_v1, b_q2f0, b_q2lf0, b_q2f1, b_q2lf1, b_kmraw, b_khraw, b_tf0, b_sf0, b_tf1,
                   b_sf1, b_uf0, b_vf0, b_uf1, b_vf1, b_uf2, b_vf2})
        zq(p, N3);
    for (T* p : {b_adx2d_raw, b_ady2d_raw, b_advua0, b_advva0, b_egf0, b_utf0, b_vtf0}) zq(p, N2);
    if (S.imp)
      for (T* p : {b_us, b_vs, b_dus, b_dvs, b_raw}) zq(p, N2);
    zflush();
    // ---- shifts reversed: egb<-egf, etb<-et, et<-etf, dt<-h+et, utb<-utf, vtb<-vtf
    // (the final substep's cotangents assigned, et's accumulated, in one launch)
    X<T>& cur = bx[S.isplit & 1];
    {
      SumTab<T> st{};
      int ne = 0;
      auto put = [&](T* d, const T* a, const T* b = nullptr, const T* e = nullptr) {
        st.d[ne] = d;
        st.s[ne][0] = a;
        st.s[ne][1] = b;
        st.s[ne][2] = e;
        ++ne;
      };
      put(cur.ua, bo[POM_UA]); put(cur.uab, bo[POM_UAB]); put(cur.va, bo[POM_VA]); put(cur.vab, bo[POM_VAB]);
      put(cur.el, bo[POM_EL]); put(cur.elb, bo[POM_ELB]); put(cur.d, bo[POM_D]);
      put(cur.egf, bo[POM_EGF], bo[POM_EGB]); put(cur.utf, bo[POM_UTF], bo[POM_UTB]);
      put(cur.vtf, bo[POM_VTF], bo[POM_VTB]); put(cur.etf, bo[POM_ETF], bo[POM_ET], bo[POM_DT]);
      put(cur.uaf, bo[POM_UAF]); put(cur.vaf, bo[POM_VAF]); put(cur.elf, bo[POM_ELF]);
      put(cur.advua, bo[POM_ADVUA]); put(cur.advva, bo[POM_ADVVA]);
      put(bi[POM_ET], bi[POM_ET], bo[POM_ETB]);
      st.n = N2;
      hipLaunchKernelGGL(k_sum_multi<T>, dim3(flat_blocks(N2), ne), dim3(256), 0, c.stream, st);
    }
    // cotangents of the final substep fields used by the internal mode
    T* b_etf30 = cur.etf;
    T* b_egf30 = cur.egf;
    T* b_utf30 = cur.utf;
    T* b_vtf30 = cur.vtf;
    if (S.internal) internal_rev(b_etf30, b_egf30, b_utf30, b_vtf30, cur.d);
    else {
      for (int f : {POM_U, POM_UB, POM_V, POM_VB, POM_T, POM_TB, POM_S, POM_SB, POM_Q2, POM_Q2B, POM_Q2L, POM_Q2LB,
                    POM_KM, POM_KH, POM_KQ})
        add(bi[f], bo[f], N3);
      add(b_rho1, bo[POM_RHO], N3);
      add(bi[POM_WUBOT], bo[POM_WUBOT], N2);
      add(bi[POM_WVBOT], bo[POM_WVBOT], N2);
    }
    if (S.imp) {
      external_implicit_rev(cur);
    } else {
      // ---- substeps reversed
      for (int n = S.isplit; n >= 1; --n) {
        X<T>& bn = bx[n & 1];
        X<T>& bp = bx[(n - 1) & 1];
        // No zeroing of bp: k_ext_rev_uv assigns every buffer of bp it
        // writes (ua uab va vab el elb d uaf vaf) and completes bn.elf
        // (assigned for n < isplit, where nothing else wrote it in this pass);
        // egf utf vtf etf are shared by both buffers, and the shared advua /
        // advva accumulator is consumed and cleared below.
        const X<T>& xp = n == 1 ? S.x0_ : S.xs[n - 1];
        const X<T>& xn = S.xs[n];
        Ctx<T> cn = c;
        cn.prm.iext = n;
        {   // partials of the depth-mean momentum outputs (with bcond 2), once per point, then their sums
            // at each input point with the filter's own terms
          ExtRevFacArgs<T> a{};
          a.c_uaf = bn.uaf; a.c_ua = bn.ua; a.c_uab = bn.uab; a.c_utf = bn.utf;
          a.c_vaf = bn.vaf; a.c_va = bn.va; a.c_vab = bn.vab; a.c_vtf = bn.vtf;
          a.h = S.g[POM_G_H]; a.elf = xn.elf; a.aru = S.g[POM_G_ARU]; a.arv = S.g[POM_G_ARV];
          a.dum = S.g[POM_G_DUM]; a.dvm = S.g[POM_G_DVM]; a.dx = S.g[POM_G_DX]; a.dy = S.g[POM_G_DY];
          a.d = xp.d; a.el = xp.el; a.elb = xp.elb; a.eatm = S.out[POM_E_ATMOS]; a.uab = xp.uab; a.vab = xp.vab;
          a.uaf = xn.uaf; a.vaf = xn.vaf;
          a.fac = xfac; a.n2 = N2;
          a.p_uab = bp.uab; a.p_vab = bp.vab; a.p_uaf = bp.uaf; a.p_vaf = bp.vaf;
          a.a_advua = bn.advua; a.a_advva = bn.advva;
          ExtRevGatArgs<T> g{};
          g.fac = xfac; g.n2 = N2;
          g.cor = S.g[POM_G_COR]; g.d = xp.d; g.ua = xp.ua; g.va = xp.va; g.fsm = S.g[POM_G_FSM];
          g.c_uab = bn.uab; g.c_vab = bn.vab; g.c_elb = bn.elb; g.c_el = bn.el; g.c_d = bn.d; g.c_egf = bn.egf;
          g.c_etf = bn.etf; g.c_elf = bn.elf;
          g.p_ua = bp.ua; g.p_va = bp.va; g.p_el = bp.el; g.p_elb = bp.elb; g.p_d = bp.d;
          g.elf_first = n < S.isplit;
          launch_ext_rev_fac(cn, a);
          launch_ext_rev_gat(cn, g);
        }
        if (n % c.prm.ispadv == 0) {   // advave produced this substep's advua / advva
          advave_rev(xp, bn.advua, bn.advva, bp.ua, bp.va, bp.uab, bp.vab, bp.d);
          fold_forcing(bn, true);   // the period's forcing cotangents; the earlier value's starts from zero
        }
        {   // bcond 1 and continuity
          ExtRevElfArgs<T> a{bn.elf, xp.d, xp.ua, xp.va, S.g[POM_G_DX], S.g[POM_G_DY], S.g[POM_G_ART],
                             S.g[POM_G_FSM], bp.ua, bp.va, bp.d, bp.elb};
          launch_ext_rev_elf(cn, a);
        }
      }
      X<T>& b0 = bx[0];
      fold_forcing(b0, false);   // the first period's forcing cotangents (its accumulator is read below)
      // X_0 -> inputs and pre-loop intermediates, in one launch (no substep
      // reads the entering elf: its cotangent is zero and is not added)
      SumTab<T> st{};
      int ne = 0;
      auto acc = [&](T* d, const T* a) {
        st.d[ne] = d;
        st.s[ne][0] = d;
        st.s[ne][1] = a;
        st.s[ne][2] = nullptr;
        ++ne;
      };
      acc(bi[POM_UA], b0.ua); acc(bi[POM_UAB], b0.uab); acc(bi[POM_VA], b0.va); acc(bi[POM_VAB], b0.vab);
      acc(bi[POM_EL], b0.el); acc(bi[POM_ELB], b0.elb); acc(bi[POM_D], b0.d); acc(bi[POM_ETF], b0.etf);
      acc(bi[POM_UAF], b0.uaf); acc(bi[POM_VAF], b0.vaf);
      acc(b_egf0, b0.egf); acc(b_utf0, b0.utf); acc(b_vtf0, b0.vtf); acc(b_advua0, b0.advua);
      acc(b_advva0, b0.advva);
      st.n = N2;
      hipLaunchKernelGGL(k_sum_multi<T>, dim3(flat_blocks(N2), ne), dim3(256), 0, c.stream, st);
    }
    // pre-loop: adx2d = raw - advua0 etc.
    {
      COutTab<T> ob{{bo[POM_ADX2D], bo[POM_ADY2D], b_egf0, b_utf0, b_vtf0}};
      T* ib[OpExtPrep::NIN] = {b_adx2d_raw, b_ady2d_raw, b_advua0, b_advva0, bi[POM_EL], bi[POM_UA], bi[POM_VA],
                               bi[POM_D], bi[POM_UTF], bi[POM_VTF]};
      launch_gather<T, OpExtPrep>(c, S.t_prep(), ob, ib);
    }
    {
      X<T> xin = S.x0_;
      xin.fluxua = (T*)S.in[POM_FLUXUA];
      xin.fluxva = (T*)S.in[POM_FLUXVA];
      xin.tps = (T*)S.in[POM_TPS];
      advave_rev(xin, b_advua0, b_advva0, bi[POM_UA], bi[POM_VA], bi[POM_UAB], bi[POM_VAB], bi[POM_D]);
    }
    {
      ColsumAdjArgs<T> a{};
      a.bout[0] = b_adx2d_raw; a.bf[0] = w<POM_ADVX>();
      a.bout[1] = b_ady2d_raw; a.bf[1] = w<POM_ADVY>();
      a.bout[2] = bo[POM_DRX2D]; a.bf[2] = w<POM_DRHOX>();
      a.bout[3] = bo[POM_DRY2D]; a.bf[3] = w<POM_DRHOY>();
      a.bout[4] = bo[POM_AAM2D]; a.bf[4] = w<POM_AAM>();
      a.n = 5;
      hipLaunchKernelGGL(k_colsum_adj<T>, g2, tb, 0, c.stream, c.B, c.vert, a);
    }
    {
      COutTab<T> ob{{bo[POM_AAM]}};
      T* ib[OpSmag::NIN] = {bi[POM_U], bi[POM_V], bi[POM_AAM]};
      if (S.imp) {
        launch_gather<T, OpSmagImp>(c, S.t_smag_imp(), ob, ib);
      } else {   // transposed stencil through four factor planes
        SmagAdjArgs<T> a{S.in[POM_U], S.in[POM_V], S.g[POM_G_DX], S.g[POM_G_DY], bo[POM_AAM], s3a, s3b, s3c, s3d,
                         bi[POM_U], bi[POM_V], bi[POM_AAM]};
        launch_adj_smag(c, a);
      }
    }
    {
      BaropgAdjArgs<T> a{S.in[POM_RHO], S.g[POM_G_RMEAN], S.in[POM_DT], S.g[POM_G_DX], S.g[POM_G_DY],
                         S.g[POM_G_DUM], S.g[POM_G_DVM], b_rho1, bo[POM_DRHOX], bo[POM_DRHOY], bi[POM_RHO],
                         bi[POM_DT]};
      hipLaunchKernelGGL(k_baropg_adj<T>, g2, tb, 0, c.stream, c.B, c.vert, c.ramp, c.prm, a);
    }
    {   // advct: transposed stencil over its face fluxes
      AdvctAdjArgs<T> a{S.in[POM_U], S.in[POM_V], S.in[POM_UB], S.in[POM_VB], S.in[POM_AAM], S.in[POM_DT],
                        S.g[POM_G_DX], S.g[POM_G_DY], S.g[POM_G_ARU], S.g[POM_G_ARV], bo[POM_ADVX], bo[POM_ADVY],
                        bi[POM_U], bi[POM_V], bi[POM_UB], bi[POM_VB], bi[POM_AAM], bi[POM_DT]};
      launch_adj_advct(c, a);
    }
  }

  // semi-implicit barotropic step reversed (implicit_p0_fixed/time_loop.py:247-398;
  // eta-solve adjoint as core.py:45-59, 231-252, 645-675)
  void external_implicit_rev(X<T>& cur) {
    const int64_t N2 = S.n2();
    const dim3 g2(c.nblk2), tb(TPB), gm(c.B.M);
    {   // filter, level shift, endpoint transports
      COutTab<T> ob{{cur.ua, cur.uab, cur.va, cur.vab, cur.el, cur.elb, cur.d, cur.egf, cur.utf, cur.vtf, cur.etf}};
      T* ib[OpImpFilter::NIN] = {bi[POM_UA], bi[POM_UAB], bi[POM_VA], bi[POM_VAB], bi[POM_EL], bi[POM_ELB],
                                 cur.uaf, cur.vaf, cur.elf, b_utf0, b_vtf0};
      launch_gather<T, OpImpFilter>(c, S.t_ifilter(), ob, ib);
    }
    if (S.open) {   // radiation bcond 2 after the corrector, d = h + elf (assigns b_uc, b_vc)
      COutTab<T> ob{{cur.uaf, cur.vaf}};
      T* ib[OpImpRadNew::NIN] = {b_uc, b_vc, nullptr, nullptr, bi[POM_EL], cur.elf};
      launch_gather<T, OpImpRadNew>(c, S.t_rad(S.uac, S.vac, S.in[POM_UAF], S.in[POM_VAF], S.out[POM_ELF]), ob, ib,
                                    (1u << OpImpRadNew::UI) | (1u << OpImpRadNew::VI));
    }
    {   // corrector
      COutTab<T> ob{{S.open ? b_uc : cur.uaf, S.open ? b_vc : cur.vaf}};
      T* ib[OpImpCorr::NIN] = {b_us, b_vs, b_dus, b_dvs, cur.elf};
      launch_gather<T, OpImpCorr>(c, S.t_corr(), ob, ib);
    }
    if (S.open) {   // post-solve surface boundary with the tide at the nodes
      COutTab<T> ob{{cur.elf}};
      T* ib[OpImpSurfTide::NIN] = {b_raw};
      launch_gather<T, OpImpSurfTide>(c, S.t_surf_tide(), ob, ib);
    } else {   // post-solve surface boundary
      COutTab<T> ob{{cur.elf}};
      T* ib[OpImpSurf::NIN] = {b_raw};
      launch_gather<T, OpImpSurf>(c, S.t_surf(), ob, ib);
    }
    {   // eta solve: b_rhs = area A^-1 (b_raw / area), b_d -= (dA/dD elfraw)^T b_rhs
      Pcg<T>& w = S.pw;
      hipLaunchKernelGGL(k_eta_adj_in<T>, g2, tb, 0, c.stream, c.B, b_raw, S.g[POM_G_ART], arhs, w.part);
      hipLaunchKernelGGL(k_eta_adj_scale, gm, tb, 0, c.stream, c.B, w.part, w.sc, w.tol, S.hp->eta_tol);
      hipLaunchKernelGGL(k_eta_adj_norm<T>, g2, tb, 0, c.stream, c.B, arhs, w.sc);
      w.solve(c, S.ell(), arhs, S.in[POM_D], nullptr, xadj, S.hp->eta_rel_tol, S.hp->eta_max_iter);
      if (w.status != POM_OK) return;
      hipLaunchKernelGGL(k_eta_adj_out<T>, g2, tb, 0, c.stream, c.B, xadj, S.g[POM_G_ART], w.sc, b_rhs, nlam);
      COutTab<T> ob{{nlam}};
      T* ib[OpEll::NIN] = {nullptr, bi[POM_D]};
      launch_gather<T, OpEll>(c, S.t_ell(), ob, ib);
    }
    {   // transports and right-hand side
      COutTab<T> ob{{nullptr, nullptr, b_rhs}};
      T* ib[OpImpRhs::NIN] = {b_us, b_vs, bi[POM_D], bi[POM_ELB]};
      launch_gather<T, OpImpRhs>(c, S.t_rhs(), ob, ib);
    }
    if (S.open) {   // radiation bcond 2 after the predictor (assigns b_us0, b_vs0)
      COutTab<T> ob{{b_us, b_vs}};
      T* ib[OpImpRad::NIN] = {b_us0, b_vs0, bi[POM_UA], bi[POM_VA], bi[POM_EL], bi[POM_D]};
      launch_gather<T, OpImpRad>(c, S.t_rad(S.us, S.vs, S.in[POM_UA], S.in[POM_VA], S.in[POM_D]), ob, ib,
                                 (1u << OpImpRad::UI) | (1u << OpImpRad::VI));
    }
    {   // predictor
      COutTab<T> ob{{S.open ? b_us0 : b_us, S.open ? b_vs0 : b_vs, b_dus, b_dvs}};
      T* ib[OpImpPred::NIN] = {bi[POM_UA], bi[POM_VA], bi[POM_EL], bi[POM_ELB], bi[POM_D], bi[POM_UAB], bi[POM_VAB],
                               w<POM_ADX2D>(), w<POM_ADY2D>(), b_advua0, b_advva0, w<POM_DRX2D>(), w<POM_DRY2D>(),
                               bi[POM_WUBOT], bi[POM_WVBOT]};
      launch_gather<T, OpImpPred>(c, S.t_pred(), ob, ib);
    }
    // the step's advua / advva are copies of the single advave call's
    add(b_advua0, cur.advua, N2);
    add(b_advva0, cur.advva, N2);
  }

  void internal_rev(T* b_etf30, T* b_egf30, T* b_utf30, T* b_vtf30, T* b_d30) {
    const dim3 g2(c.nblk2), g3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)), tb(TPB);
    Step<T>& s = S;
    // u/v Asselin filter and shift (with the Shapiro filter reversed first when on)
    if (s.shap_uv) {
      COutTab<T> ou{{bo[POM_UB]}};
      T* iu[OpShapiro<true>::NIN] = {s3c};
      launch_gather<T, OpShapiro<true>>(c, s.t_shap(s.ush, s.g[POM_G_DUM]), ou, iu, 1u);
      COutTab<T> ov{{bo[POM_VB]}};
      T* iv[OpShapiro<true>::NIN] = {s3d};
      launch_gather<T, OpShapiro<true>>(c, s.t_shap(s.vsh, s.g[POM_G_DVM]), ov, iv, 1u);
      UVFiltAdjArgs<T> a{s3c, bo[POM_U], b_uf2, bi[POM_UB], b_u1, nullptr, 5u};
      hipLaunchKernelGGL(k_uvfilt_adj<T>, g3, tb, 0, c.stream, c.B, c.vert, c.prm, a);
      UVFiltAdjArgs<T> bv{s3d, bo[POM_V], b_vf2, bi[POM_VB], b_v1, nullptr, 5u};
      hipLaunchKernelGGL(k_uvfilt_adj<T>, g3, tb, 0, c.stream, c.B, c.vert, c.prm, bv);
    } else {
      UVFiltAdjArgs<T> a{bo[POM_UB], bo[POM_U], b_uf2, bi[POM_UB], b_u1, s.imp ? s.g[POM_G_DUM] : nullptr, 5u};
      hipLaunchKernelGGL(k_uvfilt_adj<T>, g3, tb, 0, c.stream, c.B, c.vert, c.prm, a);
      UVFiltAdjArgs<T> bv{bo[POM_VB], bo[POM_V], b_vf2, bi[POM_VB], b_v1, s.imp ? s.g[POM_G_DVM] : nullptr, 5u};
      hipLaunchKernelGGL(k_uvfilt_adj<T>, g3, tb, 0, c.stream, c.B, c.vert, c.prm, bv);
    }
    if (s.open) {   // open bcond 3, edge weight from d
      COutTab<T> ob{{b_uf2, b_vf2}};
      T* ib[OpBc3D::NIN] = {b_uf1, b_vf1, b_u1, b_v1, b_d30};
      launch_gather<T, OpBc3D>(c, s.t_bc3d(), ob, ib, 3u);
    } else if (s.imp) {   // closed-basin bcond 3
      COutTab<T> ob{{b_uf2, b_vf2}};
      T* ib[OpImpBc3::NIN] = {b_uf1, b_vf1};
      launch_gather<T, OpImpBc3>(c, s.t_ibc3(), ob, ib, 3u);
    } else {   // bcond 3
      COutTab<T> ob{{b_uf2, b_vf2}};
      T* ib[OpBc3::NIN] = {b_uf1, b_vf1, b_u1, b_v1};
      launch_gather<T, OpBc3>(c, s.t_bc3(), ob, ib, 3u);
    }
    for (int dir = 0; dir < 2; ++dir) {   // profu / profv
      ProfuvAdjArgs<T> a{};
      a.fw = s.a_profuv(dir == 1);
      a.b_out = dir ? b_vf1 : b_uf1;
      a.b_wbot = bo[dir ? POM_WVBOT : POM_WUBOT];
      a.b_f = dir ? b_vf0 : b_uf0;
      a.s_km0 = s3a; a.s_km1 = s3b; a.s_etf0 = s2a; a.s_etf1 = s2b; a.s_mag = s2c;
      a.b_wbot_in = bi[dir ? POM_WVBOT : POM_WUBOT];
      a.assign_f = 1;   // first writer of b_uf0 / b_vf0
      ProfuvAdj2Args<T> b2{s.in[POM_UB], s.in[POM_VB], s.g[POM_G_CBC], s3a, s3b, s2a, s2b, s2c,
                           w<POM_KM>(), b_etf30, bi[POM_UB], bi[POM_VB]};
      if (dir == 0) {
        hipLaunchKernelGGL((k_profuv_adj1<T, 0>), g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        hipLaunchKernelGGL((k_profuv_adj2<T, 0>), g2, tb, 0, c.stream, c.B, c.vert, b2);
      } else {
        hipLaunchKernelGGL((k_profuv_adj1<T, 1>), g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        hipLaunchKernelGGL((k_profuv_adj2<T, 1>), g2, tb, 0, c.stream, c.B, c.vert, b2);
      }
    }
    {   // advu / advv: one transposed stencil for both equations
      MomAdjArgs<T> a{};
      a.u = s.u1; a.v = s.v1; a.w = s.out[POM_W]; a.ub = s.in[POM_UB]; a.vb = s.in[POM_VB]; a.uf = s.uf0;
      a.vf = s.vf0; a.dt = s.in[POM_DT]; a.egf = s.xf().egf; a.egb = s.in[POM_EGB]; a.etf = s.xf().etf;
      a.etb = s.in[POM_ETB]; a.aru = s.g[POM_G_ARU]; a.arv = s.g[POM_G_ARV]; a.cor = s.g[POM_G_COR];
      a.dx = s.g[POM_G_DX]; a.dy = s.g[POM_G_DY]; a.h = s.g[POM_G_H]; a.eatm = s.out[POM_E_ATMOS];
      a.ob_u = b_uf0; a.ob_v = b_vf0;
      a.c_advx = w<POM_ADVX>(); a.c_drhox = w<POM_DRHOX>(); a.c_advy = w<POM_ADVY>(); a.c_drhoy = w<POM_DRHOY>();
      a.c_u = b_u1; a.c_v = b_v1; a.c_w = w<POM_W>(); a.c_ub = bi[POM_UB]; a.c_vb = bi[POM_VB];
      a.c_dt = bi[POM_DT]; a.c_egf = b_egf30; a.c_egb = bi[POM_EGB]; a.c_etf = b_etf30; a.c_etb = bi[POM_ETB];
      launch_adj_advuv(c, a);
    }
    {   // dens
      COutTab<T> ob{{bo[POM_RHO]}};
      T* ib[OpDens::NIN] = {w<POM_T>(), w<POM_S>(), b_rho1};
      launch_gather<T, OpDens>(c, s.t_dens(), ob, ib, 4u);   // first writer of b_rho1
    }
    if (s.open) {   // open bcond 4 with exterior values + T/S filter
      COutTab<T> ob{{bo[POM_T], bo[POM_TB], bo[POM_S], bo[POM_SB]}};
      T* ib[OpBc4FilterExt::NIN] = {b_tf1, b_sf1, bi[POM_T], bi[POM_S], bi[POM_TB], bi[POM_SB], b_u1, b_v1,
                                    w<POM_W>(), nullptr, nullptr, bi[POM_DT]};
      launch_gather<T, OpBc4FilterExt>(c, s.t_bc4ext(), ob, ib, 3u);
    } else if (s.imp && s.shap_ts) {   // Shapiro filter of the new T, S, then bcond 4 + T/S filter
      COutTab<T> ot{{bo[POM_T]}};
      T* it[OpShapiro<false>::NIN] = {s3c};
      launch_gather<T, OpShapiro<false>>(c, s.t_shap(s.tsh, s.g[POM_G_FSM]), ot, it, 1u);
      COutTab<T> os{{bo[POM_S]}};
      T* is[OpShapiro<false>::NIN] = {s3d};
      launch_gather<T, OpShapiro<false>>(c, s.t_shap(s.ssh, s.g[POM_G_FSM]), os, is, 1u);
      COutTab<T> ob{{s3c, bo[POM_TB], s3d, bo[POM_SB]}};
      T* ib[OpImpBc4FilterG<false>::NIN] = {b_tf1, b_sf1, bi[POM_T], bi[POM_S], bi[POM_TB], bi[POM_SB]};
      launch_gather<T, OpImpBc4FilterG<false>>(c, s.t_ibc4(), ob, ib, 3u);
    } else if (s.imp) {   // closed-basin bcond 4 + T/S filter + tracer mask
      COutTab<T> ob{{bo[POM_T], bo[POM_TB], bo[POM_S], bo[POM_SB]}};
      T* ib[OpImpBc4Filter::NIN] = {b_tf1, b_sf1, bi[POM_T], bi[POM_S], bi[POM_TB], bi[POM_SB]};
      launch_gather<T, OpImpBc4Filter>(c, s.t_ibc4(), ob, ib, 3u);
    } else {   // bcond 4 + T/S filter
      COutTab<T> ob{{bo[POM_T], bo[POM_TB], bo[POM_S], bo[POM_SB]}};
      T* ib[OpBc4Filter::NIN] = {b_tf1, b_sf1, bi[POM_T], bi[POM_S], bi[POM_TB], bi[POM_SB], b_u1, b_v1,
                                 w<POM_W>(), bi[POM_DT]};
      launch_gather<T, OpBc4Filter>(c, s.t_bc4(), ob, ib, 3u);
    }
    for (int sal = 0; sal < 2; ++sal) {   // proft
      ProftAdjArgs<T> a{};
      ProftArgs<T> f = s.a_proft(sal == 1);
      a.f = f.f; a.kh = f.kh; a.etf = f.etf; a.h = f.h; a.wfsurf = f.wfsurf; a.fsurf = f.fsurf;
      a.b_out = sal ? b_sf1 : b_tf1;
      a.b_f = sal ? b_sf0 : b_tf0;
      a.b_kh = w<POM_KH>();
      a.b_etf = b_etf30;
      a.nbc = f.nbc;
      a.assign_f = 1;   // first writer of b_tf0 / b_sf0
      hipLaunchKernelGGL(k_proft_adj<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
    }
    {   // advt1 of T and S
      const FieldTab<T> tab[2] = {s.t_advt1(false), s.t_advt1(true)};
      const COutTab<T> ob[2] = {{{b_tf0}}, {{b_sf0}}};
      T* ibt[OpAdvt1::NIN] = {bi[POM_TB], bi[POM_T], nullptr, b_u1, b_v1, w<POM_W>(), w<POM_AAM>(), bi[POM_DT],
                              bi[POM_ETB], b_etf30};
      T* ibs[OpAdvt1::NIN] = {bi[POM_SB], bi[POM_S], nullptr, b_u1, b_v1, w<POM_W>(), w<POM_AAM>(), bi[POM_DT],
                              bi[POM_ETB], b_etf30};
      T* const* ib[2] = {ibt, ibs};
      launch_flux_advt1<T>(c, tab, ob, ib);
    }
    if (s.imp && !s.open) {   // closed-basin bcond 6 + q2 filter
      COutTab<T> ob{{bo[POM_Q2], bo[POM_Q2B], bo[POM_Q2L], bo[POM_Q2LB]}};
      T* ib[OpImpBc6Filter::NIN] = {b_q2f1, b_q2lf1, bi[POM_Q2], bi[POM_Q2L], bi[POM_Q2B], bi[POM_Q2LB]};
      launch_gather<T, OpImpBc6Filter>(c, s.t_ibc6(), ob, ib, 3u);
    } else {   // bcond 6 + q2 filter
      COutTab<T> ob{{bo[POM_Q2], bo[POM_Q2B], bo[POM_Q2L], bo[POM_Q2LB]}};
      T* ib[OpBc6Filter::NIN] = {b_q2f1, b_q2lf1, bi[POM_Q2], bi[POM_Q2L], bi[POM_Q2B], bi[POM_Q2LB], b_u1, b_v1};
      launch_gather<T, OpBc6Filter>(c, s.t_bc6(), ob, ib, 3u);
    }
    {   // km/kh edge copy
      COutTab<T> ob{{bo[POM_KM], bo[POM_KH]}};
      T* ib[OpKEdge::NIN] = {b_kmraw, b_khraw};
      launch_gather<T, OpKEdge>(c, s.t_kedge(), ob, ib, 3u);
    }
    {   // profq
      ProfqAdjArgs<T> a{};
      a.fw = s.a_profq();
      a.b_q2o = b_q2f1; a.b_q2lo = b_q2lf1; a.b_lo = bo[POM_L]; a.b_kqo = bo[POM_KQ];
      a.b_kmo = b_kmraw; a.b_kho = b_khraw;
      a.b_q2f = b_q2f0; a.b_q2lf = b_q2lf0; a.b_q2 = bi[POM_Q2]; a.b_q2b = bi[POM_Q2B]; a.b_q2lb = bi[POM_Q2LB];
      a.b_km = bi[POM_KM]; a.b_kh = bi[POM_KH]; a.b_kq = bi[POM_KQ]; a.b_u = b_u1; a.b_v = b_v1;
      a.b_t = bi[POM_T]; a.b_s = bi[POM_S]; a.b_rho = b_rho1; a.b_wubot = bi[POM_WUBOT];
      a.b_wvbot = bi[POM_WVBOT]; a.b_etf = b_etf30; a.b_dt = bi[POM_DT];
      a.s_unb = s3a; a.s_vnb = s3b; a.s_wubnb = s2a; a.s_wvbnb = s2b;
      a.r_bl = q_bl; a.r_bgh = q_bgh; a.r_bq2ln = q_bq2ln; a.r_ba = q_ba; a.r_bc = q_bc; a.r_bkn = q_bkn;
      a.r_bdtef = q_bdtef; a.r_bboy = q_bboy; a.r_bq2bot = q_bq2bot; a.r_bdh = q_bdh;
      a.assign_q2f = 1;   // first writer of b_q2f0 / b_q2lf0
      {
        ProfScope ps(POM_STAGE_PROFQ_ADJ, c.stream);
        hipLaunchKernelGGL(k_profq_adj_a<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        hipLaunchKernelGGL(k_profq_adj_b<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        hipLaunchKernelGGL(k_profq_adj_c1<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
        hipLaunchKernelGGL(k_profq_adj_c2<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
      }
      hipLaunchKernelGGL(k_profq_adj2<T>, dim3(c.nblk3), tb, 0, c.stream, c.B, a);
    }
    {   // advq of q2 and q2l
      const FieldTab<T> tab[2] = {s.t_advq(false), s.t_advq(true)};
      const COutTab<T> ob[2] = {{{b_q2f0}}, {{b_q2lf0}}};
      T* ibq[OpAdvq::NIN] = {bi[POM_Q2B], bi[POM_Q2], b_u1, b_v1, w<POM_W>(), w<POM_AAM>(), bi[POM_DT], bi[POM_ETB],
                             b_etf30};
      T* ibl[OpAdvq::NIN] = {bi[POM_Q2LB], bi[POM_Q2L], b_u1, b_v1, w<POM_W>(), w<POM_AAM>(), bi[POM_DT],
                             bi[POM_ETB], b_etf30};
      T* const* ib[2] = {ibq, ibl};
      launch_flux_advq<T>(c, tab, ob, ib);
    }
    {   // vertvl (w input is not differentiated)
      VertvlAdjArgs<T> a{s.u1, s.v1, s.in[POM_DT], s.g[POM_G_DX], s.g[POM_G_DY], s.g[POM_G_FSM], bo[POM_W], s3a,
                         b_u1, b_v1, bi[POM_DT], b_etf30, bi[POM_ETB], s.imp && !s.open};
      hipLaunchKernelGGL(k_vertvl_adj1<T>, g2, tb, 0, c.stream, c.B, c.vert, c.prm, a);
      hipLaunchKernelGGL(k_vertvl_adj2<T>, dim3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)), tb, 0,
                         c.stream, c.B, a);
    }
    {   // transport adjustment
      AdjustAdjArgs<T> a{s.in[POM_UTB], s.xf().utf, s.in[POM_VTB], s.xf().vtf, s.in[POM_DT],
                         b_u1, b_v1, bi[POM_U], bi[POM_V], bi[POM_UTB], b_utf30, bi[POM_VTB], b_vtf30,
                         bi[POM_DT]};
      hipLaunchKernelGGL(k_adjust_adj<T>, g2, tb, 0, c.stream, c.B, c.vert, a);
    }
  }
};

template <class T> Ctx<T> make_ctx(const pom_batch* b, const pom_params_t* p, const void* ramp, const void* vg,
                                    hipStream_t st) {
  Ctx<T> c{};
  c.B = batch_view(b);
  c.vert = Vert<T>{static_cast<const T*>(vg)};
  c.ramp = static_cast<const T*>(ramp);
  Params& q = c.prm;
  q.dti = p->dti; q.dte = p->dte; q.dti2 = p->dti2; q.dte2 = p->dte2; q.ispi = p->ispi; q.isp2i = p->isp2i;
  q.alpha = p->alpha; q.smoth = p->smoth; q.horcon = p->horcon; q.grav = p->grav; q.rhoref = p->rhoref;
  q.kappa = p->kappa; q.umol = p->umol; q.tprni = p->tprni; q.small = p->small; q.tbias = p->tbias;
  q.sbias = p->sbias; q.hmax = p->hmax; q.rfe = p->rfe; q.rfw = p->rfw; q.rfn = p->rfn; q.rfs = p->rfs;
  q.isplit = p->isplit; q.ispadv = p->ispadv; q.ntp = p->ntp; q.nbct = p->nbct; q.nbcs = p->nbcs; q.iext = 0;
  q.core = p->core;
  q.gamma_baro = p->core == POM_CORE_IMPLICIT ? p->gamma_baro : 0.0;
  const bool closed_imp = p->core == POM_CORE_IMPLICIT && p->boundary == POM_BOUNDARY_CLOSED;
  q.gamma_ts = closed_imp ? p->tracer_shapiro : 0.0;
  q.gamma_uv = closed_imp ? p->velocity_shapiro : 0.0;
  q.aam_min = p->core == POM_CORE_IMPLICIT ? p->aam_min : 0.0;
  c.nblk3 = b->nblk3;
  c.nblk2 = b->nblk2;
  c.stream = st;
  c.tune = b->tune;
  return c;
}

int check_params(const pom_params_t* p, int core) {
  if (!p || p->core != core) return POM_E_ARG;
  if (p->boundary != POM_BOUNDARY_CLOSED && p->boundary != POM_BOUNDARY_TIDE) return POM_E_ARG;
  if (p->boundary == POM_BOUNDARY_TIDE && core == POM_CORE_EXPLICIT &&
      (!std::isfinite(p->tide_coef) || !std::isfinite(p->tide_sec) || p->tide_substep0 < 0))
    return POM_E_ARG;
  if (p->isplit < 3 || p->isplit > MAXSPLIT || p->ispadv < 1) return POM_E_ARG;
  if (!((p->nbct == 1 || p->nbct == 3) && (p->nbcs == 1 || p->nbcs == 3))) return POM_E_UNSUPPORTED;
  if (core == POM_CORE_IMPLICIT) {
    if (p->eta_max_iter < 1 || p->eta_max_iter > 100000 || !(p->eta_rel_tol >= 0.0) || !(p->eta_tol >= 0.0))
      return POM_E_ARG;
    if (p->eta_precond != POM_PRECOND_JACOBI && p->eta_precond != POM_PRECOND_SPECTRAL) return POM_E_ARG;
  }
  return POM_OK;
}

// ---------------------------------------------------------------- spectral cache
// [256 B head][SpecMember x M][data: per member S_ni, S_nj, inverse eigenvalues]
constexpr size_t CACHE_HEAD = 256;
inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
inline size_t cache_data_off(int M) { return CACHE_HEAD + align256(sizeof(SpecMember) * (size_t)M); }
inline size_t cache_elems(const pom_batch* b) {
  size_t n = 0;
  for (const Member& m : b->h_mem) {
    const size_t ni = m.im - 2, nj = m.jm - 2;
    n += ni * ni + nj * nj + ni * nj;
  }
  return n;
}
inline size_t cache_bytes(const pom_batch* b) {
  return cache_data_off(b->M) + cache_elems(b) * (b->dtype == POM_F32 ? 4 : 8) + 256;
}

// lower median of the interior values of a member's 2-D field (torch.median)
inline double interior_median(const std::vector<double>& f, const Member& m) {
  std::vector<double> v;
  v.reserve((size_t)(m.im - 2) * (m.jm - 2));
  for (int i = 1; i <= m.im - 2; ++i)
    for (int j = 1; j <= m.jm - 2; ++j) v.push_back(f[m.off2 + (int64_t)i * m.jm + j]);
  const size_t k = (v.size() - 1) / 2;
  std::nth_element(v.begin(), v.begin() + k, v.end());
  return v[k];
}

template <class T>
int prepare_cache(const pom_batch* b, const pom_params_t* prm, const void* const* geom, void* cache,
                  hipStream_t st) {
  const int64_t N2 = b->n2;
  std::vector<T> hx(N2), hy(N2), hh(N2);
  if (hipMemcpyAsync(hx.data(), geom[POM_G_DX], sizeof(T) * N2, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hy.data(), geom[POM_G_DY], sizeof(T) * N2, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hh.data(), geom[POM_G_H], sizeof(T) * N2, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return POM_E_HIP;
  const std::vector<double> dx(hx.begin(), hx.end()), dy(hy.begin(), hy.end()), h(hh.begin(), hh.end());
  const double coeff = 4.0 * (prm->dti * prm->dti) * prm->grav * prm->alpha;   // core.py:175
  const double pi = 3.14159265358979323846;
  std::vector<SpecMember> tab(b->M);
  std::vector<T> data(cache_elems(b));
  int64_t o = 0;
  for (int m = 0; m < b->M; ++m) {
    const Member& mm = b->h_mem[m];
    const int ni = mm.im - 2, nj = mm.jm - 2;
    SpecMember& sm = tab[m];
    sm.ni = ni;
    sm.nj = nj;
    sm.scale = 4.0 * (ni + 1) * (nj + 1);
    auto sine = [&](int n, int64_t at) {   // S_n[k][q] = 2 sin(pi (k+1)(q+1) / (n+1))
      for (int k = 0; k < n; ++k)
        for (int q = 0; q < n; ++q) {
          const int64_t r = ((int64_t)(k + 1) * (q + 1)) % (2 * (n + 1));   // exact argument reduction
          data[at + (int64_t)k * n + q] = T(2.0 * std::sin(pi * double(r) / double(n + 1)));
        }
    };
    sm.si = o;
    sine(ni, o);
    o += (int64_t)ni * ni;
    sm.sj = o;
    sine(nj, o);
    o += (int64_t)nj * nj;
    sm.l = o;
    // inverse eigenvalues of the constant-coefficient operator at the interior
    // medians of dx, dy, h (core.py:730-760)
    const double dx0 = interior_median(dx, mm), dy0 = interior_median(dy, mm), h0 = interior_median(h, mm);
    for (int a = 0; a < ni; ++a) {
      const double sx = std::sin(pi * (a + 1) / (2.0 * (ni + 1)));
      const double lx = 4.0 * (sx * sx) / (dx0 * dx0 + 1e-30);
      for (int c2 = 0; c2 < nj; ++c2) {
        const double sy = std::sin(pi * (c2 + 1) / (2.0 * (nj + 1)));
        const double ly = 4.0 * (sy * sy) / (dy0 * dy0 + 1e-30);
        data[o + (int64_t)a * nj + c2] = T(1.0 / (1.0 + coeff * h0 * (lx + ly)));
      }
    }
    o += (int64_t)ni * nj;
  }
  char* base = static_cast<char*>(cache);
  if (hipMemcpyAsync(base + CACHE_HEAD, tab.data(), sizeof(SpecMember) * b->M, hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemcpyAsync(base + cache_data_off(b->M), data.data(), sizeof(T) * data.size(), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return POM_E_HIP;
  return POM_OK;
}

// preconditioner choice and spectral cache of a PCG driver
template <class T> void setup_pcg(Pcg<T>& w, const pom_batch* b, const pom_params_t* prm, const void* cache) {
  w.spectral = prm->eta_precond == POM_PRECOND_SPECTRAL;
  w.capture = prm->eta_capture != 0;
  w.status = POM_OK;
  w.maxni = w.maxnj = 0;
  for (const Member& m : b->h_mem) {
    w.maxni = std::max(w.maxni, m.im - 2);
    w.maxnj = std::max(w.maxnj, m.jm - 2);
  }
  if (cache) {
    w.sp = reinterpret_cast<const SpecMember*>(static_cast<const char*>(cache) + CACHE_HEAD);
    w.cache = reinterpret_cast<const T*>(static_cast<const char*>(cache) + cache_data_off(b->M));
  }
}

// attach the core, the boundary policy and the semi-implicit settings and cache to a step
template <class T> void setup_implicit(Step<T>& S, const pom_params_t* prm, const void* cache) {
  S.imp = prm->core == POM_CORE_IMPLICIT;
  S.open = prm->boundary == POM_BOUNDARY_TIDE;
  S.shap_ts = S.imp && !S.open && prm->tracer_shapiro != 0.0;
  S.shap_uv = S.imp && !S.open && prm->velocity_shapiro != 0.0;
  S.tide_wt = prm->tide_wt;
  S.hp = prm;
  if (S.imp) setup_pcg(S.pw, S.b, prm, cache);
}

template <class T>
size_t ws_size(const pom_batch* b, const pom_params_t* p, bool adjoint) {
  Step<T> S{};
  S.b = b;
  S.isplit = p->isplit;
  S.c.prm.ispadv = p->ispadv;
  S.tape = adjoint;
  setup_implicit(S, p, nullptr);
  Arena A{nullptr, 0, 0, true};
  T* dummy[POM_NSTATE] = {};
  for (int f = 0; f < POM_NSTATE; ++f) S.out[f] = dummy[f];
  S.plan(A, adjoint);
  if (adjoint) {
    Adj<T> J(S, S.c);
    J.plan(A, nullptr);
  }
  return A.off + 256;
}

template <class T>
int run_step(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vg, const void* const* in,
             void* const* out, const void* const* geom, int internal, void* vamax, void* ws, size_t wsb,
             hipStream_t st, const void* cache = nullptr, int32_t* iters = nullptr) {
  Step<T> S{};
  S.b = b;
  S.c = make_ctx<T>(b, prm, ramp, vg, st);
  setup_implicit(S, prm, cache);
  S.iters_out = iters;
  S.isplit = prm->isplit;
  S.internal = internal != 0;
  S.vamax = static_cast<T*>(vamax);
  for (int f = 0; f < POM_NSTATE; ++f) {
    S.in[f] = static_cast<const T*>(in[f]);
    S.out[f] = static_cast<T*>(out[f]);
    if (!S.in[f] || !S.out[f]) return POM_E_ARG;
  }
  for (int f = 0; f < POM_NGEOM; ++f) {
    S.g[f] = static_cast<const T*>(geom[f]);
    if (!S.g[f] && (f < POM_NGEOM_BASE || S.open)) return POM_E_ARG;
  }
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  S.plan(A, false);
  if (A.off > wsb) return POM_E_ARG;
  S.forward();
  if (S.pw.status != POM_OK) return S.pw.status;
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

template <class T>
int run_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vg, const void* const* in,
            const void* const* geom, int internal, const void* const* cot, void* const* grad, void* ws, size_t wsb,
            hipStream_t st, const void* cache = nullptr) {
  Step<T> S{};
  S.b = b;
  S.c = make_ctx<T>(b, prm, ramp, vg, st);
  setup_implicit(S, prm, cache);
  S.isplit = prm->isplit;
  S.internal = internal != 0;
  for (int f = 0; f < POM_NSTATE; ++f) {
    S.in[f] = static_cast<const T*>(in[f]);
    if (!S.in[f]) return POM_E_ARG;
  }
  for (int f = 0; f < POM_NGEOM; ++f) {
    S.g[f] = static_cast<const T*>(geom[f]);
    if (!S.g[f] && (f < POM_NGEOM_BASE || S.open)) return POM_E_ARG;
  }
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  S.tape = true;
  S.plan(A, true);
  S.vamax = nullptr;   // the recompute needs no depth-mean velocity check
  Adj<T> J(S, S.c);
  J.plan(A, reinterpret_cast<T* const*>(grad));
  if (A.off > wsb) return POM_E_ARG;
  S.forward();
  if (S.pw.status != POM_OK) return S.pw.status;
  J.run(reinterpret_cast<const T* const*>(cot), reinterpret_cast<T* const*>(grad));
  if (S.pw.status != POM_OK) return S.pw.status;
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// The forward of run_adj with its tape kept in the workspace and in the
// caller's output buffers (the last barotropic substep and the step's new
// fields live there, as in the plain step); run_reverse then runs the reverse
// of that forward from the same workspace and outputs without recomputing it.
// The two plan the workspace identically (the plan depends on the batch and
// the parameters only), and the reverse only reads the outputs.
template <class T>
int run_taped(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vg, const void* const* in,
              void* const* out, const void* const* geom, int internal, void* vamax, void* ws, size_t wsb,
              hipStream_t st, const void* cache = nullptr, int32_t* iters = nullptr) {
  Step<T> S{};
  S.b = b;
  S.c = make_ctx<T>(b, prm, ramp, vg, st);
  setup_implicit(S, prm, cache);
  S.iters_out = iters;
  S.isplit = prm->isplit;
  S.internal = internal != 0;
  for (int f = 0; f < POM_NSTATE; ++f) {
    S.in[f] = static_cast<const T*>(in[f]);
    S.out[f] = static_cast<T*>(out[f]);
    if (!S.in[f] || !S.out[f]) return POM_E_ARG;
  }
  for (int f = 0; f < POM_NGEOM; ++f) {
    S.g[f] = static_cast<const T*>(geom[f]);
    if (!S.g[f] && (f < POM_NGEOM_BASE || S.open)) return POM_E_ARG;
  }
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  S.tape = true;
  S.plan(A, false);
  S.vamax = static_cast<T*>(vamax);
  Adj<T> J(S, S.c);
  J.plan(A, nullptr);   // the reverse's share of the workspace must fit as well
  if (A.off > wsb) return POM_E_ARG;
  S.forward();
  if (S.pw.status != POM_OK) return S.pw.status;
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

template <class T>
int run_reverse(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vg, const void* const* in,
                const void* const* out, const void* const* geom, int internal, const void* const* cot,
                void* const* grad, void* ws, size_t wsb, hipStream_t st, const void* cache = nullptr) {
  Step<T> S{};
  S.b = b;
  S.c = make_ctx<T>(b, prm, ramp, vg, st);
  setup_implicit(S, prm, cache);
  S.isplit = prm->isplit;
  S.internal = internal != 0;
  for (int f = 0; f < POM_NSTATE; ++f) {
    S.in[f] = static_cast<const T*>(in[f]);
    S.out[f] = const_cast<T*>(static_cast<const T*>(out[f]));   // read only in the reverse
    if (!S.in[f] || !S.out[f]) return POM_E_ARG;
  }
  for (int f = 0; f < POM_NGEOM; ++f) {
    S.g[f] = static_cast<const T*>(geom[f]);
    if (!S.g[f] && (f < POM_NGEOM_BASE || S.open)) return POM_E_ARG;
  }
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  S.tape = true;
  S.plan(A, false);
  S.vamax = nullptr;
  Adj<T> J(S, S.c);
  J.plan(A, reinterpret_cast<T* const*>(grad));
  if (A.off > wsb) return POM_E_ARG;
  S.replay_host();
  J.run(reinterpret_cast<const T* const*>(cot), reinterpret_cast<T* const*>(grad));
  if (S.pw.status != POM_OK) return S.pw.status;
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// ---------------------------------------------------------------- operator-level entry points
// The reference's operator methods that callers invoke on their own (dens from
// the corrector hook, baropg for the pressure-velocity error); each runs the
// same kernels as the step.
template <class T>
int run_dens(const pom_batch* b, const pom_params_t* prm, const void* vg, const void* t, const void* s,
             const void* rho_in, const void* h, const void* fsm, const void* b_rho_out, void* rho_out, void* b_t,
             void* b_s, void* b_rho_in, hipStream_t st, bool adjoint) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, vg, st);
  FieldTab<T> tab{};
  tab.f[OpDens::TT] = static_cast<const T*>(t);
  tab.f[OpDens::SS] = static_cast<const T*>(s);
  tab.f[OpDens::RHO] = static_cast<const T*>(rho_in);
  tab.g[OpDens::GH] = static_cast<const T*>(h);
  tab.g[OpDens::GFSM] = static_cast<const T*>(fsm);
  if (!adjoint) {
    launch_map<T, OpDens>(c, tab, OutTab<T>{{static_cast<T*>(rho_out)}});
  } else {
    COutTab<T> ob{{static_cast<const T*>(b_rho_out)}};
    T* ib[OpDens::NIN] = {static_cast<T*>(b_t), static_cast<T*>(b_s), static_cast<T*>(b_rho_in)};
    launch_gather<T, OpDens>(c, tab, ob, ib);
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// bcond(idx) of the explicit core alone (baroclinic.py:496-928), forward or
// adjoint, on the boundary point operators the fused step uses.  fields and
// outputs per idx as documented in the header.
template <class T, class Op>
void bcond_op(const Ctx<T>& c, const FieldTab<T>& tab, const int* fid, int nf, const int* oid, int no,
              void* const* out, const void* const* cot, void* const* grad, bool adjoint) {
  if (!adjoint) {
    OutTab<T> o{};
    for (int q = 0; q < no; ++q) o.o[oid[q]] = static_cast<T*>(out[q]);
    launch_map<T, Op>(c, tab, o);
  } else {
    COutTab<T> ob{};
    for (int q = 0; q < no; ++q) ob.o[oid[q]] = static_cast<const T*>(cot[q]);
    T* ib[Op::NIN] = {};
    for (int q = 0; q < nf; ++q) ib[fid[q]] = static_cast<T*>(grad[q]);
    launch_gather<T, Op>(c, tab, ob, ib);
  }
}

int bcond_nfields(int idx) {
  static const int n[7] = {0, 1, 3, 4, 8, 1, 6};
  return (idx >= 1 && idx <= 6) ? n[idx] : -1;
}
int bcond_nout(int idx) { return (idx == 1 || idx == 5) ? 1 : 2; }

// the semi-implicit core's closed-basin policy (boundary/closed_basin.py:27-87):
// reads the first one or two entries of the same field lists
template <class T, int IDX>
void imp_bcond(const Ctx<T>& c, const void* const* fields, const void* const* geom, void* const* out,
               const void* const* cot, void* const* grad, bool adjoint) {
  using O = OpImpBcond<IDX>;
  FieldTab<T> t{};
  for (int q = 0; q < O::NIN; ++q) t.f[q] = static_cast<const T*>(fields[q]);
  t.g[O::GFSM] = static_cast<const T*>(geom[POM_G_FSM]);
  t.g[O::GDUM] = static_cast<const T*>(geom[POM_G_DUM]);
  t.g[O::GDVM] = static_cast<const T*>(geom[POM_G_DVM]);
  const int fid[2] = {0, 1}, oid[2] = {0, 1};
  bcond_op<T, O>(c, t, fid, O::NIN, oid, O::NOUT, out, cot, grad, adjoint);
}

template <class T>
int run_bcond(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vg, int idx,
              const void* const* fields, const void* const* geom, void* const* out, const void* const* cot,
              void* const* grad, hipStream_t st, bool adjoint) {
  Ctx<T> c = make_ctx<T>(b, prm, ramp, vg, st);
  auto F = [&](int q) { return static_cast<const T*>(fields[q]); };
  auto G = [&](int q) { return static_cast<const T*>(geom[q]); };
  FieldTab<T> t{};
  if (prm->core == POM_CORE_IMPLICIT) {
    switch (idx) {
      case 1: imp_bcond<T, 1>(c, fields, geom, out, cot, grad, adjoint); break;
      case 2: imp_bcond<T, 2>(c, fields, geom, out, cot, grad, adjoint); break;
      case 3: {
        t.f[OpImpBc3::UF] = F(0);
        t.f[OpImpBc3::VF] = F(1);
        t.g[OpImpBc3::GDUM] = G(POM_G_DUM);
        t.g[OpImpBc3::GDVM] = G(POM_G_DVM);
        const int fid[2] = {OpImpBc3::UF, OpImpBc3::VF}, oid[2] = {0, 1};
        bcond_op<T, OpImpBc3>(c, t, fid, 2, oid, 2, out, cot, grad, adjoint);
        break;
      }
      case 4: imp_bcond<T, 4>(c, fields, geom, out, cot, grad, adjoint); break;
      case 5: imp_bcond<T, 5>(c, fields, geom, out, cot, grad, adjoint); break;
      default: imp_bcond<T, 6>(c, fields, geom, out, cot, grad, adjoint); break;
    }
    return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
  }
  switch (idx) {
    case 1: {
      t.f[OpBc1::RAW] = F(0);
      t.g[OpBc1::GFSM] = G(POM_G_FSM);
      const int fid[1] = {OpBc1::RAW}, oid[1] = {0};
      bcond_op<T, OpBc1>(c, t, fid, 1, oid, 1, out, cot, grad, adjoint);
      break;
    }
    case 2: {
      t.f[OpBcond2::UAFIN] = F(0);
      t.f[OpBcond2::VAFIN] = F(1);
      t.f[OpBcond2::EL] = F(2);
      t.g[OpBcond2::GH] = G(POM_G_H);
      t.g[OpBcond2::GDUM] = G(POM_G_DUM);
      t.g[OpBcond2::GDVM] = G(POM_G_DVM);
      const int fid[3] = {OpBcond2::UAFIN, OpBcond2::VAFIN, OpBcond2::EL}, oid[2] = {0, 1};
      bcond_op<T, OpBcond2>(c, t, fid, 3, oid, 2, out, cot, grad, adjoint);
      break;
    }
    case 3: {
      t.f[OpBc3::UF] = F(0);
      t.f[OpBc3::VF] = F(1);
      t.f[OpBc3::U] = F(2);
      t.f[OpBc3::V] = F(3);
      t.g[OpBc3::GH] = G(POM_G_H);
      t.g[OpBc3::GDUM] = G(POM_G_DUM);
      t.g[OpBc3::GDVM] = G(POM_G_DVM);
      const int fid[4] = {OpBc3::UF, OpBc3::VF, OpBc3::U, OpBc3::V}, oid[2] = {0, 1};
      bcond_op<T, OpBc3>(c, t, fid, 4, oid, 2, out, cot, grad, adjoint);
      break;
    }
    case 4: {
      using O = OpBcond4;
      const int fid[8] = {O::TF, O::SF, O::TT, O::SS, O::U, O::V, O::W, O::DT}, oid[2] = {0, 1};
      for (int q = 0; q < 8; ++q) t.f[fid[q]] = F(q);
      t.g[O::GDX] = G(POM_G_DX);
      t.g[O::GDY] = G(POM_G_DY);
      t.g[O::GFSM] = G(POM_G_FSM);
      bcond_op<T, O>(c, t, fid, 8, oid, 2, out, cot, grad, adjoint);
      break;
    }
    case 5: {
      t.f[OpBcond5::W] = F(0);
      t.g[OpBcond5::GFSM] = G(POM_G_FSM);
      const int fid[1] = {OpBcond5::W}, oid[1] = {0};
      bcond_op<T, OpBcond5>(c, t, fid, 1, oid, 1, out, cot, grad, adjoint);
      break;
    }
    default: {
      using O = OpBcond6;
      const int fid[6] = {O::Q2F, O::Q2LF, O::Q2, O::Q2L, O::U, O::V}, oid[2] = {0, 1};
      for (int q = 0; q < 6; ++q) t.f[fid[q]] = F(q);
      t.g[O::GDX] = G(POM_G_DX);
      t.g[O::GDY] = G(POM_G_DY);
      t.g[O::GFSM] = G(POM_G_FSM);
      bcond_op<T, O>(c, t, fid, 6, oid, 2, out, cot, grad, adjoint);
      break;
    }
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

template <class T>
int run_baropg(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vg, const void* const* in,
               void* const* out, const void* const* cot, void* const* grad, hipStream_t st, bool adjoint) {
  Ctx<T> c = make_ctx<T>(b, prm, ramp, vg, st);
  const dim3 g2(c.nblk2), tb(TPB);
  if (c.nblk2 == 0) return POM_OK;
  auto I = [&](int k) { return static_cast<const T*>(in[k]); };
  if (!adjoint) {
    BaropgArgs<T> a{I(0), I(1), I(2), I(3), I(4), I(5), I(6), static_cast<T*>(out[0]), static_cast<T*>(out[1]),
                    static_cast<T*>(out[2])};
    hipLaunchKernelGGL(k_baropg<T>, g2, tb, 0, c.stream, c.B, c.vert, c.ramp, c.prm, a);
  } else {
    BaropgAdjArgs<T> a{I(0), I(1), I(2), I(3), I(4), I(5), I(6),
                       static_cast<const T*>(cot[0]), static_cast<const T*>(cot[1]), static_cast<const T*>(cot[2]),
                       static_cast<T*>(grad[0]), static_cast<T*>(grad[1])};
    hipLaunchKernelGGL(k_baropg_adj<T>, g2, tb, 0, c.stream, c.B, c.vert, c.ramp, c.prm, a);
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// ---------------------------------------------------------------- profq / proft on their own
// The Mellor-Yamada closure and the tracer vertical diffusion as the
// reference's operator methods (tracers.py:158-431, 890-1010), on the step's
// kernels: profq is k_profq followed by the km / kh edge copy (OpKEdge);
// its reverse recomputes the forward with the tape, gathers the edge copy
// and runs the four reverse kernels and the neighbour gather.
enum { PQ_Q2F, PQ_Q2LF, PQ_Q2, PQ_Q2B, PQ_Q2LB, PQ_KM, PQ_KH, PQ_KQ, PQ_U, PQ_V, PQ_T, PQ_S, PQ_RHO,
       PQ_WUSURF, PQ_WVSURF, PQ_WUBOT, PQ_WVBOT, PQ_ETF, PQ_DT, PQ_H, PQ_FSM, PQ_NIN };
enum { PQ_O_Q2, PQ_O_Q2L, PQ_O_L, PQ_O_KQ, PQ_O_KM, PQ_O_KH, PQ_O_SM, PQ_O_SH, PQ_O_CC, PQ_O_A, PQ_O_C, PQ_O_DH,
       PQ_NOUT };
constexpr int PQ_NDIFF_OUT = 6;   // cotangents: q2, q2l, l, kq, km, kh

template <class T> ProfqArgs<T> profq_args(const void* const* in) {
  auto I = [&](int k) { return static_cast<const T*>(in[k]); };
  ProfqArgs<T> a{};
  a.q2f = I(PQ_Q2F); a.q2lf = I(PQ_Q2LF); a.q2 = I(PQ_Q2); a.q2b = I(PQ_Q2B); a.q2lb = I(PQ_Q2LB);
  a.km = I(PQ_KM); a.kh = I(PQ_KH); a.kq = I(PQ_KQ); a.u = I(PQ_U); a.v = I(PQ_V); a.t = I(PQ_T); a.s = I(PQ_S);
  a.rho = I(PQ_RHO); a.wusurf = I(PQ_WUSURF); a.wvsurf = I(PQ_WVSURF); a.wubot = I(PQ_WUBOT);
  a.wvbot = I(PQ_WVBOT); a.etf = I(PQ_ETF); a.dt = I(PQ_DT); a.h = I(PQ_H);
  return a;
}
template <class T> FieldTab<T> kedge_tab(const T* km, const T* kh, const void* fsm) {
  FieldTab<T> t{};
  t.f[OpKEdge::KM] = km;
  t.f[OpKEdge::KH] = kh;
  t.g[OpKEdge::GFSM] = static_cast<const T*>(fsm);
  return t;
}
// workspace of pom_profq_adj: the forward's six outputs, its tape, the
// reverse kernels' scratch, and the cotangents of km / kh before the edge copy
template <class T> struct ProfqWs {
  T *q2o, *q2lo, *lo, *kqo, *kmo, *kho, *b_kmo, *b_kho, *unb, *vnb, *wubnb, *wvbnb;
  T *r3[8], *r2[2];
  ProfqTape<T> tape;
  void plan(Arena& A, int64_t N3, int64_t N2) {
    q2o = A.get<T>(N3); q2lo = A.get<T>(N3); lo = A.get<T>(N3); kqo = A.get<T>(N3); kmo = A.get<T>(N3);
    kho = A.get<T>(N3); b_kmo = A.get<T>(N3); b_kho = A.get<T>(N3); unb = A.get<T>(N3); vnb = A.get<T>(N3);
    wubnb = A.get<T>(N2); wvbnb = A.get<T>(N2);
    for (T*& r : r3) r = A.get<T>(N3);
    for (T*& r : r2) r = A.get<T>(N2);
    tape.ee1 = A.get<T>(N3); tape.gg1 = A.get<T>(N3); tape.q2n = A.get<T>(N3); tape.ee2 = A.get<T>(N3);
    tape.gg2 = A.get<T>(N3); tape.q2ln = A.get<T>(N3); tape.ghr = A.get<T>(N3); tape.kn = A.get<T>(N3);
  }
};
template <class T> size_t profq_ws_size(const pom_batch* b) {
  Arena A{nullptr, 0, 0, true};
  ProfqWs<T> w{};
  w.plan(A, b->n3, b->n2);
  return A.off + 256;
}

template <class T>
int run_profq(const pom_batch* b, const pom_params_t* prm, const void* vg, const void* const* in, void* const* out,
              void* ws, size_t wsb, hipStream_t st) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, vg, st);
  if (c.nblk2 == 0) return POM_OK;
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  ProfqWs<T> w{};
  w.plan(A, b->n3, b->n2);
  if (A.off > wsb) return POM_E_ARG;
  ProfqArgs<T> a = profq_args<T>(in);
  auto O = [&](int k) { return static_cast<T*>(out[k]); };
  a.q2o = O(PQ_O_Q2); a.q2lo = O(PQ_O_Q2L); a.lo = O(PQ_O_L); a.kqo = O(PQ_O_KQ);
  a.kmo = w.kmo; a.kho = w.kho;   // before the edge copy
  a.smo = O(PQ_O_SM); a.sho = O(PQ_O_SH); a.cco = O(PQ_O_CC); a.ao = O(PQ_O_A); a.co = O(PQ_O_C);
  a.dho = O(PQ_O_DH);
  hipLaunchKernelGGL(k_profq<T>, dim3(c.nblk2), dim3(TPB), 0, st, c.B, c.vert, c.prm, a);
  launch_map<T, OpKEdge>(c, kedge_tab<T>(w.kmo, w.kho, in[PQ_FSM]), OutTab<T>{{O(PQ_O_KM), O(PQ_O_KH)}});
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// grad[k] += (d outputs / d in[k])^T cot for the 17 differentiable inputs
// (q2f .. rho, wubot, wvbot, etf, dt; null entries skipped)
template <class T>
int run_profq_adj(const pom_batch* b, const pom_params_t* prm, const void* vg, const void* const* in,
                  const void* const* cot, void* const* grad, void* ws, size_t wsb, hipStream_t st) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, vg, st);
  if (c.nblk2 == 0) return POM_OK;
  const dim3 g2(c.nblk2), tb(TPB);
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  ProfqWs<T> w{};
  w.plan(A, b->n3, b->n2);
  if (A.off > wsb) return POM_E_ARG;
  ProfqArgs<T> f = profq_args<T>(in);
  f.q2o = w.q2o; f.q2lo = w.q2lo; f.lo = w.lo; f.kqo = w.kqo; f.kmo = w.kmo; f.kho = w.kho;
  f.tape = w.tape;
  hipLaunchKernelGGL(k_profq<T>, g2, tb, 0, st, c.B, c.vert, c.prm, f);
  auto C = [&](int k) { return static_cast<const T*>(cot[k]); };
  auto G = [&](int k) { return static_cast<T*>(grad[k]); };
  {   // the edge copy reversed: cotangents of the raw km, kh (assigned at every point)
    COutTab<T> ob{{C(PQ_O_KM), C(PQ_O_KH)}};
    T* ib[OpKEdge::NIN] = {w.b_kmo, w.b_kho};
    if (C(PQ_O_KM) || C(PQ_O_KH)) {
      launch_gather<T, OpKEdge>(c, kedge_tab<T>(w.kmo, w.kho, in[PQ_FSM]), ob, ib,
                                (1u << OpKEdge::KM) | (1u << OpKEdge::KH));
    }
  }
  ProfqAdjArgs<T> a{};
  a.fw = f;
  a.b_q2o = C(PQ_O_Q2); a.b_q2lo = C(PQ_O_Q2L); a.b_lo = C(PQ_O_L); a.b_kqo = C(PQ_O_KQ);
  const bool kcot = C(PQ_O_KM) || C(PQ_O_KH);
  a.b_kmo = kcot ? w.b_kmo : nullptr;
  a.b_kho = kcot ? w.b_kho : nullptr;
  a.b_q2f = G(PQ_Q2F); a.b_q2lf = G(PQ_Q2LF); a.b_q2 = G(PQ_Q2); a.b_q2b = G(PQ_Q2B); a.b_q2lb = G(PQ_Q2LB);
  a.b_km = G(PQ_KM); a.b_kh = G(PQ_KH); a.b_kq = G(PQ_KQ); a.b_u = G(PQ_U); a.b_v = G(PQ_V); a.b_t = G(PQ_T);
  a.b_s = G(PQ_S); a.b_rho = G(PQ_RHO); a.b_wubot = G(PQ_WUBOT); a.b_wvbot = G(PQ_WVBOT); a.b_etf = G(PQ_ETF);
  a.b_dt = G(PQ_DT);
  a.s_unb = w.unb; a.s_vnb = w.vnb; a.s_wubnb = w.wubnb; a.s_wvbnb = w.wvbnb;
  a.r_bl = w.r3[0]; a.r_bgh = w.r3[1]; a.r_bq2ln = w.r3[2]; a.r_ba = w.r3[3]; a.r_bc = w.r3[4]; a.r_bkn = w.r3[5];
  a.r_bdtef = w.r3[6]; a.r_bboy = w.r3[7]; a.r_bq2bot = w.r2[0]; a.r_bdh = w.r2[1];
  hipLaunchKernelGGL(k_profq_adj_a<T>, g2, tb, 0, st, c.B, c.vert, c.prm, a);
  hipLaunchKernelGGL(k_profq_adj_b<T>, g2, tb, 0, st, c.B, c.vert, c.prm, a);
  hipLaunchKernelGGL(k_profq_adj_c1<T>, g2, tb, 0, st, c.B, c.vert, c.prm, a);
  hipLaunchKernelGGL(k_profq_adj_c2<T>, g2, tb, 0, st, c.B, c.vert, c.prm, a);
  hipLaunchKernelGGL(k_profq_adj2<T>, dim3(c.nblk3), tb, 0, st, c.B, a);
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

enum { PT_F, PT_KH, PT_WFSURF, PT_FSURF, PT_ETF, PT_H, PT_NIN };
enum { PT_O_F, PT_O_A, PT_O_C, PT_O_EE, PT_O_GG, PT_O_DH, PT_NOUT };

template <class T>
int run_proft(const pom_batch* b, const pom_params_t* prm, const void* vg, const void* const* in, int nbc,
              void* const* out, const void* cot, void* const* grad, hipStream_t st, bool adjoint) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, vg, st);
  if (c.nblk2 == 0) return POM_OK;
  auto I = [&](int k) { return static_cast<const T*>(in[k]); };
  if (!adjoint) {
    auto O = [&](int k) { return static_cast<T*>(out[k]); };
    ProftArgs<T> a{I(PT_F), I(PT_KH), I(PT_ETF), I(PT_H), I(PT_WFSURF), I(PT_FSURF), O(PT_O_F), nbc,
                   O(PT_O_A), O(PT_O_C), O(PT_O_EE), O(PT_O_GG), O(PT_O_DH)};
    hipLaunchKernelGGL(k_proft<T>, dim3(c.nblk2), dim3(TPB), 0, st, c.B, c.vert, c.prm, a);
  } else {
    ProftAdjArgs<T> a{};
    a.f = I(PT_F); a.kh = I(PT_KH); a.etf = I(PT_ETF); a.h = I(PT_H); a.wfsurf = I(PT_WFSURF);
    a.fsurf = I(PT_FSURF); a.b_out = static_cast<const T*>(cot);
    a.b_f = static_cast<T*>(grad[0]); a.b_kh = static_cast<T*>(grad[1]); a.b_etf = static_cast<T*>(grad[2]);
    a.nbc = nbc;
    hipLaunchKernelGGL(k_proft_adj<T>, dim3(c.nblk2), dim3(TPB), 0, st, c.B, c.vert, c.prm, a);
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// ---------------------------------------------------------------- eta solve entry points
// The semi-implicit core's elliptic operator and eta solve as stand-alone
// calls (solve_implicit_eta_cg / apply_elliptic_operator and the adjoint of
// _ImplicitEtaSolveFunction), on the same kernels as the step.
template <class T> struct EtaWs {
  Pcg<T> w;
  T *a, *x, *nlam, *lift;
  void plan(Arena& A, const pom_batch* b) {
    w.plan(A, b->n2, b->nblk2, b->M);
    a = A.get<T>(b->n2);
    x = A.get<T>(b->n2);
    nlam = A.get<T>(b->n2);
    lift = A.get<T>(b->n2);
  }
};

template <class T> size_t eta_ws_size(const pom_batch* b) {
  Arena A{nullptr, 0, 0, true};
  EtaWs<T> e{};
  e.plan(A, b);
  return A.off + 256;
}

// ---- static scratch of pom_eta_solve_noalloc
// The eta solve's work arrays in a file-scope device array instead of a
// caller workspace.  Each GPU holds its own copy until the process ends.  In
// host code `scratch` names a host-side stand-in, never device memory: the
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static_assert(NOALLOC_SCRATCH_BYTES == POM_NOALLOC_SCRATCH_BYTES, "header and library disagree");
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// device address of the current device's scratch; POM_E_HIP if it cannot be found
int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return POM_E_HIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return POM_E_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return POM_OK;
}

template <class T> EllGeom<T> ell_geom(const Ctx<T>& c, const void* const* geom, const pom_params_t* prm) {
  auto G = [&](int f) { return static_cast<const T*>(geom[f]); };
  return EllGeom<T>{G(POM_G_DX), G(POM_G_DY), G(POM_G_ART), G(POM_G_DUM), G(POM_G_DVM),
                    4.0 * (c.prm.dti * c.prm.dti) * c.prm.grav * c.prm.alpha,
                    prm->boundary == POM_BOUNDARY_TIDE ? G(POM_G_ETA_FIXED) : nullptr};
}

template <class T> FieldTab<T> ell_tab(const void* const* geom, const void* eta, const void* d) {
  FieldTab<T> t{};
  t.f[OpEll::ETA] = static_cast<const T*>(eta);
  t.f[OpEll::D_] = static_cast<const T*>(d);
  t.g[OpEll::GDX] = static_cast<const T*>(geom[POM_G_DX]);
  t.g[OpEll::GDY] = static_cast<const T*>(geom[POM_G_DY]);
  t.g[OpEll::GART] = static_cast<const T*>(geom[POM_G_ART]);
  t.g[OpEll::GDUM] = static_cast<const T*>(geom[POM_G_DUM]);
  t.g[OpEll::GDVM] = static_cast<const T*>(geom[POM_G_DVM]);
  return t;
}

template <class T>
int run_elliptic(const pom_batch* b, const pom_params_t* prm, const void* const* geom, const void* eta, const void* d,
                 void* out, const void* gout, void* b_eta, void* b_d, hipStream_t st, bool adjoint) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, nullptr, st);
  if (c.nblk2 == 0) return POM_OK;
  const FieldTab<T> t = ell_tab<T>(geom, eta, d);
  if (!adjoint) {
    launch_map<T, OpEll>(c, t, OutTab<T>{{static_cast<T*>(out)}});
  } else {
    COutTab<T> ob{{static_cast<const T*>(gout)}};
    T* ib[OpEll::NIN] = {static_cast<T*>(b_eta), static_cast<T*>(b_d)};
    launch_gather<T, OpEll>(c, t, ob, ib);
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

template <class T>
int run_eta_solve(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                  const void* rhs, const void* d, const void* x0, void* x, int32_t* iters, void* ws, size_t wsb,
                  hipStream_t st) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, nullptr, st);
  if (c.nblk2 == 0) return POM_OK;
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  EtaWs<T> e{};
  e.plan(A, b);
  if (A.off > wsb) return POM_E_ARG;
  setup_pcg(e.w, b, prm, cache);
  const bool tide = prm->boundary == POM_BOUNDARY_TIDE;
  if (tide)
    hipLaunchKernelGGL(k_tide_lift<T>, dim3(c.nblk2), dim3(TPB), 0, st, c.B,
                       static_cast<const T*>(geom[POM_G_TIDE_NODE]), static_cast<const T*>(geom[POM_G_TIDE_AMP]),
                       static_cast<const T*>(geom[POM_G_TIDE_PHASE]), prm->tide_wt, e.lift);
  hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(256), 0, st, e.w.tol, (int64_t)b->M, prm->eta_tol);
  e.w.solve(c, ell_geom<T>(c, geom, prm), static_cast<const T*>(rhs), static_cast<const T*>(d),
            static_cast<const T*>(x0), static_cast<T*>(x), prm->eta_rel_tol, prm->eta_max_iter,
            tide ? e.lift : nullptr);
  if (e.w.status != POM_OK) return e.w.status;
  if (iters) hipLaunchKernelGGL(k_copy<int32_t>, dim3(1), dim3(256), 0, st, iters, e.w.its, (int64_t)b->M);
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// lambda = area A^-1 (g / area) on the max-normalised right-hand side
// (core.py:231-252); b_rhs = lambda; b_d += -(d(A x)/dd)^T lambda (core.py:645-675)
template <class T>
int run_eta_solve_adj(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                      const void* x, const void* d, const void* gx, void* b_rhs, void* b_d, int32_t* iters, void* ws,
                      size_t wsb, hipStream_t st) {
  Ctx<T> c = make_ctx<T>(b, prm, nullptr, nullptr, st);
  if (c.nblk2 == 0) return POM_OK;
  Arena A{static_cast<char*>(ws), 0, wsb, false};
  EtaWs<T> e{};
  e.plan(A, b);
  if (A.off > wsb) return POM_E_ARG;
  setup_pcg(e.w, b, prm, cache);
  const dim3 g2(c.nblk2), tb(TPB), gm(c.B.M);
  const T* art = static_cast<const T*>(geom[POM_G_ART]);
  hipLaunchKernelGGL(k_eta_adj_in<T>, g2, tb, 0, st, c.B, static_cast<const T*>(gx), art, e.a, e.w.part);
  hipLaunchKernelGGL(k_eta_adj_scale, gm, tb, 0, st, c.B, e.w.part, e.w.sc, e.w.tol, prm->eta_tol);
  hipLaunchKernelGGL(k_eta_adj_norm<T>, g2, tb, 0, st, c.B, e.a, e.w.sc);
  e.w.solve(c, ell_geom<T>(c, geom, prm), e.a, static_cast<const T*>(d), nullptr, e.x, prm->eta_rel_tol,
            prm->eta_max_iter);
  if (e.w.status != POM_OK) return e.w.status;
  hipLaunchKernelGGL(k_eta_adj_out<T>, g2, tb, 0, st, c.B, e.x, art, e.w.sc, static_cast<T*>(b_rhs), e.nlam);
  if (iters) hipLaunchKernelGGL(k_copy<int32_t>, dim3(1), dim3(256), 0, st, iters, e.w.its, (int64_t)b->M);
  if (b_d) {
    COutTab<T> ob{{e.nlam}};
    T* ib[OpEll::NIN] = {nullptr, static_cast<T*>(b_d)};
    launch_gather<T, OpEll>(c, ell_tab<T>(geom, x, d), ob, ib);
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

bool ell_geom_ok(const void* const* geom) {
  if (!geom) return false;
  for (int f : {POM_G_DX, POM_G_DY, POM_G_ART, POM_G_DUM, POM_G_DVM})
    if (!geom[f]) return false;
  return true;
}

// the tide policy's fixed set (and, for the forward solve, its nodes)
bool tide_geom_ok(const pom_params_t* prm, const void* const* geom, bool nodes) {
  if (prm->boundary != POM_BOUNDARY_TIDE) return true;
  if (!geom[POM_G_ETA_FIXED]) return false;
  return !nodes || (geom[POM_G_TIDE_NODE] && geom[POM_G_TIDE_AMP] && geom[POM_G_TIDE_PHASE]);
}

}  // namespace

extern "C" {

int pom_batch_create(const pom_member_t* members, int32_t n, int32_t dtype, pom_batch** out) {
  if (!members || n < 1 || !out || (dtype != POM_F32 && dtype != POM_F64)) return POM_E_ARG;
  std::vector<Member> hm(n);
  std::vector<int32_t> b3(n + 1), b2(n + 1);
  b3[0] = b2[0] = 0;
  int64_t n2 = 0, n3 = 0;
  for (int m = 0; m < n; ++m) {
    const pom_member_t& s = members[m];
    if (s.im < 4 || s.jm < 4) return POM_E_ARG;
    hm[m] = Member{s.im, s.jm, s.off2, s.off3};
    const int64_t c2 = (int64_t)s.im * s.jm, c3 = c2 * KB;
    // kernels address a member's field with 32-bit byte offsets
    if (c3 * (dtype == POM_F64 ? 8 : 4) >= (int64_t(1) << 31)) return POM_E_ARG;
    b3[m + 1] = b3[m] + (int32_t)((c3 + TPB - 1) / TPB);
    b2[m + 1] = b2[m] + (int32_t)((c2 + TPB - 1) / TPB);
    n2 = std::max(n2, s.off2 + c2);
    n3 = std::max(n3, s.off3 + c3);
  }
  // and a field of the whole batch with 32-bit byte offsets from its pointer
  if (n3 * (dtype == POM_F64 ? 8 : 4) >= (int64_t(1) << 32)) return POM_E_ARG;
  pom_batch* b = new pom_batch{};
  b->h_mem = hm;
  b->dtype = dtype;
  b->M = n;
  b->nblk3 = b3[n];
  b->nblk2 = b2[n];
  b->n2 = n2;
  b->n3 = n3;
  if (hipMalloc(&b->d_mem, sizeof(Member) * n) != hipSuccess ||
      hipMalloc(&b->d_blk3, sizeof(int32_t) * (n + 1)) != hipSuccess ||
      hipMalloc(&b->d_blk2, sizeof(int32_t) * (n + 1)) != hipSuccess) {
    pom_batch_destroy(b);
    return POM_E_HIP;
  }
  if (hipMemcpy(b->d_mem, hm.data(), sizeof(Member) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_blk3, b3.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_blk2, b2.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice) != hipSuccess) {
    pom_batch_destroy(b);
    return POM_E_HIP;
  }
  // edge-band columns, by distance to the nearest edge (Batch::band)
  std::vector<int2> bucket[BAND_W];
  for (int m = 0; m < n; ++m) {
    const int im = hm[m].im, jm = hm[m].jm;
    for (int i = 0; i < im; ++i)
      for (int j = 0; j < jm; ++j) {
        const int d = std::min(std::min(i, im - 1 - i), std::min(j, jm - 1 - j));
        if (d < BAND_W) bucket[d].push_back(make_int2(m, i * jm + j));
      }
  }
  std::vector<int2> band;
  b->nband[0] = 0;
  for (int d = 0; d < BAND_W; ++d) {
    band.insert(band.end(), bucket[d].begin(), bucket[d].end());
    b->nband[d + 1] = (int32_t)band.size();
  }
  if (hipMalloc(&b->d_band, sizeof(int2) * band.size()) != hipSuccess ||
      hipMemcpy(b->d_band, band.data(), sizeof(int2) * band.size(), hipMemcpyHostToDevice) != hipSuccess) {
    pom_batch_destroy(b);
    return POM_E_HIP;
  }
  *out = b;
  return POM_OK;
}

int pom_batch_tune(pom_batch* b, const pom_tuning_t* t) {
  if (!b || !t) return POM_E_ARG;
  const int32_t v[6] = {t->map2_blocks_per_group, t->map2_min_groups, t->flat2_blocks_per_group,
                        t->flat2_min_groups, t->flat3_blocks_per_group, t->flat3_min_groups};
  for (int q = 0; q < 6; ++q)
    if (v[q] < 1 || v[q] > (1 << 20)) return POM_E_ARG;
  b->tune.map2_bpg = v[0]; b->tune.map2_min = v[1];
  b->tune.flat2_bpg = v[2]; b->tune.flat2_min = v[3];
  b->tune.flat3_bpg = v[4]; b->tune.flat3_min = v[5];
  return POM_OK;
}

int pom_batch_get_tuning(const pom_batch* b, pom_tuning_t* t) {
  if (!b || !t) return POM_E_ARG;
  t->map2_blocks_per_group = b->tune.map2_bpg; t->map2_min_groups = b->tune.map2_min;
  t->flat2_blocks_per_group = b->tune.flat2_bpg; t->flat2_min_groups = b->tune.flat2_min;
  t->flat3_blocks_per_group = b->tune.flat3_bpg; t->flat3_min_groups = b->tune.flat3_min;
  return POM_OK;
}

int pom_profile_stage(int32_t stage) {
  if (stage < -1 || stage >= POM_NSTAGE) return POM_E_ARG;
  // The events are created and recorded once here, outside any stream
  // capture: an event whose only records were captured into a HIP graph is
  // not a valid handle for hipEventElapsedTime, even after the graph ran.
  while (stage >= 0 && g_prof.pool.size() < PROF_POOL) {
    hipEvent_t x;
    if (hipEventCreate(&x) != hipSuccess) return POM_E_HIP;
    g_prof.pool.push_back(x);
    if (hipEventRecord(x, nullptr) != hipSuccess || hipEventSynchronize(x) != hipSuccess) return POM_E_HIP;
  }
  g_prof.stage = stage;
  if (stage >= 0) g_prof.used = 0;   // stopping (-1) keeps the recorded events readable
  return POM_OK;
}

int pom_profile_read(double* total_ms, int32_t* count) {
  if (!total_ms || !count) return POM_E_ARG;
  double t = 0.0;
  for (size_t i = 0; i + 1 < g_prof.used; i += 2) {
    if (hipEventSynchronize(g_prof.pool[i + 1]) != hipSuccess) return POM_E_HIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_prof.pool[i], g_prof.pool[i + 1]) != hipSuccess) return POM_E_HIP;
    t += ms;
  }
  *total_ms = t;
  *count = (int32_t)(g_prof.used / 2);
  g_prof.used = 0;
  return POM_OK;
}

int pom_batch_destroy(pom_batch* b) {
  if (!b) return POM_E_ARG;
  (void)hipFree(b->d_mem);
  (void)hipFree(b->d_blk3);
  (void)hipFree(b->d_blk2);
  (void)hipFree(b->d_band);
  delete b;
  return POM_OK;
}

int pom_explicit_workspace(const pom_batch* b, const pom_params_t* prm, int32_t adjoint, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_EXPLICIT)) return e;
  if (adjoint && prm->boundary == POM_BOUNDARY_TIDE) return POM_E_UNSUPPORTED;   // forward-only family
  *bytes = b->dtype == POM_F32 ? ws_size<float>(b, prm, adjoint != 0) : ws_size<double>(b, prm, adjoint != 0);
  return POM_OK;
}

int pom_explicit_step(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                      const void* const* state_in, void* const* state_out, const void* const* geom, int32_t internal,
                      void* vamax_out, void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !ramp || !vgrid || !state_in || !state_out || !geom || !vamax_out || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_EXPLICIT)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_step<float>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes, st);
  return run_step<double>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes, st);
}

int pom_explicit_step_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                          const void* const* state_in, const void* const* geom, int32_t internal,
                          const void* const* cot_out, void* const* grad_in, void* workspace, size_t ws_bytes,
                          void* stream) {
  if (!b || !ramp || !vgrid || !state_in || !geom || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_EXPLICIT)) return e;
  if (prm->boundary == POM_BOUNDARY_TIDE) return POM_E_UNSUPPORTED;   // forward-only family
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_adj<float>(b, prm, ramp, vgrid, state_in, geom, internal, cot_out, grad_in, workspace, ws_bytes, st);
  return run_adj<double>(b, prm, ramp, vgrid, state_in, geom, internal, cot_out, grad_in, workspace, ws_bytes, st);
}

int pom_explicit_step_taped(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                            const void* const* state_in, void* const* state_out, const void* const* geom,
                            int32_t internal, void* vamax_out, void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !ramp || !vgrid || !state_in || !state_out || !geom || !vamax_out || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_EXPLICIT)) return e;
  if (prm->boundary == POM_BOUNDARY_TIDE) return POM_E_UNSUPPORTED;   // forward-only family
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_taped<float>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes,
                            st);
  return run_taped<double>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes,
                           st);
}

int pom_explicit_step_reverse(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                              const void* const* state_in, const void* const* state_out, const void* const* geom,
                              int32_t internal, const void* const* cot_out, void* const* grad_in, void* workspace,
                              size_t ws_bytes, void* stream) {
  if (!b || !ramp || !vgrid || !state_in || !state_out || !geom || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_EXPLICIT)) return e;
  if (prm->boundary == POM_BOUNDARY_TIDE) return POM_E_UNSUPPORTED;   // forward-only family
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_reverse<float>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, cot_out, grad_in, workspace,
                              ws_bytes, st);
  return run_reverse<double>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, cot_out, grad_in, workspace,
                             ws_bytes, st);
}

int pom_dens(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* t, const void* s,
             const void* rho_in, const void* h, const void* fsm, void* rho_out, void* stream) {
  if (!b || !prm || !vgrid || !t || !s || !rho_in || !h || !fsm || !rho_out) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_dens<float>(b, prm, vgrid, t, s, rho_in, h, fsm, nullptr, rho_out, nullptr, nullptr, nullptr, st, false);
  return run_dens<double>(b, prm, vgrid, t, s, rho_in, h, fsm, nullptr, rho_out, nullptr, nullptr, nullptr, st, false);
}

int pom_dens_adj(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* t, const void* s,
                 const void* rho_in, const void* h, const void* fsm, const void* b_rho_out, void* b_t, void* b_s,
                 void* b_rho_in, void* stream) {
  if (!b || !prm || !vgrid || !t || !s || !rho_in || !h || !fsm || !b_rho_out) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_dens<float>(b, prm, vgrid, t, s, rho_in, h, fsm, b_rho_out, nullptr, b_t, b_s, b_rho_in, st, true);
  return run_dens<double>(b, prm, vgrid, t, s, rho_in, h, fsm, b_rho_out, nullptr, b_t, b_s, b_rho_in, st, true);
}

int pom_bcond(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid, int32_t idx,
              const void* const* fields, const void* const* geom, void* const* out, void* stream) {
  const int nf = bcond_nfields(idx);
  if (!b || !prm || !ramp || !vgrid || !fields || !geom || !out || nf < 0) return POM_E_ARG;
  if (prm->boundary != POM_BOUNDARY_CLOSED) return POM_E_UNSUPPORTED;
  for (int q = 0; q < nf; ++q)
    if (!fields[q]) return POM_E_ARG;
  for (int q = 0; q < bcond_nout(idx); ++q)
    if (!out[q]) return POM_E_ARG;
  for (int q = 0; q < POM_NGEOM_BASE; ++q)
    if (!geom[q]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_bcond<float>(b, prm, ramp, vgrid, idx, fields, geom, out, nullptr, nullptr, st, false);
  return run_bcond<double>(b, prm, ramp, vgrid, idx, fields, geom, out, nullptr, nullptr, st, false);
}

int pom_bcond_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid, int32_t idx,
                  const void* const* fields, const void* const* geom, const void* const* cot_out,
                  void* const* grad_fields, void* stream) {
  const int nf = bcond_nfields(idx);
  if (!b || !prm || !ramp || !vgrid || !fields || !geom || !cot_out || !grad_fields || nf < 0) return POM_E_ARG;
  if (prm->boundary != POM_BOUNDARY_CLOSED) return POM_E_UNSUPPORTED;
  for (int q = 0; q < nf; ++q)
    if (!fields[q]) return POM_E_ARG;
  for (int q = 0; q < POM_NGEOM_BASE; ++q)
    if (!geom[q]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_bcond<float>(b, prm, ramp, vgrid, idx, fields, geom, nullptr, cot_out, grad_fields, st, true);
  return run_bcond<double>(b, prm, ramp, vgrid, idx, fields, geom, nullptr, cot_out, grad_fields, st, true);
}

int pom_baropg(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
               const void* const* in, void* const* out, void* stream) {
  if (!b || !prm || !ramp || !vgrid || !in || !out) return POM_E_ARG;
  for (int k = 0; k < 7; ++k)
    if (!in[k]) return POM_E_ARG;
  for (int k = 0; k < 3; ++k)
    if (!out[k]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_baropg<float>(b, prm, ramp, vgrid, in, out, nullptr, nullptr, st, false);
  return run_baropg<double>(b, prm, ramp, vgrid, in, out, nullptr, nullptr, st, false);
}

int pom_baropg_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                   const void* const* in, const void* const* cot_out, void* const* grad_in, void* stream) {
  if (!b || !prm || !ramp || !vgrid || !in || !cot_out || !grad_in) return POM_E_ARG;
  for (int k = 0; k < 7; ++k)
    if (!in[k]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_baropg<float>(b, prm, ramp, vgrid, in, nullptr, cot_out, grad_in, st, true);
  return run_baropg<double>(b, prm, ramp, vgrid, in, nullptr, cot_out, grad_in, st, true);
}


int pom_profq_workspace(const pom_batch* b, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  *bytes = b->dtype == POM_F32 ? profq_ws_size<float>(b) : profq_ws_size<double>(b);
  return POM_OK;
}

int pom_profq(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in,
              void* const* out, void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !prm || !vgrid || !in || !out || !workspace) return POM_E_ARG;
  for (int k = 0; k < PQ_NIN; ++k)
    if (!in[k]) return POM_E_ARG;
  for (int k = 0; k < PQ_NDIFF_OUT; ++k)
    if (!out[k]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_profq<float>(b, prm, vgrid, in, out, workspace, ws_bytes, st);
  return run_profq<double>(b, prm, vgrid, in, out, workspace, ws_bytes, st);
}

int pom_profq_adj(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in,
                  const void* const* cot_out, void* const* grad_in, void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !prm || !vgrid || !in || !cot_out || !grad_in || !workspace) return POM_E_ARG;
  for (int k = 0; k < PQ_NIN; ++k)
    if (!in[k]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_profq_adj<float>(b, prm, vgrid, in, cot_out, grad_in, workspace, ws_bytes, st);
  return run_profq_adj<double>(b, prm, vgrid, in, cot_out, grad_in, workspace, ws_bytes, st);
}

int pom_proft(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in, int32_t nbc,
              void* const* out, void* stream) {
  if (!b || !prm || !vgrid || !in || !out || !out[0]) return POM_E_ARG;
  if (nbc == 2 || nbc == 4) return POM_E_UNSUPPORTED;   // shortwave absorption: not built
  if (nbc != 1 && nbc != 3) return POM_E_ARG;
  for (int k = 0; k < PT_NIN; ++k)
    if (!in[k]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_proft<float>(b, prm, vgrid, in, nbc, out, nullptr, nullptr, st, false);
  return run_proft<double>(b, prm, vgrid, in, nbc, out, nullptr, nullptr, st, false);
}

int pom_proft_adj(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in,
                  int32_t nbc, const void* cot_f, void* const* grad_in, void* stream) {
  if (!b || !prm || !vgrid || !in || !cot_f || !grad_in) return POM_E_ARG;
  if (nbc == 2 || nbc == 4) return POM_E_UNSUPPORTED;
  if (nbc != 1 && nbc != 3) return POM_E_ARG;
  for (int k = 0; k < PT_NIN; ++k)
    if (!in[k]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_proft<float>(b, prm, vgrid, in, nbc, nullptr, cot_f, grad_in, st, true);
  return run_proft<double>(b, prm, vgrid, in, nbc, nullptr, cot_f, grad_in, st, true);
}

int pom_elliptic_apply(const pom_batch* b, const pom_params_t* prm, const void* const* geom, const void* eta,
                       const void* d, void* out, void* stream) {
  if (!b || !prm || !ell_geom_ok(geom) || !eta || !d || !out) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_elliptic<float>(b, prm, geom, eta, d, out, nullptr, nullptr, nullptr, st, false);
  return run_elliptic<double>(b, prm, geom, eta, d, out, nullptr, nullptr, nullptr, st, false);
}

int pom_elliptic_apply_adj(const pom_batch* b, const pom_params_t* prm, const void* const* geom, const void* eta,
                           const void* d, const void* cot_out, void* grad_eta, void* grad_d, void* stream) {
  if (!b || !prm || !ell_geom_ok(geom) || !eta || !d || !cot_out) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_elliptic<float>(b, prm, geom, eta, d, nullptr, cot_out, grad_eta, grad_d, st, true);
  return run_elliptic<double>(b, prm, geom, eta, d, nullptr, cot_out, grad_eta, grad_d, st, true);
}

int pom_eta_workspace(const pom_batch* b, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  *bytes = b->dtype == POM_F32 ? eta_ws_size<float>(b) : eta_ws_size<double>(b);
  return POM_OK;
}

int pom_eta_solve(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                  const void* rhs, const void* d, const void* x0, void* x, int32_t* iters_out, void* workspace,
                  size_t ws_bytes, void* stream) {
  if (!b || !ell_geom_ok(geom) || !rhs || !d || !x || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  if (!tide_geom_ok(prm, geom, true)) return POM_E_ARG;
  if (prm->eta_precond == POM_PRECOND_SPECTRAL && !cache) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_eta_solve<float>(b, prm, cache, geom, rhs, d, x0, x, iters_out, workspace, ws_bytes, st);
  return run_eta_solve<double>(b, prm, cache, geom, rhs, d, x0, x, iters_out, workspace, ws_bytes, st);
}

int pom_eta_solve_noalloc(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                          const void* rhs, const void* d, const void* x0, void* x, int32_t* iters_out, void* stream) {
  if (!b || !ell_geom_ok(geom) || !rhs || !d || !x) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  if (!tide_geom_ok(prm, geom, true)) return POM_E_ARG;
  if (prm->eta_precond == POM_PRECOND_SPECTRAL && !cache) return POM_E_ARG;
  const size_t need = b->dtype == POM_F32 ? eta_ws_size<float>(b) : eta_ws_size<double>(b);
  if (need > NOALLOC_SCRATCH_BYTES) return POM_E_SCRATCH;
  void* ws = nullptr;
  if (int e = scratch_address(&ws)) return e;
  pom_params_t q = *prm;
  q.eta_capture = 1;   // no pinned host word, no host synchronisation
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_eta_solve<float>(b, &q, cache, geom, rhs, d, x0, x, iters_out, ws, NOALLOC_SCRATCH_BYTES, st);
  return run_eta_solve<double>(b, &q, cache, geom, rhs, d, x0, x, iters_out, ws, NOALLOC_SCRATCH_BYTES, st);
}

int pom_eta_solve_adj(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                      const void* x, const void* d, const void* cot_x, void* grad_rhs, void* grad_d,
                      int32_t* iters_out, void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !ell_geom_ok(geom) || !x || !d || !cot_x || !grad_rhs || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  if (!tide_geom_ok(prm, geom, false)) return POM_E_ARG;
  if (prm->eta_precond == POM_PRECOND_SPECTRAL && !cache) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_eta_solve_adj<float>(b, prm, cache, geom, x, d, cot_x, grad_rhs, grad_d, iters_out, workspace,
                                    ws_bytes, st);
  return run_eta_solve_adj<double>(b, prm, cache, geom, x, d, cot_x, grad_rhs, grad_d, iters_out, workspace,
                                   ws_bytes, st);
}

int pom_implicit_cache_bytes(const pom_batch* b, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  *bytes = cache_bytes(b);
  return POM_OK;
}

int pom_implicit_prepare(const pom_batch* b, const pom_params_t* prm, const void* const* geom, void* cache,
                         size_t cache_bytes_, void* stream) {
  if (!b || !geom || !cache) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  if (cache_bytes_ < cache_bytes(b)) return POM_E_ARG;
  for (int f : {POM_G_DX, POM_G_DY, POM_G_H})
    if (!geom[f]) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return prepare_cache<float>(b, prm, geom, cache, st);
  return prepare_cache<double>(b, prm, geom, cache, st);
}

int pom_implicit_workspace(const pom_batch* b, const pom_params_t* prm, int32_t adjoint, size_t* bytes) {
  if (!b || !bytes) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  *bytes = b->dtype == POM_F32 ? ws_size<float>(b, prm, adjoint != 0) : ws_size<double>(b, prm, adjoint != 0);
  return POM_OK;
}

int pom_implicit_step(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                      const void* vgrid, const void* const* state_in, void* const* state_out, const void* const* geom,
                      int32_t internal, void* vamax_out, int32_t* eta_iters_out, void* workspace, size_t ws_bytes,
                      void* stream) {
  if (!b || !cache || !ramp || !vgrid || !state_in || !state_out || !geom || !vamax_out || !workspace)
    return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_step<float>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes,
                           st, cache, eta_iters_out);
  return run_step<double>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes,
                          st, cache, eta_iters_out);
}

int pom_implicit_step_adj(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                          const void* vgrid, const void* const* state_in, const void* const* geom, int32_t internal,
                          const void* const* cot_out, void* const* grad_in, void* workspace, size_t ws_bytes,
                          void* stream) {
  if (!b || !cache || !ramp || !vgrid || !state_in || !geom || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_adj<float>(b, prm, ramp, vgrid, state_in, geom, internal, cot_out, grad_in, workspace, ws_bytes, st,
                          cache);
  return run_adj<double>(b, prm, ramp, vgrid, state_in, geom, internal, cot_out, grad_in, workspace, ws_bytes, st,
                         cache);
}

int pom_implicit_step_taped(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                            const void* vgrid, const void* const* state_in, void* const* state_out,
                            const void* const* geom, int32_t internal, void* vamax_out, int32_t* eta_iters_out,
                            void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !cache || !ramp || !vgrid || !state_in || !state_out || !geom || !vamax_out || !workspace)
    return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_taped<float>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes,
                            st, cache, eta_iters_out);
  return run_taped<double>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, vamax_out, workspace, ws_bytes,
                           st, cache, eta_iters_out);
}

int pom_implicit_step_reverse(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                              const void* vgrid, const void* const* state_in, const void* const* state_out,
                              const void* const* geom, int32_t internal, const void* const* cot_out,
                              void* const* grad_in, void* workspace, size_t ws_bytes, void* stream) {
  if (!b || !cache || !ramp || !vgrid || !state_in || !state_out || !geom || !workspace) return POM_E_ARG;
  if (int e = check_params(prm, POM_CORE_IMPLICIT)) return e;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32)
    return run_reverse<float>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, cot_out, grad_in, workspace,
                              ws_bytes, st, cache);
  return run_reverse<double>(b, prm, ramp, vgrid, state_in, state_out, geom, internal, cot_out, grad_in, workspace,
                             ws_bytes, st, cache);
}

}  // extern "C"
