// Batched preconditioned conjugate gradients for the eta equation
//   A(D) eta = rhs,  A eta = eta - 4 dti^2 g alpha div(D grad eta) / art,
// eta = 0 on every edge (closed basin) or, with the tide policy, eta held at
// the tide at the fixed points, one independent system per member.
//
// reference: src/neuralpom/cores/implicit_p0_fixed/core.py:460-589 (eager PCG: update
// order, stopping rule rs <= tol^2 | rs <= rel_tol^2 rs0 | rs <= 1e-30, p edges zeroed),
// :591-639 (operator), :677-708 (Jacobi), :710-773 (spectral DST-I preconditioner).
//
// Each member keeps its own scalars and an active flag; a member that has
// converged is frozen (its kernels skip), so running extra iterations for the
// other members leaves its solution exactly as if it had stopped.  Dot products
// are reduced per block and then per member in a fixed order (deterministic),
// accumulated in double.  The spectral preconditioner is two DST-I transforms
// per axis, written as dense products with the sine matrices
//   S_n[k][m] = 2 sin(pi (k+1)(m+1) / (n+1))
// (the reference's FFT-based 2 x DST-I) in an LDS-tiled batched GEMM.
#pragma once
#include "pom_common.h"

namespace pom {

// per-member scalar slots
// PC_RHO + 0 / 1: r.z, by iteration parity (an iteration's p update writes
// the next rho while other blocks of the same launch still read this one)
enum { PC_RHO = 0, PC_RHO1, PC_THR, PC_TOL2, PC_RS, PC_RS0, PC_AUX, PC_N };

template <class T> struct EllGeom {
  const T *dx, *dy, *art, *dum, *dvm;
  double coeff;             // 4 dti^2 g alpha
  const T* fixed = nullptr; // nonzero at interior points held fixed (tide nodes, dry cells); null: none
};

// A(D) x at an interior point (same arithmetic as OpEll)
template <class T>
__device__ __forceinline__ T ell_point(const Pos& p, const EllGeom<T>& g, const T* __restrict__ d,
                                       const T* __restrict__ x) {
  const int i = p.i, j = p.j;
  const int64_t c = ix2(p, i, j), e = ix2(p, i + 1, j), w = ix2(p, i - 1, j), n = ix2(p, i, j + 1),
                s = ix2(p, i, j - 1);
  const T h = T(0.5);
  const T ex = (h * (g.dy[e] + g.dy[c])) / (h * (g.dx[e] + g.dx[c]) + T(1e-20));
  const T wx = (h * (g.dy[c] + g.dy[w])) / (h * (g.dx[c] + g.dx[w]) + T(1e-20));
  const T ny = (h * (g.dx[n] + g.dx[c])) / (h * (g.dy[n] + g.dy[c]) + T(1e-20));
  const T sy = (h * (g.dx[c] + g.dx[s])) / (h * (g.dy[c] + g.dy[s]) + T(1e-20));
  const T fe = h * (d[e] + d[c]) * (x[e] - x[c]) * ex * g.dum[e];
  const T fw = h * (d[c] + d[w]) * (x[c] - x[w]) * wx * g.dum[c];
  const T fn = h * (d[n] + d[c]) * (x[n] - x[c]) * ny * g.dvm[n];
  const T fs = h * (d[c] + d[s]) * (x[c] - x[s]) * sy * g.dvm[c];
  const T div = (fe - fw + fn - fs) * (T(1) / (g.art[c] + T(1e-20)));
  return x[c] - T(g.coeff) * div;
}

// Jacobi diagonal at an interior point (core.py:677-699)
template <class T>
__device__ __forceinline__ T ell_diag(const Pos& p, const EllGeom<T>& g, const T* __restrict__ d) {
  const int i = p.i, j = p.j;
  const int64_t c = ix2(p, i, j), e = ix2(p, i + 1, j), w = ix2(p, i - 1, j), n = ix2(p, i, j + 1),
                s = ix2(p, i, j - 1);
  const T h = T(0.5);
  const T ex = (h * (g.dy[e] + g.dy[c])) / (h * (g.dx[e] + g.dx[c]) + T(1e-20));
  const T wx = (h * (g.dy[c] + g.dy[w])) / (h * (g.dx[c] + g.dx[w]) + T(1e-20));
  const T ny = (h * (g.dx[n] + g.dx[c])) / (h * (g.dy[n] + g.dy[c]) + T(1e-20));
  const T sy = (h * (g.dx[c] + g.dx[s])) / (h * (g.dy[c] + g.dy[s]) + T(1e-20));
  const T he = h * (d[e] + d[c]), hw = h * (d[c] + d[w]), hn = h * (d[n] + d[c]), hs = h * (d[c] + d[s]);
  return T(1) + T(g.coeff) * (T(1) / (g.art[c] + T(1e-20))) *
                    (he * ex * g.dum[e] + hw * wx * g.dum[c] + hn * ny * g.dvm[n] + hs * sy * g.dvm[c]);
}

__device__ __forceinline__ bool interior_p(const Pos& p) {
  return p.i >= 1 && p.i <= p.im - 2 && p.j >= 1 && p.j <= p.jm - 2;
}

// an unknown of the solve: interior and not held fixed (case_adapters/
// coastal_boundary.py:44-49 holds the edges, dry cells and tide nodes)
template <class T> __device__ __forceinline__ bool free_p(const Pos& p, const T* fixed, int64_t c) {
  return interior_p(p) && !(fixed && fixed[c] != T(0));
}

// block sum of a double (TPB threads), result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// member of a 2-D block (the logical block of pos2); block partial sums are
// stored at the logical block index, so a member's partials are contiguous
__device__ __forceinline__ int block_member(const Batch& B) { return member_of(B.blk2, B.M, logical_block()); }

// ---------------------------------------------------------------- tide
// the tide at the nodes, amp cos(wt - phase), zero elsewhere
// (case_adapters/coastal_boundary.py:9-10; wt = (speed pi / 180) t / 3600 and
// phase in radians are formed by the caller)
template <class T>
__global__ void __launch_bounds__(TPB) k_tide_lift(Batch B, const T* __restrict__ node, const T* __restrict__ amp,
                                                   const T* __restrict__ phase, double wt, T* __restrict__ lift) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix2(p, p.i, p.j);
  lift[c] = node[c] != T(0) ? T(double(amp[c]) * cos(wt - double(phase[c]))) : T(0);
}

// ---------------------------------------------------------------- set-up
// x = x0 at the unknowns, the lift (zero when null) at the edges and fixed
// points; r = rhs - A x at the unknowns, 0 elsewhere (rhs edges are zeroed by
// the reference); z = 0
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_x0(Batch B, const T* __restrict__ x0, const T* __restrict__ fixed,
                                                const T* __restrict__ lift, T* __restrict__ x, T* __restrict__ z) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix2(p, p.i, p.j);
  x[c] = free_p(p, fixed, c) ? (x0 ? x0[c] : T(0)) : (lift ? lift[c] : T(0));
  z[c] = T(0);
}

template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_resid(Batch B, EllGeom<T> g, const T* __restrict__ d,
                                                   const T* __restrict__ x, const T* __restrict__ rhs,
                                                   T* __restrict__ r, T* __restrict__ p_out) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix2(p, p.i, p.j);
  r[c] = free_p(p, g.fixed, c) ? rhs[c] - ell_point(p, g, d, x) : T(0);
  p_out[c] = T(0);
}

// Jacobi: z = r / (diag + 1e-30) on the interior (active members only), with
// partial sums of r.z and r.r
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_jacobi(Batch B, EllGeom<T> g, const T* __restrict__ d,
                                                    const T* __restrict__ r, T* __restrict__ z,
                                                    const int* __restrict__ act, double* __restrict__ part) {
  __shared__ double sh[TPB];
  const int m = block_member(B);
  if (!act[m]) {
    if (threadIdx.x == 0) part[2 * logical_block()] = part[2 * logical_block() + 1] = 0.0;
    return;
  }
  Pos p;
  double rz = 0.0;
  if (pos2(B, p) && free_p(p, g.fixed, ix2(p, p.i, p.j))) {
    const int64_t c = ix2(p, p.i, p.j);
    const T zv = r[c] / (ell_diag(p, g, d) + T(1e-30));
    z[c] = zv;
    rz = double(r[c]) * double(zv);
  }
  const double s = block_sum(rz, sh);
  if (threadIdx.x == 0) part[2 * logical_block()] = s;
}

// partial sums of a.b over the interior (active members only) into part[2 b + slot]
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_dot(Batch B, const T* __restrict__ a, const T* __restrict__ b,
                                                 const int* __restrict__ act, double* __restrict__ part, int slot) {
  __shared__ double sh[TPB];
  const int m = block_member(B);
  if (!act[m]) {
    if (threadIdx.x == 0) part[2 * logical_block() + slot] = 0.0;
    return;
  }
  Pos p;
  double v = 0.0;
  if (pos2(B, p) && interior_p(p)) {
    const int64_t c = ix2(p, p.i, p.j);
    v = double(a[c]) * double(b[c]);
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[2 * logical_block() + slot] = s;
}

// per-member sum of part[2 b + slot] over the member's blocks (one block per member)
__device__ __forceinline__ double member_sum(const Batch& B, const double* part, int m, int slot, double* sh) {
  const int b0 = B.blk2[m], b1 = B.blk2[m + 1];
  double v = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += TPB) v += part[2 * b + slot];
  return block_sum(v, sh);
}

// every member starts active
__global__ void k_pcg_act1(int* act, int M) {
  for (int m = threadIdx.x; m < M; m += blockDim.x) act[m] = 1;
}

// after set-up: rho = r.z, rs0 = r.r, thresholds; all members active
__global__ void __launch_bounds__(TPB) k_pcg_start(Batch B, const double* __restrict__ part, double* sc,
                                                   int* act, int* its, const double* __restrict__ tol,
                                                   double rel_tol) {
  __shared__ double sh[TPB];
  const int m = blockIdx.x;
  const double rz = member_sum(B, part, m, 0, sh);
  __syncthreads();
  const double rr = member_sum(B, part, m, 1, sh);
  if (threadIdx.x == 0) {
    double* s = sc + PC_N * m;
    s[PC_RHO] = rz;
    s[PC_RS0] = rr;
    s[PC_THR] = rel_tol * rel_tol * fmax(rr, 1e-30);
    s[PC_TOL2] = tol[m] * tol[m];
    act[m] = 1;
    its[m] = 0;
  }
}

// ---------------------------------------------------------------- iteration
// ap = A p (interior), partial sums of p.ap
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_ap(Batch B, EllGeom<T> g, const T* __restrict__ d,
                                                const T* __restrict__ p_, T* __restrict__ ap,
                                                const int* __restrict__ act, double* __restrict__ part, int* nact) {
  __shared__ double sh[TPB];
  if (logical_block() == 0 && threadIdx.x == 0) *nact = 0;   // counted again by this iteration's test
  const int m = block_member(B);
  if (!act[m]) {
    if (threadIdx.x == 0) part[2 * logical_block()] = 0.0;
    return;
  }
  Pos p;
  double v = 0.0;
  if (pos2(B, p)) {
    const int64_t c = ix2(p, p.i, p.j);
    if (free_p(p, g.fixed, c)) {
      const T a = ell_point(p, g, d, p_);
      ap[c] = a;
      v = double(p_[c]) * double(a);
    } else {
      ap[c] = p_[c];
    }
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[2 * logical_block()] = s;
}

// Each block of a member forms the member's scalars (alpha here, the
// convergence test and beta below) itself, from the member's block partial
// sums in one fixed order: every block gets the same value, and the
// iteration needs no launch per reduction.  The member's first block keeps
// the values the next launches read (its[], act[], rho).
__device__ __forceinline__ bool member_first(const Batch& B, int m) { return logical_block() == B.blk2[m]; }

// alpha = rho / (p.ap + 1e-30); x += alpha p, r -= alpha ap; partial sums of r.r
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_xr(Batch B, T* __restrict__ x, T* __restrict__ r,
                                                const T* __restrict__ p_, const T* __restrict__ ap,
                                                const double* __restrict__ sc, const int* __restrict__ act,
                                                double* __restrict__ part, int par) {
  __shared__ double sh[TPB];
  const int m = block_member(B);
  if (!act[m]) {
    if (threadIdx.x == 0) part[2 * logical_block() + 1] = 0.0;
    return;
  }
  const double den = member_sum(B, part, m, 0, sh);
  __syncthreads();   // sh is reused below
  const T alpha = T(sc[PC_N * m + PC_RHO + par] / (den + 1e-30));
  Pos p;
  double v = 0.0;
  if (pos2(B, p)) {
    const int64_t c = ix2(p, p.i, p.j);
    if (interior_p(p)) {
      x[c] = x[c] + alpha * p_[c];
      const T rv = r[c] - alpha * ap[c];
      r[c] = rv;
      v = double(rv) * double(rv);
    }
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[2 * logical_block() + 1] = s;
}

// convergence: rs = r.r; a member stops when rs <= tol^2, rs <= rel_tol^2 rs0
// or rs <= 1e-30 (its x and r keep this iteration's update).  Called by every
// block of an active member (act[m] read as set when the launch began, or
// cleared by the member's first block in this launch, which happens only
// when every block finds the member converged); true if it goes on.
__device__ __forceinline__ bool pcg_test(const Batch& B, const double* part, double* sc, int* act, int* its,
                                         int* nact, int m, double* sh) {
  const double rs = member_sum(B, part, m, 1, sh);
  __syncthreads();   // sh is reused by the caller
  const double* s = sc + PC_N * m;
  const bool go = !(rs <= s[PC_TOL2] || rs <= s[PC_THR] || rs <= 1e-30);
  if (member_first(B, m) && threadIdx.x == 0) {
    sc[PC_N * m + PC_RS] = rs;
    its[m] += 1;
    if (go) atomicAdd(nact, 1);
    else act[m] = 0;
  }
  return go;
}

// the convergence test, then partial sums of r.z of the (spectral) preconditioned z
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_dot_test(Batch B, const T* __restrict__ r, const T* __restrict__ z,
                                                      int* act, double* __restrict__ part, double* sc, int* its,
                                                      int* nact) {
  __shared__ double sh[TPB];
  const int m = block_member(B);
  if (!act[m] || !pcg_test(B, part, sc, act, its, nact, m, sh)) {
    if (threadIdx.x == 0) part[2 * logical_block()] = 0.0;
    return;
  }
  Pos p;
  double v = 0.0;
  if (pos2(B, p) && interior_p(p)) {
    const int64_t c = ix2(p, p.i, p.j);
    v = double(r[c]) * double(z[c]);
  }
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[2 * logical_block()] = s;
}

// the convergence test, then the Jacobi preconditioner z = r / diag and partial sums of r.z
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_jacobi_test(Batch B, EllGeom<T> g, const T* __restrict__ d,
                                                         const T* __restrict__ r, T* __restrict__ z, int* act,
                                                         double* __restrict__ part, double* sc, int* its,
                                                         int* nact) {
  __shared__ double sh[TPB];
  const int m = block_member(B);
  if (!act[m] || !pcg_test(B, part, sc, act, its, nact, m, sh)) {
    if (threadIdx.x == 0) part[2 * logical_block()] = 0.0;
    return;
  }
  Pos p;
  double rz = 0.0;
  if (pos2(B, p) && free_p(p, g.fixed, ix2(p, p.i, p.j))) {
    const int64_t c = ix2(p, p.i, p.j);
    const T zv = r[c] / (ell_diag(p, g, d) + T(1e-30));
    z[c] = zv;
    rz = double(r[c]) * double(zv);
  }
  const double s = block_sum(rz, sh);
  if (threadIdx.x == 0) part[2 * logical_block()] = s;
}

// p = z at the start (first), else beta = (r.z) / (rho + 1e-30), p = p beta + z
// and rho = r.z for the next iteration (slot par ^ 1); 0 at the edges and fixed points
template <class T>
__global__ void __launch_bounds__(TPB) k_pcg_p(Batch B, T* __restrict__ p_, const T* __restrict__ z,
                                               const T* __restrict__ fixed, double* __restrict__ sc,
                                               const int* __restrict__ act, const double* __restrict__ part,
                                               int first, int par) {
  __shared__ double sh[TPB];
  const int m = block_member(B);
  if (!act[m]) return;
  T beta = T(0);
  if (!first) {
    const double rz = member_sum(B, part, m, 0, sh);
    beta = T(rz / (sc[PC_N * m + PC_RHO + par] + 1e-30));
    if (member_first(B, m) && threadIdx.x == 0) sc[PC_N * m + PC_RHO + (par ^ 1)] = rz;
  }
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix2(p, p.i, p.j);
  if (!free_p(p, fixed, c)) {
    p_[c] = T(0);
    return;
  }
  p_[c] = first ? z[c] : p_[c] * beta + z[c];
}

// ---------------------------------------------------------------- adjoint right-hand side
// The eta-solve adjoint (core.py:231-252) solves A^T lam = g as
// lam = area A^-1 (g / area) with area = art / art(1,1) on the interior (1 on
// the edges), the right-hand side divided by its max-norm (clamped below at 1)
// and the absolute tolerance scaled alike.
template <class T>
__device__ __forceinline__ T adj_area(const Pos& p, const T* __restrict__ art) {
  return interior_p(p) ? art[ix2(p, p.i, p.j)] / art[ix2(p, 1, 1)] : T(1);
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  return sh[0];
}

// a = g / area; partial max |a|
template <class T>
__global__ void __launch_bounds__(TPB) k_eta_adj_in(Batch B, const T* __restrict__ g, const T* __restrict__ art,
                                                    T* __restrict__ a, double* __restrict__ part) {
  __shared__ double sh[TPB];
  Pos p;
  double v = 0.0;
  if (pos2(B, p)) {
    const int64_t c = ix2(p, p.i, p.j);
    const T x = g[c] / adj_area(p, art);
    a[c] = x;
    v = fabs(double(x));
  }
  const double m = block_max(v, sh);
  if (threadIdx.x == 0) part[2 * logical_block()] = m;
}

// scale = max(max|a|, 1); tol = eta_tol / scale
__global__ void __launch_bounds__(TPB) k_eta_adj_scale(Batch B, const double* __restrict__ part, double* sc,
                                                       double* tol, double eta_tol) {
  __shared__ double sh[TPB];
  const int m = blockIdx.x;
  const int b0 = B.blk2[m], b1 = B.blk2[m + 1];
  double v = 0.0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += TPB) v = fmax(v, part[2 * b]);
  const double mx = block_max(v, sh);
  if (threadIdx.x == 0) {
    const double scale = mx < 1.0 ? 1.0 : mx;
    sc[PC_N * m + PC_AUX] = scale;
    tol[m] = eta_tol / scale;
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_eta_adj_norm(Batch B, T* __restrict__ a, const double* __restrict__ sc) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix2(p, p.i, p.j);
  a[c] = a[c] / T(sc[PC_N * p.m + PC_AUX]);
}

// lam = x area scale (the rhs cotangent), nlam = -lam
template <class T>
__global__ void __launch_bounds__(TPB) k_eta_adj_out(Batch B, const T* __restrict__ x, const T* __restrict__ art,
                                                     const double* __restrict__ sc, T* __restrict__ lam,
                                                     T* __restrict__ nlam) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix2(p, p.i, p.j);
  const T l = x[c] * adj_area(p, art) * T(sc[PC_N * p.m + PC_AUX]);
  lam[c] = l;
  nlam[c] = -l;
}

// ---------------------------------------------------------------- spectral preconditioner
// per-member table of the sine matrices and inverse eigenvalues in the cache
struct SpecMember {
  int64_t si, sj, l;   // element offsets of S_ni, S_nj, inverse eigenvalues (ni x nj)
  int32_t ni, nj;
  double scale;        // 4 (ni + 1)(nj + 1)
};

// C = A B (+ epilogue) for every active member: A is M x K (lda), B is K x N (ldb),
// C is M x N (ldc).  mode 0: plain; 1: C *= L (elementwise, ld = N); 2: C /= scale
template <class T> struct GemmArg {
  const T* a;        // base pointers; per-member offsets from the descriptors below
  const T* b;
  T* c;
  int a_kind, b_kind, c_kind;   // 0: S_ni, 1: S_nj, 2: member field interior (ld jm), 3: dense scratch (ld nj)
  int mode;
};

template <class T>
__device__ __forceinline__ const T* gemm_ptr(int kind, const T* base, const SpecMember& s, const Member& mm,
                                             int64_t& ld) {
  switch (kind) {
    case 0: ld = s.ni; return base + s.si;
    case 1: ld = s.nj; return base + s.sj;
    case 2: ld = mm.jm; return base + mm.off2 + mm.jm + 1;
    default: ld = s.nj; return base + mm.off2;
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_spec_gemm(const Member* __restrict__ mem, const SpecMember* __restrict__ sp,
                                                   const T* __restrict__ cache, const int* __restrict__ act,
                                                   GemmArg<T> g, int left) {
  // left: C = S_ni X (M = ni, N = nj, K = ni); else C = X S_nj (K = nj)
  constexpr int TS = 16;
  __shared__ T As[TS][TS + 1], Bs[TS][TS + 1];
  const int m = blockIdx.z;
  if (!act[m]) return;
  const SpecMember s = sp[m];
  const Member mm = mem[m];
  const int M = s.ni, N = s.nj, K = left ? s.ni : s.nj;
  const int row0 = blockIdx.y * TS, col0 = blockIdx.x * TS;
  if (row0 >= M || col0 >= N) return;
  int64_t lda, ldb, ldc;
  const T* A = gemm_ptr<T>(g.a_kind, g.a_kind < 2 ? cache : g.a, s, mm, lda);
  const T* Bm = gemm_ptr<T>(g.b_kind, g.b_kind < 2 ? cache : g.b, s, mm, ldb);
  T* Cm = const_cast<T*>(gemm_ptr<T>(g.c_kind, g.c, s, mm, ldc));
  const int tr = threadIdx.x / TS, tc = threadIdx.x % TS;
  const int row = row0 + tr, col = col0 + tc;
  T acc = T(0);
  for (int k0 = 0; k0 < K; k0 += TS) {
    As[tr][tc] = (row < M && k0 + tc < K) ? A[(int64_t)row * lda + k0 + tc] : T(0);
    Bs[tr][tc] = (k0 + tr < K && col < N) ? Bm[(int64_t)(k0 + tr) * ldb + col] : T(0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TS; ++k) acc += As[tr][k] * Bs[k][tc];
    __syncthreads();
  }
  if (row < M && col < N) {
    if (g.mode == 1) acc *= cache[s.l + (int64_t)row * N + col];
    else if (g.mode == 2) acc /= T(s.scale);
    Cm[(int64_t)row * ldc + col] = acc;
  }
}

// fp32 products of the DST-I preconditioner on the matrix cores: a 32 x 32
// tile of C per workgroup of 256 threads.  A and B are staged through LDS 64
// columns of K at a time; wave w multiplies K sub-steps 16 w .. 16 w + 15 of
// each stage as 2 x 2 tiles of v_mfma_f32_16x16x4f32 (fp32 products, fp32
// accumulation), and the four partial tiles are summed through LDS at the end
// (split K: four times the workgroups of one 64 x 64 tile per workgroup, which
// left under two waves per SIMD at 352 x 320 and fewer for small members).
// Operand layout of the instruction (wave64): lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; result register r of lane l is C[4 (l / 16) + r][l % 16].
// Same arguments, operands and epilogue as k_spec_gemm; the grid is 32 x 32 tiles.
typedef float pom_floatx4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(TPB) k_spec_gemm_mfma(const Member* __restrict__ mem,
                                                        const SpecMember* __restrict__ sp,
                                                        const float* __restrict__ cache, const int* __restrict__ act,
                                                        GemmArg<float> g, int left) {
  constexpr int BT = 32, KT = 64;
  static_assert(TPB == 256 && KT == 16 * (TPB / 64), "four waves, 16 K sub-steps each per stage");
  __shared__ float As[BT][KT + 1];
  __shared__ float Bs[KT][BT + 1];
  __shared__ float Cs[TPB / 64 - 1][BT * BT];
  const int m = blockIdx.z;
  if (!act[m]) return;
  const SpecMember s = sp[m];
  const Member mm = mem[m];
  const int M = s.ni, N = s.nj, K = left ? s.ni : s.nj;
  const int row0 = blockIdx.y * BT, col0 = blockIdx.x * BT;
  if (row0 >= M || col0 >= N) return;
  int64_t lda, ldb, ldc;
  const float* A = gemm_ptr<float>(g.a_kind, g.a_kind < 2 ? cache : g.a, s, mm, lda);
  const float* Bm = gemm_ptr<float>(g.b_kind, g.b_kind < 2 ? cache : g.b, s, mm, ldb);
  float* Cm = const_cast<float*>(gemm_ptr<float>(g.c_kind, g.c, s, mm, ldc));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  pom_floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = pom_floatx4{0.f, 0.f, 0.f, 0.f};
  // the next stage's operands are loaded into registers while this stage's
  // products run, and stored to LDS after the stage's closing barrier
  constexpr int PER = BT * KT / TPB;
  float ra[PER], rb[PER];
  auto load_stage = [&](int k0) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x + q * TPB;
      const int r = e / KT, kk = e % KT;
      ra[q] = (row0 + r < M && k0 + kk < K) ? A[(int64_t)(row0 + r) * lda + k0 + kk] : 0.f;
      const int kb = e / BT, cc = e % BT;
      rb[q] = (k0 + kb < K && col0 + cc < N) ? Bm[(int64_t)(k0 + kb) * ldb + col0 + cc] : 0.f;
    }
  };
  load_stage(0);
  for (int k0 = 0; k0 < K; k0 += KT) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int e = threadIdx.x + q * TPB;
      As[e / KT][e % KT] = ra[q];
      Bs[e / BT][e % BT] = rb[q];
    }
    __syncthreads();
    if (k0 + KT < K) load_stage(k0 + KT);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int kk = 16 * wave + 4 * t + lk;
      const float a0 = As[li][kk], a1 = As[16 + li][kk];
      const float b0 = Bs[kk][li], b1 = Bs[kk][16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  if (wave > 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) Cs[wave - 1][(16 * i + 4 * lk + r) * BT + 16 * j + li] = acc[i][j][r];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int tr = 16 * i + 4 * lk + r, tc = 16 * j + li;
          const int row = row0 + tr, col = col0 + tc;
          if (row < M && col < N) {
            float v = acc[i][j][r];
#pragma unroll
            for (int w = 0; w < TPB / 64 - 1; ++w) v += Cs[w][tr * BT + tc];
            if (g.mode == 1) v *= cache[s.l + (int64_t)row * N + col];
            else if (g.mode == 2) v /= float(s.scale);
            Cm[(int64_t)row * ldc + col] = v;
          }
        }
  }
}

}  // namespace pom
