// Column operators of the explicit POM step: one thread per water column,
// the 16 levels held in registers.  Forward and hand-written reverse sweeps.
//
//   baropg   cumulative sigma-coordinate pressure gradient (momentum.py:543-700)
//   colsum   depth integrals sum_k f*dz                      (time_loop.py:239-258)
//   adjust   barotropic transport correction of u, v          (time_loop.py:440-477)
//   vertvl   continuity -> w, with bcond 5                    (baroclinic.py:930-1040, :849-851)
//   proft    implicit vertical diffusion of T or S (nbc 1/3)  (tracers.py:890-1010)
//   profuv   implicit vertical viscosity + bottom drag        (momentum.py:250-465)
//   profq    Mellor-Yamada 2.5 closure                        (tracers.py:158-431)
//   uvfilt   Asselin filter of u, v minus its depth mean      (time_loop.py:632-678)
//   forcing  surface forcing reset                            (time_loop.py:118-164)
#pragma once
#include "pom_common.h"

namespace pom {

template <class T> __device__ __forceinline__ void ld_col(const T* base, T* x) {
#pragma unroll
  for (int k = 0; k < KB; ++k) x[k] = base[k];
}
template <class T> __device__ __forceinline__ void st_col(T* base, const T* x) {
#pragma unroll
  for (int k = 0; k < KB; ++k) base[k] = x[k];
}
// the whole column is read before the first store: otherwise every store may
// alias the next load and the KB read-modify-writes run one memory latency each
template <class T> __device__ __forceinline__ void add_col(T* base, const T* x) {
  T y[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) y[k] = base[k];
#pragma unroll
  for (int k = 0; k < KB; ++k) base[k] = y[k] + x[k];
}
// add_col, or a plain store when the kernel is the array's first writer
template <class T> __device__ __forceinline__ void put_col(T* base, const T* x, bool assign) {
  if (assign) {
#pragma unroll
    for (int k = 0; k < KB; ++k) base[k] = x[k];
  } else {
    add_col(base, x);
  }
}

// ---------------------------------------------------------------------
// Coalesced column I/O.  In a 2-D launch the TPB threads of a block own TPB
// consecutive columns (i*jm + j) of one member, and a 3-D field stores those
// columns back to back, k fastest.  Moving a thread's column through an LDS
// tile replaces 16 accesses at a 64 B lane stride (each reaching HBM as its
// own partial-line request) by 16 unit-stride wavefront accesses.
// Every thread of the block calls these (they contain block barriers):
// col_pos keeps a thread past the member's last column on that last column
// with valid = false, instead of returning.
// ---------------------------------------------------------------------
constexpr int TILE = TPB * (KB + 1);   // LDS elements of one column tile (padded rows)

template <class T> struct ColIO {
  T* lds;          // TILE elements of __shared__ memory
  int64_t base;    // element offset (3-D) of the block's first column
  int64_t first;   // member-local index of the block's first column
  int64_t ncol;    // columns of the member (im * jm)
  int n;           // columns of this block inside the member
  bool valid;      // this thread owns a column of its own
};

template <class T>
__device__ __forceinline__ ColIO<T> col_pos(const Batch& B, Pos& p, T* lds) {
  const int b = logical_block();
  const int m = member_of(B.blk2, B.M, b);
  const Member mm = B.mem[m];
  const int64_t n2 = (int64_t)mm.im * mm.jm;
  const int64_t first = (int64_t)(b - B.blk2[m]) * TPB;   // < n2 by construction of blk2
  int64_t local = first + threadIdx.x;
  ColIO<T> io;
  io.lds = lds;
  io.valid = local < n2;
  if (!io.valid) local = n2 - 1;
  io.n = (int)(n2 - first < TPB ? n2 - first : TPB);
  io.base = mm.off3 + first * KB;
  io.first = first;
  io.ncol = n2;
  p.m = m; p.im = mm.im; p.jm = mm.jm; p.o2 = mm.off2; p.o3 = mm.off3;
  p.k = 0;
  p.i = (int)(local / mm.jm);
  p.j = (int)(local - (int64_t)p.i * mm.jm);
  return io;
}

// padded LDS slot of tile element e (column e / KB, level e % KB)
__device__ __forceinline__ int tile_ix(int e) { return e + e / KB; }

// Loads are issued for all KB rows before the first LDS store (indices past
// the block's last column re-read its last element), so a thread's loads
// overlap instead of each waiting for the previous store.
template <class T> __device__ __forceinline__ void tile_fill(const ColIO<T>& io, const T* __restrict__ g, int tot) {
  if (tot <= 0) return;   // uniform over the block
  T v[KB];
  #pragma unroll
  for (int r = 0; r < KB; ++r) {
    const int e = r * TPB + (int)threadIdx.x;
    v[r] = g[e < tot ? e : tot - 1];
  }
  #pragma unroll
  for (int r = 0; r < KB; ++r) io.lds[tile_ix(r * TPB + (int)threadIdx.x)] = v[r];
}

template <class T> __device__ __forceinline__ void col_ld(const ColIO<T>& io, const T* __restrict__ f, T* x) {
  tile_fill(io, f + io.base, io.n * KB);
  __syncthreads();
  #pragma unroll
  for (int k = 0; k < KB; ++k) x[k] = io.lds[threadIdx.x * (KB + 1) + k];
  __syncthreads();
}

// column of the thread's neighbour `shift` columns further on (shift = 1:
// (i, j+1); shift = jm: (i+1, j)); neighbours past the member's end are not
// read (their slots hold stale data, which callers must not use)
template <class T>
__device__ __forceinline__ void col_ld_shift(const ColIO<T>& io, const T* __restrict__ f, int64_t shift, T* x) {
  const int64_t avail = io.ncol - io.first - shift;
  const int tot = (int)(avail < io.n ? (avail > 0 ? avail : 0) : io.n) * KB;
  tile_fill(io, f + io.base + shift * KB, tot);
  __syncthreads();
  #pragma unroll
  for (int k = 0; k < KB; ++k) x[k] = io.lds[threadIdx.x * (KB + 1) + k];
  __syncthreads();
}

template <class T> __device__ __forceinline__ void col_put(const ColIO<T>& io, const T* x) {
  #pragma unroll
  for (int k = 0; k < KB; ++k) io.lds[threadIdx.x * (KB + 1) + k] = x[k];
  __syncthreads();
}

template <class T> __device__ __forceinline__ void col_st(const ColIO<T>& io, T* __restrict__ f, const T* x) {
  col_put(io, x);
  T* g = f + io.base;
  const int tot = io.n * KB;
  #pragma unroll
  for (int r = 0; r < KB; ++r) {
    const int e = r * TPB + (int)threadIdx.x;
    if (e < tot) g[e] = io.lds[tile_ix(e)];
  }
  __syncthreads();
}

// f += x (read-modify-write, both sides unit-stride)
template <class T> __device__ __forceinline__ void col_add(const ColIO<T>& io, T* __restrict__ f, const T* x) {
  col_put(io, x);
  T* g = f + io.base;
  const int tot = io.n * KB;
  #pragma unroll
  for (int r = 0; r < KB; ++r) {
    const int e = r * TPB + (int)threadIdx.x;
    if (e < tot) g[e] += io.lds[tile_ix(e)];
  }
  __syncthreads();
}

// =====================================================================
// baropg
// =====================================================================
template <class T> struct BaropgArgs {
  const T *rho, *rmean, *dt, *dx, *dy, *dum, *dvm;
  T *rho_out, *drhox, *drhoy;
};

// One face (x: DIR 0, between (i-1,j) and (i,j); y: DIR 1) of the owner
// column (i,j) (interior).  Writes the scaled column (KBM1 levels).
template <class T, int DIR>
__device__ __forceinline__ void baropg_face(const Pos& p, const Vert<T>& vg, const Params& prm, T ramp, const T* rho,
                                            const T* rmean, const T* dt, const T* dd2, const T* mask, T* outc) {
  const int i = p.i, j = p.j;
  const int ni = DIR == 0 ? i - 1 : i, nj = DIR == 0 ? j : j - 1;
  const T* r0 = rho + ix3(p, i, j, 0);
  const T* m0 = rmean + ix3(p, i, j, 0);
  const T* r1 = rho + ix3(p, ni, nj, 0);
  const T* m1 = rmean + ix3(p, ni, nj, 0);
  const T dt0 = dt[ix2(p, i, j)], dt1 = dt[ix2(p, ni, nj)];
  const T tdt = dt0 + dt1;
  const T grav = T(prm.grav);
  const T c0 = T(0.5 * prm.grav);
  T ra0p = r0[0] - m0[0], ra1p = r1[0] - m1[0];
  const T surf = c0 * (-vg.zz(p.m, 0)) * tdt * (ra0p - ra1p);
  const T scale = T(0.25) * tdt * mask[ix2(p, i, j)] * (dd2[ix2(p, i, j)] + dd2[ix2(p, ni, nj)]);
  outc[0] = surf * scale * ramp;
  T acc = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T ra0 = r0[k] - m0[k], ra1 = r1[k] - m1[k];
    const T zk = vg.zz(p.m, k), zkm = vg.zz(p.m, k - 1);
    const T diff = (ra0 - ra1) + (ra0p - ra1p);
    const T t1 = grav * T(0.25) * (zkm - zk) * tdt * diff;
    const T sm = (ra0 + ra1) - (ra0p + ra1p);
    const T t2 = grav * T(0.25) * (zkm + zk) * (dt0 - dt1) * sm;
    acc = acc + (t1 + t2);
    outc[k] = (surf + acc) * scale * ramp;
    ra0p = ra0;
    ra1p = ra1;
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) k_baropg(Batch B, Vert<T> vg, const T* ramp, Params prm, BaropgArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0);
  T ox[KB], oy[KB];
  // both columns loaded before the first store (rho_out may alias rho)
  #pragma unroll
  for (int k = 0; k < KB; ++k) {
    ox[k] = a.rho[c + k];
    oy[k] = a.rmean[c + k];
  }
  #pragma unroll
  for (int k = 0; k < KB; ++k) ox[k] = (ox[k] - oy[k]) + oy[k];
  st_col(a.rho_out + c, ox);
  #pragma unroll
  for (int k = 0; k < KB; ++k) ox[k] = oy[k] = T(0);
  if (interior_pos(p)) {
    const T rm = ramp[p.m];
    baropg_face<T, 0>(p, vg, prm, rm, a.rho, a.rmean, a.dt, a.dy, a.dum, ox);
    baropg_face<T, 1>(p, vg, prm, rm, a.rho, a.rmean, a.dt, a.dx, a.dvm, oy);
  }
  st_col(a.drhox + c, ox);
  st_col(a.drhoy + c, oy);
}

template <class T> struct BaropgAdjArgs {
  const T *rho, *rmean, *dt, *dx, *dy, *dum, *dvm;
  const T *b_rho_out, *b_drhox, *b_drhoy;   // output cotangents (may be null)
  T *b_rho, *b_dt;                          // accumulated input cotangents
};

// reverse of baropg_face for owner column p: partials w.r.t. the owner's
// column (g0, gdt0) and the neighbour's column (g1, gdt1)
template <class T, int DIR>
__device__ __forceinline__ void baropg_face_rev(const Pos& p, const Vert<T>& vg, const Params& prm, T ramp,
                                                const T* rho, const T* rmean, const T* dt, const T* dd2,
                                                const T* mask, const T* bout, T* g0, T* g1, T& gdt0, T& gdt1) {
  const int i = p.i, j = p.j;
  const int ni = DIR == 0 ? i - 1 : i, nj = DIR == 0 ? j : j - 1;
  const T* r0 = rho + ix3(p, i, j, 0);
  const T* m0 = rmean + ix3(p, i, j, 0);
  const T* r1 = rho + ix3(p, ni, nj, 0);
  const T* m1 = rmean + ix3(p, ni, nj, 0);
  const T* ob = bout + ix3(p, i, j, 0);
  const T dt0 = dt[ix2(p, i, j)], dt1 = dt[ix2(p, ni, nj)];
  const T tdt = dt0 + dt1;
  const T grav = T(prm.grav);
  const T c0 = T(0.5 * prm.grav);
  const T mk = mask[ix2(p, i, j)], dd = dd2[ix2(p, i, j)] + dd2[ix2(p, ni, nj)];
  const T scale = T(0.25) * tdt * mk * dd;
  T ra0[KB], ra1[KB];
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) {
    ra0[k] = r0[k] - m0[k];
    ra1[k] = r1[k] - m1[k];
  }
  // recompute the unscaled column
  const T surf = c0 * (-vg.zz(p.m, 0)) * tdt * (ra0[0] - ra1[0]);
  T col[KB];
  col[0] = surf;
  T acc = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T zk = vg.zz(p.m, k), zkm = vg.zz(p.m, k - 1);
    const T t1 = grav * T(0.25) * (zkm - zk) * tdt * ((ra0[k] - ra1[k]) + (ra0[k - 1] - ra1[k - 1]));
    const T t2 = grav * T(0.25) * (zkm + zk) * (dt0 - dt1) * ((ra0[k] + ra1[k]) - (ra0[k - 1] + ra1[k - 1]));
    acc = acc + (t1 + t2);
    col[k] = surf + acc;
  }
  // out(k) = col(k) * scale * ramp
  T b_scale = T(0), b_tdt = T(0), b_ddt = T(0);
  T bcol[KB];
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) {
    bcol[k] = ob[k] * scale * ramp;
    b_scale += ob[k] * col[k] * ramp;
  }
  b_tdt += b_scale * T(0.25) * mk * dd;
  // cumulative sum reverse: inc(k) feeds col(k..kbm1-1); surf feeds all
  T b_surf = T(0);
  T suffix = T(0);
  #pragma unroll
  for (int k = 0; k < KB; ++k) g0[k] = g1[k] = T(0);
  #pragma unroll
  for (int k = KBM1 - 1; k >= 1; --k) {
    suffix += bcol[k];
    const T zk = vg.zz(p.m, k), zkm = vg.zz(p.m, k - 1);
    const T A = grav * T(0.25) * (zkm - zk);
    const T Bc = grav * T(0.25) * (zkm + zk);
    const T binc = suffix;
    const T ddt = dt0 - dt1;
    // t1 = A*tdt*diff, diff = (ra0k - ra1k) + (ra0km - ra1km)
    const T bdiff = binc * A * tdt;
    b_tdt += binc * A * ((ra0[k] - ra1[k]) + (ra0[k - 1] - ra1[k - 1]));
    // t2 = B*ddt*sm, sm = (ra0k + ra1k) - (ra0km + ra1km)
    const T bsm = binc * Bc * ddt;
    b_ddt += binc * Bc * ((ra0[k] + ra1[k]) - (ra0[k - 1] + ra1[k - 1]));
    g0[k] += bdiff + bsm;
    g1[k] += -bdiff + bsm;
    g0[k - 1] += bdiff - bsm;
    g1[k - 1] += -bdiff - bsm;
  }
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) b_surf += bcol[k];
  const T cz = c0 * (-vg.zz(p.m, 0));
  g0[0] += b_surf * cz * tdt;
  g1[0] -= b_surf * cz * tdt;
  b_tdt += b_surf * cz * (ra0[0] - ra1[0]);
  gdt0 = b_tdt + b_ddt;
  gdt1 = b_tdt - b_ddt;
}

template <class T>
__global__ void __launch_bounds__(TPB) k_baropg_adj(Batch B, Vert<T> vg, const T* ramp, Params prm,
                                                     BaropgAdjArgs<T> a) {
  Pos q;
  if (!pos2(B, q)) return;
  const int64_t c = ix3(q, q.i, q.j, 0);
  T br[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) br[k] = a.b_rho_out ? a.b_rho_out[c + k] : T(0);
  T bdt = T(0);
  T g0[KB], g1[KB], d0, d1;
  const T rm = ramp[q.m];
  // owner = q itself (x and y faces): take own parts
  if (interior_pos(q)) {
    if (a.b_drhox) {
      baropg_face_rev<T, 0>(q, vg, prm, rm, a.rho, a.rmean, a.dt, a.dy, a.dum, a.b_drhox, g0, g1, d0, d1);
      #pragma unroll
      for (int k = 0; k < KB; ++k) br[k] += g0[k];
      bdt += d0;
    }
    if (a.b_drhoy) {
      baropg_face_rev<T, 1>(q, vg, prm, rm, a.rho, a.rmean, a.dt, a.dx, a.dvm, a.b_drhoy, g0, g1, d0, d1);
      #pragma unroll
      for (int k = 0; k < KB; ++k) br[k] += g0[k];
      bdt += d0;
    }
  }
  // owner = (i+1, j): q is its x-neighbour
  {
    Pos o = q;
    o.i = q.i + 1;
    if (a.b_drhox && interior_pos(o)) {
      baropg_face_rev<T, 0>(o, vg, prm, rm, a.rho, a.rmean, a.dt, a.dy, a.dum, a.b_drhox, g0, g1, d0, d1);
      #pragma unroll
      for (int k = 0; k < KB; ++k) br[k] += g1[k];
      bdt += d1;
    }
  }
  // owner = (i, j+1): q is its y-neighbour
  {
    Pos o = q;
    o.j = q.j + 1;
    if (a.b_drhoy && interior_pos(o)) {
      baropg_face_rev<T, 1>(o, vg, prm, rm, a.rho, a.rmean, a.dt, a.dx, a.dvm, a.b_drhoy, g0, g1, d0, d1);
      #pragma unroll
      for (int k = 0; k < KB; ++k) br[k] += g1[k];
      bdt += d1;
    }
  }
  if (a.b_rho) add_col(a.b_rho + c, br);
  if (a.b_dt) a.b_dt[ix2(q, q.i, q.j)] += bdt;
}

// =====================================================================
// colsum: out2[n] = sum_{k<kbm1} f_n(k) dz(k) for up to 5 (f, out) pairs
// =====================================================================
template <class T> struct ColsumArgs {
  const T* f[5];
  T* out[5];
  int n;
};
template <class T>
__global__ void __launch_bounds__(TPB) k_colsum(Batch B, Vert<T> vg, ColsumArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0);
  for (int n = 0; n < a.n; ++n) {
    T s = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) s += a.f[n][c + k] * vg.dz(p.m, k);
    a.out[n][ix2(p, p.i, p.j)] = s;
  }
}
// reverse: fbar(k) += outbar * dz(k)
template <class T> struct ColsumAdjArgs {
  const T* bout[5];
  T* bf[5];
  int n;
};
template <class T>
__global__ void __launch_bounds__(TPB) k_colsum_adj(Batch B, Vert<T> vg, ColsumAdjArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0);
  T dz[KBM1];
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) dz[k] = vg.dz(p.m, k);
  for (int n = 0; n < a.n; ++n) {
    if (!a.bout[n] || !a.bf[n]) continue;
    // whole column loaded before the first store (the stores may alias the loads)
    const T g = a.bout[n][ix2(p, p.i, p.j)];
    T x[KBM1];
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) x[k] = a.bf[n][c + k];
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) a.bf[n][c + k] = x[k] + g * dz[k];
  }
}

// =====================================================================
// adjust: u1 = (u - sum_k u dz) + (utb+utf)/(dt(i)+dt(i-1)+eps) for i>=1,
// v1 likewise in j; levels k < kbm1
// =====================================================================
template <class T> struct AdjustArgs {
  const T *u, *v, *utb, *utf, *vtb, *vtf, *dt;
  T *u1, *v1;
};
template <class T>
__global__ void __launch_bounds__(TPB) k_adjust(Batch B, Vert<T> vg, AdjustArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0);
  const int64_t c2 = ix2(p, p.i, p.j);
  T x[KB];
  ld_col(a.u + c, x);
  if (p.i >= 1) {
    T s = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) s += x[k] * vg.dz(p.m, k);
    const T corr = (a.utb[c2] + a.utf[c2]) / (a.dt[c2] + a.dt[ix2(p, p.i - 1, p.j)] + T(1e-20));
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) x[k] = (x[k] - s) + corr;
  }
  st_col(a.u1 + c, x);
  ld_col(a.v + c, x);
  if (p.j >= 1) {
    T s = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) s += x[k] * vg.dz(p.m, k);
    const T corr = (a.vtb[c2] + a.vtf[c2]) / (a.dt[c2] + a.dt[ix2(p, p.i, p.j - 1)] + T(1e-20));
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) x[k] = (x[k] - s) + corr;
  }
  st_col(a.v1 + c, x);
}

template <class T> struct AdjustAdjArgs {
  const T *utb, *utf, *vtb, *vtf, *dt;
  const T *b_u1, *b_v1;
  T *b_u, *b_v, *b_utb, *b_utf, *b_vtb, *b_vtf, *b_dt;
};
template <class T>
__global__ void __launch_bounds__(TPB) k_adjust_adj(Batch B, Vert<T> vg, AdjustAdjArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0);
  const int64_t c2 = ix2(p, p.i, p.j);
  T bdt = T(0);
  // sum of the level cotangents of column (ii, jj), k < kbm1
  auto colS = [&](const T* b, int ii, int jj) {
    T s = T(0);
    const int64_t cc = ix3(p, ii, jj, 0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) s += b[cc + k];
    return s;
  };
  // u part
  if (a.b_u1) {
    T g[KB];
    ld_col(a.b_u1 + c, g);
    if (p.i >= 1) {
      T S = T(0);
      #pragma unroll
      for (int k = 0; k < KBM1; ++k) S += g[k];
      const T den = a.dt[c2] + a.dt[ix2(p, p.i - 1, p.j)] + T(1e-20);
      const T corr = (a.utb[c2] + a.utf[c2]) / den;
      if (a.b_utb) a.b_utb[c2] += S / den;
      if (a.b_utf) a.b_utf[c2] += S / den;
      bdt += -S * corr / den;
      #pragma unroll
      for (int k = 0; k < KBM1; ++k) g[k] = g[k] - S * vg.dz(p.m, k);
    }
    if (a.b_u) add_col(a.b_u + c, g);
    if (p.i + 1 < p.im) {   // dt(i) is dt(i'-1) of column i' = i+1
      const int64_t n2 = ix2(p, p.i + 1, p.j);
      const T S = colS(a.b_u1, p.i + 1, p.j);
      const T den = a.dt[n2] + a.dt[c2] + T(1e-20);
      bdt += -S * ((a.utb[n2] + a.utf[n2]) / den) / den;
    }
  }
  if (a.b_v1) {
    T g[KB];
    ld_col(a.b_v1 + c, g);
    if (p.j >= 1) {
      T S = T(0);
      #pragma unroll
      for (int k = 0; k < KBM1; ++k) S += g[k];
      const T den = a.dt[c2] + a.dt[ix2(p, p.i, p.j - 1)] + T(1e-20);
      const T corr = (a.vtb[c2] + a.vtf[c2]) / den;
      if (a.b_vtb) a.b_vtb[c2] += S / den;
      if (a.b_vtf) a.b_vtf[c2] += S / den;
      bdt += -S * corr / den;
      #pragma unroll
      for (int k = 0; k < KBM1; ++k) g[k] = g[k] - S * vg.dz(p.m, k);
    }
    if (a.b_v) add_col(a.b_v + c, g);
    if (p.j + 1 < p.jm) {
      const int64_t n2 = ix2(p, p.i, p.j + 1);
      const T S = colS(a.b_v1, p.i, p.j + 1);
      const T den = a.dt[n2] + a.dt[c2] + T(1e-20);
      bdt += -S * ((a.vtb[n2] + a.vtf[n2]) / den) / den;
    }
  }
  if (a.b_dt) a.b_dt[c2] += bdt;
}

// =====================================================================
// vertvl + bcond 5
// =====================================================================
template <class T> struct VertvlArgs {
  const T *u, *v, *w, *dt, *etf, *etb, *vfluxb, *vfluxf, *dx, *dy, *fsm;
  T* wout;
  int all_levels;   // bcond 5 masks every level (closed basin of the semi-implicit core), else k < kbm1
};
template <class T> __device__ __forceinline__ T vv_xf(const Pos& p, const VertvlArgs<T>& a, int i, int j, int k) {
  return T(0.25) * (a.dy[ix2(p, i, j)] + a.dy[ix2(p, i - 1, j)]) * (a.dt[ix2(p, i, j)] + a.dt[ix2(p, i - 1, j)]) *
         a.u[ix3(p, i, j, k)];
}
template <class T> __device__ __forceinline__ T vv_yf(const Pos& p, const VertvlArgs<T>& a, int i, int j, int k) {
  return T(0.25) * (a.dx[ix2(p, i, j)] + a.dx[ix2(p, i, j - 1)]) * (a.dt[ix2(p, i, j)] + a.dt[ix2(p, i, j - 1)]) *
         a.v[ix3(p, i, j, k)];
}
template <class T>
__global__ void __launch_bounds__(TPB) k_vertvl(Batch B, Vert<T> vg, Params prm, VertvlArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0), c2 = ix2(p, p.i, p.j);
  const T fsm = a.fsm[c2];
  T w[KB];
  if (interior_pos(p)) {
    const int i = p.i, j = p.j;
    w[0] = T(0.5) * (a.vfluxb[c2] + a.vfluxf[c2]);
    const T area = a.dx[c2] * a.dy[c2];
    const T eta = (a.etf[c2] - a.etb[c2]) / T(prm.dti2);
    T acc = T(0);
    #pragma unroll
    for (int k = 0; k < KBM1; ++k) {
      const T div = vv_xf(p, a, i + 1, j, k) - vv_xf(p, a, i, j, k) + vv_yf(p, a, i, j + 1, k) - vv_yf(p, a, i, j, k);
      const T inc = (div / area + eta) * vg.dz(p.m, k);
      acc = acc + inc;
      w[k + 1] = w[0] + acc;
    }
  } else {
    ld_col(a.w + c, w);
  }
  // bcond 5: w * fsm on k < kbm1 (explicit, baroclinic.py:849-851, and the open
  // forms of the tide policy); on every level for the closed-basin policy of
  // the semi-implicit core (closed_basin.py:77-79)
  #pragma unroll
  for (int k = 0; k < KB; ++k)
    if (k < KBM1 || a.all_levels) w[k] = w[k] * fsm;
  st_col(a.wout + c, w);
}

// reverse, pass 1 (per interior column): div cotangent bdiv(k) and eta
template <class T> struct VertvlAdjArgs {
  const T *u, *v, *dt, *dx, *dy, *fsm;
  const T* b_w;
  T* bdiv;   // scratch [3-D]: cotangent of div(k) / area at interior columns, else 0
  T *b_u, *b_v, *b_dt, *b_etf, *b_etb;
  int all_levels;   // as VertvlArgs::all_levels
};
template <class T>
__global__ void __launch_bounds__(TPB) k_vertvl_adj1(Batch B, Vert<T> vg, Params prm, VertvlAdjArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0), c2 = ix2(p, p.i, p.j);
  T bd[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) bd[k] = T(0);
  if (interior_pos(p)) {
    const T fsm = a.fsm[c2];
    const T area = a.dx[c2] * a.dy[c2];
    T suffix = T(0), beta = T(0);
    // w(k+1) = w0 + sum_{k'<=k} inc(k'); wout(k) = w(k)*fsm for k<kbm1
    #pragma unroll
    for (int k = KBM1 - 1; k >= 0; --k) {
      const int kw = k + 1;
      suffix += a.b_w[c + kw] * ((kw < KBM1 || a.all_levels) ? fsm : T(1));
      const T binc = suffix * vg.dz(p.m, k);
      bd[k] = binc / area;
      beta += binc;
    }
    beta = beta / T(prm.dti2);
    if (a.b_etf) a.b_etf[c2] += beta;
    if (a.b_etb) a.b_etb[c2] -= beta;
  }
  st_col(a.bdiv + c, bd);
}
// reverse, pass 2 (one thread per cell, a column's 16 levels in 16 adjacent
// lanes): the u, v and dt cotangents from bdiv.  Face i of cell (i, j, k)
// (i, j >= 1) carries xflux = dy4 (dt(i) + dt(i-1)) u(i) with
// dy4 = (dy(i) + dy(i-1)) / 4, entering div(i-1) with + and div(i) with -;
// face j likewise with v.  dt's partial is summed over the column's levels
// k < kbm1 by lane shuffles.
template <class T>
__global__ void __launch_bounds__(TPB) k_vertvl_adj2(Batch B, VertvlAdjArgs<T> a) {
  const int nblk = B.blk3[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q;
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    T pdt = T(0);
    if (live && k < KBM1) {
      const int i = q.i, j = q.j, im = q.im, jm = q.jm;
      auto bdv = [&](int ii, int jj) -> T {   // bdiv at column (ii, jj), level k; 0 outside the member
        const bool in = ii >= 0 && ii < im && jj >= 0 && jj < jm;
        const T x = ld_at(a.bdiv, q.o3, lix3(q, clampi(ii, im), clampi(jj, jm), k));
        return in ? x : T(0);
      };
      auto L2 = [&](const T* f, int ii, int jj) { return ld_at(f, q.o2, lix2(q, clampi(ii, im), clampi(jj, jm))); };
      const uint32_t ix = lix3(q, i, j, k);
      const T b0 = bdv(i, j), bw = bdv(i - 1, j), bs = bdv(i, j - 1), be = bdv(i + 1, j), bn = bdv(i, j + 1);
      const T uq = ld_at(a.u, q.o3, ix), vq = ld_at(a.v, q.o3, ix);
      const T ue = ld_at(a.u, q.o3, lix3(q, clampi(i + 1, im), j, k));
      const T vn = ld_at(a.v, q.o3, lix3(q, i, clampi(j + 1, jm), k));
      const T dyq = L2(a.dy, i, j), dxq = L2(a.dx, i, j), dtq = L2(a.dt, i, j);
      if (i >= 1 && j >= 1) {   // faces i and j of this cell
        const T cx = T(0.25) * (dyq + L2(a.dy, i - 1, j)), dts = dtq + L2(a.dt, i - 1, j);
        const T cy = T(0.25) * (dxq + L2(a.dx, i, j - 1)), dty = dtq + L2(a.dt, i, j - 1);
        const T fx = bw - b0, fy = bs - b0;
        pdt += cx * uq * fx + cy * vq * fy;
        if (a.b_u) {
          T& r = ref_at(a.b_u, q.o3, ix);
          r = r + cx * dts * fx;
        }
        if (a.b_v) {
          T& r = ref_at(a.b_v, q.o3, ix);
          r = r + cy * dty * fy;
        }
      }
      // dt(i, j) as dt(i'-1) of face i' = i + 1 and dt(j'-1) of face j' = j + 1
      if (i + 1 < im && j >= 1) pdt += T(0.25) * (L2(a.dy, i + 1, j) + dyq) * ue * (b0 - be);
      if (j + 1 < jm && i >= 1) pdt += T(0.25) * (L2(a.dx, i, j + 1) + dxq) * vn * (b0 - bn);
    }
#pragma unroll
    for (int o = KB / 2; o > 0; o >>= 1) pdt += __shfl_xor(pdt, o, KB);
    if (live && k == 0 && a.b_dt) {
      T& r = ref_at(a.b_dt, q.o2, lix2(q, q.i, q.j));
      r = r + pdt;
    }
  }
}

// =====================================================================
// proft (nbc 1: flux surface BC; nbc 3: prescribed surface value)
// =====================================================================
template <class T> struct ProftArgs {
  const T *f, *kh, *etf, *h, *wfsurf, *fsurf;
  T* out;
  int nbc;
  // the reference operator's work arrays (tracers.py:1000-1010 returns them):
  // a, c, ee, gg (3-D) and dh = h + etf (2-D); each may be null
  T *ao, *co, *eeo, *ggo, *dho;
};
template <class T>
__device__ __forceinline__ void proft_coef(const Pos& p, const Vert<T>& vg, const Params& prm, const T* kh, T dh,
                                           T* a, T* c) {
  const T dh2 = dh * dh;
  #pragma unroll
  for (int k = 0; k < KB; ++k) a[k] = c[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T num = T(-prm.dti2) * (kh[k] + T(prm.umol));
    a[k - 1] = num / (vg.dz(p.m, k - 1) * vg.dzz(p.m, k - 1) * dh2);
    c[k] = num / (vg.dz(p.m, k) * vg.dzz(p.m, k - 1) * dh2);
  }
}
template <class T>
__global__ void __launch_bounds__(TPB) k_proft(Batch B, Vert<T> vg, Params prm, ProftArgs<T> s) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0), c2 = ix2(p, p.i, p.j);
  T f[KB], kh[KB], a[KB], cc[KB], ee[KB], gg[KB];
  ld_col(s.f + c, f);
  ld_col(s.kh + c, kh);
  const T dh = s.h[c2] + s.etf[c2];
  proft_coef(p, vg, prm, kh, dh, a, cc);
  if (s.nbc == 1) {
    const T den = a[0] - T(1);
    ee[0] = a[0] / den;
    gg[0] = (T(-prm.dti2) * s.wfsurf[c2] / (-vg.dz(p.m, 0) * dh) - f[0]) / den;
  } else {
    ee[0] = T(0);
    gg[0] = s.fsurf[c2];
  }
  #pragma unroll
  for (int k = 1; k < KBM2; ++k) {
    const T den = a[k] + cc[k] * (T(1) - ee[k - 1]) - T(1);
    const T inv = T(1) / (den + T(1e-20));
    ee[k] = a[k] * inv;
    gg[k] = (cc[k] * gg[k - 1] - f[k]) * inv;
  }
  const int kb = KBM1 - 1;
  T x[KB];
  x[KB - 1] = f[KB - 1];
  x[kb] = (cc[kb] * gg[kb - 1] - f[kb]) / (cc[kb] * (T(1) - ee[kb - 1]) - T(1) + T(1e-20));
  #pragma unroll
  for (int k = kb - 1; k >= 0; --k) x[k] = ee[k] * x[k + 1] + gg[k];
  st_col(s.out + c, x);
  if (s.eeo || s.ggo) {   // levels past the sweep hold zeros, as the reference's buffers do
    #pragma unroll
    for (int k = KBM2; k < KB; ++k) ee[k] = gg[k] = T(0);
    if (s.eeo) st_col(s.eeo + c, ee);
    if (s.ggo) st_col(s.ggo + c, gg);
  }
  if (s.ao) st_col(s.ao + c, a);
  if (s.co) st_col(s.co + c, cc);
  if (s.dho) s.dho[c2] = dh;
}

template <class T> struct ProftAdjArgs {
  const T *f, *kh, *etf, *h, *wfsurf, *fsurf;
  const T* b_out;
  T *b_f, *b_kh, *b_etf;
  int nbc;
  int assign_f;   // 1: b_f is assigned (this kernel is its first writer), 0: accumulated
};
template <class T>
__global__ void __launch_bounds__(TPB) k_proft_adj(Batch B, Vert<T> vg, Params prm, ProftAdjArgs<T> s) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0), c2 = ix2(p, p.i, p.j);
  T f[KB], kh[KB], a[KB], cc[KB], ee[KB], gg[KB], inv[KB], x[KB];
  ld_col(s.f + c, f);
  ld_col(s.kh + c, kh);
  const T dh = s.h[c2] + s.etf[c2];
  const T dti2 = T(prm.dti2);
  proft_coef(p, vg, prm, kh, dh, a, cc);
  T den0 = a[0] - T(1), tflux = T(0);
  if (s.nbc == 1) {
    ee[0] = a[0] / den0;
    tflux = -dti2 * s.wfsurf[c2] / (-vg.dz(p.m, 0) * dh);
    gg[0] = (tflux - f[0]) / den0;
  } else {
    ee[0] = T(0);
    gg[0] = s.fsurf[c2];   // prescribed surface value: no dependence on the inputs
  }
  #pragma unroll
  for (int k = 1; k < KBM2; ++k) {
    const T den = a[k] + cc[k] * (T(1) - ee[k - 1]) - T(1);
    inv[k] = T(1) / (den + T(1e-20));
    ee[k] = a[k] * inv[k];
    gg[k] = (cc[k] * gg[k - 1] - f[k]) * inv[k];
  }
  const int kb = KBM1 - 1;
  const T denb = cc[kb] * (T(1) - ee[kb - 1]) - T(1) + T(1e-20);
  x[kb] = (cc[kb] * gg[kb - 1] - f[kb]) / denb;
  #pragma unroll
  for (int k = kb - 1; k >= 0; --k) x[k] = ee[k] * x[k + 1] + gg[k];
  // ---- reverse
  T bx[KB], bee[KB], bgg[KB], ba[KB], bc[KB], bf[KB];
  ld_col(s.b_out + c, bx);
  #pragma unroll
  for (int k = 0; k < KB; ++k) bee[k] = bgg[k] = ba[k] = bc[k] = bf[k] = T(0);
  bf[KB - 1] += bx[KB - 1];
  #pragma unroll
  for (int k = 0; k <= kb - 1; ++k) {   // x_k = ee_k x_{k+1} + gg_k (forward ran k = kb-1 .. 0)
    bee[k] += bx[k] * x[k + 1];
    bx[k + 1] += bx[k] * ee[k];
    bgg[k] += bx[k];
  }
  {
    const T bn = bx[kb] / denb;
    const T bd = -bx[kb] * x[kb] / denb;
    bc[kb] += bn * gg[kb - 1] + bd * (T(1) - ee[kb - 1]);
    bgg[kb - 1] += bn * cc[kb];
    bf[kb] -= bn;
    bee[kb - 1] -= bd * cc[kb];
  }
  #pragma unroll
  for (int k = KBM2 - 1; k >= 1; --k) {
    const T sk = cc[k] * gg[k - 1] - f[k];
    const T bs = bgg[k] * inv[k];
    const T binv = bgg[k] * sk + bee[k] * a[k];
    ba[k] += bee[k] * inv[k];
    bc[k] += bs * gg[k - 1];
    bgg[k - 1] += bs * cc[k];
    bf[k] -= bs;
    const T bden = -binv * inv[k] * inv[k];
    ba[k] += bden;
    bc[k] += bden * (T(1) - ee[k - 1]);
    bee[k - 1] -= bden * cc[k];
  }
  T bdh = T(0);
  if (s.nbc == 1) {
    const T btf = bgg[0] / den0;
    bf[0] -= bgg[0] / den0;
    const T bden0 = -bgg[0] * gg[0] / den0 - bee[0] * ee[0] / den0;
    ba[0] += bee[0] / den0 + bden0;
    bdh += -btf * tflux / dh;
  }
  // coefficients: a(k-1) = num_k / (P dh2), c(k) = num_k / (Q dh2)
  const T dh2 = dh * dh;
  T bdh2 = T(0);
  T bkh[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) bkh[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T P = vg.dz(p.m, k - 1) * vg.dzz(p.m, k - 1);
    const T Q = vg.dz(p.m, k) * vg.dzz(p.m, k - 1);
    const T bnum = ba[k - 1] / (P * dh2) + bc[k] / (Q * dh2);
    bdh2 -= ba[k - 1] * a[k - 1] / dh2 + bc[k] * cc[k] / dh2;
    bkh[k] += bnum * (-dti2);
  }
  bdh += T(2) * dh * bdh2;
  if (s.b_f) put_col(s.b_f + c, bf, s.assign_f != 0);
  if (s.b_kh) add_col(s.b_kh + c, bkh);
  if (s.b_etf) s.b_etf[c2] += bdh;
}

// =====================================================================
// profu / profv (DIR 0: u at (i-1/2, j), DIR 1: v at (i, j-1/2))
// =====================================================================
template <class T> struct ProfuvArgs {
  const T *f, *km, *etf, *h, *wsurf, *wbot, *cbc, *ub, *vb, *mask;
  T *out, *wbot_out, *tps_out;
};
template <class T, int DIR>
__device__ __forceinline__ T profuv_mag(const Pos& p, const T* ub, const T* vb) {
  const int i = p.i, j = p.j, kb = KBM1 - 1;
  if (DIR == 0) {
    const T vs = vb[ix3(p, i, j, kb)] + vb[ix3(p, i, j + 1, kb)] + vb[ix3(p, i - 1, j, kb)] + vb[ix3(p, i - 1, j + 1, kb)];
    const T u0 = ub[ix3(p, i, j, kb)];
    const T q = T(0.25) * vs;
    return sqrt(u0 * u0 + q * q + T(1e-8));
  } else {
    const T us = ub[ix3(p, i, j, kb)] + ub[ix3(p, i + 1, j, kb)] + ub[ix3(p, i, j - 1, kb)] + ub[ix3(p, i + 1, j - 1, kb)];
    const T v0 = vb[ix3(p, i, j, kb)];
    const T q = T(0.25) * us;
    return sqrt(q * q + v0 * v0 + T(1e-8));
  }
}

// forward sweep and back substitution of one interior column; keeps the
// intermediates needed by the reverse sweep
template <class T, int DIR> struct ProfuvCol {
  T a[KB], c[KB], ee[KB], gg[KB], inv[KB], f[KB], x[KB], kmf[KB];
  T dhf, den0, tflux, mag, cbca, tps, fric, denb, m;
};
template <class T, int DIR>
__device__ __forceinline__ void profuv_col(const Pos& p, const Vert<T>& vg, const Params& prm, const ProfuvArgs<T>& s,
                                           ProfuvCol<T, DIR>& C) {
  const int i = p.i, j = p.j;
  const int ni = DIR == 0 ? i - 1 : i, nj = DIR == 0 ? j : j - 1;
  const int64_t c = ix3(p, i, j, 0), cn = ix3(p, ni, nj, 0);
  const int64_t c2 = ix2(p, i, j), n2 = ix2(p, ni, nj);
  const T dti2 = T(prm.dti2);
  ld_col(s.f + c, C.f);
  C.dhf = T(0.5) * (s.h[c2] + s.etf[c2] + s.h[n2] + s.etf[n2]);
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.kmf[k] = T(0.5) * (s.km[c + k] + s.km[cn + k]);
  const T dh2 = C.dhf * C.dhf;
  #pragma unroll
  for (int k = 0; k < KB; ++k) C.a[k] = C.c[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T num = -dti2 * (C.kmf[k] + T(prm.umol));
    C.a[k - 1] = num / (vg.dz(p.m, k - 1) * vg.dzz(p.m, k - 1) * dh2);
    C.c[k] = num / (vg.dz(p.m, k) * vg.dzz(p.m, k - 1) * dh2);
  }
  C.den0 = C.a[0] - T(1);
  C.ee[0] = C.a[0] / (C.den0 + T(1e-20));
  C.tflux = -dti2 * s.wsurf[c2] / (-vg.dz(p.m, 0) * C.dhf);
  C.gg[0] = (C.tflux - C.f[0]) / (C.den0 + T(1e-20));
  #pragma unroll
  for (int k = 1; k < KBM2; ++k) {
    const T den = C.a[k] + C.c[k] * (T(1) - C.ee[k - 1]) - T(1);
    C.inv[k] = T(1) / (den + T(1e-20));
    C.ee[k] = C.a[k] * C.inv[k];
    C.gg[k] = (C.c[k] * C.gg[k - 1] - C.f[k]) * C.inv[k];
  }
  const int kb = KBM1 - 1;
  C.cbca = T(0.5) * (s.cbc[c2] + s.cbc[n2]);
  C.mag = profuv_mag<T, DIR>(p, s.ub, s.vb);
  C.tps = C.cbca * C.mag;
  C.fric = C.tps * dti2 / (-vg.dz(p.m, kb) * C.dhf);
  C.denb = C.fric - T(1) - (C.ee[kb - 1] - T(1)) * C.c[kb];
  C.m = s.mask[c2];
  const T numb = C.c[kb] * C.gg[kb - 1] - C.f[kb];
  C.x[kb] = numb / (C.denb + T(1e-20)) * C.m;
  #pragma unroll
  for (int k = kb - 1; k >= 0; --k) C.x[k] = (C.ee[k] * C.x[k + 1] + C.gg[k]) * C.m;
  C.x[KB - 1] = C.f[KB - 1];
}

template <class T, int DIR>
__global__ void __launch_bounds__(TPB) k_profuv(Batch B, Vert<T> vg, Params prm, ProfuvArgs<T> s) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0), c2 = ix2(p, p.i, p.j);
  if (interior_pos(p)) {
    ProfuvCol<T, DIR> C;
    profuv_col<T, DIR>(p, vg, prm, s, C);
    st_col(s.out + c, C.x);
    s.wbot_out[c2] = -C.tps * C.x[KBM1 - 1];
    if (s.tps_out) s.tps_out[c2] = C.tps;
  } else {
    T x[KB];
    ld_col(s.f + c, x);
    st_col(s.out + c, x);
    s.wbot_out[c2] = s.wbot[c2];
    if (s.tps_out) s.tps_out[c2] = T(0);
  }
}

// reverse pass 1 (per interior column): own-column cotangents of f, and
// partial cotangents of km, etf (own and neighbour) and of mag
template <class T> struct ProfuvAdjArgs {
  ProfuvArgs<T> fw;
  const T *b_out, *b_wbot;
  T* b_f;                 // accumulated [3-D]
  T *s_km0, *s_km1;       // scratch [3-D]: km cotangent for the own / neighbour column
  T *s_etf0, *s_etf1;     // scratch [2-D]
  T* s_mag;               // scratch [2-D]: cotangent of mag
  T* b_wbot_in;           // accumulated [2-D] (pass-through at non-interior columns)
  int assign_f;           // 1: b_f is assigned (this kernel is its first writer), 0: accumulated
};
template <class T, int DIR>
__global__ void __launch_bounds__(TPB) k_profuv_adj1(Batch B, Vert<T> vg, Params prm, ProfuvAdjArgs<T> s) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0), c2 = ix2(p, p.i, p.j);
  T zero[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) zero[k] = T(0);
  if (!interior_pos(p)) {
    T g[KB];
    ld_col(s.b_out + c, g);
    if (s.b_f) put_col(s.b_f + c, g, s.assign_f != 0);
    if (s.b_wbot_in && s.b_wbot) s.b_wbot_in[c2] += s.b_wbot[c2];
    st_col(s.s_km0 + c, zero);
    st_col(s.s_km1 + c, zero);
    s.s_etf0[c2] = s.s_etf1[c2] = s.s_mag[c2] = T(0);
    return;
  }
  ProfuvCol<T, DIR> C;
  profuv_col<T, DIR>(p, vg, prm, s.fw, C);
  const T dti2 = T(prm.dti2);
  const int kb = KBM1 - 1;
  T bx[KB], bee[KB], bgg[KB], ba[KB], bc[KB], bf[KB];
  ld_col(s.b_out + c, bx);
  #pragma unroll
  for (int k = 0; k < KB; ++k) bee[k] = bgg[k] = ba[k] = bc[k] = bf[k] = T(0);
  bf[KB - 1] += bx[KB - 1];
  // wbot = -tps * x(kb)
  const T bw = s.b_wbot ? s.b_wbot[c2] : T(0);
  bx[kb] += -bw * C.tps;
  T btps = -bw * C.x[kb];
  // x_k = (ee_k x_{k+1} + gg_k) * m
  #pragma unroll
  for (int k = 0; k <= kb - 1; ++k) {
    const T g = bx[k] * C.m;
    bee[k] += g * C.x[k + 1];
    bx[k + 1] += g * C.ee[k];
    bgg[k] += g;
  }
  // x_kb = numb / (denb + eps) * m
  const T denbe = C.denb + T(1e-20);
  const T xkb = (C.c[kb] * C.gg[kb - 1] - C.f[kb]) / denbe;
  const T bn = bx[kb] * C.m / denbe;
  const T bd = -bx[kb] * C.m * xkb / denbe;
  bc[kb] += bn * C.gg[kb - 1];
  bgg[kb - 1] += bn * C.c[kb];
  bf[kb] -= bn;
  // denb = fric - 1 - (ee_{kb-1} - 1) c_kb
  T bfric = bd;
  bee[kb - 1] -= bd * C.c[kb];
  bc[kb] -= bd * (C.ee[kb - 1] - T(1));
  // fric = tps dti2 / (-dz_kb dhf)
  const T dzk = -vg.dz(p.m, kb);
  btps += bfric * dti2 / (dzk * C.dhf);
  T bdhf = -bfric * C.fric / C.dhf;
  // tps = cbca * mag
  s.s_mag[c2] = btps * C.cbca;
  #pragma unroll
  for (int k = KBM2 - 1; k >= 1; --k) {
    const T sk = C.c[k] * C.gg[k - 1] - C.f[k];
    const T bs = bgg[k] * C.inv[k];
    const T binv = bgg[k] * sk + bee[k] * C.a[k];
    ba[k] += bee[k] * C.inv[k];
    bc[k] += bs * C.gg[k - 1];
    bgg[k - 1] += bs * C.c[k];
    bf[k] -= bs;
    const T bden = -binv * C.inv[k] * C.inv[k];
    ba[k] += bden;
    bc[k] += bden * (T(1) - C.ee[k - 1]);
    bee[k - 1] -= bden * C.c[k];
  }
  {
    const T d0 = C.den0 + T(1e-20);
    const T btf = bgg[0] / d0;
    bf[0] -= bgg[0] / d0;
    const T bden0 = -bgg[0] * C.gg[0] / d0 - bee[0] * C.ee[0] / d0;
    ba[0] += bee[0] / d0 + bden0;
    bdhf += -btf * C.tflux / C.dhf;
  }
  const T dh2 = C.dhf * C.dhf;
  T bdh2 = T(0);
  T bkmf[KB];
  #pragma unroll
  for (int k = 0; k < KB; ++k) bkmf[k] = T(0);
  #pragma unroll
  for (int k = 1; k < KBM1; ++k) {
    const T P = vg.dz(p.m, k - 1) * vg.dzz(p.m, k - 1);
    const T Q = vg.dz(p.m, k) * vg.dzz(p.m, k - 1);
    const T bnum = ba[k - 1] / (P * dh2) + bc[k] / (Q * dh2);
    bdh2 -= ba[k - 1] * C.a[k - 1] / dh2 + bc[k] * C.c[k] / dh2;
    bkmf[k] += bnum * (-dti2);
  }
  bdhf += T(2) * C.dhf * bdh2;
  #pragma unroll
  for (int k = 0; k < KB; ++k) bkmf[k] *= T(0.5);
  if (s.b_f) put_col(s.b_f + c, bf, s.assign_f != 0);
  st_col(s.s_km0 + c, bkmf);
  st_col(s.s_km1 + c, bkmf);
  s.s_etf0[c2] = T(0.5) * bdhf;
  s.s_etf1[c2] = T(0.5) * bdhf;
}

// reverse pass 2 (per column): gather km, etf, ub, vb cotangents
template <class T> struct ProfuvAdj2Args {
  const T *ub, *vb, *cbc;
  const T *s_km0, *s_km1, *s_etf0, *s_etf1, *s_mag;
  T *b_km, *b_etf, *b_ub, *b_vb;
};
template <class T, int DIR>
__global__ void __launch_bounds__(TPB) k_profuv_adj2(Batch B, Vert<T> vg, ProfuvAdj2Args<T> s) {
  Pos p;
  if (!pos2(B, p)) return;
  const int i = p.i, j = p.j, kb = KBM1 - 1;
  const int64_t c = ix3(p, i, j, 0), c2 = ix2(p, i, j);
  // the column whose neighbour is (i, j): (i+1, j) for DIR 0, (i, j+1) for DIR 1
  const int oi = DIR == 0 ? i + 1 : i, oj = DIR == 0 ? j : j + 1;
  const bool has_o = oi < p.im && oj < p.jm;
  T g[KB];
  ld_col(s.s_km0 + c, g);
  if (has_o) {
    const int64_t oc = ix3(p, oi, oj, 0);
    #pragma unroll
    for (int k = 0; k < KB; ++k) g[k] += s.s_km1[oc + k];
  }
  if (s.b_km) add_col(s.b_km + c, g);
  T be = s.s_etf0[c2];
  if (has_o) be += s.s_etf1[ix2(p, oi, oj)];
  if (s.b_etf) s.b_etf[c2] += be;
  // bottom-level velocity cotangents through mag of interior columns
  auto inter = [&](int ii, int jj) { return ii >= 1 && ii <= p.im - 2 && jj >= 1 && jj <= p.jm - 2; };
  auto magv = [&](int ii, int jj) {
    Pos o = p;
    o.i = ii;
    o.j = jj;
    return profuv_mag<T, DIR>(o, s.ub, s.vb);
  };
  T bu = T(0), bv = T(0);
  if (DIR == 0) {
    // mag(i,j) = sqrt(ub(i,j)^2 + (0.25 vs)^2 + 1e-8), vs over vb(i,j), vb(i,j+1), vb(i-1,j), vb(i-1,j+1)
    if (inter(i, j)) bu += s.s_mag[c2] * s.ub[ix3(p, i, j, kb)] / magv(i, j);
    const int oi4[4] = {i, i, i + 1, i + 1}, oj4[4] = {j, j - 1, j, j - 1};
    for (int n = 0; n < 4; ++n) {
      const int ii = oi4[n], jj = oj4[n];
      if (!inter(ii, jj)) continue;
      const T vs = s.vb[ix3(p, ii, jj, kb)] + s.vb[ix3(p, ii, jj + 1, kb)] + s.vb[ix3(p, ii - 1, jj, kb)] +
                   s.vb[ix3(p, ii - 1, jj + 1, kb)];
      bv += s.s_mag[ix2(p, ii, jj)] * (T(0.25) * vs) * T(0.25) / magv(ii, jj);
    }
  } else {
    if (inter(i, j)) bv += s.s_mag[c2] * s.vb[ix3(p, i, j, kb)] / magv(i, j);
    const int oi4[4] = {i, i - 1, i, i - 1}, oj4[4] = {j, j, j + 1, j + 1};
    for (int n = 0; n < 4; ++n) {
      const int ii = oi4[n], jj = oj4[n];
      if (!inter(ii, jj)) continue;
      const T us = s.ub[ix3(p, ii, jj, kb)] + s.ub[ix3(p, ii + 1, jj, kb)] + s.ub[ix3(p, ii, jj - 1, kb)] +
                   s.ub[ix3(p, ii + 1, jj - 1, kb)];
      bu += s.s_mag[ix2(p, ii, jj)] * (T(0.25) * us) * T(0.25) / magv(ii, jj);
    }
  }
  if (s.b_ub) s.b_ub[c + kb] += bu;
  if (s.b_vb) s.b_vb[c + kb] += bv;
}

// =====================================================================
// uvfilt: ub_new = u1 + f (term - sum term dz), term = uf + ub - 2 u1 (k < kbm1)
// =====================================================================
template <class T> struct UVFiltArgs {
  const T *uf, *ub, *u1;
  T *ub_new, *u_new;
  const T* mask;   // dum / dvm applied to the filtered level (semi-implicit core), or null
};
template <class T>
__global__ void __launch_bounds__(TPB) k_uvfilt(Batch B, Vert<T> vg, Params prm, UVFiltArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, 0);
  T uf[KB], ub[KB], u[KB], term[KB];
  ld_col(a.uf + c, uf);
  ld_col(a.ub + c, ub);
  ld_col(a.u1 + c, u);
  T s = T(0);
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) {
    term[k] = uf[k] + ub[k] - T(2) * u[k];
    s += term[k] * vg.dz(p.m, k);
  }
  const T f = T(0.5 * prm.smoth);
  T o[KB];
  #pragma unroll
  for (int k = 0; k < KBM1; ++k) o[k] = u[k] + f * (term[k] - s);
  o[KB - 1] = u[KB - 1];
  if (a.mask) {   // numerics/double_gyre.py:132-184 with gamma 0
    const T mk = a.mask[ix2(p, p.i, p.j)];
    #pragma unroll
    for (int k = 0; k < KB; ++k) o[k] = o[k] * mk;
  }
  st_col(a.ub_new + c, o);
  st_col(a.u_new + c, uf);
}
template <class T> struct UVFiltAdjArgs {
  const T *b_ub_new, *b_u_new;
  T *b_uf, *b_ub, *b_u1;
  const T* mask;
  uint32_t assign;   // bit 0: b_uf, bit 2: b_u1 are assigned (their first writer), not accumulated
};
// one thread per cell, a column's 16 levels in 16 adjacent lanes; the column
// sum of the masked ub cotangent by lane shuffles
template <class T>
__global__ void __launch_bounds__(TPB) k_uvfilt_adj(Batch B, Vert<T> vg, Params prm, UVFiltAdjArgs<T> a) {
  const int nblk = B.blk3[B.M];
  const T f = T(0.5 * prm.smoth);
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q;
    const bool live = pos3_at(B, blk, q);   // a water column's 16 lanes are all live or all not
    const int k = threadIdx.x & (KB - 1);
    const uint32_t ix = live ? lix3(q, q.i, q.j, k) : 0u;
    // every load issued before the column sum waits on the first
    T g = T(0), un = T(0), mk = T(1), ouf = T(0), oub = T(0), ou1 = T(0), dz = T(0);
    if (live) {
      if (a.b_ub_new) g = ld_at(a.b_ub_new, q.o3, ix);
      if (a.b_u_new) un = ld_at(a.b_u_new, q.o3, ix);
      if (a.mask) mk = ld_at(a.mask, q.o2, lix2(q, q.i, q.j));
      if (a.b_uf && !(a.assign & 1u)) ouf = ld_at(a.b_uf, q.o3, ix);
      if (a.b_ub) oub = ld_at(a.b_ub, q.o3, ix);
      if (a.b_u1 && !(a.assign & 4u)) ou1 = ld_at(a.b_u1, q.o3, ix);
      dz = vg.dz(q.m, k);
    }
    if (a.mask) g = g * mk;
    T S = k < KBM1 ? g : T(0);
#pragma unroll
    for (int o = KB / 2; o > 0; o >>= 1) S += __shfl_xor(S, o, KB);
    if (!live) continue;
    const T gt = k < KBM1 ? f * g - f * S * dz : T(0);
    const T gu = k < KBM1 ? g - T(2) * gt : g;
    const T gf = gt + un;
    if (a.b_uf) ref_at(a.b_uf, q.o3, ix) = ouf + gf;
    if (a.b_ub) ref_at(a.b_ub, q.o3, ix) = oub + gt;
    if (a.b_u1) ref_at(a.b_u1, q.o3, ix) = ou1 + gu;
  }
}

// =====================================================================
// surface forcing reset (passive fields): interior e_atmos, vfluxf, swrad = 0;
// wtsurf += vfluxf (tatm - t - tbias) with tatm = t + tbias; wssurf = vfluxf (0 - s - sbias)
// =====================================================================
template <class T> struct ForcingArgs {
  const T *e_atmos, *vfluxf, *swrad, *wtsurf, *wssurf, *t, *s;
  T *e_atmos_o, *vfluxf_o, *swrad_o, *wtsurf_o, *wssurf_o;
};
template <class T>
__global__ void __launch_bounds__(TPB) k_forcing(Batch B, Params prm, ForcingArgs<T> a) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t c2 = ix2(p, p.i, p.j);
  if (interior_pos(p)) {
    a.e_atmos_o[c2] = T(0);
    a.vfluxf_o[c2] = T(0);
    a.swrad_o[c2] = T(0);
    const T ts = a.t[ix3(p, p.i, p.j, 0)];
    const T tb = T(prm.tbias), sb = T(prm.sbias);
    a.wtsurf_o[c2] = a.wtsurf[c2] + T(0) * (((ts + tb) - ts) - tb);
    a.wssurf_o[c2] = T(0) * (T(0) - a.s[ix3(p, p.i, p.j, 0)] - sb);
  } else {
    a.e_atmos_o[c2] = a.e_atmos[c2];
    a.vfluxf_o[c2] = a.vfluxf[c2];
    a.swrad_o[c2] = a.swrad[c2];
    a.wtsurf_o[c2] = a.wtsurf[c2];
    a.wssurf_o[c2] = a.wssurf[c2];
  }
}

// w after the surface reset when the internal mode is skipped: w(I,J,0) = vfluxf
template <class T>
__global__ void __launch_bounds__(TPB) k_wsurf(Batch B, const T* w, const T* vfluxf, T* wout) {
  Pos p;
  if (!pos3(B, p)) return;
  const int64_t c = ix3(p, p.i, p.j, p.k);
  wout[c] = (p.k == 0 && interior_pos(p)) ? vfluxf[ix2(p, p.i, p.j)] : w[c];
}

}  // namespace pom
