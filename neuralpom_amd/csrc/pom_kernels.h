// Generic map (forward) and gather (adjoint) kernels over the point
// functions of pom_ops.h.
//
// Forward: one thread per output point (3-D: k fastest, so a wavefront covers
// 4 whole water columns and every load is a contiguous 256-byte run).
// Adjoint: for each input field F the thread owning input point q sums
//   sum_O sum_{o in box} obar_O(q - o) * d out_O(q - o) / d in_F(q)
// where the partials come from the DS-typed instantiation of eval() whose
// active accesses are the fields seeded at offset o.  Offsets whose
// instantiation is passive are removed at compile time.  No atomics: every
// cotangent element is written by exactly one thread, so the result is
// deterministic.
#pragma once
#include <utility>
#include "pom_ops.h"

namespace pom {

constexpr int MAXO = 12;
template <class T> struct OutTab { T* o[MAXO]; };
template <class T> struct COutTab { const T* o[MAXO]; };

// Launch shapes of one batch (pom_batch_tune; the defaults are the values
// measured by hand on the 64-member seamount batch).  A launch of nblk
// 256-point blocks takes groups(nblk, bpg, min) workgroups, each striding
// over the blocks by the grid size (XCD-aware order kept): bpg blocks per
// workgroup once the launch has more than min blocks.  The arithmetic per
// point does not depend on them.
struct Tuning {
  int32_t map2_bpg = 4, map2_min = 768;     // k_map2 (the barotropic substep)
  int32_t flat2_bpg = 1, flat2_min = 768;   // 2-D k_map and k_gather_all2
  int32_t flat3_bpg = 1, flat3_min = 768;   // 3-D k_map and k_gather_all3
};
inline int groups(int nblk, int bpg, int mn) {
  if (bpg <= 1 || nblk <= mn) return nblk;
  const int g = (nblk + bpg - 1) / bpg;
  return g > mn ? g : mn;
}

template <class T> struct Ctx {
  Batch B;
  Vert<T> vert;
  const T* ramp;
  Params prm;
  int32_t nblk3, nblk2;
  hipStream_t stream;
  Tuning tune;
};

// Boundary operators read horizontal neighbours only at evaluation points
// within EDGE_BAND of the member's edge; elsewhere they are pointwise in
// (i, j).  An input point q farther than EDGE_BAND + R from every edge then
// receives contributions from its own column only, and the gather skips the
// other offsets there.
template <class Op, class = void> struct edge_band { static constexpr int value = -1; };
template <class Op> struct edge_band<Op, std::void_t<decltype(Op::EDGE_BAND)>> {
  static constexpr int value = Op::EDGE_BAND;
};
template <class Op> __device__ __forceinline__ bool deep_point(const Pos& q) {
  constexpr int EB = edge_band<Op>::value;
  if constexpr (EB < 0) {
    return false;
  } else {
    return q.i >= EB + Op::RI && q.i <= q.im - 1 - EB - Op::RI && q.j >= EB + Op::RJ && q.j <= q.jm - 1 - EB - Op::RJ;
  }
}
// The forward accessor of an evaluation point away from every edge (deep_point):
// operators with an edge band then take their interior formula alone (deep_v).
template <class T> struct AccFwdDeep : Acc<T> {
  static constexpr bool deep = true;
};

template <class T, class Op>
__global__ void __launch_bounds__(TPB) k_map(Batch B, Vert<T> vert, const T* ramp, Params prm, FieldTab<T> tab,
                                              OutTab<T> out) {
  const int nblk = Op::OUT3 ? B.blk3[B.M] : B.blk2[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
  Pos p;
  if constexpr (Op::OUT3) {
    if (!pos3_at(B, blk, p)) continue;
  } else {
    if (!pos2_at(B, blk, p)) continue;
  }
  Acc<T> A{&tab, p, vert, ramp};
  const int64_t o_ = Op::OUT3 ? p.o3 : p.o2;   // the member's first element
  const uint32_t li = Op::OUT3 ? lix3(p, p.i, p.j, p.k) : lix2(p, p.i, p.j);
  // every output evaluated before the first store: a store may alias the
  // inputs, so loads after it could neither be shared between outputs nor
  // issued early; points away from the edges of an edge-band operator take
  // its interior formula alone
  T res[Op::NOUT];
  bool done = false;
  if constexpr (edge_band<Op>::value >= 0) {
    if (deep_point<Op>(p)) {
      const AccFwdDeep<T> AD{A};
      static_for<0, Op::NOUT - 1>([&](auto O) {
        constexpr int o = decltype(O)::value;
        res[o] = val(Op::template eval<o>(AD, prm));
      });
      done = true;
    }
  }
  if (!done) {
    static_for<0, Op::NOUT - 1>([&](auto O) {
      constexpr int o = decltype(O)::value;
      res[o] = val(Op::template eval<o>(A, prm));
    });
  }
  static_for<0, Op::NOUT - 1>([&](auto O) {
    constexpr int o = decltype(O)::value;
    if (out.o[o]) ref_at(out.o[o], o_, li) = res[o];
  });
  }
}

// ---- running max of |x| over a launch (the depth-mean velocity check).  NaN
// propagates, as torch.max does: fabs(NaN) is a positive NaN, whose bits
// exceed every finite value's, so the unsigned atomic max keeps it.
__device__ __forceinline__ void amax(float* f, float v) { atomicMax(reinterpret_cast<unsigned int*>(f), __float_as_uint(v)); }
__device__ __forceinline__ void amax(double* f, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(f), (unsigned long long)__double_as_longlong(v));
}
template <class T> __device__ __forceinline__ T nanmax(T a, T b) { return (a != a || a > b) ? a : b; }
// block reduction of per-thread |x| values (TPB threads), one atomic per block
template <class T> __device__ __forceinline__ void block_amax(T v, T* out) {
  __shared__ T red[TPB / 64];
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = red[0];
#pragma unroll
    for (int w = 1; w < TPB / 64; ++w) v = nanmax(v, red[w]);
    if (!(v <= T(0))) amax(out, v);
  }
}

// Workgroups of k_map2 (Tuning::map2_*).  One workgroup per 256-point block
// would give the seamount batch 3 051 workgroups for 1 792 resident slots
// (7 waves per SIMD), one block-max atomic each and a last, partly filled wave
// of workgroups.  A large launch instead takes a quarter as many workgroups,
// each striding over the blocks by the grid size, so that every XCD still
// sweeps one contiguous range of blocks at a time (measured per launch:
// 36.2 us with one block per workgroup, 30.1 us with four, 32.0 with six; the
// HBM fetch fell slightly).

// Two 2-D operators in one pass over the points: B reads A's outputs 0 and 1
// at the point itself as its fields FB0 and FB1 (from registers, through
// AccSub) and nothing of A's at other points.  All outputs are evaluated
// before the first store.  With amax non-null, the running max of |A output
// 1| is accumulated into *amax.
template <class T, class OpA, class OpB, int FB0, int FB1>
__global__ void __launch_bounds__(TPB) k_map2(Batch B, Vert<T> vert, const T* ramp, Params prm, FieldTab<T> ta,
                                               OutTab<T> oa, FieldTab<T> tb, OutTab<T> ob, T* amax) {
  static_assert(!OpA::OUT3 && !OpB::OUT3 && OpA::NOUT == 2, "two 2-D outputs feeding a 2-D operator");
  T vmax = T(0);
  for (int blk = logical_block(); blk < B.blk2[B.M]; blk += (int)gridDim.x) {
  Pos p;
  if (pos2_at(B, blk, p)) {
    const uint32_t li = lix2(p, p.i, p.j);
    T ra[2], rb[OpB::NOUT];
    {
      Acc<T> A{&ta, p, vert, ramp};
      ra[0] = val(OpA::template eval<0>(A, prm));
      ra[1] = val(OpA::template eval<1>(A, prm));
    }
    {
      AccSub<T, FB0, FB1> A{{&tb, p, vert, ramp}, ra[0], ra[1]};
      static_for<0, OpB::NOUT - 1>([&](auto O) {
        constexpr int o = decltype(O)::value;
        rb[o] = val(OpB::template eval<o>(A, prm));
      });
    }
    vmax = nanmax(vmax, fabs(ra[1]));
    if (oa.o[0]) ref_at(oa.o[0], p.o2, li) = ra[0];
    if (oa.o[1]) ref_at(oa.o[1], p.o2, li) = ra[1];
    static_for<0, OpB::NOUT - 1>([&](auto O) {
      constexpr int o = decltype(O)::value;
      if (ob.o[o]) ref_at(ob.o[o], p.o2, li) = rb[o];
    });
  }
  }
  if (amax) block_amax(vmax, amax);   // uniform over the launch
}

// ------------------------------------------------------------ fused gathers
// All seeded input fields of an operator in one pass: at each stencil offset
// the point function is evaluated once with every field seeded there (DS
// tangents with one lane per field), and the lanes go to per-field
// accumulators.  Same sums, in the same order, as one gather per field.
// input cotangents: o[f] += partial sums, or o[f] = when bit f of assign is set
// (the gather is then the first writer of that buffer and it need not be zeroed)
template <class T> struct InTab {
  T* o[MAXF];
  uint32_t assign;
};
template <class T> __device__ __forceinline__ void put_in(const InTab<T>& t, int f, int64_t o, uint32_t li, T v) {
  T& r = ref_at(t.o[f], o, li);
  if ((t.assign >> f) & 1u) r = v;
  else r += v;
}

template <class T, class Op, int DI, int DJ, int DK, uint32_t MASK, bool DEEP = false>
__device__ __forceinline__ void gather_all_term(const FieldTab<T>& tab, const COutTab<T>& obar, const Pos& q, int pk,
                                                const Vert<T>& vert, const T* ramp, const Params& prm,
                                                const NbrIdx& nb, T* acc) {
  if constexpr (MASK != 0u) {
    Pos p = q;
    p.i = q.i - DI;
    p.j = q.j - DJ;
    p.k = pk;
    // Operators without a vertical stencil are evaluated by every thread and
    // the term discarded (select, not multiply: the discarded value may be
    // inf or NaN) where the evaluation point lies outside the member: all
    // reads are clamped into the member, and without a branch per offset the
    // loads of successive offsets overlap.  With a vertical stencil the out-of-
    // column evaluations are a larger share and are skipped (measured faster).
    const bool valid = p.i >= 0 && p.i < p.im && p.j >= 0 && p.j < p.jm && p.k >= 0 && p.k < KB;
    if constexpr (Op::RK > 0) {
      if (!valid) return;
    }
    using ACC = std::conditional_t<DEEP, AccTDeep<T, MASK, DI, DJ, DK>, AccT<T, MASK, DI, DJ, DK>>;
    const AccT<T, MASK, DI, DJ, DK> base{&tab, p, vert, ramp, &nb};
    const ACC A{base};
    const int ci = clampi(p.i, p.im), cj = clampi(p.j, p.jm);
    const uint32_t idx = Op::OUT3 ? lix3(p, ci, cj, clampi(p.k, KB)) : lix2(p, ci, cj);
    // the output cotangents first, each under its own (uniform) null check and
    // with no use in between, so that their loads are in flight together
    T ob[Op::NOUT];
    static_for<0, Op::NOUT - 1>([&](auto O) {
      constexpr int o = decltype(O)::value;
      if constexpr (mask_of<decltype(Op::template eval<o>(A, prm))>::value != 0u) {
        ob[o] = T(0);
        if (obar.o[o]) ob[o] = ld_at(obar.o[o], Op::OUT3 ? p.o3 : p.o2, idx);
      }
    });
    static_for<0, Op::NOUT - 1>([&](auto O) {
      constexpr int o = decltype(O)::value;
      using R = decltype(Op::template eval<o>(A, prm));
      constexpr uint32_t MR = mask_of<R>::value;
      if constexpr (MR != 0u) {
        if (obar.o[o]) {
          const R r = Op::template eval<o>(A, prm);
          for_bits<MR>([&](auto I) {
            constexpr int f = decltype(I)::value;
            const T t = ob[o] * r.d[slot(MR, f)];
            acc[f] += valid ? t : T(0);
          });
        }
      }
    });
  }
}

// distance of the water column of q to its member's nearest edge is below w
__device__ __forceinline__ bool in_band(const Pos& q, int w) {
  return q.i < w || q.i > q.im - 1 - w || q.j < w || q.j > q.jm - 1 - w;
}

// The inputs the own-column path (PATH 1) can reach: the union of the tangent
// lanes of every output at each vertical offset it evaluates.  An input
// outside it receives no contribution from that path, so its cotangent is
// neither read nor written there (an assigned one is written zero).
template <class T, class Op, uint32_t M, int DK, int... O>
constexpr uint32_t deep_reach_o(std::integer_sequence<int, O...>) {
  using A = AccTDeep<T, M, 0, 0, DK>;
  return (0u | ... | mask_of<decltype(Op::template eval<O>(std::declval<const A&>(), std::declval<const Params&>()))>::value);
}
template <class T, class Op, uint32_t M3, uint32_t M2, int... K>
constexpr uint32_t deep_reach_k(std::integer_sequence<int, K...>) {
  return (0u | ... | deep_reach_o<T, Op, (M3 | ((K - Op::RK) == 0 ? M2 : 0u)), K - Op::RK>(
                         std::make_integer_sequence<int, Op::NOUT>{}));
}
template <class T, class Op, uint32_t M3, uint32_t M2>
constexpr uint32_t deep_reach() {
  return deep_reach_k<T, Op, M3, M2>(std::make_integer_sequence<int, 2 * Op::RK + 1>{});
}

// The gather at 3-D point q: 3-D fields accumulate at q, 2-D fields per
// level (offsets with dk = 0 only) and are reduced over the 16 lanes of the
// water column (every lane of the column calls this, live or not).  PATH 0
// picks the path per point; 1 takes the own-column path of points away from
// the edges (deep_point), 2 the full stencil.
template <class T, class Op, uint32_t M3, uint32_t M2, int PATH>
__device__ __forceinline__ void gather3_at(const Pos& q, bool live, const Vert<T>& vert, const T* ramp,
                                           const Params& prm, const FieldTab<T>& tab, const COutTab<T>& obar,
                                           const InTab<T>& inbar) {
  T acc[Op::NIN];
#pragma unroll
  for (int f = 0; f < Op::NIN; ++f) acc[f] = T(0);
  if (live) {
    NbrIdx nb;
    make_nbr(q, nb);
    const bool deep = PATH == 1 || (PATH == 0 && edge_band<Op>::value >= 0 && deep_point<Op>(q));
    if (deep) {
      // own column only: a compact path of the vertical offsets
      static_for<-Op::RK, Op::RK>([&](auto K) {
        constexpr int dk = decltype(K)::value;
        gather_all_term<T, Op, 0, 0, dk, (M3 | (dk == 0 ? M2 : 0u)), true>(tab, obar, q, q.k - dk, vert, ramp, prm,
                                                                          nb, acc);
      });
    } else if constexpr (PATH != 1) {
      static_for<-Op::RI, Op::RI>([&](auto I) {
        static_for<-Op::RJ, Op::RJ>([&](auto J) {
          static_for<-Op::RK, Op::RK>([&](auto K) {
            constexpr int di = decltype(I)::value, dj = decltype(J)::value, dk = decltype(K)::value;
            gather_all_term<T, Op, di, dj, dk, (M3 | (dk == 0 ? M2 : 0u))>(tab, obar, q, q.k - dk, vert, ramp,
                                                                           prm, nb, acc);
          });
        });
      });
    }
  }
  // column sums of the 2-D fields' partials
  static_for<0, Op::NIN - 1>([&](auto Fc) {
    constexpr int f = decltype(Fc)::value;
    if constexpr (!((M3 >> f) & 1u) && ((M2 >> f) & 1u)) {
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) acc[f] += __shfl_xor(acc[f], o, 16);
    }
  });
  // accumulate: every old value is read before the first store (the input
  // cotangent arrays of one launch are distinct), so the reads are in flight
  // together; on the own-column path only the inputs it reaches (RM)
  constexpr uint32_t RM = PATH == 1 ? deep_reach<T, Op, M3, M2>() : ~0u;
  T old[Op::NIN];
  static_for<0, Op::NIN - 1>([&](auto Fc) {
    constexpr int f = decltype(Fc)::value;
    old[f] = T(0);
    if constexpr (((M3 >> f) & 1u) && ((RM >> f) & 1u)) {
      if (live && inbar.o[f] && !((inbar.assign >> f) & 1u)) old[f] = ld_at(inbar.o[f], q.o3, lix3(q, q.i, q.j, q.k));
    } else if constexpr (!((M3 >> f) & 1u) && ((M2 >> f) & 1u) && ((RM >> f) & 1u)) {
      if (live && q.k == 0 && inbar.o[f] && !((inbar.assign >> f) & 1u))
        old[f] = ld_at(inbar.o[f], q.o2, lix2(q, q.i, q.j));
    }
  });
  static_for<0, Op::NIN - 1>([&](auto Fc) {
    constexpr int f = decltype(Fc)::value;
    constexpr bool reach = (RM >> f) & 1u;
    if constexpr ((M3 >> f) & 1u) {
      if (live && inbar.o[f] && (reach || ((inbar.assign >> f) & 1u)))
        ref_at(inbar.o[f], q.o3, lix3(q, q.i, q.j, q.k)) = old[f] + acc[f];
    } else if constexpr ((M2 >> f) & 1u) {
      if (live && q.k == 0 && inbar.o[f] && (reach || ((inbar.assign >> f) & 1u)))
        ref_at(inbar.o[f], q.o2, lix2(q, q.i, q.j)) = old[f] + acc[f];
    }
  });
}

// 3-D operators: one thread per 3-D point
template <class T, class Op, uint32_t M3, uint32_t M2>
__global__ void __launch_bounds__(TPB) k_gather_all3(Batch B, Vert<T> vert, const T* ramp, Params prm,
                                                      FieldTab<T> tab, COutTab<T> obar, InTab<T> inbar) {
  const int nblk = B.blk3[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {   // uniform over the workgroup
    Pos q;
    const bool live = pos3_at(B, blk, q);
    gather3_at<T, Op, M3, M2, 0>(q, live, vert, ramp, prm, tab, obar, inbar);
  }
}

// Operators with an edge band take two launches: the points off the band of
// width W = EDGE_BAND + R (own-column path only), then the band's columns
// from the batch's band list (full stencil only), so that no wave carries
// both paths: with one launch, every wave holding one band column (a wave
// covers four) ran the full stencil as well.
template <class T, class Op, uint32_t M3, uint32_t M2, int W>
__global__ void __launch_bounds__(TPB) k_gather_all3_off(Batch B, Vert<T> vert, const T* ramp, Params prm,
                                                          FieldTab<T> tab, COutTab<T> obar, InTab<T> inbar) {
  const int nblk = B.blk3[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
    Pos q;
    bool live = pos3_at(B, blk, q);
    live = live && !in_band(q, W);
    gather3_at<T, Op, M3, M2, 1>(q, live, vert, ramp, prm, tab, obar, inbar);
  }
}
template <class T, class Op, uint32_t M3, uint32_t M2, int W>
__global__ void __launch_bounds__(TPB) k_gather_all3_band(Batch B, Vert<T> vert, const T* ramp, Params prm,
                                                           FieldTab<T> tab, COutTab<T> obar, InTab<T> inbar) {
  const int n = B.nband[W] * KB;   // 16 lanes per listed column
  for (int base = blockIdx.x * TPB; base < n; base += (int)gridDim.x * TPB) {   // uniform over the workgroup
    const int t = base + (int)threadIdx.x;
    const bool live = t < n;
    Pos q{};
    if (live) {
      const int2 e = B.band[t >> 4];
      const Member mm = B.mem[e.x];
      q.m = e.x; q.im = mm.im; q.jm = mm.jm; q.o2 = mm.off2; q.o3 = mm.off3;
      q.i = e.y / mm.jm;
      q.j = e.y - q.i * mm.jm;
      q.k = t & (KB - 1);
    }
    gather3_at<T, Op, M3, M2, 2>(q, live, vert, ramp, prm, tab, obar, inbar);
  }
}

// 2-D operators: one thread per column
template <class T, class Op, uint32_t M2>
__global__ void __launch_bounds__(TPB) k_gather_all2(Batch B, Vert<T> vert, const T* ramp, Params prm,
                                                      FieldTab<T> tab, COutTab<T> obar, InTab<T> inbar) {
  const int nblk = B.blk2[B.M];
  for (int blk = logical_block(); blk < nblk; blk += (int)gridDim.x) {
  Pos q;
  if (!pos2_at(B, blk, q)) continue;
  T acc[Op::NIN];
#pragma unroll
  for (int f = 0; f < Op::NIN; ++f) acc[f] = T(0);
  NbrIdx nb;
  make_nbr(q, nb);
  if (edge_band<Op>::value >= 0 && deep_point<Op>(q)) {   // own column only
    gather_all_term<T, Op, 0, 0, 0, M2, true>(tab, obar, q, 0, vert, ramp, prm, nb, acc);
  } else {
    static_for<-Op::RI, Op::RI>([&](auto I) {
      static_for<-Op::RJ, Op::RJ>([&](auto J) {
        constexpr int di = decltype(I)::value, dj = decltype(J)::value;
        gather_all_term<T, Op, di, dj, 0, M2>(tab, obar, q, 0, vert, ramp, prm, nb, acc);
      });
    });
  }
  const uint32_t li = lix2(q, q.i, q.j);
  T old[Op::NIN];   // every old value read before the first store (as gather3_at)
  static_for<0, Op::NIN - 1>([&](auto Fc) {
    constexpr int f = decltype(Fc)::value;
    old[f] = T(0);
    if (inbar.o[f] && !((inbar.assign >> f) & 1u)) old[f] = ld_at(inbar.o[f], q.o2, li);
  });
  static_for<0, Op::NIN - 1>([&](auto Fc) {
    constexpr int f = decltype(Fc)::value;
    if (inbar.o[f]) ref_at(inbar.o[f], q.o2, li) = old[f] + acc[f];
  });
  }
}

// ------------------------------------------------------------ host launchers
template <class T, class Op>
inline void launch_map(const Ctx<T>& c, const FieldTab<T>& tab, const OutTab<T>& out) {
  const int nb = Op::OUT3 ? c.nblk3 : c.nblk2;
  if (nb == 0) return;
  const int g = Op::OUT3 ? groups(nb, c.tune.flat3_bpg, c.tune.flat3_min) : groups(nb, c.tune.flat2_bpg, c.tune.flat2_min);
  hipLaunchKernelGGL((k_map<T, Op>), dim3(g), dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp, c.prm, tab, out);
}

template <class T, class OpA, class OpB, int FB0, int FB1>
inline void launch_map2(const Ctx<T>& c, const FieldTab<T>& ta, const OutTab<T>& oa, const FieldTab<T>& tb,
                        const OutTab<T>& ob, T* amax) {
  if (c.nblk2 == 0) return;
  hipLaunchKernelGGL((k_map2<T, OpA, OpB, FB0, FB1>), dim3(groups(c.nblk2, c.tune.map2_bpg, c.tune.map2_min)), dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp,
                     c.prm, ta, oa, tb, ob, amax);
}

// accumulate d(outputs)/d(inputs)^T obar into inbar[F] for every F whose
// inbar pointer is non-null.  Defined in pom_gather_impl.h and instantiated
// once per operator in gather_<op>.hip (parallel compilation units).
template <class T, class Op>
void launch_gather(const Ctx<T>& c, const FieldTab<T>& tab, const COutTab<T>& obar, T* const* inbar,
                   uint32_t assign = 0u);

// the same products for OpAdvt1 and OpAdvq by their flux-form transposed
// stencils (pom_flux.h; instantiated in flux_adv.hip), for a pair of fields
// (T and S; q2 and q2l): tab, obar and inbar hold two entries each, and one
// launch serves both when they share the velocity, aam, w and depth cotangents
template <class T>
void launch_flux_advt1(const Ctx<T>& c, const FieldTab<T>* tab, const COutTab<T>* obar, T* const* const* inbar,
                       uint32_t assign = 0u);
template <class T>
void launch_flux_advq(const Ctx<T>& c, const FieldTab<T>* tab, const COutTab<T>* obar, T* const* const* inbar,
                      uint32_t assign = 0u);

}  // namespace pom
