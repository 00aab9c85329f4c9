// Definition of launch_gather (see pom_kernels.h); included only by the
// per-operator instantiation units gather_<op>.hip.
#pragma once
#include <cstdio>
#include <cstdlib>

#include "pom_kernels.h"

namespace pom {

// accumulate d(outputs)/d(inputs)^T obar into inbar[F] for every F whose
// inbar pointer is non-null
template <class T, class Op>
void launch_gather(const Ctx<T>& c, const FieldTab<T>& tab, const COutTab<T>& obar, T* const* inbar,
                   uint32_t assign) {
  InTab<T> it{};
  it.assign = assign;
  bool any = false;
  for (int f = 0; f < Op::NIN; ++f) {
    it.o[f] = inbar[f];
    any = any || inbar[f] != nullptr;
    // the kernels read every old value before the first store
    for (int g = 0; g < f; ++g)
      if (inbar[f] && inbar[g] == inbar[f]) {
        fprintf(stderr, "launch_gather: input cotangents %d and %d share one array\n", g, f);
        abort();
      }
  }
  if (!any) return;
  constexpr uint32_t ALL = (Op::NIN >= 32) ? ~0u : ((1u << Op::NIN) - 1u);
  constexpr uint32_t M3 = (1u << Op::N3) - 1u;
  constexpr uint32_t M2 = ALL & ~M3;
  if constexpr (Op::OUT3) {
    if (c.nblk3 == 0) return;
    const dim3 g3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min));
    if constexpr (edge_band<Op>::value >= 0) {
      constexpr int W = edge_band<Op>::value + Op::RI;
      static_assert(Op::RI == Op::RJ && W <= BAND_W, "edge band beyond the batch's band list");
      if (c.B.band) {   // the points off the band, then the band's columns
        hipLaunchKernelGGL((k_gather_all3_off<T, Op, M3, M2, W>), g3, dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp,
                           c.prm, tab, obar, it);
        const int nb = (c.B.nband[W] * KB + TPB - 1) / TPB;
        if (nb > 0)
          hipLaunchKernelGGL((k_gather_all3_band<T, Op, M3, M2, W>), dim3(nb), dim3(TPB), 0, c.stream, c.B, c.vert,
                             c.ramp, c.prm, tab, obar, it);
        return;
      }
    }
    hipLaunchKernelGGL((k_gather_all3<T, Op, M3, M2>), g3, dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp, c.prm, tab,
                       obar, it);
  } else {
    if (c.nblk2 == 0) return;
    hipLaunchKernelGGL((k_gather_all2<T, Op, M2>), dim3(groups(c.nblk2, c.tune.flat2_bpg, c.tune.flat2_min)),
                       dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp, c.prm, tab, obar, it);
  }
}

}  // namespace pom

#define POM_INSTANTIATE_GATHER(OP)                                                                      \
  template void pom::launch_gather<float, pom::OP>(const pom::Ctx<float>&, const pom::FieldTab<float>&, \
                                                   const pom::COutTab<float>&, float* const*, uint32_t); \
  template void pom::launch_gather<double, pom::OP>(const pom::Ctx<double>&, const pom::FieldTab<double>&, \
                                                    const pom::COutTab<double>&, double* const*, uint32_t);
