// Definition of launch_gather (see pom_kernels.h); included only by the
// per-operator instantiation units gather_<op>.hip.
#pragma once
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
  }
  if (!any) return;
  constexpr uint32_t ALL = (Op::NIN >= 32) ? ~0u : ((1u << Op::NIN) - 1u);
  constexpr uint32_t M3 = (1u << Op::N3) - 1u;
  constexpr uint32_t M2 = ALL & ~M3;
  if constexpr (Op::OUT3) {
    if (c.nblk3 == 0) return;
    hipLaunchKernelGGL((k_gather_all3<T, Op, M3, M2>), dim3(groups(c.nblk3, c.tune.flat3_bpg, c.tune.flat3_min)),
                       dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp, c.prm, tab, obar, it);
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
