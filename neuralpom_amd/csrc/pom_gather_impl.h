// Definition of launch_gather (see pom_kernels.h); included only by the
// per-operator instantiation units gather_<op>.hip.
#pragma once
#include "pom_kernels.h"

namespace pom {

// accumulate d(outputs)/d(inputs)^T obar into inbar[F] for every F whose
// inbar pointer is non-null
template <class T, class Op>
void launch_gather(const Ctx<T>& c, const FieldTab<T>& tab, const COutTab<T>& obar, T* const* inbar) {
  static_for_host<0, Op::NIN - 1>([&](auto Fc) {
    constexpr int F = decltype(Fc)::value;
    if (!inbar[F]) return;
    if constexpr (F < Op::N3) {
      hipLaunchKernelGGL((k_gather3<T, Op, F>), dim3(c.nblk3), dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp, c.prm,
                         tab, obar, inbar[F]);
    } else if constexpr (Op::OUT3) {
      hipLaunchKernelGGL((k_gather2z<T, Op, F>), dim3(c.nblk3), dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp,
                         c.prm, tab, obar, inbar[F]);
    } else {
      hipLaunchKernelGGL((k_gather2<T, Op, F>), dim3(c.nblk2), dim3(TPB), 0, c.stream, c.B, c.vert, c.ramp, c.prm,
                         tab, obar, inbar[F]);
    }
  });
}

}  // namespace pom

#define POM_INSTANTIATE_GATHER(OP)                                                                      \
  template void pom::launch_gather<float, pom::OP>(const pom::Ctx<float>&, const pom::FieldTab<float>&, \
                                                   const pom::COutTab<float>&, float* const*);          \
  template void pom::launch_gather<double, pom::OP>(const pom::Ctx<double>&, const pom::FieldTab<double>&, \
                                                    const pom::COutTab<double>&, double* const*);
