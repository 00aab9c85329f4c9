// Launchers of the flux-form advt1 and advq adjoints (pom_flux.h).
#include "pom_flux.h"

namespace pom {

template <class T>
void launch_flux_advt1(const Ctx<T>& c, const FieldTab<T>& tab, const COutTab<T>& obar, T* const* inbar,
                       uint32_t assign) {
  InTab<T> it{};
  it.assign = assign;
  bool any = false;
  for (int f = 0; f < OpAdvt1::NIN; ++f) {
    it.o[f] = inbar[f];
    any = any || inbar[f] != nullptr;
  }
  if (!any || c.nblk3 == 0) return;
  hipLaunchKernelGGL(k_flux_advt1<T>, dim3(c.nblk3), dim3(TPB), 0, c.stream, c.B, c.vert, c.prm, tab, obar, it);
}

template <class T>
void launch_flux_advq(const Ctx<T>& c, const FieldTab<T>& tab, const COutTab<T>& obar, T* const* inbar,
                      uint32_t assign) {
  InTab<T> it{};
  it.assign = assign;
  bool any = false;
  for (int f = 0; f < OpAdvq::NIN; ++f) {
    it.o[f] = inbar[f];
    any = any || inbar[f] != nullptr;
  }
  if (!any || c.nblk3 == 0) return;
  hipLaunchKernelGGL(k_flux_advq<T>, dim3(c.nblk3), dim3(TPB), 0, c.stream, c.B, c.vert, c.prm, tab, obar, it);
}

template void launch_flux_advq<float>(const Ctx<float>&, const FieldTab<float>&, const COutTab<float>&,
                                      float* const*, uint32_t);
template void launch_flux_advq<double>(const Ctx<double>&, const FieldTab<double>&, const COutTab<double>&,
                                       double* const*, uint32_t);
template void launch_flux_advt1<float>(const Ctx<float>&, const FieldTab<float>&, const COutTab<float>&,
                                       float* const*, uint32_t);
template void launch_flux_advt1<double>(const Ctx<double>&, const FieldTab<double>&, const COutTab<double>&,
                                        double* const*, uint32_t);

}  // namespace pom
