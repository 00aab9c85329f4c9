// Launchers of the flux-form advt1 and advq adjoints (pom_flux.h).
#include "pom_flux.h"

namespace pom {
namespace {

// the fields that a launch's tracers share (their cotangents are written once)
template <class Op> constexpr bool shared_field(int f) {
  return f == Op::U || f == Op::V || f == Op::W || f == Op::AAM || f == Op::DT || f == Op::ETB || f == Op::ETF;
}

template <class T, class Op, int NT>
FluxSet<T, NT> flux_set(const FieldTab<T>* tab, const COutTab<T>* obar, T* const* const* inbar, uint32_t assign,
                        int first) {
  FluxSet<T, NT> fs{};
  for (int t = 0; t < NT; ++t) {
    fs.tab[t] = tab[first + t];
    fs.ob[t] = obar[first + t].o[0];
    fs.ib[t].assign = assign;
    for (int f = 0; f < Op::NIN; ++f) fs.ib[t].o[f] = inbar[first + t][f];
  }
  return fs;
}

template <class T, class Op>
bool any_input(T* const* inbar) {
  for (int f = 0; f < Op::NIN; ++f)
    if (inbar[f]) return true;
  return false;
}

// one launch for both tracers when they name the same shared arrays, else one per tracer
template <class T, class Op, class K1, class K2>
void launch_pair(const Ctx<T>& c, const FieldTab<T>* tab, const COutTab<T>* obar, T* const* const* inbar,
                 uint32_t assign, K1 k1, K2 k2) {
  if (c.nblk3 == 0) return;
  bool same = true;
  for (int f = 0; f < Op::NIN; ++f)
    if (shared_field<Op>(f) && inbar[0][f] != inbar[1][f]) same = false;
  if (same) {
    if (!any_input<T, Op>(inbar[0]) && !any_input<T, Op>(inbar[1])) return;
    hipLaunchKernelGGL(k2, dim3(c.nblk3), dim3(TPB), 0, c.stream, c.B, c.vert, c.prm,
                       flux_set<T, Op, 2>(tab, obar, inbar, assign, 0));
    return;
  }
  // the second launch adds to the arrays that it has in common with the first
  uint32_t common = 0u;
  for (int f = 0; f < Op::NIN; ++f)
    if (inbar[0][f] && inbar[0][f] == inbar[1][f]) common |= 1u << f;
  for (int t = 0; t < 2; ++t) {
    if (!any_input<T, Op>(inbar[t])) continue;
    const uint32_t as = t == 1 && any_input<T, Op>(inbar[0]) ? assign & ~common : assign;
    hipLaunchKernelGGL(k1, dim3(c.nblk3), dim3(TPB), 0, c.stream, c.B, c.vert, c.prm,
                       flux_set<T, Op, 1>(tab, obar, inbar, as, t));
  }
}

}  // namespace

template <class T>
void launch_flux_advt1(const Ctx<T>& c, const FieldTab<T>* tab, const COutTab<T>* obar, T* const* const* inbar,
                       uint32_t assign) {
  launch_pair<T, OpAdvt1>(c, tab, obar, inbar, assign, k_flux_advt1<T, 1>, k_flux_advt1<T, 2>);
}

template <class T>
void launch_flux_advq(const Ctx<T>& c, const FieldTab<T>* tab, const COutTab<T>* obar, T* const* const* inbar,
                      uint32_t assign) {
  launch_pair<T, OpAdvq>(c, tab, obar, inbar, assign, k_flux_advq<T, 1>, k_flux_advq<T, 2>);
}

template void launch_flux_advq<float>(const Ctx<float>&, const FieldTab<float>*, const COutTab<float>*,
                                      float* const* const*, uint32_t);
template void launch_flux_advq<double>(const Ctx<double>&, const FieldTab<double>*, const COutTab<double>*,
                                       double* const* const*, uint32_t);
template void launch_flux_advt1<float>(const Ctx<float>&, const FieldTab<float>*, const COutTab<float>*,
                                       float* const* const*, uint32_t);
template void launch_flux_advt1<double>(const Ctx<double>&, const FieldTab<double>*, const COutTab<double>*,
                                        double* const* const*, uint32_t);

}  // namespace pom
