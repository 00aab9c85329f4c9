// Adjoint gathers of the semi-implicit core's spatial Shapiro filters and of
// the unmasked bcond-4 + T/S filter that precedes the tracer one (float and double).
#include "pom_ops_imp.h"
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpShapiro<false>)
POM_INSTANTIATE_GATHER(OpShapiro<true>)
POM_INSTANTIATE_GATHER(OpImpBc4FilterG<false>)
