// Adjoint gathers of OpImpCorr (float and double).
#include "pom_ops_imp.h"
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpImpCorr)
