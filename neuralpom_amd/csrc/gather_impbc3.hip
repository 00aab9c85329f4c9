// Adjoint gathers of OpImpBc3 (float and double).
#include "pom_ops_imp.h"
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpImpBc3)
