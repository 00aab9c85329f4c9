// Adjoint gathers of OpImpRad (float and double).
#include "pom_ops_imp.h"
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpImpRad)
