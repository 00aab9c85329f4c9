// Adjoint gathers of the semi-implicit closed-basin bcond operators (float and double).
#include "pom_ops_imp.h"
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpImpBcond<1>)
POM_INSTANTIATE_GATHER(OpImpBcond<2>)
POM_INSTANTIATE_GATHER(OpImpBcond<4>)
POM_INSTANTIATE_GATHER(OpImpBcond<5>)
POM_INSTANTIATE_GATHER(OpImpBcond<6>)
