// Adjoint gathers of OpBc4FilterExt (float and double).
#include "pom_ops_imp.h"
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBc4FilterExt)
