// Adjoint gathers of OpBc1 (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBc1)
