// Adjoint gathers of OpBcond2 (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBcond2)
