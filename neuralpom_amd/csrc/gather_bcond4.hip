// Adjoint gathers of OpBcond4 (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBcond4)
