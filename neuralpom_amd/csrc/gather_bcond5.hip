// Adjoint gathers of OpBcond5 (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBcond5)
