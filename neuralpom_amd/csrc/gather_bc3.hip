// Adjoint gathers of OpBc3 (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBc3)
