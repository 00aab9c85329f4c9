// Adjoint gathers of OpBc6Filter (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBc6Filter)
