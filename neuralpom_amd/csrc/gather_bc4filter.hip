// Adjoint gathers of OpBc4Filter (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpBc4Filter)
