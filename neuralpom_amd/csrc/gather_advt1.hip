// Adjoint gathers of OpAdvt1 (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpAdvt1)
