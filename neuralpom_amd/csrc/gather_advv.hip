// Adjoint gathers of OpAdvv (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpAdvv)
