// Adjoint gathers of OpAdvq (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpAdvq)
