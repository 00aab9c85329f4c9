// Adjoint gathers of OpExtUV (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpExtUV)
