// Adjoint gathers of OpExtElfRaw (float and double).
#include "pom_gather_impl.h"

POM_INSTANTIATE_GATHER(OpExtElfRaw)
