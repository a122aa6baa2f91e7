// Transport kernels instantiated for the general path, fast mode, chunks of 5-8 tracers.
#include "transport_kernels.h"

MGPU_DEFINE_CHUNK_SET(false, true, 1)
