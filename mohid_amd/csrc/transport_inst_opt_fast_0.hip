// Transport kernels instantiated for the optimised path, fast mode, chunks of 1-4 tracers.
#include "transport_kernels.h"

MGPU_DEFINE_CHUNK_SET(true, true, 0)
