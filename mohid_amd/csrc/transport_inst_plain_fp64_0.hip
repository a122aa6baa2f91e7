// Transport kernels instantiated for the general path, FP64, chunks of 1-4 tracers.
#include "transport_kernels.h"

MGPU_DEFINE_CHUNK_SET(false, false, 0)
