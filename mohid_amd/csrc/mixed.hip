// Mixed-precision transport (placeholder until the FP32 sweep lands).
#include "mgpu_internal.h"

int launch_transport_mixed(const Dims &, const StepScalars &, const SharedFields &, int, long,
                           const double *, double *, double *, double *, unsigned long long *,
                           int, int, int, hipStream_t, int) {
    return MGPU_ERR_UNSUPPORTED;
}
