// Self-check of the arithmetic identity the FP64 kernels rely on: xdiv(a,
// recip(b)) (face.h) returns the bits of the IEEE quotient a / b for every
// pair, taking the shared-reciprocal path inside its safe range and the
// division itself outside it.
#include <stdint.h>

#include "face.h"
#include "mgpu_internal.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    // splitmix64 finaliser (Steele, Lea and Flood, 2014)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// One operand from 64 random bits: sign, a significand that is random, all
// ones, all zeros or one ulp off either, and an exponent that is mostly in
// the range the transport meets, sometimes near the safe range's ends, and
// sometimes anywhere (zeros, subnormals, infinities and NaNs included).
__device__ __forceinline__ double operand(uint64_t h, uint64_t g) {
    uint64_t mant = h & ((1ull << 52) - 1);
    switch ((g >> 8) & 7) {
        case 0: mant = (1ull << 52) - 1; break;
        case 1: mant = 0; break;
        case 2: mant = 1; break;
        case 3: mant = (1ull << 52) - 2; break;
        default: break;
    }
    int e;
    const int cls = (int)(g & 15);
    if (cls < 10) e = 1023 + (int)((g >> 16) % 121) - 60;          // 2^-60 .. 2^60
    else if (cls < 13) e = 1023 + (int)((g >> 16) % 41) - 520 + ((g >> 24) & 1) * 1000;
    else e = (int)((g >> 16) % 2048);                                // anywhere
    const uint64_t bits = ((g >> 40) & 1ull) << 63 | ((uint64_t)e << 52) | mant;
    return __builtin_bit_cast(double, bits);
}

__global__ void xdiv_check_kernel(long n, uint64_t seed, unsigned long long *bad,
                                  unsigned long long *fast) {
    unsigned long long nb = 0, nf = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const uint64_t h0 = mix64(seed ^ (4 * (uint64_t)i + 0));
        const uint64_t h1 = mix64(seed ^ (4 * (uint64_t)i + 1));
        const uint64_t h2 = mix64(seed ^ (4 * (uint64_t)i + 2));
        const uint64_t h3 = mix64(seed ^ (4 * (uint64_t)i + 3));
        const double a = operand(h0, h1), b = operand(h2, h3);
        const Recip R = recip(b);
        const double q = xdiv(a, R);
        const double ref = a / b;
        const uint64_t qb = __builtin_bit_cast(uint64_t, q);
        const uint64_t rb = __builtin_bit_cast(uint64_t, ref);
        const bool same = qb == rb || (q != q && ref != ref);
        nb += same ? 0 : 1;
        const double m = fabs(a * R.r);
        nf += (R.ok && m >= 0x1p-400 && m < 0x1p400) ? 1 : 0;
    }
    if (nb) atomicAdd(bad, nb);
    if (nf) atomicAdd(fast, nf);
}

}  // namespace

extern "C" int mgpu_selfcheck_xdiv(int64_t n, uint64_t seed, uint64_t *mismatches,
                                   uint64_t *fast_path) {
    if (n < 0 || !mismatches || !fast_path) return MGPU_ERR_BADARG;
    unsigned long long *cnt = nullptr;
    HIPCHK(hipMalloc(&cnt, 2 * sizeof(unsigned long long)));
    hipError_t e = hipMemset(cnt, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(xdiv_check_kernel, dim3(2048), dim3(256), 0, nullptr, (long)n, seed,
                           cnt, cnt + 1);
        e = hipGetLastError();
    }
    unsigned long long h[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(h, cnt, sizeof(h), hipMemcpyDeviceToHost);
    const hipError_t ef = hipFree(cnt);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) { mgpu_set_error("mgpu_selfcheck_xdiv", e); return MGPU_ERR_HIP; }
    *mismatches = h[0];
    *fast_path = h[1];
    return MGPU_SUCCESS;
}
