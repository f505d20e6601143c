// Device self-checks of arithmetic the kernels rely on for bit-exact results.
#include "rsh_internal.hpp"
#include "rsh_kernels.hpp"

namespace rsh {
namespace {

// rcp_rn against the IEEE quotient 1.0f / x for every float whose exponent
// field lies in rcp_rn's estimate range (2..252): 2^32 bit patterns, a
// grid-stride walk; counts[0] tested inputs, counts[1] mismatches.
__global__ void __launch_bounds__(256) k_check_rcp(unsigned long long* counts) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long tested = 0, bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
        const uint32_t bits = (uint32_t)i;
        const uint32_t ex = (bits >> 23) & 0xFFu;
        if (ex < 2u || ex > 252u) continue;
        const float x = __uint_as_float(bits);
        tested++;
        if (__float_as_uint(rcp_rn(x)) != __float_as_uint(1.0f / x)) bad++;
    }
    for (int d = 32; d > 0; d >>= 1) {
        tested += __shfl_xor(tested, d, 64);
        bad += __shfl_xor(bad, d, 64);
    }
    if ((threadIdx.x & 63u) == 0u) {
        atomicAdd(&counts[0], tested);
        atomicAdd(&counts[1], bad);
    }
}

}  // namespace
}  // namespace rsh

extern "C" int rsh_selfcheck_rcp(int device, uint64_t* tested, uint64_t* mismatches) {
    using namespace rsh;
    if (!tested || !mismatches) return (int)hipErrorInvalidValue;
    hipError_t err = hipSetDevice(device);
    if (err != hipSuccess) return (int)err;
    unsigned long long* d = nullptr;
    err = hipMalloc(&d, 2 * sizeof(unsigned long long));
    if (err != hipSuccess) return (int)err;
    unsigned long long h[2] = {0, 0};
    err = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        k_check_rcp<<<4096, 256>>>(d);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *tested = h[0];
    *mismatches = h[1];
    return (int)err;
}
