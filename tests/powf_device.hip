// SPDX-License-Identifier: MIT
// powf_device.hip — TEST INFRASTRUCTURE ONLY: the device build of the Lp
// kernels' powers (turbotops_amd/csrc/powf.hpp) swept over every float of a
// range on the GPU, against the double-precision power of the ROCm device
// library rounded once to float. Not part of the product.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "powf.hpp"

namespace {

// lo..hi: float bit patterns; half: n for pow_half_f32 (x^(n + 1/2)), else 0
// with exponent p for pow_f32. out[0] += mismatches, out[1] = max ulp.
__global__ void sweep(uint32_t lo, uint32_t hi, float p, unsigned half, unsigned long long* out) {
    unsigned long long mism = 0;
    unsigned maxulp = 0;
    for (uint64_t t = lo + blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; t <= hi;
         t += uint64_t(gridDim.x) * blockDim.x) {
        const float x = __uint_as_float(uint32_t(t));
        const float got = half ? ttg::pow_half_f32(x, half) : ttg::pow_f32(x, p);
        const float want = float(pow(double(x), half ? double(half) + 0.5 : double(p)));
        const int d = int(__float_as_uint(got)) - int(__float_as_uint(want));
        const unsigned ad = unsigned(d < 0 ? -d : d);
        mism += ad != 0;
        maxulp = ad > maxulp ? ad : maxulp;
    }
    atomicAdd(&out[0], mism);
    atomicMax(reinterpret_cast<unsigned*>(&out[1]), maxulp);
}

}  // namespace

extern "C" int pd_sweep(uint32_t lo, uint32_t hi, float p, unsigned half, unsigned long long* mismatches,
                        unsigned* max_ulp) {
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, 16) != hipSuccess) return 1;
    unsigned long long h[2] = {0, 0};
    if (hipMemcpy(d, h, 16, hipMemcpyHostToDevice) != hipSuccess) return 2;
    sweep<<<8192, 256>>>(lo, hi, p, half, d);
    if (hipGetLastError() != hipSuccess) return 3;
    if (hipMemcpy(h, d, 16, hipMemcpyDeviceToHost) != hipSuccess) return 4;
    (void)hipFree(d);
    *mismatches = h[0];
    *max_ulp = unsigned(h[1]);
    return 0;
}
