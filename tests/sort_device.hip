// SPDX-License-Identifier: MIT
// sort_device.hip — TEST INFRASTRUCTURE ONLY: the library's radix sort
// (turbotops_amd/csrc/sort.hip, compiled into this test library) behind a
// small C ABI, so tests/test_sort_gpu.py can check it directly against a
// stable sort on the host. Not part of the product.
#include "sort.hip"

using namespace ttg;

extern "C" {

// Stable sort of n keys (and optional 32-bit payloads) by bits [lo, hi):
// counted = 0 runs radix_sort_u64; counted = 1 runs radix_sort_u64_counted
// on a buffer of `capacity` >= n keys whose count n is read on the device
// (payloads are not supported there). Returns 0 on success.
int sd_sort(const uint64_t* keys, const uint32_t* vals, uint64_t n, int lo, int hi, int counted,
            uint64_t capacity, uint64_t* out_keys, uint32_t* out_vals) {
    try {
        Workspace ws;
        hipStream_t s = nullptr;
        if (counted) {
            if (vals || capacity < n || capacity >= (uint64_t(1) << 32)) return 10;
            u64* k = ws.get<u64>(std::max<u64>(capacity, 1));
            u64* alt = ws.get<u64>(std::max<u64>(capacity, 1));
            u64* scratch = ws.get<u64>(radix_scratch_entries(std::max<u64>(capacity, 1)));
            u32* dc = ws.get<u32>(1);
            const u32 c32 = u32(n);
            if (n) TTG_HIP(hipMemcpy(k, keys, n * 8, hipMemcpyHostToDevice));
            TTG_HIP(hipMemcpy(dc, &c32, 4, hipMemcpyHostToDevice));
            u64* res = radix_sort_u64_counted(k, capacity, dc, lo, hi, alt, scratch, s);
            TTG_HIP(hipStreamSynchronize(s));
            if (n) TTG_HIP(hipMemcpy(out_keys, res, n * 8, hipMemcpyDeviceToHost));
            return 0;
        }
        u64* k = ws.get<u64>(std::max<u64>(n, 1));
        u64* alt = ws.get<u64>(std::max<u64>(n, 1));
        u32* v = vals ? ws.get<u32>(std::max<u64>(n, 1)) : nullptr;
        u32* valt = vals ? ws.get<u32>(std::max<u64>(n, 1)) : nullptr;
        if (n) TTG_HIP(hipMemcpy(k, keys, n * 8, hipMemcpyHostToDevice));
        if (n && vals) TTG_HIP(hipMemcpy(v, vals, n * 4, hipMemcpyHostToDevice));
        radix_sort_u64(k, v, n, lo, hi, alt, valt, ws, s);
        TTG_HIP(hipStreamSynchronize(s));
        if (n) TTG_HIP(hipMemcpy(out_keys, k, n * 8, hipMemcpyDeviceToHost));
        if (n && vals) TTG_HIP(hipMemcpy(out_vals, v, n * 4, hipMemcpyDeviceToHost));
        return 0;
    } catch (...) {
        return 1;
    }
}

}  // extern "C"
