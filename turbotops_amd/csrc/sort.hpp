// SPDX-License-Identifier: MIT
// sort.hpp — device primitives shared by the front-end kernels: exclusive
// prefix sums and a stable LSD radix sort of 64-bit keys (optional 32-bit
// payload). All stream-ordered; temporaries come from a Workspace.
#pragma once

#include <vector>

#include "common.hpp"

namespace ttg {

// A grow-only bag of device allocations released together.
struct Workspace {
    std::vector<void*> ptrs;
    Workspace() = default;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() { release(); }
    // frees every allocation (the work that used them must have finished)
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
    template <typename T>
    T* get(u64 count) {
        void* p = nullptr;
        if (count) TTG_HIP(hipMalloc(&p, count * sizeof(T)));
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
};

// out[i] = sum_{t < i} in[i]; out may alias in. Returns nothing; the total
// is written to *d_total when non-null.
void exclusive_scan_u32(const u32* in, u32* out, u64 count, u32* d_total, Workspace& ws,
                        hipStream_t s);
void exclusive_scan_u64(const u64* in, u64* out, u64 count, u64* d_total, Workspace& ws,
                        hipStream_t s);

// Stable ascending sort of keys[0, count) by bits [lo_bit, hi_bit) carrying
// vals (may be null). Uses keys_alt / vals_alt as ping-pong buffers; the
// sorted data ends up in keys / vals. Any count.
void radix_sort_u64(u64* keys, u32* vals, u64 count, int lo_bit, int hi_bit, u64* keys_alt,
                    u32* vals_alt, Workspace& ws, hipStream_t s);

// Capturable forms (no allocation; fixed grids; counts read on the device):
// temporaries of scan_tmp_entries(count) entries for the scan, and for the
// sort radix_scratch_entries(capacity) 64-bit entries (the digit histograms
// and their scan). The sort orders keys[0, *d_count) (count <= capacity <
// 2^32) by bits [lo_bit, hi_bit) and returns the buffer that holds the
// result: keys after an even number of 8-bit passes, keys_alt after an odd
// one (no copy back).
u64 scan_tmp_entries(u64 count);
u64 radix_scratch_entries(u64 capacity);
void exclusive_scan_u32_buf(const u32* in, u32* out, u64 count, u32* d_total, u32* tmp,
                            hipStream_t s);
u64* radix_sort_u64_counted(u64* keys, u64 capacity, const u32* d_count, int lo_bit, int hi_bit,
                            u64* keys_alt, u64* scratch, hipStream_t s);

// Bits needed to represent v (0 for v == 0).
inline int bit_length(u64 v) { return v ? 64 - __builtin_clzll(v) : 0; }

}  // namespace ttg
