// SPDX-License-Identifier: MIT
// hostcopy.hip — large device-to-host copies into ordinary (pageable) host
// memory: the distance matrix and the edge lists handed to the host reducers.
//
// A copy into pageable memory goes through the runtime's own staging at a
// few GB/s. Here it runs through two pinned 64 MiB buffers per device: the
// DMA of chunk c + 1 overlaps the copy of chunk c out of the pinned buffer,
// which host threads split between them. The buffers are allocated on a
// device's first large copy and kept until the process ends; calls on one
// device take turns on them.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "frontend.hpp"

namespace ttg {
namespace {

constexpr u64 kChunk = 64ull << 20;            // bytes per pinned buffer
constexpr u64 kStagedFrom = 32ull << 20;       // smaller copies go straight through the runtime

struct Staging {
    std::mutex mu;
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
};

Staging& staging(int device) {
    static std::mutex m;
    static std::map<int, std::unique_ptr<Staging>> per_device;
    std::lock_guard<std::mutex> lk(m);
    auto& p = per_device[device];
    if (!p) p = std::make_unique<Staging>();
    return *p;
}

// memcpy split over up to 8 host threads (slices of whole 4 KiB pages)
void host_copy(char* dst, const char* src, u64 bytes) {
    const u32 hw = std::max(1u, std::thread::hardware_concurrency());
    const u32 t = bytes >= (8ull << 20) ? std::min(8u, hw) : 1u;
    if (t == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const u64 per = ((bytes + t - 1) / t + 4095) & ~u64(4095);
    std::vector<std::thread> pool;
    for (u32 k = 0; k < t; ++k) {
        const u64 off = u64(k) * per;
        if (off >= bytes) break;
        const u64 len = std::min(per, bytes - off);
        pool.emplace_back([=] { std::memcpy(dst + off, src + off, len); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

void copy_to_host(void* dst, const void* src, u64 bytes, hipStream_t s) {
    if (bytes == 0) return;
    if (bytes < kStagedFrom) {
        TTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
        return;
    }
    int device = 0;
    TTG_HIP(hipGetDevice(&device));
    Staging& st = staging(device);
    std::lock_guard<std::mutex> lk(st.mu);
    if (!st.buf[0]) {
        for (int k = 0; k < 2; ++k) {
            TTG_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.buf[k]), kChunk, hipHostMallocDefault));
            TTG_HIP(hipEventCreateWithFlags(&st.done[k], hipEventDisableTiming));
        }
    }
    const char* from = static_cast<const char*>(src);
    char* to = static_cast<char*>(dst);
    const u64 chunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](u64 c) {
        const u64 off = c * kChunk, len = std::min(kChunk, bytes - off);
        TTG_HIP(hipMemcpyAsync(st.buf[c & 1], from + off, len, hipMemcpyDeviceToHost, s));
        TTG_HIP(hipEventRecord(st.done[c & 1], s));
    };
    issue(0);
    if (chunks > 1) issue(1);
    for (u64 c = 0; c < chunks; ++c) {
        TTG_HIP(hipEventSynchronize(st.done[c & 1]));
        const u64 off = c * kChunk, len = std::min(kChunk, bytes - off);
        host_copy(to + off, st.buf[c & 1], len);
        if (c + 2 < chunks) issue(c + 2);   // buffer c & 1 is free again
    }
}

}  // namespace ttg
