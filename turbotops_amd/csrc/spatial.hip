// SPDX-License-Identifier: MIT
// spatial.hip — a spatial order of a point cloud: points sorted by a
// Morton-style code of their grid cell and copied in that order, so that
// nearby positions hold nearby points. Used by the searches that visit
// candidates near a query first (Boruvka's nearest-outside-component search,
// the exact kNN scan); their results never depend on the order.
#include <algorithm>

#include "frontend.hpp"

namespace ttg {

namespace {

constexpr u32 SB = 256;

// Order-preserving map of float bits to unsigned (finite values).
__device__ __forceinline__ u32 ord_bits(float f) {
    const u32 b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord_bits(u32 u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// ext[2k], ext[2k+1]: order bits of the smallest and largest finite
// coordinate k, reduced per wave and per workgroup before one atomic per
// workgroup and coordinate (a grid of at most 64 workgroups: atomics on one
// address serialise).
__global__ __launch_bounds__(SB) void sp_extent(const float* __restrict__ x, u32 n, u32 dim, u32 kdim,
                                                u32* __restrict__ ext) {
    __shared__ u32 slo[SB / 64], shi[SB / 64];
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (u32 k = 0; k < kdim; ++k) {
        u32 lo = 0xffffffffu, hi = 0u;
        for (u32 v = blockIdx.x * SB + threadIdx.x; v < n; v += gridDim.x * SB) {
            const float f = x[u64(v) * dim + k];
            if (__builtin_isfinite(f)) {
                lo = min(lo, ord_bits(f));
                hi = max(hi, ord_bits(f));
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, u32(__shfl_xor(int(lo), off)));
            hi = max(hi, u32(__shfl_xor(int(hi), off)));
        }
        if (lane == 0) {
            slo[w] = lo;
            shi[w] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (u32 q = 1; q < SB / 64; ++q) {
                lo = min(lo, slo[q]);
                hi = max(hi, shi[q]);
            }
            atomicMin(ext + 2 * k, lo);
            atomicMax(ext + 2 * k + 1, hi);
        }
        __syncthreads();
    }
}

__global__ void sp_init_extent(u32* __restrict__ ext, u32 kdim) {
    for (u32 t = threadIdx.x; t < 2 * kdim; t += blockDim.x) ext[t] = (t & 1) ? 0u : 0xffffffffu;
}

// keys[v] = (code << 32) | v: the code interleaves `bits` bits of the grid
// cell of each of the first kdim coordinates, most significant level first.
__global__ void sp_codes(const float* __restrict__ x, u32 n, u32 dim, u32 kdim, u32 bits,
                         const u32* __restrict__ ext, u64* __restrict__ keys) {
    const float cells = float(1u << bits);
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        u32 cell[32];
        for (u32 k = 0; k < kdim; ++k) {
            const float lo = unord_bits(ext[2 * k]), hi = unord_bits(ext[2 * k + 1]);
            const float f = x[u64(v) * dim + k];
            u32 c = 0;
            if (hi > lo && f > lo) {   // false for NaN and for an empty extent
                const float r = (f - lo) / (hi - lo) * cells;
                c = r >= cells ? (1u << bits) - 1u : u32(r);
            }
            cell[k] = c;
        }
        u64 code = 0;
        for (int b = int(bits) - 1; b >= 0; --b)
            for (u32 k = 0; k < kdim; ++k) code = (code << 1) | ((cell[k] >> b) & 1u);
        keys[v] = (code << 32) | v;
    }
}

__global__ void sp_gather(const float* __restrict__ x, u32 n, u32 dim, const u64* __restrict__ sorted,
                          float* __restrict__ xs, u32* __restrict__ perm) {
    const u64 total = u64(n) * dim;
    for (u64 t = blockIdx.x * u64(blockDim.x) + threadIdx.x; t < total; t += u64(gridDim.x) * blockDim.x) {
        const u64 pos = t / dim, k = t - pos * dim;
        const u32 v = u32(sorted[pos]);
        xs[t] = x[u64(v) * dim + k];
        if (perm && k == 0) perm[pos] = v;
    }
}

// One wave per box: the min and max of each coordinate over its members
// (points of a block, or the boxes of a super-block's blocks).
__global__ __launch_bounds__(SB) void sp_boxes(const float* __restrict__ src, u32 count, u32 per, u32 dim,
                                                int from_boxes, float* __restrict__ out) {
    const u32 lane = threadIdx.x & 63;
    const u32 nbox = (count + per - 1) / per;
    for (u32 b = (blockIdx.x * SB + threadIdx.x) / 64; b < nbox; b += gridDim.x * (SB / 64)) {
        const u32 m = min(per, count - b * per);
        const u32 e = b * per + (lane < m ? lane : 0u);
        for (u32 k = 0; k < dim; ++k) {
            float lo, hi;
            if (from_boxes) {
                lo = src[u64(e) * 2 * dim + k];
                hi = src[u64(e) * 2 * dim + dim + k];
            } else {
                lo = hi = src[u64(e) * dim + k];
            }
            for (int off = 32; off > 0; off >>= 1) {
                lo = __builtin_fminf(lo, __shfl_xor(lo, off));
                hi = __builtin_fmaxf(hi, __shfl_xor(hi, off));
            }
            if (lane == 0) {
                out[u64(b) * 2 * dim + k] = lo;
                out[u64(b) * 2 * dim + dim + k] = hi;
            }
        }
    }
}

u32 grid_of(u64 count) { return u32(std::max<u64>(1, std::min<u64>((count + SB - 1) / SB, 16384))); }

}  // namespace

u64 spatial_ext_entries(u32 dim) { return 2 * u64(std::min<u32>(std::max<u32>(dim, 1), 32)); }

void spatial_boxes(const float* src, u32 count, u32 per, u32 dim, bool from_boxes, float* out, hipStream_t s) {
    if (!count) return;
    if (per == 0 || per > 64) fail(TT_EINVAL, "spatial_boxes: groups of 1 to 64 members (one wave each)");
    const u64 nbox = (u64(count) + per - 1) / per;
    sp_boxes<<<grid_of(nbox * 64), SB, 0, s>>>(src, count, per, dim, from_boxes ? 1 : 0, out);
    TTG_LAUNCH_CHECK();
}

void spatial_order(const float* d_x, u32 n, u32 dim, const u32* d_n, u64* keys, u64* alt, u64* sort_scratch,
                   u32* ext, float* xs, u32* perm, hipStream_t s) {
    if (!n) return;
    if (dim == 0) fail(TT_EINVAL, "spatial_order: dim must be >= 1");
    // grid cells over the extent of the first kdim coordinates, `bits` bits
    // each (code < 2^32), sorted stably (ties keep index order)
    const u32 kdim = std::min<u32>(dim, 32);
    const u32 bits = std::max<u32>(1, std::min<u32>(10, 32 / kdim));
    sp_init_extent<<<1, 64, 0, s>>>(ext, kdim);
    sp_extent<<<std::min<u32>(grid_of(n), 64), SB, 0, s>>>(d_x, n, dim, kdim, ext);
    sp_codes<<<grid_of(n), SB, 0, s>>>(d_x, n, dim, kdim, bits, ext, keys);
    TTG_LAUNCH_CHECK();
    const u64* order = radix_sort_u64_counted(keys, n, d_n, 32, 32 + int(bits * kdim), alt, sort_scratch, s);
    sp_gather<<<grid_of(u64(n) * dim), SB, 0, s>>>(d_x, n, dim, order, xs, perm);
    TTG_LAUNCH_CHECK();
}

}  // namespace ttg
