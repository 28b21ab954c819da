// SPDX-License-Identifier: MIT
// sort.hip — exclusive prefix sums and a stable LSD radix sort (64-bit keys,
// optional 32-bit payload) for gfx950.
//
// Radix sort, one pass per 8-bit digit, on a persistent grid (at most 2048
// workgroups, each owning a contiguous range of 4096-key tiles; the ranges
// are taken from the key count on the device, so a sort sized for a large
// capacity costs little when few keys are present):
//   upsweep    per-block digit histogram over its range (LDS atomics, one
//              counter row per wave64) -> hist[digit][block] (64-bit, so the
//              global offsets of the scan below may pass 2^32)
//   scan       exclusive scan of hist in digit-major order: the global start
//              of every (digit, block) run
//   downsweep  each block walks its tiles in order, ranks a tile's keys
//              stably — the lanes sharing a digit come from a per-wave LDS
//              match slot (or are the whole wave), mbcnt gives the rank among
//              them, per-wave running counters keep (t, lane) order and give
//              the tile's digit counts — reorders them in LDS, and appends
//              each digit's run at the block's running offset for that digit.
#include <algorithm>

#include "sort.hpp"

namespace ttg {

u64 scan_tmp_entries(u64 count) {
    constexpr u64 tile = 256 * 8;
    u64 t = 0;
    while (count > tile) {
        count = (count + tile - 1) / tile;
        t += count;
    }
    return t;
}

namespace {

constexpr u32 SB = 256;        // scan block
constexpr u32 SI = 8;          // items per thread
constexpr u32 STILE = SB * SI;

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v, u32 lane) {
#pragma unroll
    for (u32 off = 1; off < 64; off <<= 1) {
        const T o = __shfl_up(v, off);
        if (lane >= off) v += o;
    }
    return v;
}

// Exclusive scan of one value per thread across a 256-thread block; returns
// the exclusive prefix and writes the block total to *total.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* total) {
    __shared__ T wsum[SB / 64];
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T inc = wave_inclusive_scan(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    T before = 0, all = 0;
#pragma unroll
    for (u32 t = 0; t < SB / 64; ++t) {
        if (t < w) before += wsum[t];
        all += wsum[t];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

template <typename T>
__global__ __launch_bounds__(SB) void scan_reduce(const T* __restrict__ in, u64 count,
                                                  T* __restrict__ sums) {
    const u64 base = u64(blockIdx.x) * STILE;
    T s = 0;
#pragma unroll
    for (u32 t = 0; t < SI; ++t) {
        const u64 i = base + u64(t) * SB + threadIdx.x;
        if (i < count) s += in[i];
    }
    T total;
    (void)block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// in and out may alias: a block reads its whole tile before writing.
template <typename T>
__global__ __launch_bounds__(SB) void scan_apply(const T* in, T* out, u64 count,
                                                 const T* __restrict__ offsets,
                                                 T* __restrict__ d_total) {
    __shared__ T tile[STILE];
    const u64 base = u64(blockIdx.x) * STILE;
#pragma unroll
    for (u32 t = 0; t < SI; ++t) {
        const u64 i = base + u64(t) * SB + threadIdx.x;
        tile[t * SB + threadIdx.x] = i < count ? in[i] : T(0);
    }
    __syncthreads();
    T v[SI], s = 0;
#pragma unroll
    for (u32 t = 0; t < SI; ++t) {
        v[t] = tile[threadIdx.x * SI + t];
        s += v[t];
    }
    T total;
    T run = block_exclusive_scan(s, &total) + (offsets ? offsets[blockIdx.x] : T(0));
    __syncthreads();
#pragma unroll
    for (u32 t = 0; t < SI; ++t) {
        tile[threadIdx.x * SI + t] = run;
        run += v[t];
    }
    __syncthreads();
#pragma unroll
    for (u32 t = 0; t < SI; ++t) {
        const u64 i = base + u64(t) * SB + threadIdx.x;
        if (i < count) out[i] = tile[t * SB + threadIdx.x];
    }
    if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SB - 1)
        *d_total = tile[STILE - 1] + v[SI - 1];
}

template <typename T>
__global__ void set_scalar(T* p, T v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Exclusive scan with caller-provided temporaries (scan_tmp_entries(count)
// entries): no allocation, so it can be recorded into a graph.
template <typename T>
void exclusive_scan_buf(const T* in, T* out, u64 count, T* d_total, T* tmp, hipStream_t s) {
    if (count == 0) {
        if (d_total) set_scalar<T><<<1, 64, 0, s>>>(d_total, T(0));
        return;
    }
    const u64 blocks = (count + STILE - 1) / STILE;
    if (blocks == 1) {
        scan_apply<T><<<1, SB, 0, s>>>(in, out, count, nullptr, d_total);
        TTG_LAUNCH_CHECK();
        return;
    }
    T* sums = tmp;
    scan_reduce<T><<<u32(blocks), SB, 0, s>>>(in, count, sums);
    TTG_LAUNCH_CHECK();
    exclusive_scan_buf<T>(sums, sums, blocks, nullptr, tmp + blocks, s);
    scan_apply<T><<<u32(blocks), SB, 0, s>>>(in, out, count, sums, d_total);
    TTG_LAUNCH_CHECK();
}

template <typename T>
void exclusive_scan(const T* in, T* out, u64 count, T* d_total, Workspace& ws, hipStream_t s) {
    T* tmp = ws.get<T>(std::max<u64>(1, scan_tmp_entries(count)));
    exclusive_scan_buf<T>(in, out, count, d_total, tmp, s);
}

// ---- radix sort ---------------------------------------------------------
constexpr u32 RB = 256;             // threads (4 wave64s)
constexpr u32 RI = 16;              // keys per thread and tile
constexpr u32 RTILE = RB * RI;      // keys per tile
constexpr u32 RADIX = 256;
constexpr u32 WSEG = RTILE / (RB / 64);   // keys per wave segment of a tile
constexpr u32 MAX_SORT_BLOCKS = 2048;     // persistent workgroups (8 per CU)

__device__ __forceinline__ u32 lane_rank(u64 mask) {
    return __builtin_amdgcn_mbcnt_hi(u32(mask >> 32), __builtin_amdgcn_mbcnt_lo(u32(mask), 0u));
}

// The contiguous range of tiles block b of nb owns once `count` keys are
// known on the device: ranges follow block order, so every digit's keys of
// block b land in one contiguous run after those of blocks below b.
__device__ __forceinline__ void block_tiles(u64 count, u32 nb, u32 b, u64& t0, u64& t1) {
    const u64 tiles = (count + RTILE - 1) / RTILE;
    t0 = tiles * b / nb;
    t1 = tiles * (b + 1) / nb;
}

__global__ __launch_bounds__(RB) void rs_upsweep(const u64* __restrict__ keys, u64 count,
                                                 const u32* __restrict__ d_count, u32 shift,
                                                 u32 dmask, u64* __restrict__ hist) {
    __shared__ u32 h[RB / 64][RADIX];
    if (d_count) count = min<u64>(count, *d_count);
    u64 t0, t1;
    block_tiles(count, gridDim.x, blockIdx.x, t0, t1);
    for (u32 t = threadIdx.x; t < (RB / 64) * RADIX; t += RB) (&h[0][0])[t] = 0;
    __syncthreads();
    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (u64 tile = t0; tile < t1; ++tile) {
        const u64 base = tile * RTILE + u64(w) * WSEG;
#pragma unroll 4
        for (u32 t = 0; t < RI; ++t) {
            const u64 i = base + t * 64 + lane;
            if (i < count) atomicAdd(&h[w][u32(keys[i] >> shift) & dmask], 1u);
        }
    }
    __syncthreads();
    for (u32 d = threadIdx.x; d < RADIX; d += RB) {
        u32 c = 0;
#pragma unroll
        for (u32 v = 0; v < RB / 64; ++v) c += h[v][d];
        hist[u64(d) * gridDim.x + blockIdx.x] = c;
    }
}

// Each block walks its tiles in order and ranks a tile's keys stably in one
// pass: for each step of 64 keys a wave finds the lanes sharing each lane's
// digit (all of them when the digit is uniform across the wave, else by
// OR-ing lane bits into a per-wave LDS slot per digit and reading the slot
// back), takes the rank among them from mbcnt, and keeps a per-wave running
// count per digit. The counts, scanned over (digit, wave), give each key's
// place in the tile, which is reordered in LDS by digit; then every digit's
// run is written at the block's running offset for that digit, which
// advances. A block's writes to one digit are therefore one contiguous
// stream in time as in memory, and partial cache lines of one tile are
// completed by the next in the same L2. (Ranks taken from eight ballots per
// key, with a separate LDS-atomic histogram, cost 8.2 ms per 8-bit pass on
// 1.14e9 keys; this form 6.7 ms.)
template <bool VALS>
__global__ __launch_bounds__(RB) void rs_downsweep(const u64* __restrict__ keys,
                                                   const u32* __restrict__ vals, u64 count,
                                                   const u32* __restrict__ d_count,
                                                   u32 shift, u32 dmask, u32 dbits,
                                                   const u64* __restrict__ offs,
                                                   u64* __restrict__ keys_out,
                                                   u32* __restrict__ vals_out) {
    (void)dbits;
    __shared__ u32 wrun[RB / 64][RADIX];     // per wave: digit counts, then the wave's start per digit
    __shared__ u64 slot[RB / 64][RADIX];     // per wave: lanes holding each digit in one step
    __shared__ u32 bstart[RADIX];
    __shared__ u64 run[RADIX];      // this block's next output position per digit
    __shared__ u64 skeys[RTILE];
    __shared__ u32 svals[VALS ? RTILE : 1];

    const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (d_count) count = min<u64>(count, *d_count);
    u64 t0, t1;
    block_tiles(count, gridDim.x, blockIdx.x, t0, t1);
    if (t0 >= t1) return;   // uniform: the block owns no tile
    run[threadIdx.x] = offs[u64(threadIdx.x) * gridDim.x + blockIdx.x];   // RB == RADIX
    for (u32 t = threadIdx.x; t < (RB / 64) * RADIX; t += RB) (&slot[0][0])[t] = 0;

    for (u64 tile = t0; tile < t1; ++tile) {
        const u64 tile0 = tile * RTILE;
        for (u32 t = threadIdx.x; t < (RB / 64) * RADIX; t += RB) (&wrun[0][0])[t] = 0;
        const u64 base = tile0 + u64(w) * WSEG;
        u64 k[RI];
        u32 v[RI];
#pragma unroll
        for (u32 t = 0; t < RI; ++t) {
            const u64 i = base + t * 64 + lane;
            const bool ok = i < count;
            k[t] = ok ? keys[i] : 0;
            if constexpr (VALS) v[t] = ok ? vals[i] : 0;
        }
        __syncthreads();

        // rank within (wave, digit), in (step, lane) order
        u32 lr[RI];
#pragma unroll
        for (u32 t = 0; t < RI; ++t) {
            const u64 i = base + t * 64 + lane;
            const bool ok = i < count;
            const u32 d = u32(k[t] >> shift) & dmask;
            const u64 act = __ballot(ok);
            const u32 d0 = __builtin_amdgcn_readfirstlane(d);
            const bool uniform = __ballot(ok && d != d0) == 0;   // wave-uniform
            u64 peers;
            if (uniform) {
                peers = act;
            } else {
                if (ok) __hip_atomic_fetch_or(&slot[w][d], u64(1) << lane, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WAVEFRONT);
                __builtin_amdgcn_wave_barrier();
                peers = ok ? slot[w][d] : 0ull;
                __builtin_amdgcn_wave_barrier();
            }
            const u32 rank = lane_rank(peers);
            const u32 b0 = ok ? wrun[w][d] : 0u;
            lr[t] = b0 + rank;
            __builtin_amdgcn_wave_barrier();
            if (ok && rank == 0) {
                wrun[w][d] = b0 + u32(__popcll(peers));
                if (!uniform) slot[w][d] = 0;
            }
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();

        // tile-local digit starts, and each wave's start per digit
        u32 tot = 0;
        {
            const u32 d = threadIdx.x;
            u32 c[RB / 64];
#pragma unroll
            for (u32 q = 0; q < RB / 64; ++q) {
                c[q] = wrun[q][d];
                tot += c[q];
            }
            u32 all;
            const u32 start = block_exclusive_scan(tot, &all);
            bstart[d] = start;
            u32 r = start;
#pragma unroll
            for (u32 q = 0; q < RB / 64; ++q) {
                wrun[q][d] = r;
                r += c[q];
            }
        }
        __syncthreads();

#pragma unroll
        for (u32 t = 0; t < RI; ++t) {
            const u64 i = base + t * 64 + lane;
            if (i < count) {
                const u32 pos = wrun[w][u32(k[t] >> shift) & dmask] + lr[t];
                skeys[pos] = k[t];
                if constexpr (VALS) svals[pos] = v[t];
            }
        }
        __syncthreads();

        const u32 tile_count = u32(min<u64>(RTILE, count - tile0));
#pragma unroll 4
        for (u32 t = 0; t < RI; ++t) {
            const u32 q = t * RB + threadIdx.x;
            if (q < tile_count) {
                const u64 key = skeys[q];
                const u32 d = u32(key >> shift) & dmask;
                const u64 g = run[d] + (q - bstart[d]);
                keys_out[g] = key;
                if constexpr (VALS) vals_out[g] = svals[q];
            }
        }
        __syncthreads();
        run[threadIdx.x] += tot;   // this tile's keys of digit threadIdx.x
    }
}

u32 sort_blocks(u64 capacity) {
    return u32(std::max<u64>(1, std::min<u64>(MAX_SORT_BLOCKS, (capacity + RTILE - 1) / RTILE)));
}

}  // namespace

void exclusive_scan_u32_buf(const u32* in, u32* out, u64 count, u32* d_total, u32* tmp,
                            hipStream_t s) {
    exclusive_scan_buf<u32>(in, out, count, d_total, tmp, s);
}

void exclusive_scan_u32(const u32* in, u32* out, u64 count, u32* d_total, Workspace& ws,
                        hipStream_t s) {
    exclusive_scan<u32>(in, out, count, d_total, ws, s);
}

void exclusive_scan_u64(const u64* in, u64* out, u64 count, u64* d_total, Workspace& ws,
                        hipStream_t s) {
    exclusive_scan<u64>(in, out, count, d_total, ws, s);
}


u64* radix_sort_u64_counted(u64* keys, u64 capacity, const u32* d_count, int lo_bit, int hi_bit,
                            u64* keys_alt, u64* scratch, hipStream_t s) {
    if (capacity <= 1 || hi_bit <= lo_bit) return keys;
    if (capacity >= (u64(1) << 32)) fail(TT_ELENGTH, "radix_sort: a counted sort takes fewer than 2^32 keys");
    const u32 nb = sort_blocks(capacity);
    u64* hist = scratch;
    u64* scan_tmp = scratch + u64(RADIX) * nb;
    u64 *ks = keys, *kd = keys_alt;
    for (int shift = lo_bit; shift < hi_bit; shift += 8) {
        const u32 dbits = u32(std::min(8, hi_bit - shift));
        const u32 dmask = (1u << dbits) - 1u;
        rs_upsweep<<<nb, RB, 0, s>>>(ks, capacity, d_count, u32(shift), dmask, hist);
        TTG_LAUNCH_CHECK();
        exclusive_scan_buf<u64>(hist, hist, u64(RADIX) * nb, nullptr, scan_tmp, s);
        rs_downsweep<false><<<nb, RB, 0, s>>>(ks, nullptr, capacity, d_count, u32(shift), dmask, dbits, hist,
                                              kd, nullptr);
        TTG_LAUNCH_CHECK();
        std::swap(ks, kd);
    }
    return ks;
}

u64 radix_scratch_entries(u64 capacity) {
    const u64 h = u64(RADIX) * sort_blocks(capacity);
    return h + std::max<u64>(1, scan_tmp_entries(h));
}

void radix_sort_u64(u64* keys, u32* vals, u64 count, int lo_bit, int hi_bit, u64* keys_alt,
                    u32* vals_alt, Workspace& ws, hipStream_t s) {
    if (count <= 1 || hi_bit <= lo_bit) return;
    const u32 nb = sort_blocks(count);
    u64* hist = ws.get<u64>(u64(RADIX) * nb);
    u64* stmp = ws.get<u64>(std::max<u64>(1, scan_tmp_entries(u64(RADIX) * nb)));
    u64 *ks = keys, *kd = keys_alt;
    u32 *vs = vals, *vd = vals_alt;
    for (int shift = lo_bit; shift < hi_bit; shift += 8) {
        const u32 dbits = u32(std::min(8, hi_bit - shift));
        const u32 dmask = (1u << dbits) - 1u;
        rs_upsweep<<<nb, RB, 0, s>>>(ks, count, nullptr, u32(shift), dmask, hist);
        TTG_LAUNCH_CHECK();
        exclusive_scan_buf<u64>(hist, hist, u64(RADIX) * nb, nullptr, stmp, s);
        if (vals)
            rs_downsweep<true><<<nb, RB, 0, s>>>(ks, vs, count, nullptr, u32(shift), dmask, dbits, hist, kd, vd);
        else
            rs_downsweep<false><<<nb, RB, 0, s>>>(ks, nullptr, count, nullptr, u32(shift), dmask, dbits, hist, kd,
                                                  nullptr);
        TTG_LAUNCH_CHECK();
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    if (ks != keys) {
        TTG_HIP(hipMemcpyAsync(keys, ks, count * sizeof(u64), hipMemcpyDeviceToDevice, s));
        if (vals) TTG_HIP(hipMemcpyAsync(vals, vs, count * sizeof(u32), hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace ttg
