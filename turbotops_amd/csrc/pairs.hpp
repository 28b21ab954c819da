// SPDX-License-Identifier: MIT
// pairs.hpp — per-pair Minkowski evaluators shared by every kernel that
// measures distances (matrix tiles, radius graph, kNN, Boruvka). Each is the
// reference's scalar kernel (include/turbotops/metric.hpp:92-207, NO_SIMD
// branches) operation by operation; the library builds with
// -ffp-contract=off, so the only fused multiply-adds are explicit.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace ttg {

constexpr u32 ROWS = 16;      // rows per wave in the tile kernels
constexpr u32 KC = 16;        // coordinates per register chunk (runtime dim)

struct MetricArgs {
    float p;       // Minkowski exponent
    float inv_p;   // 1.0f / p in float, as the reference computes it (metric.hpp:206)
    u32 ip;        // integral exponent for square-and-multiply; n for p = n + 1/2
    u32 half;      // p = n + 1/2, 1 <= n <= 16: terms as x^n sqrt(x) (lp_term)
};

// One Lp term x^p, x in [0, 1] (the reference's powf(x, p), metric.hpp:204):
// for half-integral p (ma.half, uniform across the grid) pow_half_f32, as
// accurate as pow_f32 in a third of its instructions.
__device__ __forceinline__ float lp_term(float x, const MetricArgs& ma) {
    return ma.half ? pow_half_f32(x, ma.ip) : pow_f32(x, ma.p);
}

// ---- per-pair kernels, compile-time dimension -----------------------------
// a: row point (wave-uniform, scalar loads); b: column point in VGPRs.
// KEY = true returns, for L2, the sum before the square root (the ordering
// key of pdist<..., KEY>); the other metrics return the distance either way.
template <int V, int DIM, bool KEY = false>
__device__ __forceinline__ float pair_dist(const float* __restrict__ a, const float (&b)[DIM],
                                           const MetricArgs& ma) {
    if constexpr (V == MV_L2) {                       // metric.hpp:107-111
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            const float d = a[k] - b[k];
            s = __builtin_fmaf(d, d, s);
        }
        if constexpr (KEY) return s;
        else return __builtin_sqrtf(s);
    } else if constexpr (V == MV_L1) {                // metric.hpp:157
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < DIM; ++k) s += __builtin_fabsf(a[k] - b[k]);
        return s;
    } else {
        float m = 0.0f;                               // metric.hpp:172
#pragma unroll
        for (int k = 0; k < DIM; ++k) m = fmax_acc(m, __builtin_fabsf(a[k] - b[k]));
        if constexpr (V == MV_LINF) {
            return m;
        } else {                                      // metric.hpp:183-207
            if (m == 0.0f) return 0.0f;
            const float inv = 1.0f / m;
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                const float v = __builtin_fabsf(a[k] - b[k]) * inv;
                const float x = (v < 1.0f) ? v : 1.0f;
                if constexpr (V == MV_LP_INT) s += ipow_ref(x, ma.ip);
                else s += lp_term(x, ma);
            }
            return m * pow_f32(s, ma.inv_p);
        }
    }
}

// ---- per-pair kernels, runtime dimension (chunks of KC coordinates) -------
// The chunking keeps each pair's accumulation strictly sequential in k.
template <int V, int R>
__device__ __forceinline__ void rows_runtime(const float* __restrict__ x, u32 n, u32 dim,
                                             const MetricArgs& ma, u32 ibase, u32 jc,
                                             float (&val)[R]) {
    const float* __restrict__ xb = x + u64(jc) * dim;
    float acc[R];
#pragma unroll
    for (u32 r = 0; r < R; ++r) acc[r] = 0.0f;
    // pass 1: L1 / L2 sums, or the L-inf maximum (also Lp's M)
    for (u32 k0 = 0; k0 < dim; k0 += KC) {
        const u32 len = min(KC, dim - k0);
        float b[KC];
#pragma unroll
        for (u32 t = 0; t < KC; ++t) b[t] = t < len ? xb[k0 + t] : 0.0f;
#pragma unroll
        for (u32 r = 0; r < R; ++r) {
            const u32 i = min(ibase + r, n - 1);
            const float* __restrict__ a = x + u64(i) * dim + k0;
            float s = acc[r];
#pragma unroll
            for (u32 t = 0; t < KC; ++t) {
                if (t < len) {   // uniform guard: no dynamic register indexing
                    const float d = a[t] - b[t];
                    if constexpr (V == MV_L2) s = __builtin_fmaf(d, d, s);
                    else if constexpr (V == MV_L1) s += __builtin_fabsf(d);
                    else s = fmax_acc(s, __builtin_fabsf(d));
                }
            }
            acc[r] = s;
        }
    }
    if constexpr (V == MV_L2) {
#pragma unroll
        for (u32 r = 0; r < R; ++r) val[r] = __builtin_sqrtf(acc[r]);
    } else if constexpr (V == MV_L1 || V == MV_LINF) {
#pragma unroll
        for (u32 r = 0; r < R; ++r) val[r] = acc[r];
    } else {
        // pass 2 (Lp): sum of normalised powers, then M * s^(1/p)
        float inv[R], s2[R];
#pragma unroll
        for (u32 r = 0; r < R; ++r) {
            inv[r] = acc[r] == 0.0f ? 0.0f : 1.0f / acc[r];
            s2[r] = 0.0f;
        }
        for (u32 k0 = 0; k0 < dim; k0 += KC) {
            const u32 len = min(KC, dim - k0);
            float b[KC];
#pragma unroll
            for (u32 t = 0; t < KC; ++t) b[t] = t < len ? xb[k0 + t] : 0.0f;
#pragma unroll
            for (u32 r = 0; r < R; ++r) {
                const u32 i = min(ibase + r, n - 1);
                const float* __restrict__ a = x + u64(i) * dim + k0;
                float s = s2[r];
#pragma unroll
                for (u32 t = 0; t < KC; ++t) {
                    if (t < len) {
                        const float v = __builtin_fabsf(a[t] - b[t]) * inv[r];
                        const float xx = (v < 1.0f) ? v : 1.0f;
                        if constexpr (V == MV_LP_INT) s += ipow_ref(xx, ma.ip);
                        else s += lp_term(xx, ma);
                    }
                }
                s2[r] = s;
            }
        }
#pragma unroll
        for (u32 r = 0; r < R; ++r)
            val[r] = acc[r] == 0.0f ? 0.0f : acc[r] * pow_f32(s2[r], ma.inv_p);
    }
}

// L1, L2 and L-inf rows from DIM >= 8 (DIM a multiple of 4), four coordinates
// at a time: the chunk loop is kept rolled, so only one chunk of the R row
// points (4R scalars) is live in SGPRs. Fully unrolled, R x DIM row scalars
// exceed the 106 SGPRs and spill through v_writelane / v_readlane, which
// nearly doubles the VALU work per pair at DIM = 16. The column chunk is
// re-read per step (64 lanes x 16 B, cache hits). Each pair still accumulates
// strictly in k order.
template <int V, int DIM, int R, bool KEY, bool CLAMP>
__device__ __forceinline__ void rows_chunked_run(const float* __restrict__ x, u32 n, u32 ibase,
                                                 u32 jc, float (&acc)[R]) {
    constexpr u32 C = 4;
    const float* __restrict__ xb = x + u64(jc) * DIM;
#pragma unroll 1
    for (u32 k0 = 0; k0 < u32(DIM); k0 += C) {
        float b[C];
#pragma unroll
        for (u32 t = 0; t < C; ++t) b[t] = xb[k0 + t];
#pragma unroll
        for (u32 r = 0; r < R; ++r) {
            const u32 i = CLAMP ? min(ibase + r, n - 1) : ibase + r;
            const float* __restrict__ a = x + u64(i) * DIM + k0;
            float s = acc[r];
#pragma unroll
            for (u32 t = 0; t < C; ++t) {
                const float d = a[t] - b[t];
                if constexpr (V == MV_L2) s = __builtin_fmaf(d, d, s);
                else if constexpr (V == MV_L1) s += __builtin_fabsf(d);
                else s = fmax_acc(s, __builtin_fabsf(d));
            }
            acc[r] = s;
        }
    }
}

template <int V, int DIM, int R, bool KEY>
__device__ __forceinline__ void rows_chunked(const float* __restrict__ x, u32 n, u32 ibase, u32 jc,
                                             float (&val)[R]) {
    float acc[R];
#pragma unroll
    for (u32 r = 0; r < R; ++r) acc[r] = 0.0f;
    // rows past n (the last row block) read row n - 1; their values are dropped
    if (ibase + R <= n) rows_chunked_run<V, DIM, R, KEY, false>(x, n, ibase, jc, acc);
    else rows_chunked_run<V, DIM, R, KEY, true>(x, n, ibase, jc, acc);
#pragma unroll
    for (u32 r = 0; r < R; ++r) {
        if constexpr (V == MV_L2 && !KEY) val[r] = __builtin_sqrtf(acc[r]);
        else val[r] = acc[r];
    }
}

template <int V, int DIM, int R, bool KEY = false>
__device__ __forceinline__ void rows_static(const float* __restrict__ x, u32 n,
                                            const MetricArgs& ma, u32 ibase, u32 jc,
                                            float (&val)[R]) {
    if constexpr (DIM >= 8 && DIM % 4 == 0 && (V == MV_L1 || V == MV_L2 || V == MV_LINF)) {
        rows_chunked<V, DIM, R, KEY>(x, n, ibase, jc, val);
        return;
    }
    float b[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) b[k] = x[u64(jc) * DIM + k];
#pragma unroll
    for (u32 r = 0; r < R; ++r) {
        const u32 i = min(ibase + r, n - 1);
        val[r] = pair_dist<V, DIM, KEY>(x + u64(i) * DIM, b, ma);
    }
}


// ---- one pair from two pointers (global or LDS), any dimension --------------
// DIM > 0: compile-time dimension; DIM == 0: runtime `dim`. Sequential in k
// exactly as the reference. KEY = true returns the ordering key instead of
// the distance: for L2 the fma sum before the square root (sqrtf is monotone,
// so key order refines distance order); the distance itself for the others.
template <int V, int DIM, bool KEY = false>
__device__ __forceinline__ float pdist(const float* __restrict__ a, const float* __restrict__ b,
                                       u32 dim_rt, const MetricArgs& ma) {
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    if constexpr (V == MV_L2) {
        float s = 0.0f;
#pragma unroll
        for (u32 k = 0; k < dim; ++k) {
            const float d = a[k] - b[k];
            s = __builtin_fmaf(d, d, s);
        }
        if constexpr (KEY) return s;
        else return __builtin_sqrtf(s);
    } else if constexpr (V == MV_L1) {
        float s = 0.0f;
#pragma unroll
        for (u32 k = 0; k < dim; ++k) s += __builtin_fabsf(a[k] - b[k]);
        return s;
    } else {
        float m = 0.0f;
#pragma unroll
        for (u32 k = 0; k < dim; ++k) m = fmax_acc(m, __builtin_fabsf(a[k] - b[k]));
        if constexpr (V == MV_LINF) {
            return m;
        } else {
            if (m == 0.0f) return 0.0f;
            const float inv = 1.0f / m;
            float s = 0.0f;
#pragma unroll
            for (u32 k = 0; k < dim; ++k) {
                const float v = __builtin_fabsf(a[k] - b[k]) * inv;
                const float x = (v < 1.0f) ? v : 1.0f;
                if constexpr (V == MV_LP_INT) s += ipow_ref(x, ma.ip);
                else s += lp_term(x, ma);
            }
            return m * pow_f32(s, ma.inv_p);
        }
    }
}

// A key above which the distance is certainly above d: for L2 the keys are
// sums before the square root, and any s > (d^2 (1 + 4e-7)) rounded up has
// sqrtf(s) >= d + 1 ulp (d (1 + 2e-7) exceeds d by at least 1.68 ulp, more
// than the half ulp of rounding). For the other metrics the key is the
// distance itself. +inf stays +inf. Used to skip the exact square root for
// candidates that cannot pass a "<= d" test.
template <int V>
__device__ __forceinline__ float key_reject_above(float d) {
    if constexpr (V == MV_L2) {
        if (!(d < INFINITY)) return INFINITY;
        const float b = float(double(d) * double(d) * (1.0 + 4e-7));   // b >= 0
        return b < INFINITY ? __uint_as_float(__float_as_uint(b) + 1u) : INFINITY;   // next float up
    } else {
        return d;
    }
}

// Distance from an ordering key (see pdist<..., KEY = true>).
template <int V>
__device__ __forceinline__ float key_to_dist(float key) {
    if constexpr (V == MV_L2) return __builtin_sqrtf(key);
    else return key;
}

// Dispatches a functor templated on <V, DIM> over the metric variant and the
// specialised dimensions {1, 2, 3, 4, 8, 16}; other dimensions run DIM = 0.
template <template <int, int> class F, typename... Args>
void dispatch_metric_dim(MetricVariant v, u32 dim, Args&&... args) {
    auto by_dim = [&](auto vtag) {
        constexpr int VV = decltype(vtag)::value;
        switch (dim) {
            case 1: F<VV, 1>::run(args...); break;
            case 2: F<VV, 2>::run(args...); break;
            case 3: F<VV, 3>::run(args...); break;
            case 4: F<VV, 4>::run(args...); break;
            case 8: F<VV, 8>::run(args...); break;
            case 16: F<VV, 16>::run(args...); break;
            default: F<VV, 0>::run(args...); break;
        }
    };
    switch (v) {
        case MV_L1: by_dim(std::integral_constant<int, MV_L1>{}); break;
        case MV_L2: by_dim(std::integral_constant<int, MV_L2>{}); break;
        case MV_LINF: by_dim(std::integral_constant<int, MV_LINF>{}); break;
        case MV_LP_INT: by_dim(std::integral_constant<int, MV_LP_INT>{}); break;
        default: by_dim(std::integral_constant<int, MV_LP_POW>{}); break;
    }
}

MetricArgs metric_args(int kind, float p);

}  // namespace ttg
