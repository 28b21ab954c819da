// SPDX-License-Identifier: MIT
// grad.hip — gradient of a persistence diagram with respect to the points.
//
// The reference computes diagrams only (rips.hpp:52-67 keeps bar values,
// not simplices); this is the backward of the differentiable form. Every
// finite bar value of a Rips diagram is the length of one edge — the longest
// edge (the "critical edge") of its birth or death simplex — so dL/dx is a
// scatter of dL/d(value) times the derivative of the Minkowski distance of
// that edge:
//   L2     d/dx_a = (x_a - x_b) / d          (0 when d == 0)
//   L1     d/dx_a = sign(x_a - x_b)
//   L-inf  d/dx_a = sign(x_a - x_b) on the first coordinate of largest |diff|
//   Lp     d/dx_a = sign(diff) |diff|^(p-1) / d^(p-1)
// and d/dx_b = -d/dx_a.
//
// Deterministic: each (edge, endpoint) contribution gets a key
// (point, contribution id); one radix sort groups them by point in a fixed
// order, and one lane per point coordinate sums its run in FP64 and rounds
// once. No floating-point atomics.
#include <algorithm>

#include "frontend.hpp"

namespace ttg {
namespace {

// An endpoint >= n is clamped to n: such keys sort after every real point,
// no segment start is recorded for them, and their edges contribute nothing.
__global__ void grad_keys(const u32* __restrict__ edges, u32 m, u32 n, int cb, u64* __restrict__ keys) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * m; t += gridDim.x * blockDim.x)
        keys[t] = (u64(min(edges[t], n)) << cb) | t;   // edges[2e + side] is the endpoint
}

__global__ void grad_count(const u64* __restrict__ keys, u32 total, u32 n, int cb, u32* __restrict__ start) {
    // start[p] = first sorted slot of point p (segment starts, rest untouched)
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const u32 p = u32(keys[t] >> cb);
        if (p < n && (t == 0 || u32(keys[t - 1] >> cb) != p)) start[p] = t;
    }
}

__device__ __forceinline__ double sgn(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); }

// derivative of d(a, b) with respect to x_a[k]
__device__ double ddist(int kind, double p, const float* __restrict__ xa, const float* __restrict__ xb,
                        u32 dim, u32 k) {
    if (kind == TT_METRIC_L2) {
        double s = 0;
        for (u32 q = 0; q < dim; ++q) {
            const double t = double(xa[q]) - double(xb[q]);
            s += t * t;
        }
        return s > 0 ? (double(xa[k]) - double(xb[k])) / sqrt(s) : 0.0;
    }
    if (kind == TT_METRIC_L1) return sgn(double(xa[k]) - double(xb[k]));
    if (kind == TT_METRIC_LINF) {
        u32 arg = 0;
        double best = -1;
        for (u32 q = 0; q < dim; ++q) {
            const double t = fabs(double(xa[q]) - double(xb[q]));
            if (t > best) {
                best = t;
                arg = q;
            }
        }
        return arg == k ? sgn(double(xa[k]) - double(xb[k])) : 0.0;
    }
    double s = 0;
    for (u32 q = 0; q < dim; ++q) s += pow(fabs(double(xa[q]) - double(xb[q])), p);
    if (!(s > 0)) return 0.0;
    const double d = pow(s, 1.0 / p);
    const double t = double(xa[k]) - double(xb[k]);
    return sgn(t) * pow(fabs(t), p - 1.0) / pow(d, p - 1.0);
}

__global__ void grad_gather(const float* __restrict__ x, u32 n, u32 dim, int kind, float p,
                            const u32* __restrict__ edges, const float* __restrict__ g,
                            const u64* __restrict__ keys, u32 total, int cb, const u32* __restrict__ start,
                            float* __restrict__ dx) {
    const u64 cells = u64(n) * dim;
    const u64 mask = (u64(1) << cb) - 1;
    for (u64 c = blockIdx.x * u64(blockDim.x) + threadIdx.x; c < cells; c += u64(gridDim.x) * blockDim.x) {
        const u32 pnt = u32(c / dim), k = u32(c % dim);
        double acc = 0;
        for (u32 t = start[pnt]; t < total && u32(keys[t] >> cb) == pnt; ++t) {
            const u32 slot = u32(keys[t] & mask);
            const u32 e = slot >> 1, side = slot & 1;
            const u32 a = edges[2 * e + side], b = edges[2 * e + (side ^ 1)];
            if (a == b || b >= n) continue;   // a == pnt < n
            acc += double(g[e]) * ddist(kind, double(p), x + u64(a) * dim, x + u64(b) * dim, dim, k);
        }
        dx[c] = float(acc);
    }
}

}  // namespace

void diagram_backward(const float* d_x, u32 n, u32 dim, int kind, float p, const u32* d_edges,
                      const float* d_g, u32 m, float* d_dx, Workspace& ws, hipStream_t s) {
    check_metric(kind, p);
    const u64 cells = u64(n) * dim;
    fill_device_u32(reinterpret_cast<u32*>(d_dx), cells, 0u, s);
    if (!m || !cells) return;
    const int cb = std::max(1, bit_length(2 * u64(m) - 1));
    const int pb = std::max(1, bit_length(u64(n)));   // endpoints clamped to n
    if (cb + pb > 64) fail(TT_ELENGTH, "diagram_backward: too many contributions");
    u64* keys = ws.get<u64>(2 * u64(m));
    u64* alt = ws.get<u64>(2 * u64(m));
    u32* start = ws.get<u32>(n);
    fill_device_u32(start, n, 0xffffffffu, s);
    const u32 blocks = u32(std::min<u64>((2 * u64(m) + 255) / 256, 4096));
    grad_keys<<<blocks, 256, 0, s>>>(d_edges, m, n, cb, keys);
    TTG_LAUNCH_CHECK();
    radix_sort_u64(keys, nullptr, 2 * u64(m), 0, cb + pb, alt, nullptr, ws, s);
    grad_count<<<blocks, 256, 0, s>>>(keys, 2 * m, n, cb, start);
    TTG_LAUNCH_CHECK();
    const u32 gblocks = u32(std::min<u64>((cells + 255) / 256, 16384));
    grad_gather<<<gblocks, 256, 0, s>>>(d_x, n, dim, kind, p, d_edges, d_g, keys, 2 * m, cb, start, d_dx);
    TTG_LAUNCH_CHECK();
}

}  // namespace ttg

using namespace ttg;

extern "C" {

tt_status tt_diagram_backward_dev(const float* d_x, uint32_t n, uint32_t dim, int kind, float p,
                                  const uint32_t* d_edges, const float* d_grad, uint32_t m, float* d_dx,
                                  void* stream) {
    return guarded([&] {
        if (u64(n) * dim && (!d_x || !d_dx)) fail(TT_EINVAL, "tt_diagram_backward_dev: null buffer");
        if (m && (!d_edges || !d_grad)) fail(TT_EINVAL, "tt_diagram_backward_dev: null edges");
        hipStream_t s = as_stream(stream);
        Workspace ws;
        diagram_backward(d_x, n, dim, kind, p, d_edges, d_grad, m, d_dx, ws, s);
        TTG_HIP(hipStreamSynchronize(s));   // ws is released on return
    });
}

tt_status tt_diagram_backward(const float* x, uint32_t n, uint32_t dim, int kind, float p,
                              const uint32_t* edges, const float* grad, uint32_t m, float* dx) {
    return guarded([&] {
        check_points(x, u64(n) * dim, "tt_diagram_backward");
        for (u64 t = 0; t < 2 * u64(m); ++t)
            if (edges[t] >= n) fail(TT_EINVAL, "tt_diagram_backward: edge endpoint out of range");
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx_d = ws.get<float>(u64(n) * dim);
        float* x_d = ws.get<float>(u64(n) * dim);
        u32* e_d = ws.get<u32>(2 * u64(m));
        float* g_d = ws.get<float>(m);
        if (u64(n) * dim) TTG_HIP(hipMemcpyAsync(x_d, x, u64(n) * dim * 4, hipMemcpyHostToDevice, s));
        if (m) {
            TTG_HIP(hipMemcpyAsync(e_d, edges, 2 * u64(m) * 4, hipMemcpyHostToDevice, s));
            TTG_HIP(hipMemcpyAsync(g_d, grad, u64(m) * 4, hipMemcpyHostToDevice, s));
        }
        diagram_backward(x_d, n, dim, kind, p, e_d, g_d, m, dx_d, ws, s);
        if (u64(n) * dim) TTG_HIP(hipMemcpyAsync(dx, dx_d, u64(n) * dim * 4, hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
