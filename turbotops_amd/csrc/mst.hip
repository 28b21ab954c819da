// SPDX-License-Identifier: MIT
// mst.hip — H0 of the Rips filtration of a point cloud: the minimum spanning
// tree of the complete graph by Boruvka rounds.
//
// Replaces mst::h0's Prim loop (include/turbotops/mst.hpp:25-99). The H0
// deaths are the MST edge weights, a multiset every MST shares, so the
// sorted deaths equal Prim's whatever the algorithm.
//
// Edge order: (key, min id, max id), a strict total order, so the MST is
// unique and Boruvka never closes a cycle. key is the distance for L1, L-inf
// and Lp, and for L2 the fma sum before the square root: sqrtf is monotone, so
// that order refines the distance order and the tree is an MST for the
// distances too; the deaths are sqrtf(key) of its edges.
//
// Per round:
//   nearest     one lane per vertex scans every other vertex (points and
//               component labels staged in LDS) for its smallest (key, id)
//               edge into another component; a vertex whose previous best
//               still leads outside keeps it (the outside set only shrinks)
//   comp_min    per component, atomicMin of the key bits, then atomicMin of
//               (min id << 32 | max id) among the vertices that reach it
//   hook        each component points at the one across its edge; of a
//               mutual pair the smaller root stays a root; each chosen edge is
//               recorded once
//   jump        pointer jumping until every label names its root
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "pairs.hpp"
#include "frontend.hpp"

namespace ttg {

namespace {

constexpr u32 MB = 256;

template <int V, int DIM>
__global__ __launch_bounds__(MB) void mst_nearest(const float* __restrict__ x, u32 n, u32 dim_rt,
                                                  MetricArgs ma, const u32* __restrict__ comp,
                                                  const u32* __restrict__ todo, u32 ntodo,
                                                  float* __restrict__ best_key,
                                                  u32* __restrict__ best_j) {
    constexpr u32 DS = DIM > 0 ? DIM : 1;
    __shared__ float tile[DIM > 0 ? MB * DS : 1];
    __shared__ u32 tcomp[MB];
    const u32 dim = DIM > 0 ? u32(DIM) : dim_rt;
    const u32 t = blockIdx.x * MB + threadIdx.x;
    const u32 i = t < ntodo ? todo[t] : todo[ntodo - 1];
    const u32 ci = comp[i];
    float q[DS];
    const float* __restrict__ qp = x + u64(i) * dim;
    if constexpr (DIM > 0) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) q[k] = qp[k];
    }
    float bk = INFINITY;
    u32 bj = 0xffffffffu;
    for (u32 j0 = 0; j0 < n; j0 += MB) {
        const u32 len = min(MB, n - j0);
        __syncthreads();
        if constexpr (DIM > 0)
            for (u32 e = threadIdx.x; e < len * DS; e += MB) tile[e] = x[u64(j0) * DS + e];
        if (threadIdx.x < len) tcomp[threadIdx.x] = comp[j0 + threadIdx.x];
        __syncthreads();
        for (u32 jj = 0; jj < len; ++jj) {
            if (tcomp[jj] == ci) continue;
            const u32 j = j0 + jj;
            float kv;
            if constexpr (DIM > 0) kv = pdist<V, DIM, true>(q, tile + jj * DS, dim, ma);
            else kv = pdist<V, 0, true>(qp, x + u64(j) * dim, dim, ma);
            // (key, j) order: for a fixed i this is the edge order (key, min, max)
            if (kv < bk) {
                bk = kv;
                bj = j;
            }
        }
    }
    if (t < ntodo) {
        best_key[i] = bk;
        best_j[i] = bj;
    }
}

template <int V, int DIM>
struct NearestLaunch {
    static void run(u32 blocks, hipStream_t s, const float* x, u32 n, u32 dim, MetricArgs ma,
                    const u32* comp, const u32* todo, u32 ntodo, float* bk, u32* bj) {
        mst_nearest<V, DIM><<<blocks, MB, 0, s>>>(x, n, dim, ma, comp, todo, ntodo, bk, bj);
    }
};

// Vertices that need a fresh scan: every vertex in round 0, afterwards those
// whose best neighbour has joined their own component.
__global__ void mst_todo(const u32* __restrict__ comp, const u32* __restrict__ best_j, u32 n,
                         int first, u32* __restrict__ todo, u32* __restrict__ ntodo) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const bool need = first || best_j[v] == 0xffffffffu || comp[best_j[v]] == comp[v];
        if (need) todo[atomicAdd(ntodo, 1u)] = v;
    }
}

__global__ void mst_comp_key(const u32* __restrict__ comp, const float* __restrict__ best_key,
                             u32 n, u32* __restrict__ cw) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        atomicMin(cw + comp[v], __float_as_uint(best_key[v]));
}

__global__ void mst_comp_pair(const u32* __restrict__ comp, const float* __restrict__ best_key,
                              const u32* __restrict__ best_j, u32 n, const u32* __restrict__ cw,
                              unsigned long long* __restrict__ cp) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const u32 c = comp[v];
        if (__float_as_uint(best_key[v]) != cw[c] || best_j[v] == 0xffffffffu) continue;
        const u32 j = best_j[v];
        const u64 a = min(v, j), b = max(v, j);
        atomicMin(cp + c, (unsigned long long)((a << 32) | b));
    }
}

__global__ void mst_hook(const u32* __restrict__ comp, u32 n, const u32* __restrict__ cw,
                         const unsigned long long* __restrict__ cp, u32* __restrict__ parent,
                         float* __restrict__ out_keys, u32* __restrict__ out_count) {
    for (u32 c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        if (comp[c] != c) continue;            // not a root
        const u64 e = cp[c];
        if (e == ~0ull) continue;              // no outgoing edge (single component)
        const u32 u = u32(e >> 32), v = u32(e);
        const u32 cu = comp[u], cv = comp[v];
        const u32 other = cu == c ? cv : cu;
        const bool mutual = cp[other] == e && cw[other] == cw[c];
        if (mutual && c < other) {
            out_keys[atomicAdd(out_count, 1u)] = __uint_as_float(cw[c]);
        } else {
            parent[c] = other;
            if (!mutual) out_keys[atomicAdd(out_count, 1u)] = __uint_as_float(cw[c]);
        }
    }
}

__global__ void mst_jump(u32* __restrict__ parent, u32 n, u32* __restrict__ changed) {
    for (u32 c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const u32 p = parent[c];
        const u32 gp = parent[p];
        if (gp != p) {
            parent[c] = gp;
            *changed = 1u;
        }
    }
}

__global__ void mst_relabel(u32* __restrict__ comp, const u32* __restrict__ parent, u32 n,
                            u32* __restrict__ cw, unsigned long long* __restrict__ cp) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        comp[v] = parent[comp[v]];
        cw[v] = 0xffffffffu;
        cp[v] = ~0ull;
    }
}

template <int V>
__global__ void mst_finish(float* __restrict__ keys, u32 m, u64* __restrict__ sortkeys) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x)
        sortkeys[t] = __float_as_uint(key_to_dist<V>(keys[t]));
}

__global__ void mst_unpack(const u64* __restrict__ sortkeys, u32 m, float* __restrict__ out) {
    for (u32 t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x)
        out[t] = __uint_as_float(u32(sortkeys[t]));
}

__global__ void iota_u32(u32* __restrict__ p, u32 n) {
    for (u32 v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) p[v] = v;
}

u32 blocks_of(u64 count) {
    return u32(std::max<u64>(1, std::min<u64>((count + MB - 1) / MB, 16384)));
}

}  // namespace


struct MstStats {
    u32 rounds = 0;
    u64 scanned = 0;    // vertex scans (each n - 1 candidates)
};

namespace {

// Sub-buffers of one device region, handed out in order at 256-byte
// alignment. With base == nullptr it only measures (the bytes a run needs).
struct Carve {
    char* base = nullptr;
    u64 cap = ~0ull, used = 0;
    template <typename T>
    T* get(u64 count) {
        const u64 off = (used + 255) & ~u64(255);
        const u64 need = count * sizeof(T);
        if (off + need > cap) fail(TT_ELENGTH, "mst: scratch region too small");
        used = off + need;
        return base ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

// Every device buffer a Boruvka run on n vertices uses, carved from `c` in
// a fixed order (a measuring Carve gives the total).
struct MstBuffers {
    u32 *comp, *parent, *cw, *bj, *todo, *scal;
    unsigned long long* cp;
    float *bk, *keys;
    u64 *sk, *alt, *sort_scratch;
    MstBuffers(Carve& c, u32 n) {
        comp = c.get<u32>(n);
        parent = c.get<u32>(n);
        cw = c.get<u32>(n);
        cp = c.get<unsigned long long>(n);
        bk = c.get<float>(n);
        bj = c.get<u32>(n);
        todo = c.get<u32>(n);
        scal = c.get<u32>(4);      // [0] ntodo, [1] changed, [2] out count, [3] n - 1
        keys = c.get<float>(n);
        sk = c.get<u64>(n);
        alt = c.get<u64>(n);
        sort_scratch = c.get<u64>(radix_scratch_entries(n));
    }
};

}  // namespace

u64 mst_scratch_bytes(u32 n) {
    Carve c;
    MstBuffers b(c, std::max<u32>(n, 2));
    (void)b;
    return c.used;
}

// Sorted MST edge weights (n - 1 floats, device) of the complete graph, with
// every temporary carved from the device region [base, base + cap).
void mst_deaths_in(const float* d_x, u32 n, u32 dim, int kind, float p, float* d_deaths, char* base, u64 cap,
                   hipStream_t s, MstStats* st) {
    if (n < 2) return;
    const MetricVariant v = variant_of(kind, p);
    const MetricArgs ma = metric_args(kind, p);
    Carve c{base, cap, 0};
    const MstBuffers B(c, n);
    u32 *comp = B.comp, *parent = B.parent, *cw = B.cw, *bj = B.bj, *todo = B.todo, *scal = B.scal;
    unsigned long long* cp = B.cp;
    float *bk = B.bk, *keys = B.keys;
    iota_u32<<<blocks_of(n), MB, 0, s>>>(comp, n);
    iota_u32<<<blocks_of(n), MB, 0, s>>>(parent, n);
    fill_device_u32(cw, n, 0xffffffffu, s);
    fill_device_u32(reinterpret_cast<u32*>(cp), 2 * u64(n), 0xffffffffu, s);
    fill_device_u32(bj, n, 0xffffffffu, s);
    fill_device_u32(scal, 3, 0u, s);
    fill_device_u32(scal + 3, 1, n - 1, s);
    TTG_LAUNCH_CHECK();
    u32 edges = 0;
    for (u32 round = 0; edges < n - 1; ++round) {
        if (round > 64) fail(TT_EDEVICE, "mst: Boruvka did not converge");
        fill_device_u32(scal, 1, 0u, s);
        mst_todo<<<blocks_of(n), MB, 0, s>>>(comp, bj, n, round == 0, todo, scal);
        u32 ntodo = 0;
        TTG_HIP(hipMemcpyAsync(&ntodo, scal, sizeof(u32), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
        if (ntodo) {
            dispatch_metric_dim<NearestLaunch>(v, dim, (ntodo + MB - 1) / MB, s, d_x, n, dim, ma, comp, todo,
                                               ntodo, bk, bj);
            TTG_LAUNCH_CHECK();
        }
        if (st) {
            st->rounds = round + 1;
            st->scanned += ntodo;
        }
        mst_comp_key<<<blocks_of(n), MB, 0, s>>>(comp, bk, n, cw);
        mst_comp_pair<<<blocks_of(n), MB, 0, s>>>(comp, bk, bj, n, cw, cp);
        mst_hook<<<blocks_of(n), MB, 0, s>>>(comp, n, cw, cp, parent, keys, scal + 2);
        TTG_LAUNCH_CHECK();
        for (int it = 0; it < 64; ++it) {
            fill_device_u32(scal + 1, 1, 0u, s);
            mst_jump<<<blocks_of(n), MB, 0, s>>>(parent, n, scal + 1);
            u32 changed = 0;
            TTG_HIP(hipMemcpyAsync(&changed, scal + 1, sizeof(u32), hipMemcpyDeviceToHost, s));
            TTG_HIP(hipStreamSynchronize(s));
            if (!changed) break;
        }
        mst_relabel<<<blocks_of(n), MB, 0, s>>>(comp, parent, n, cw, cp);
        TTG_LAUNCH_CHECK();
        TTG_HIP(hipMemcpyAsync(&edges, scal + 2, sizeof(u32), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
    }
    if (edges != n - 1) fail(TT_EDEVICE, "mst: spanning tree has the wrong edge count");
    switch (v) {
        case MV_L2: mst_finish<MV_L2><<<blocks_of(n), MB, 0, s>>>(keys, n - 1, B.sk); break;
        default: mst_finish<MV_L1><<<blocks_of(n), MB, 0, s>>>(keys, n - 1, B.sk); break;
    }
    TTG_LAUNCH_CHECK();
    const u64* sorted = radix_sort_u64_counted(B.sk, n - 1, scal + 3, 0, 31, B.alt, B.sort_scratch, s);
    mst_unpack<<<blocks_of(n), MB, 0, s>>>(sorted, n - 1, d_deaths);
    TTG_LAUNCH_CHECK();
}

// The same with the region from a Workspace.
void mst_deaths_dev(const float* d_x, u32 n, u32 dim, int kind, float p, float* d_deaths, Workspace& ws,
                    hipStream_t s, MstStats* st) {
    if (n < 2) return;
    const u64 bytes = mst_scratch_bytes(n);
    mst_deaths_in(d_x, n, dim, kind, p, d_deaths, ws.get<char>(bytes), bytes, s, st);
}

// ---- the no-allocation form ---------------------------------------------
// One scratch region per device for tt_mst_deaths_dev_noalloc: a file-scope
// device array, so a call allocates nothing. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol); it passes the device address hipGetSymbolAddress returns, looked
// up once per device with that device current.
constexpr u32 kMstNoAllocMaxN = 1u << 20;
constexpr u64 kMstScratchBytes = u64(56) << 20;

static __device__ __align__(256) unsigned char scratch[kMstScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<char*> g_scratch_ptr;   // per device ordinal; nullptr until looked up

char* scratch_ptr() {
    int dev = 0;
    TTG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.size() <= size_t(dev)) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(TT_EDEVICE, std::string("hip: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

}  // namespace ttg

using namespace ttg;

extern "C" {

tt_status tt_mst_deaths(const float* pts, uint32_t n, uint32_t dim, int kind, float p,
                        float* deaths) {
    return guarded([&] {
        check_metric(kind, p);
        check_points(pts, u64(n) * dim, "tt_mst_deaths");
        for (u64 t = 0; t < u64(n) * dim; ++t)
            if (!std::isfinite(pts[t])) fail(TT_EINVAL, "mst: non-finite coordinate");
        if (n < 2) return;
        if (dim == 0) fail(TT_EINVAL, "tt_mst_deaths: dim must be >= 1");
        if (!deaths) fail(TT_EINVAL, "tt_mst_deaths: null deaths");
        hipStream_t s = nullptr;
        Workspace ws;
        float* dx = ws.get<float>(u64(n) * dim);
        float* dd = ws.get<float>(n - 1);
        TTG_HIP(hipMemcpyAsync(dx, pts, u64(n) * dim * sizeof(float), hipMemcpyHostToDevice, s));
        mst_deaths_dev(dx, n, dim, kind, p, dd, ws, s, nullptr);
        TTG_HIP(hipMemcpyAsync(deaths, dd, u64(n - 1) * sizeof(float), hipMemcpyDeviceToHost, s));
        TTG_HIP(hipStreamSynchronize(s));
    });
}

tt_status tt_mst_deaths_dev(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                            float* d_deaths, uint32_t* rounds, uint64_t* scans, void* stream) {
    return guarded([&] {
        check_metric(kind, p);
        if (n >= 2 && (!d_pts || !d_deaths)) fail(TT_EINVAL, "tt_mst_deaths_dev: null buffer");
        if (n >= 2 && dim == 0) fail(TT_EINVAL, "tt_mst_deaths_dev: dim must be >= 1");
        hipStream_t s = as_stream(stream);
        Workspace ws;
        MstStats st;
        mst_deaths_dev(d_pts, n, dim, kind, p, d_deaths, ws, s, &st);
        TTG_HIP(hipStreamSynchronize(s));
        if (rounds) *rounds = st.rounds;
        if (scans) *scans = st.scanned;
    });
}

tt_status tt_mst_deaths_dev_noalloc(const float* d_pts, uint32_t n, uint32_t dim, int kind, float p,
                                    float* d_deaths, void* stream) {
    return guarded([&] {
        check_metric(kind, p);
        if (n > kMstNoAllocMaxN)
            fail(TT_ELENGTH, "tt_mst_deaths_dev_noalloc: n exceeds " + std::to_string(kMstNoAllocMaxN) +
                                 " points (the static scratch region)");
        if (n >= 2 && (!d_pts || !d_deaths)) fail(TT_EINVAL, "tt_mst_deaths_dev_noalloc: null buffer");
        if (n >= 2 && dim == 0) fail(TT_EINVAL, "tt_mst_deaths_dev_noalloc: dim must be >= 1");
        if (n < 2) return;
        const u64 bytes = mst_scratch_bytes(n);
        if (bytes > kMstScratchBytes) fail(TT_ELENGTH, "tt_mst_deaths_dev_noalloc: scratch region too small");
        hipStream_t s = as_stream(stream);
        mst_deaths_in(d_pts, n, dim, kind, p, d_deaths, scratch_ptr(), kMstScratchBytes, s, nullptr);
        TTG_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
