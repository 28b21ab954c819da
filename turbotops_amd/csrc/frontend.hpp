// SPDX-License-Identifier: MIT
// frontend.hpp — stream-ordered device stages of the front end, shared by
// the C ABI entry points across translation units.
#pragma once

#include "sort.hpp"

namespace ttg {

// dist.hip
// ev_begin / ev_end (may be null): recorded on s right before and right after
// the tile launch, so a caller can time that kernel alone
void dist_matrix(const float* d_x, u32 n, u32 dim, int kind, float p, int layout, float* d_mat,
                 u32* d_rowmax, float* d_radius, hipStream_t s, int rpw_override = 0,
                 hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
void fill_device_u32(u32* p, u64 count, u32 v, hipStream_t s);
int tile_rows_per_wave();
void set_tile_rows_per_wave(int rpw);

// spatial.hip: xs = the n points of d_x (row-major, dim floats) sorted by a
// Morton-style grid code (stable), perm[pos] = the original index of xs's
// row pos (perm may be null). d_n holds n on the device (the counted sort).
// Temporaries: keys and alt (n entries each), sort_scratch
// (radix_scratch_entries(n)), ext (spatial_ext_entries(dim)).
u64 spatial_ext_entries(u32 dim);
void spatial_order(const float* d_x, u32 n, u32 dim, const u32* d_n, u64* keys, u64* alt, u64* sort_scratch,
                   u32* ext, float* xs, u32* perm, hipStream_t s);
// Boxes of groups of `per` consecutive members: out[b] = the coordinate
// minima (dim floats) then maxima (dim floats) of members [b per, (b+1) per)
// of src, which holds count points (from_boxes false: dim floats each) or
// count boxes (from_boxes true: 2 dim floats each).
void spatial_boxes(const float* src, u32 count, u32 per, u32 dim, bool from_boxes, float* out, hipStream_t s);

// hostcopy.hip: device -> pageable host memory, staged through pinned
// buffers for large sizes; returns when the data is in dst
void copy_to_host(void* dst, const void* src, u64 bytes, hipStream_t s);

// edges.hip
struct DenseEdges {
    u64 m = 0;
    u64* idx = nullptr;      // device, filtration order
    float* d = nullptr;      // device
    u8* apparent = nullptr;  // device (null unless requested); the witness flag in comb order
};
// comb: ties in diameter broken by ascending index (RedZeD's order)
// instead of descending (the coboundary reducer's `precedes`). want_flags:
// the apparent flags in `precedes` order, the boundary engine's ready-witness
// flags (witness_rule) in comb order.
DenseEdges dense_edges(const float* d_mat, u32 n, float thr, bool want_flags, Workspace& ws,
                       hipStream_t s, bool comb = false);
// EdgeList vertex pairs (smaller id first) of edge indices; with flags, bit
// 31 of each larger id carries the edge's flag.
void edge_vertices(const u64* d_idx, u64 m, int32_t* d_verts, hipStream_t s, const u8* d_flags = nullptr);

// graph.hip: the kNN lists of one cloud for k = kk > 64 (kk <= n - 1): one
// workgroup per query selects its kk best (distance, id) pairs, then two
// stable sorts order them; temporaries come from ws
void knn_large_k(const float* d_x, u32 n, u32 dim, int kind, float p, u32 kk, u32* d_ids, Workspace& ws,
                 hipStream_t s);

// sparse.hip
struct SparseEdges {
    u64 m = 0;
    u64* idx = nullptr;      // device, filtration order
    float* d = nullptr;      // device
    u8* apparent = nullptr;  // device (null unless requested)
};
SparseEdges sparse_edges(const u64* d_row_ptr, const u32* d_col, const float* d_w, u32 n, float thr,
                         bool want_apparent, Workspace& ws, hipStream_t s, bool comb = false);

}  // namespace ttg
