// SPDX-License-Identifier: MIT
// frontend.hpp — stream-ordered device stages of the front end, shared by
// the C ABI entry points across translation units.
#pragma once

#include "sort.hpp"

namespace ttg {

// dist.hip
void dist_matrix(const float* d_x, u32 n, u32 dim, int kind, float p, int layout, float* d_mat,
                 u32* d_rowmax, float* d_radius, hipStream_t s, int rpw_override = 0);
void fill_device_u32(u32* p, u64 count, u32 v, hipStream_t s);
int tile_rows_per_wave();
void set_tile_rows_per_wave(int rpw);

// edges.hip
struct DenseEdges {
    u64 m = 0;
    u64* idx = nullptr;      // device, filtration order
    float* d = nullptr;      // device
    u8* apparent = nullptr;  // device (null unless requested)
};
DenseEdges dense_edges(const float* d_mat, u32 n, float thr, bool want_apparent, Workspace& ws,
                       hipStream_t s);

}  // namespace ttg
