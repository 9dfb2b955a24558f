// Host-side launchers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace hpt {

void launch_cast(const DevScene &sc, const float *org3, const float *dir3, const float4 *org4,
                 const float4 *dir4, CountRef n_dev, int64_t n_host, WorkCounter *wc,
                 int *out_id, float *out_t, int *counts, unsigned long long *totals, int grid,
                 hipStream_t st);
void launch_camera(const ViewParams *views, int view_id, int64_t n, Paths out, hipStream_t st);
void launch_fold(const ViewParams *views, int view_id, int64_t n, const float *accum, float *out,
                 hipStream_t st);
void launch_misses(const ViewParams *views, int view_id, int64_t n, const int *hit_id,
                   int samples, float *accum, hipStream_t st);
void launch_primary(const ViewParams *views, int view_id, int64_t n, const int *hit_id,
                    const float *hit_t, float *accum, int2 *prim_hit, int *prim_list,
                    int *prim_count, hipStream_t st);
void launch_expand(const ViewParams *views, const Segment *segs_all, const PassDesc *passes,
                   const int *pass, int64_t max_total, const int *prim_list, const int2 *prim_hit,
                   const float4 *prim_dir, Paths out, int *out_id, float *out_t, CountRef live,
                   hipStream_t st);
void launch_shade(const DevScene &sc, const ViewParams *views, int depth_index, CountRef n_in,
                  Paths in, const int *hit_id, const float *hit_t, Paths out, CountRef n_out,
                  float *accum, int grid, hipStream_t st);
void launch_next_pass(int *pass, hipStream_t st);
void launch_zero_counter(WorkCounter *wc, hipStream_t st);
void launch_resolve(const ViewParams *views, int view_id, int width, int height, int ssx, int ssy,
                    int samples, float brightness, const float *accum, uchar4 *rgba,
                    hipStream_t st);
int occupancy_blocks_cast();
int occupancy_blocks_shade();
// Fetch forms of the lane-refill cast (hpt_tuning.fetch): node pairs (1),
// triangle prefetch (2), wave-spread leaf rounds (4) with a leaf queue of
// depth 1, 2, 4 or 8 (bits 3-4 = 0..3; with 4 only, and without 2); with
// a queue of 4 or 8, bit 5 (32) compiles the form for six waves per SIMD
// and (without node pairs) bit 6 (64) folds the leaf rounds through keys in
// LDS; bit 7 (128) reads the triangle records once per triangle through
// the leaf entries' ids; bits 8-9 (with 1) widen the node window, bit 10
// (1024) ends it at the end of the 128-byte line that holds the first node,
// and bit 11 (2048) reads the nodes, and with 4 and 128 the leaf entries'
// ids and the records, with buffer loads (fetch_fits) and skips records a
// ray has already tested (its mailbox).  These are the compiled forms.
inline bool valid_fetch(int f) {
    switch (f) {
    case 0: case 1: case 2: case 3: case 4: case 5: case 12: case 13: case 20: case 21:
    case 28: case 29: case 52: case 60: case 84: case 92: case 116: case 124: case 148:
    case 149: case 156: case 157: case 189: case 220: case 252: case 413: case 445:
    case 669: case 701: case 1469: case 1725: case 3517: case 3773:
        return true;
    default:
        return false;
    }
}

// The buffer-load forms address nodes, leaf entries' record numbers and
// per-triangle records by 32-bit byte offsets (the window's slots past the
// last node included): each array under 2 GiB.
inline bool fetch_fits(int f, const DevScene &ds) {
    return !(f & 2048) || ((ds.n_nodes + 4LL) * 32 <= (int64_t)INT32_MAX &&
                           ds.n_entries * 4LL <= (int64_t)INT32_MAX &&
                           ds.n_recs * 48LL <= (int64_t)INT32_MAX);
}

}  // namespace hpt
