// Device data layouts and kernel declarations of the wavefront path tracer.
//
// Scene in HBM (built once per scene by hpt_scene_create):
//   nodes  : 32 bytes per KD node, pre-order.  float4 {box_start, a},
//            float4 {box_end, b}; a/b are int bits.  Interior: a = index of
//            the first node after the subtree (where the reference's
//            sibling/parent walk continues, kernel.cl:153-167), b = -1, and
//            the left child is the next node.  Leaf: a = triangle count,
//            b = offset of its first triangle record (count 0 = empty leaf).
//   tris   : 48 bytes per leaf entry, in leaf-list order: float4 {pa, id},
//            float4 {e1 = pb - pa, 0}, float4 {e2 = pc - pa, 0}.
//   tri_recs, leaf_ids : the same records once per triangle, numbered in
//            the order the leaves first list them, and the leaf lists as
//            those record numbers (the indirect fetch form).
//   shading arrays keep the reference's layouts (float4 points/normals,
//   float2 texcoords, int4 mesh and lights, 96-byte materials, RGBA8 texels).
//
// Ray state is structure-of-arrays, one entry per live path:
//   org  float4 {x, y, z, view id bits}
//   dir  float4 {x, y, z, 0}
//   thr  float4 {strength xyz, global pixel-sample index bits}
//   seed uint64 (48-bit LCG state)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hptracing.h"

namespace hpt {

struct DevScene {
    const float4 *nodes;
    int32_t n_nodes;
    const float4 *tris;
    const float4 *tri_recs;   // the records once per triangle (indirect form)
    const int32_t *leaf_ids;  // record number per leaf entry (indirect form)
    int32_t n_entries;        // leaf entries (length of tris and leaf_ids; at least 1)
    int32_t n_recs;           // per-triangle records (length of tri_recs; at least 1)
    const float4 *points;
    const float4 *normals;
    int32_t n_normals;
    const float2 *texcoords;
    int32_t n_texcoords;
    const int4 *mesh;
    const hpt_material *materials;
    const int4 *lights;
    int32_t n_lights;
    const uchar4 *textures;
    int32_t n_textures;
    int32_t fast_div;  // slab quotients via certified reciprocals (same results)
    int32_t boxes_fast; // every box coordinate is within the certified range
    int32_t traversal; // 0: interleaved per lane; 1: wave phases; 2: phases + lane refill
    int32_t arith_ocl; // dot/cross as ROCm's OpenCL builtins (HPT_ARITH_ROCM_OPENCL)
    int32_t refill;     // traversal 2: idle lanes that trigger a refill
    int32_t node_steps; // traversal 2: node visits per round
    int32_t walk_min;   // traversal 2: a round's node phase ends below this many walking lanes
    int32_t fetch;      // traversal 2: bit 0 paired node fetch, bit 1 triangle prefetch,
                        // bit 2 leaf tests spread over the wave (not with bit 1),
                        // bits 3-4 (with bit 2) leaf queue depth 1 << value,
                        // bit 5 six-wave register budget, bit 6 keyed leaf fold,
                        // bit 7 per-triangle records through leaf ids,
                        // bits 8-9 (with bit 0) node window of 2 + value,
                        // bit 10 window cut at the 128-byte line end,
                        // bit 11 buffer loads (fetch_fits) and, with bits
                        // 2 and 7, the per-ray record mailbox
};

struct Paths {
    float4 *org;
    float4 *dir;
    float4 *thr;
    uint64_t *seed;
};

// Per-view constants of one render request inside a launch.
struct ViewParams {
    float view[3];
    float background[3];
    int32_t disable_diffuse;
    int32_t depth;
    int32_t sample;
    int32_t sx, sy;              // camera rays per row / column (W*ssx, H*ssy)
    int32_t vr;                  // equirectangular camera
    int32_t vr_shift, vr_sample_x, vr_sample_y;
    float up[3], right[3];
    float width, height;
    uint64_t seed;
    int64_t ray_base;            // first global pixel-sample index of the view
    int64_t accum_stride;        // camera rays between accumulator slices
};

// Sample passes run kPassSlices at a time: pass p of a camera ray adds into
// accumulator slice p % kPassSlices (slices accum_stride camera rays apart),
// so the paths of one camera ray in flight together never share an entry;
// each slice takes its passes in increasing order, and the slices are summed
// in slice order at the end (k_fold).  The sums are the same for a view
// rendered alone, in a batch or in the server.
constexpr int kPassSlices = 16;

// A run of primary hits of one view expanded into the path buffer.
struct Segment {
    int32_t view;
    int32_t pass;
    int64_t prim_offset;         // into the primary-hit list
    int64_t count;
    int64_t out_offset;          // prefix sum of counts
};

// Work-distribution counters of a persistent kernel: `next` hands out ray
// chunks, `done` counts finished waves; the last wave re-arms both.
struct WorkCounter {
    int32_t next;
    int32_t done;
};

// A live-path count in device memory.  Within a render the counts form a
// table [pass][depth + 1]; the pass is read from device memory (`pass`), so
// one captured launch sequence serves every pass.
struct CountRef {
    int *base;
    const int *pass;   // nullptr: base[offset]
    int stride;
    int offset;
    __device__ __forceinline__ int *ptr() const {
        return base + (pass ? (*pass) * stride : 0) + offset;
    }
};

// The segments and ray total of one pass: segs[seg_begin .. seg_begin+n_segs).
struct PassDesc {
    int32_t seg_begin;
    int32_t n_segs;
    int64_t total;
};

constexpr int kBlock = 256;

}  // namespace hpt
