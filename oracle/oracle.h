/*
 * CPU oracle of hpTracing's path-tracing hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference's OpenCL kernels
 * (src/hp/cl_src/kernel.cl) and of its host wavefront driver
 * (TraceRunner::run, src/hp/trace_runner.cpp:116-356).  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only to check or to time against; the product path never calls it.
 *
 * Arithmetic is IEEE float32 with contraction off (built with
 * -ffp-contract=off), correctly rounded division and square root, and the
 * operation order written in kernel.cl.  OpenCL builtins are restated as:
 *   dot(a,b)    = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   cross(a,b)  = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   length(a)   = sqrt(dot(a,a)),  normalize(a) = a / length(a) per component
 *   fmin/fmax   = IEEE minNum/maxNum (C fminf/fmaxf)
 *   read_imagef = OpenCL 1.2 bilinear filter, normalized coords, REPEAT,
 *                 on RGBA UNORM8 texels (value/255)
 *   write_imagef to UNORM8 = round-to-nearest-even of clamp(x,0,1)*255
 *
 * Parity pins (each is exercised by tests/):
 *   orc_box_intersect     - the 7 known-answer cases of the reference's
 *                           SceneCLTest.box_intersect (test_scene_cl.cpp:91-116)
 *   orc_single_intersect  - float64 brute force / closed form on random
 *                           triangles; the inert degenerate guard (NaN paths)
 *   orc_kd_cast           - equals orc_naive_cast (naive_intersect semantics,
 *                           kernel.cl:210-245) whenever the 512-leaf cap does
 *                           not bind, the reference's own invariant
 *                           (test_scene_cl.cpp:212-232)
 *   orc_rand / randf      - closed form of the 48-bit LCG
 *   cameras, resolve      - closed forms
 *   shading / bounces     - statistical: images against a float64 render
 * Functions without such a pin are marked "parity unpinned" below.
 */
#ifndef HPT_ORACLE_H
#define HPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layouts of the reference's device structs (types.h.cl:8-43). */
typedef struct { float x, y, z, w; } orc_f3;               /* cl_float3 */
typedef struct {
    orc_f3 emission, ambient, diffuse, specular;
    float optical_density, dissolve, shininess;
    int32_t diffuse_texture_id, ambient_texture_id;
    float specular_possibility, refract_possibility, diffuse_possibility;
} orc_material;                                             /* 96 bytes */
typedef struct {
    orc_f3 box_start, box_end;
    int32_t child, parent, sibling, data;
} orc_node;                                                 /* 48 bytes */

typedef struct {
    const orc_f3 *points;     int64_t n_points;
    const orc_f3 *normals;    int64_t n_normals;
    const float *texcoords;   int64_t n_texcoords;   /* 2 floats each */
    const int32_t *mesh;      int64_t n_triangles;   /* int4 each     */
    const orc_material *materials; int64_t n_materials;
    const int32_t *lights;    int64_t n_lights;      /* int4 each     */
    const orc_node *nodes;    int64_t n_nodes;
    const int32_t *leaf_data; int64_t n_leaf_data;
    const uint8_t *textures;  int64_t n_textures;    /* 512*512*4 each */
} orc_scene;

#define ORC_MAX_MATCH_DATA 512

/* Evaluation of dot/cross in the intersection test (process-wide setting;
 * set it before casting).  ROCM_OPENCL (default, the library's default
 * too): the fused forms of ROCm's OpenCL builtins, i.e. the reference
 * kernel as that library compiles it.  LITERAL: every product and sum
 * rounded, as kernel.cl writes them.  orc_dot_cross evaluates both for a pin against the library's
 * recorded outputs (4 floats per pair: dot, cross.xyz). */
#define ORC_ARITH_LITERAL 0
#define ORC_ARITH_ROCM_OPENCL 1
void orc_set_arithmetic(int mode);
void orc_dot_cross(int mode, const float *a, const float *b, int64_t n, float *out);

/* ---- kernel.cl device functions ---- */
int orc_box_intersect(const float box_start[3], const float box_end[3],
                      const float start[3], const float dir[3]);
float orc_single_intersect(const float start[3], const float dir[3],
                           const float pa[3], const float pb[3], const float pc[3]);
int64_t orc_rand(int64_t *seed);
float orc_randf(int64_t *seed);
void orc_randf3(int64_t *seed, float out[3]);

/* kdtree_intersect's search for n rays (xyz triples).  tri_id = index of the
 * hit triangle or -1, t = its distance or -42.  counts (optional,
 * ORC_CAST_COUNTS ints per ray) = nodes box-tested until the 512th non-empty
 * leaf was collected (all of them if it never was), non-empty leaves
 * collected, triangles tested, box-hit non-empty leaves dropped by the cap,
 * nodes box-tested by the whole walk.  n_threads <= 0: all cores. */
#define ORC_CAST_COUNTS 5
void orc_kd_cast(const orc_scene *s, const float *org, const float *dir, int64_t n,
                 int32_t *tri_id, float *t, int32_t *counts, int n_threads);
/* naive_intersect (kernel.cl:210-245): brute-force min over all triangles. */
void orc_naive_cast(const orc_scene *s, const float *org, const float *dir, int64_t n,
                    int32_t *tri_id, float *t, int n_threads);

/* set_viewdirs / set_viewdirs_vr (kernel.cl:559-613): n = sx*sy directions,
 * index sx*y + x.  For the VR camera pass width = sample_x argument. */
void orc_set_viewdirs(const float view[3], const float up[3], const float right[3],
                      float width, float height, int sx, int sy, float *dirs);
void orc_set_viewdirs_vr(int x_shift, int sample_x, int sample_y, int grid_x, int grid_y,
                         float *dirs);

/* get_image (kernel.cl:615-652) into RGBA8 (alpha 0). */
void orc_get_image(const float *accum, int width, int height, int ssx, int ssy, int spp,
                   float brightness, uint8_t *rgba);

/* Texture fetch of s1_run (read_imagef, kernel.cl:326-327). */
void orc_read_texture(const orc_scene *s, int layer, float u, float v, float out[3]);

/* Full render: TraceRunner::run (trace_runner.cpp:116-356), wavefront order.
 * seed_mode 0: each camera ray r owns an LCG state, reset at the start of
 *              pass p to orc_path_seed(seed, r, p) -- the scheme of the GPU
 *              path, so images are reproducible;
 * seed_mode 1: the reference's scheme: one seed per list slot, read as
 *              v_seed[slot] + slot by every stage (kernel.cl:358-360 ...),
 *              initial seeds from a 31-bit generator seeded with `seed`,
 *              list appends in sequential order.
 * accum: W*ssx*H*ssy*3 floats (zeroed here).  live (optional, depth ints):
 * rays cast per depth level summed over passes (index 0 = primaries). */
typedef struct {
    float view[3], up[3], right[3];
    float width, height;
    int32_t sample_x, sample_y, supersample_x, supersample_y;
    float background[3];
    int32_t sample, depth, disable_diffuse;
    float brightness;
    uint64_t seed;
} orc_view;
int orc_render(const orc_scene *s, const orc_view *v, int seed_mode, int n_threads,
               float *accum, uint8_t *rgba, int64_t *live);
/* The literal wavefront engine in either seed mode (orc_render runs seed
 * mode 0 depth-first per path, which reorders the same arithmetic). */
int orc_render_wavefront(const orc_scene *s, const orc_view *v, int seed_mode, float *accum,
                         uint8_t *rgba, int64_t *live);
/* Same render restricted to the camera rays [ray_begin, ray_end): the
 * bounded sample of the CPU baseline.  Returns rays cast. */
int64_t orc_render_partial(const orc_scene *s, const orc_view *v, int64_t ray_begin,
                           int64_t ray_end, int n_threads, float *accum);
/* Same render for the n camera rays listed in `rays` (any order, out-of-range
 * entries skipped), each ray's direction computed alone: the stratified
 * sample of the CPU baseline.  accum holds 3 floats per camera ray of the
 * whole view.  Returns rays cast. */
int64_t orc_render_rays(const orc_scene *s, const orc_view *v, const int64_t *rays, int64_t n,
                        int n_threads, float *accum);

uint64_t orc_path_seed(uint64_t seed, int64_t ray, int32_t pass);

/* Float64 restatement (oracle64.c, liboracle64.so): pinhole renders with
 * double sums and, optionally, sums of squared per-pass values. */
int orc64_render(const orc_scene *s, const orc_view *v, int n_threads, double *accum,
                 double *accum_sq);

#ifdef __cplusplus
}
#endif
#endif
