/*
 * hptracing.h - C ABI of the MI355X wavefront path tracer.
 *
 * This header is the drop-in boundary that replaces the OpenCL host layer of
 * hpTracing (`hp::TraceRunner`, src/hp/trace_runner.h:21-66) and the OpenCL
 * kernel argument lists of src/hp/cl_src/kernel.cl.  Everything here is plain
 * C: opaque handles, pointers and sizes, `int` status codes and a
 * thread-local error string.  No HIP or torch type appears in a signature;
 * a HIP stream is passed as `void *` (NULL = HIP's null stream).
 *
 * Two shared libraries implement it:
 *   libhpt_host.so  - CPU side: OBJ/MTL loading, light table, textures,
 *                     KD-tree build and flatten   (hpt_host_*)
 *   libhptracing.so - GPU side: HIP kernels for gfx950 (hpt_scene_*, hpt_cast,
 *                     hpt_render, hpt_batch_render, hpt_server_*)
 */
#ifndef HPTRACING_H
#define HPTRACING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    HPT_OK = 0,
    HPT_ERR_INVALID = -1, /* bad argument (null pointer, size out of range)  */
    HPT_ERR_IO = -2,      /* file missing or unreadable                       */
    HPT_ERR_SCENE = -3,   /* malformed scene (bad index, missing material)    */
    HPT_ERR_HIP = -4,     /* HIP runtime error                                */
    HPT_ERR_NOMEM = -5,   /* host or device allocation failed                 */
    HPT_ERR_BUDGET = -6,  /* request larger than the device-memory budget     */
    HPT_ERR_STATE = -7,   /* call not valid in the object's current state     */
    HPT_ERR_TIMEOUT = -8  /* wait ended before the request completed         */
};

/* Message of the last failing call on this thread ("" if none). */
const char *hpt_last_error(void);
/* Same for the CPU-side library (hpt_host_* calls). */
const char *hpt_host_last_error(void);

/* ------------------------------------------------------------ scene types */
/* Material: same fields, order and 96-byte layout as the reference's
 * `Material` (src/hp/cl_src/types.h:277-292; cl_float3 occupies 16 bytes). */
typedef struct {
    float emission[4];
    float ambient[4];
    float diffuse[4];
    float specular[4];
    float optical_density;
    float dissolve;
    float shininess;
    int32_t diffuse_texture_id;
    int32_t ambient_texture_id;
    float specular_possibility;
    float refract_possibility;
    float diffuse_possibility;
} hpt_material;

/* KD node exactly as `KDTree::getData()` emits it (KDTreeNodeHeader,
 * src/hp/cl_src/types.h:294-301, 48 bytes). */
typedef struct {
    float box_start[4];
    float box_end[4];
    int32_t child;   /* first (left) child, -1 for a leaf                    */
    int32_t parent;  /* -1 for the root                                      */
    int32_t sibling; /* right sibling, set on left children only, else -1    */
    int32_t data;    /* offset of [count, tri ids...] in leaf_data, or -1    */
} hpt_kd_node;

/* Texture layers are 512 x 512 RGBA8 (Scene::texture_width/height,
 * src/hp/scene/base.cpp:34-35). */
#define HPT_TEXTURE_SIZE 512
/* Entries of the emitter table (LIGHTS_TOTAL_NUMBER, scene/base.cpp:228). */
#define HPT_LIGHT_TABLE_SIZE 256

/* Flat scene arrays, i.e. everything `TraceRunner::TraceRunner` uploads
 * (src/hp/trace_runner.cpp:134-155).  Points and normals use the cl_float3
 * layout of the reference host vectors (4 floats, the 4th ignored);
 * texcoords are 2 floats; mesh and lights are int4 (v0, v1, v2, material). */
typedef struct {
    const float *points;       int64_t n_points;
    const float *normals;      int64_t n_normals;
    const float *texcoords;    int64_t n_texcoords;
    const int32_t *mesh;       int64_t n_triangles;
    const hpt_material *materials; int64_t n_materials;
    const int32_t *lights;     int64_t n_lights;    /* 0 or 256 entries */
    const hpt_kd_node *nodes;  int64_t n_nodes;
    const int32_t *leaf_data;  int64_t n_leaf_data;
    const uint8_t *textures;   int64_t n_textures;  /* layers of 512*512*4 */
} hpt_scene_desc;

/* ======================================================= CPU side (host) */
typedef struct hpt_host_scene hpt_host_scene;

/* Load an OBJ (+ its mtllib) and build the KD tree: the reference's
 * `KDTree(filename, mtl_basepath)` (src/hp/scene/kdtree.cpp:235-246) which
 * runs `Scene::Scene` (src/hp/scene/base.cpp:37-158).  `mtl_basepath` may be
 * NULL or "" (then the mtllib name is used as given, tiny_obj_loader.cpp:
 * 485-489).  `n_threads` <= 0 picks the hardware concurrency; the tree does
 * not depend on it. */
int hpt_host_load_obj(const char *obj_path, const char *mtl_basepath,
                      int n_threads, hpt_host_scene **out);

/* Build a scene from in-memory arrays instead of an OBJ file: points are
 * xyz triples, mesh is int4 (v0, v1, v2, material).  Runs the same light
 * table (scene/base.cpp:160-198) and KD build as hpt_host_load_obj. */
int hpt_host_from_arrays(const float *points_xyz, int64_t n_points,
                         const int32_t *mesh, int64_t n_triangles,
                         const hpt_material *materials, int64_t n_materials,
                         int n_threads, hpt_host_scene **out);

/* Size of the KD tree hpt_host_from_arrays would build over these triangles
 * (points xyz triples, mesh int4), found by the same splits and the same
 * empty-leaf collapse without keeping the tree: each subtree is released
 * once counted.  For sizing a scene before building it, where the tree may
 * not fit in memory.  n_leaf_ints = n_leaves + n_leaf_entries is the length
 * of getData()'s leaf list. */
typedef struct {
    int64_t n_nodes;
    int64_t n_leaves;        /* non-empty leaves                  */
    int64_t n_leaf_entries;  /* triangle ids over all leaf lists  */
    int64_t n_leaf_ints;
    int64_t max_leaf;        /* triangles in the largest leaf     */
    int32_t max_depth;       /* edges from the root to the deepest leaf */
    int32_t reserved;
    double seconds;
} hpt_kd_counts;
int hpt_host_kd_count(const float *points_xyz, int64_t n_points,
                      const int32_t *mesh, int64_t n_triangles, int n_threads,
                      hpt_kd_counts *out);

/* Read an image file as the texture pipeline reads it (CImg<float>(path),
 * scene/base.cpp:100): PNM natively, PNG and JPEG as ImageMagick's PNM
 * output gives them (gray one channel, colour three).  Writes the size to
 * *w, *h, *c and, when `out` is not NULL and `capacity` >= w * h * c, the
 * samples as planar floats (out[(ch * h + y) * w + x]); with `out` NULL
 * only the size is written.  HPT_ERR_INVALID when `out` is too small. */
int hpt_host_read_image(const char *path, int32_t *w, int32_t *h, int32_t *c,
                        float *out, int64_t capacity);

/* Fill `out` with pointers into the scene (valid until hpt_host_free):
 * the flattened tree of `KDTree::getData()` (kdtree.cpp:248-292) plus the
 * Scene vectors. */
int hpt_host_describe(const hpt_host_scene *scene, hpt_scene_desc *out);

/* Material probabilities as Scene::Scene computes them (base.cpp:85-95). */
int hpt_host_material_from_mtl(const float ka[3], const float kd[3],
                               const float ks[3], const float ke[3],
                               float ior, float dissolve, float shininess,
                               hpt_material *out);

void hpt_host_free(hpt_host_scene *scene);

/* Seconds spent in the KD build of the last load on this scene. */
double hpt_host_build_seconds(const hpt_host_scene *scene);

/* ====================================================== GPU side (device) */
typedef struct hpt_scene hpt_scene;

/* Device-resident scene: replaces the uploads of `TraceRunner::TraceRunner`
 * (trace_runner.cpp:41-108).  Copies and repacks every array into HBM; the
 * descriptor may be freed afterwards. */
int hpt_scene_create(const hpt_scene_desc *desc, hpt_scene **out);
/* Frees the scene's device memory.  Every server created on the scene must
 * be destroyed first, no other call on the scene may still be running, and
 * casts enqueued on the caller's streams must have completed: destroy waits
 * only for the scene's own stream (renders). */
int hpt_scene_destroy(hpt_scene *scene);

typedef struct {
    int64_t n_nodes, n_leaf_entries, n_triangles, n_textures;
    int64_t device_bytes; /* HBM held by the scene */
} hpt_scene_stats;
int hpt_scene_stats_get(const hpt_scene *scene, hpt_scene_stats *out);

/* Launch and arithmetic choices of a scene's kernels.  None changes a
 * result: every setting gives the same bits.  Defaults are set by
 * hpt_scene_create (fast_division = 1, traversal = 2, refill_threshold = 32,
 * node_steps = 24, fetch = 3517 (1469 for scenes whose node, leaf-entry or
 * per-triangle record array is 2 GiB or more),
 * walk_min = 0, sort_rays = 0, cast grid =
 * min(6, occupancy) x CUs, shade grid = occupancy x CUs). */
typedef struct {
    int32_t fast_division;       /* slab quotients via certified reciprocals */
    int32_t cast_blocks_per_cu;  /* persistent grid of the cast kernel      */
    int32_t shade_blocks_per_cu; /* persistent grid of the shade kernel     */
    int32_t traversal;           /* 0: node and leaf work interleaved per lane,
                                    1: whole-wave node and leaf phases,
                                    2: phases with per-lane ray refill      */
    int32_t refill_threshold;    /* traversal 2: idle lanes (1..64) that make
                                    a wave fetch new rays                   */
    int32_t node_steps;          /* traversal 2: node visits per round      */
    int32_t fetch;               /* traversal 2: bit 0 fetches nodes in pairs
                                    (idx, idx + 1), bit 1 fetches triangle
                                    records one ahead, bit 2 spreads each
                                    leaf round's triangle tests over the
                                    wave (not with bit 1), and with bit 2,
                                    bits 3-4 = q let a lane queue 1 << q
                                    leaves per round; with a queue of 4 or
                                    8, bit 5 builds the kernel for six
                                    waves per SIMD and (without bit 0) bit 6
                                    folds each leaf round through 64-bit
                                    (distance, position) keys in LDS; with
                                    bit 2, bit 7 reads the triangle records
                                    once per triangle through the leaf
                                    entries' ids instead of once per leaf
                                    entry; with bit 0, bits 8-9 = w widen
                                    the node fetch to 2 + w consecutive
                                    nodes, and bit 10 ends it at the end
                                    of the 128-byte line (four nodes) that
                                    holds the first; bit 11 reads the nodes
                                    (and with bit 7 the leaf entries' ids and
                                    the records) with buffer loads (arrays
                                    under 2 GiB only) and, with bits 2 and 7,
                                    skips triangle tests a ray has already
                                    made.  The compiled forms: 0-5,
                                    12, 13, 20, 21, 28, 29, 52, 60, 84, 92,
                                    116, 124, 148, 149, 156, 157, 189, 220,
                                    252, 413, 445, 669, 701, 1469, 1725,
                                    3517, 3773                              */
    int32_t walk_min;            /* traversal 2: a round's node phase ends
                                    once fewer lanes than this (0..64) can
                                    walk (0: only when none can)            */
    int32_t sort_rays;           /* hpt_cast: order a batch of at least
                                    65536 rays by the Morton code of its
                                    origins in the scene box before casting
                                    (results still in the caller's order;
                                    pays for incoherent batches).  Uses
                                    stream-ordered allocations and a radix
                                    sort on the caller's stream
                                    (hpt_cast_noalloc: a static array)    */
} hpt_tuning;
/* Setting the tuning waits for a render running on the scene; casts and
 * servers on the scene launch with a consistent copy of the settings taken
 * when each launch (or server macro pass) starts. */
int hpt_scene_get_tuning(const hpt_scene *scene, hpt_tuning *out);
int hpt_scene_set_tuning(hpt_scene *scene, const hpt_tuning *tuning);

/* Evaluation of OpenCL's dot() and cross() in the intersection test.
 * ROCM_OPENCL (default): the fused forms ROCm's OpenCL device library uses
 * for those builtins, dot = fma(a.z, b.z, fma(a.y, b.y, a.x*b.x)) and one
 * fma per cross component, so that hit triangles AND distances equal the
 * reference kernel's bit for bit (kernel.cl:85-109 compiled by that
 * library, which is where the reference's CL_DEVICE_TYPE_GPU host runs on
 * this hardware).  LITERAL: the textbook formulas with every product and
 * sum rounded, (a.x*b.x + a.y*b.y) + a.z*b.z and a.y*b.z - a.z*b.y ...;
 * hit triangles agree with the default except on exact ties, distances
 * within 1e-6 relative on 99.7 % (sliver soup) to 99.99 % (room) of hits. */
#define HPT_ARITH_LITERAL 0
#define HPT_ARITH_ROCM_OPENCL 1
int hpt_scene_set_arithmetic(hpt_scene *scene, int mode);

/* Ray cast, the parity entry point of `kdtree_intersect`
 * (kernel.cl:111-208): for each of n rays (origin, direction as xyz float
 * triples in DEVICE memory) write the index of the hit triangle (-1 on a
 * miss) and its distance t (-42 on a miss, the kernel's initial value).  The
 * reference keeps only the triangle's int4 (kernel.cl:191); the index is the
 * position of that int4 in the mesh.  n <= 2^30 per call.  Asynchronous on
 * `stream` (a hipStream_t; NULL is the null stream, ordered with the work
 * of every blocking stream).  Each call takes its work counter from the
 * stream-ordered allocator, so concurrent calls on other streams or threads
 * share nothing; callers that cannot allocate (graph capture) use
 * hpt_cast_noalloc. */
int hpt_cast(hpt_scene *scene, const float *org, const float *dir, int64_t n,
             int32_t *tri_id, float *t, void *stream);

/* Same, with traversal statistics written to `counts` (HPT_CAST_COUNTS
 * int32 per ray, device memory): nodes box-tested up to the 512th non-empty
 * leaf, non-empty leaves collected, triangles tested, box-hit non-empty
 * leaves beyond the cap, nodes box-tested by the reference's full walk.
 * Used to price the cast in algorithmic bytes. */
#define HPT_CAST_COUNTS 5
int hpt_cast_counted(hpt_scene *scene, const float *org, const float *dir,
                     int64_t n, int32_t *tri_id, float *t, int32_t *counts,
                     void *stream);

/* hpt_cast without any device allocation: same arguments, same results bit
 * for bit.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  hpt_cast takes a stream-ordered
 * work counter per call (and sort buffers when the scene's sort_rays is
 * on); this entry point uses a static device array of 64 MiB + 256 bytes
 * instead, which each GPU holds from the library's load until the process
 * ends.  n <= HPT_NOALLOC_MAX_RAYS (larger n: HPT_ERR_INVALID, nothing
 * launched).  The current device must be the scene's.  Calls that overlap
 * on one GPU share that array and race: serialise them (one stream). */
#define HPT_NOALLOC_MAX_RAYS (1 << 20)
int hpt_cast_noalloc(hpt_scene *scene, const float *org, const float *dir,
                     int64_t n, int32_t *tri_id, float *t, void *stream);

/* One render request: the argument list of `TraceRunner::run`
 * (trace_runner.h:53-65), plus an explicit RNG seed (the reference seeds
 * from std::rand() after srand(time(0)), trace_runner.cpp:63,175-178).
 * width/height are the angular extents (hptracing.cpp:74-75 passes
 * W/H*angle and angle); height < 0 selects the equirectangular camera with
 * shift = int(right[0]) (trace_runner.cpp:131-140). */
typedef struct {
    float view[3];
    float up[3];
    float right[3];
    float width;
    float height;
    int32_t sample_x, sample_y;           /* output image size            */
    int32_t supersample_x, supersample_y; /* 2 x 2 with the CLI's -x      */
    float background[3];
    int32_t sample;                       /* samples per pixel pass count */
    int32_t depth;
    int32_t disable_diffuse;
    float brightness;
    uint64_t seed;
} hpt_view;

#define HPT_MAX_DEPTH 64

typedef struct {
    int64_t primary_rays;               /* camera rays, cast once          */
    int64_t rays_cast;                  /* all intersect work incl. primary */
    int64_t live_per_depth[HPT_MAX_DEPTH]; /* rays cast at bounce i, summed over passes */
    double seconds;                     /* wall time of the render call    */
    /* With HPT_RENDER_TIME_KERNELS: summed device time (HIP events on the
     * render stream) of every cast and every shade launch. */
    double cast_ms, shade_ms;
    int64_t cast_launches, shade_launches;
    /* With HPT_RENDER_COUNT_WORK: traversal work summed over all casts
     * (nodes box-tested up to the 512th leaf, non-empty leaves collected,
     * triangles tested). */
    int64_t node_visits, leaf_visits, tri_tests;
} hpt_render_stats;

#define HPT_RENDER_TIME_KERNELS 1u
#define HPT_RENDER_COUNT_WORK 2u
/* Enqueue every pass directly instead of replaying one captured hipGraph
 * of the pass (the graph is also skipped while kernels are timed). */
#define HPT_RENDER_NO_GRAPH 4u
/* Launch one sample pass at a time, as TraceRunner::run does, instead of
 * sixteen passes per launch sequence (fewer when the path buffers of
 * sixteen passes would pass 48 GiB).  The image is the same bit for bit
 * (pass p always accumulates into slice p % 16). */
#define HPT_RENDER_PASS_AT_A_TIME 8u

/* Render one view.  `rgba` (host, W*H*4 bytes, may be NULL) receives the
 * tone-mapped image exactly as get_image writes it (kernel.cl:615-652;
 * alpha 0, row 0 = camera row 0).  `accum` (host, W*ssx*H*ssy*3 floats,
 * may be NULL) receives the raw per-sample radiance sums that the reference
 * keeps in results_mem.  `stats` may be NULL. */
int hpt_render(hpt_scene *scene, const hpt_view *view, uint8_t *rgba,
               float *accum, hpt_render_stats *stats);
/* hpt_render with HPT_RENDER_* flags. */
int hpt_render_ex(hpt_scene *scene, const hpt_view *view, uint8_t *rgba,
                  float *accum, hpt_render_stats *stats, unsigned flags);

/* Tune the scene's launch forms (hpt_tuning) on renders of `view` with the
 * passes of one launch group (min(sample, 16)), the image shrunk until
 * camera rays x passes is at most max_rays (<= 0: 2^24): coordinate descent
 * over traversal, fetch form, refill threshold, node steps, division form
 * and cast grid (two sweeps, each knob taking the value with the shortest
 * cast time, if more than 1 % shorter, while the others are held), then the
 * shade grid.  Keeps the
 * result on the scene and returns it in `chosen` (may be NULL).  No
 * reference counterpart. */
int hpt_autotune(hpt_scene *scene, const hpt_view *view, int max_rays,
                 hpt_tuning *chosen);

/* The same tuning for ray casts: the knobs of hpt_autotune (the shade grid
 * aside) timed on hpt_cast of the first min(n, max_rays) rays of a batch in
 * device memory (xyz triples; max_rays <= 0: 2^22).  Keeps the result on
 * the scene and returns it in `chosen` (may be NULL).  No reference
 * counterpart. */
int hpt_autotune_cast(hpt_scene *scene, const float *org, const float *dir,
                      int64_t n, int64_t max_rays, hpt_tuning *chosen);

/* Render several views of one scene together (no reference counterpart; the
 * reference renders one view per TraceRunner::run call).  The camera rays
 * of all views are cast in one launch, and every pass launches the
 * expand / cast / shade sequence once over the live paths of every view
 * whose spp exceeds the pass index, so views of uneven size and spp share
 * the machine.  Each view's result equals its hpt_render result bit for bit
 * (paths carry their own random streams).  rgba[k] / accum[k] may be NULL;
 * `stats` sums over the batch. */
int hpt_batch_render(hpt_scene *scene, const hpt_view *views, int n_views,
                     uint8_t *const *rgba, float *const *accum,
                     hpt_render_stats *stats, unsigned flags);

/* ------------------------------------------------------------- serving */
/* Continuous-batching render server over one resident scene (no reference
 * counterpart; the reference's GUI re-renders one view at a time on a
 * resident scene, hptracing.cpp:124-130).  A worker thread admits queued
 * requests in arrival order whenever their camera rays fit in the device
 * arena (sized from device_budget_bytes, about 1.2 KB per camera ray),
 * and runs macro passes of eight sample passes over every admitted request,
 * so requests that arrive while others render join at the next macro pass.
 * A request's result equals its hpt_render result bit for bit. */
typedef struct hpt_server hpt_server;

typedef struct {
    int64_t device_budget_bytes; /* arena for in-flight requests; <= 0: 4 GiB */
} hpt_server_opts;

typedef struct {
    int64_t id;
    int32_t status;     /* 0 queued, 1 rendering, 2 done, 3 failed */
    double queue_ms;    /* submit -> admission */
    double latency_ms;  /* submit -> result in the caller's buffers */
    int64_t camera_rays;
} hpt_request_info;

typedef struct {
    int64_t queued, active, completed;
    int64_t macro_passes;  /* passes run over the admitted set */
    int64_t rays_cast;     /* primaries + continuation rays */
    double busy_seconds;   /* worker time spent rendering */
    int64_t capacity_rays; /* camera rays the arena holds */
} hpt_server_stats;

int hpt_server_create(hpt_scene *scene, const hpt_server_opts *opts,
                      hpt_server **out);
/* Queue a request.  rgba (W*H*4 bytes) and accum (may be NULL) are written
 * when it completes and must stay valid until hpt_server_wait returns for
 * it.  HPT_ERR_BUDGET if the request can never fit the arena. */
int hpt_server_submit(hpt_server *server, const hpt_view *view, uint8_t *rgba,
                      float *accum, int64_t *request_id);

/* Progressive preview: the reference's GUI shows a view by re-rendering it
 * whenever the camera moves (src/hptracing.cpp:60-83, 124-130); here a
 * progressive request publishes its image after every macro pass, and the
 * caller reads the latest one while the request goes on. */
#define HPT_SERVE_PROGRESSIVE 1u
/* hpt_server_submit with HPT_SERVE_* flags. */
int hpt_server_submit_ex(hpt_server *server, const hpt_view *view, uint8_t *rgba,
                         float *accum, unsigned flags, int64_t *request_id);
/* For a progressive request: wait until it has published an image of at
 * least min_passes sample passes, or finished, or timeout_s seconds passed
 * (< 0: no limit); then copy its latest image (W*H*4 bytes) into rgba (may
 * be NULL) and that image's pass count into *passes (0: none yet).  The
 * image of S passes equals hpt_render's for the same view with sample = S,
 * bit for bit.  HPT_ERR_TIMEOUT if fewer passes were published in time (the
 * outputs then hold the latest image).  Valid until hpt_server_wait retires
 * the request. */
int hpt_server_peek(hpt_server *server, int64_t request_id, int32_t min_passes,
                    double timeout_s, uint8_t *rgba, int32_t *passes);

/* Wait up to timeout_s seconds (< 0: forever).  HPT_ERR_TIMEOUT if still
 * running; otherwise the request is retired and its id forgotten. */
int hpt_server_wait(hpt_server *server, int64_t request_id, double timeout_s,
                    hpt_request_info *info);
int hpt_server_stats_get(hpt_server *server, hpt_server_stats *out);
/* Finish every queued and running request, then stop the worker.  No other
 * call on the server may still be running (wait and peek included). */
int hpt_server_destroy(hpt_server *server);

#ifdef __cplusplus
}
#endif
#endif /* HPTRACING_H */
